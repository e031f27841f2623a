"""Trains the plane-proposal network (include/pm_net.h) and writes its
weights file.

Data: seeded synthetic 512 x 384 RGBA pairs from cvt_amd.synth_gpu
(8-24 cells, disparity 0-128 px, or a disparity range drawn per batch from
--dmax), generated on the GPU batch by batch.
Targets: the plane fields plain PatchMatch produces on each pair (this
library's engine, 16 iterations, 9 x 9 patch, the run's cost), LR-checked
at 1 px: disparity / 128 and the plane's normal where the check keeps the
pixel, no target elsewhere.  Training: masked L1 (disparity weight 1,
normal weight 0.05), Adam with a cosine learning-rate decay, every kernel
the library's own (pm_net_train_step).  Each pair is used once.

  python -m cvt_amd.train_proposal --pairs 20000 --out cvt_amd/weights/proposal.bin
"""
import argparse
import math
import os
import sys
import time

import numpy as np


def log(*a):
    print(*a, flush=True)


def to_normal(planes):
    """nd_state_to_normal (pmhstereo.cl:201-210) of (..., 4) plane fields"""
    import torch
    tx, ty = 1.0 - planes[..., 0], -planes[..., 1]
    nz = 1.0 / torch.sqrt(tx * tx + ty * ty + 1.0)
    nx, ny = -tx * nz, -ty * nz
    nz = torch.sqrt(torch.clamp(1.0 - nx * nx - ny * ny, min=0.0))
    return torch.stack([nx, ny, nz], -1)


class TargetMaker:
    """Plain-PatchMatch targets of a batch of pairs on `contexts` engine
    contexts whose streams run side by side."""

    def __init__(self, W, H, iters, cost_kind, cost_weight, contexts=2,
                 dmax=(128.0,)):
        import torch

        import cvt_amd
        self.W, self.H, self.iters = W, H, iters
        self.streams = [torch.cuda.Stream() for _ in range(contexts)]
        # contexts per disparity range (a context parameter)
        self.ctxs = {d: [cvt_amd.Context(radius=4, disparity_max=d,
                                         cost_kind=cost_kind,
                                         cost_weight=cost_weight,
                                         stream=s.cuda_stream)
                         for s in self.streams] for d in dmax}
        self.disp = [torch.empty((H, W), device="cuda")
                     for _ in self.streams]

    def __call__(self, left, right, dmax=128.0):
        import torch

        import cvt_amd
        B = left.shape[0]
        ctxs = self.ctxs[dmax]
        torch.cuda.synchronize()
        targets = torch.empty((B, self.H, self.W, 4), device="cuda")
        for b0 in range(0, B, len(ctxs)):
            group = list(range(b0, min(B, b0 + len(ctxs))))
            for k, b in enumerate(group):
                c = ctxs[k]
                c.set_images_device(left[b].data_ptr(), right[b].data_ptr(),
                                    self.W, self.H, 4)
                c.setup()
                c.iterate(self.iters)
                c.disparity_device(self.disp[k].data_ptr(),
                                   cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
            for k, b in enumerate(group):
                planes = torch.from_numpy(ctxs[k].planes(0)).cuda()
                torch.cuda.synchronize()
                d = self.disp[k]
                t = torch.cat([(d / 128.0)[..., None], to_normal(planes)], -1)
                t[..., 0] = torch.where(d > 0, t[..., 0],
                                        torch.full_like(d, -1.0))
                targets[b] = t
        return targets

    def close(self):
        for cs in self.ctxs.values():
            for c in cs:
                c.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--width", type=int, default=512)
    ap.add_argument("--height", type=int, default=384)
    ap.add_argument("--iters", type=int, default=16,
                    help="PatchMatch iterations of the targets")
    ap.add_argument("--cost", choices=["color_grad", "census"],
                    default="color_grad")
    ap.add_argument("--dmax", default="128",
                    help="disparity range of the pairs, or a comma list "
                         "to draw one per batch from")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", default=None)
    ap.add_argument("--time-limit", type=float, default=0.0,
                    help="stop after this many seconds (0: no limit); the "
                         "learning-rate decay then follows whichever of "
                         "the pair count and the time runs out first")
    ap.add_argument("--save-every", type=int, default=0,
                    help="also write the weights every this many steps")
    args = ap.parse_args()

    import torch

    import cvt_amd
    from cvt_amd import synth_gpu
    from cvt_amd.net import ProposalNet, WEIGHTS

    cost_kind, cost_weight = ((cvt_amd.PM_COST_COLOR_GRAD, 25.5)
                              if args.cost == "color_grad"
                              else (cvt_amd.PM_COST_CENSUS, 16.0))
    W, H, B = args.width, args.height, args.batch
    stream = torch.cuda.Stream()
    net = ProposalNet(seed=args.seed, stream=stream.cuda_stream)
    if args.init:
        net.load(args.init)
    dmaxes = tuple(float(v) for v in args.dmax.split(","))
    maker = TargetMaker(W, H, args.iters, cost_kind, cost_weight,
                        dmax=dmaxes)
    steps = (args.pairs + B - 1) // B
    rng = np.random.default_rng(args.seed)
    t0 = time.time()
    losses = []
    done = 0
    for step in range(steps):
        seed = args.seed * 1000003 + step
        cells = int(rng.integers(8, 25))
        dmax = dmaxes[int(rng.integers(0, len(dmaxes)))]
        left, right, _ = synth_gpu.make_pairs(B, W, H, cells=cells,
                                              dmax=dmax, seed=seed)
        target = maker(left, right, dmax)
        x = torch.cat([left, right], -1).contiguous()
        torch.cuda.synchronize()
        frac = step / steps
        if args.time_limit:
            frac = max(frac, (time.time() - t0) / args.time_limit)
        lr = args.lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi *
                                                          min(frac, 1.0))))
        loss = net.train_step(x.data_ptr(), target.data_ptr(), B, H, W, lr)
        losses.append(loss)
        done += B
        if step % 25 == 0 or step == steps - 1:
            log("step %d/%d pairs %d loss %.5f (mean of last 25 %.5f) lr %.2e "
                "%.1f s" % (step, steps, done, loss, float(np.mean(losses[-25:])),
                            lr, time.time() - t0))
        if args.save_every and (step + 1) % args.save_every == 0:
            net.save(args.out or WEIGHTS)
        if args.time_limit and time.time() - t0 > args.time_limit:
            log("time limit reached after %d pairs" % done)
            break
    out = args.out or WEIGHTS
    net.save(out)
    log("saved %s (%d parameters, %d pairs, %.1f s)" % (out, net.num_params,
                                                        done, time.time() - t0))
    maker.close()
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
