"""The plane-proposal network (include/pm_net.h) on the GPU.

* forward pass and parameter gradients of a training step against a torch
  float32 reference of the same U-Net on the CPU, with its activations and
  weights rounded to bf16 where the HIP kernels store bf16 (tolerances
  below: bf16 storage rounds to 2^-9 relative, and the two sides sum in
  different orders);
* a few training steps on one pair lower the loss;
* pm_setup_proposal bit for bit against the oracle: each pixel keeps the
  plane built from the network's output (the mirrored pass for the right
  view) unless the random init plane costs less.
"""
import numpy as np
import pytest

import cvt_amd
from cvt_amd.net import ProposalNet
from test_gpu_parity import first_diff, same

pytestmark = pytest.mark.gpu

FWD_RTOL = 2e-2    # relative L2 error of the head output
GRAD_RTOL = 6e-2   # relative L2 error of each layer's weight gradient

# the topology of pm_net_capi.cpp: tensors (level, channels) and the convs
# (in1, in2, up2, out, dilation, relu) in forward order, pools interleaved
CH = {"x0": (0, 8), "e0a": (0, 32), "e0": (0, 32), "p1": (1, 32),
      "e1a": (1, 64), "e1": (1, 64), "p2": (2, 64), "e2a": (2, 96),
      "e2": (2, 96), "p3": (3, 96), "b0": (3, 128), "b1": (3, 128),
      "b2": (3, 128), "b3": (3, 128), "d2a": (2, 96), "d2": (2, 96),
      "d1a": (1, 64), "d1": (1, 64), "d0a": (0, 32), "d0": (0, 32),
      "out": (0, 8), "cv": (2, 48), "cv1": (1, 64)}
CONVS = [("x0", None, "e0a", 1, 1), ("e0a", None, "e0", 1, 1),
         ("p1", "cv1", "e1a", 1, 1), ("e1a", None, "e1", 1, 1),
         ("p2", "cv", "e2a", 1, 1), ("e2a", None, "e2", 1, 1),
         ("p3", None, "b0", 1, 1), ("b0", None, "b1", 2, 1),
         ("b1", None, "b2", 4, 1), ("b2", None, "b3", 8, 1),
         ("e2", "b3", "d2a", 1, 1), ("d2a", None, "d2", 1, 1),
         ("e1", "d2", "d1a", 1, 1), ("d1a", None, "d1", 1, 1),
         ("e0", "d1", "d0a", 1, 1), ("d0a", None, "d0", 1, 1),
         ("d0", None, "out", 1, 0)]
ORDER = [0, 1, ("e0", "p1"), 2, 3, ("e1", "p2"), 4, 5, ("e2", "p3"),
         6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16]


def layout():
    """(weight offset, bias offset, cin, cout, kpad, cpad) per conv"""
    out, off = [], 0
    for in1, in2, o, _, _ in CONVS:
        cin = CH[in1][1] + (CH[in2][1] if in2 else 0)
        cout = CH[o][1]
        kpad, cpad = (9 * cin + 31) // 32 * 32, (cout + 31) // 32 * 32
        woff = off
        off = (off + cpad * kpad + 63) // 64 * 64
        boff = off
        off = (off + cpad + 63) // 64 * 64
        out.append((woff, boff, cin, cout, kpad, cpad))
    return out, off


def torch_weights(params):
    import torch
    lay, _ = layout()
    ws = []
    for woff, boff, cin, cout, kpad, cpad in lay:
        w = params[woff:woff + cpad * kpad].reshape(cpad, kpad)
        w = w[:cout, :9 * cin].reshape(cout, 3, 3, cin).transpose(0, 3, 1, 2)
        ws.append((torch.tensor(np.ascontiguousarray(w)),
                   torch.tensor(params[boff:boff + cout].copy())))
    return ws


def bf(t):
    import torch
    return t.to(torch.bfloat16).to(torch.float32)


def ref_cost_volume(x0, factor, levels):
    """a matching cost volume from the bf16 input (N, 8, H, W): factor x
    factor box average (rounded to bf16), then for d = 0..levels-1 the
    mean over the 3 x 3 window of sum_rgb |A(x + i, y + j) - B(x + i - d,
    y + j)|, coordinates clamped to the image, rounded to bf16"""
    import torch
    import torch.nn.functional as F
    d4 = bf(F.avg_pool2d(x0, factor))
    N, _, H4, W4 = d4.shape
    a, b = d4[:, 0:3], d4[:, 4:7]
    cols = torch.arange(W4)
    rows = torch.arange(H4)
    out = []
    for d in range(levels):
        acc = torch.zeros((N, H4, W4))
        for j in (-1, 0, 1):
            yy = (rows + j).clamp(0, H4 - 1)
            for t in (-1, 0, 1):
                xa = (cols + t).clamp(0, W4 - 1)
                xb = (cols + t - d).clamp(0, W4 - 1)
                da = a[:, :, yy][:, :, :, xa]
                db = b[:, :, yy][:, :, :, xb]
                acc = acc + (da - db).abs().sum(1)
        out.append(acc / 9.0)
    return bf(torch.stack(out, 1))


def ref_forward(x, ws):
    """x: (N, 8, H, W) float32 CPU tensor; returns the head (N, 8, H, W)"""
    import torch
    import torch.nn.functional as F
    act = {"x0": bf(x)}
    act["cv"] = ref_cost_volume(act["x0"], 4, 48)
    act["cv1"] = ref_cost_volume(act["x0"], 2, 64)
    for op in ORDER:
        if isinstance(op, tuple):
            act[op[1]] = bf(F.avg_pool2d(act[op[0]], 2))
            continue
        in1, in2, o, dil, relu = CONVS[op]
        inp = act[in1]
        if in2 in ("cv", "cv1"):          # same resolution, no gradient
            inp = torch.cat([inp, act[in2]], 1)
        elif in2:
            up = F.interpolate(act[in2], scale_factor=2, mode="nearest")
            inp = torch.cat([inp, up[..., :inp.shape[2], :inp.shape[3]]], 1)
        w, b = ws[op]
        y = F.conv2d(inp, bf(w), b, padding=dil, dilation=dil)
        if relu:
            y = torch.relu(y)
        act[o] = y if o == "out" else bf(y)
    return act["out"]


def ref_loss(o, t, wd, wn):
    import torch
    valid = (t[:, 0] >= 0).float()
    d = (o[:, :4] - t).abs()
    l = (wd * d[:, 0] + wn * d[:, 1:].sum(1)) * valid
    return l.sum() / (o.shape[0] * o.shape[2] * o.shape[3])


def batch(synth_gpu, N, H, W, seed):
    import torch
    L, R, gt = synth_gpu.make_pairs(N, W, H, cells=5, dmax=20.0, seed=seed)
    x = torch.cat([L, R], -1).contiguous()
    t = torch.zeros((N, H, W, 4), device="cuda")
    t[..., 0] = gt / 128.0
    t[..., 3] = 1.0
    t[:, ::5, ::3, 0] = -1.0          # pixels without a target
    return x, t


def test_param_layout_matches():
    with ProposalNet(seed=3) as net:
        assert net.num_params == layout()[1]


# sizes: whole 64 x 4 pixel tiles at the first level; partial tiles in x
# at every level and partial rows of tiles at the deeper ones
@pytest.mark.parametrize("N,H,W", [(2, 48, 64), (1, 40, 200)])
def test_forward_matches_torch_reference(N, H, W):
    import torch
    from cvt_amd import synth_gpu
    x, _ = batch(synth_gpu, N, H, W, seed=5)
    out = torch.empty((N, H, W, 4), device="cuda")
    with ProposalNet(seed=7) as net:
        # push the head off its small init so every channel matters
        p = net.params()
        lay, _ = layout()
        woff, _, cin, _, kpad, _ = lay[-1]
        p[woff:woff + 4 * kpad] *= 10.0
        net.set_params(p)
        net.forward(x.data_ptr(), N, H, W, out.data_ptr())
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        params = net.params()
    want = ref_forward(x.cpu().permute(0, 3, 1, 2).contiguous(),
                       torch_weights(params))[:, :4].permute(0, 2, 3, 1)
    want = want.numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < FWD_RTOL, err


def test_gradients_match_torch_autograd():
    import torch
    from cvt_amd import synth_gpu
    N, H, W = 2, 32, 48
    x, t = batch(synth_gpu, N, H, W, seed=9)
    wd, wn = 1.0, 0.05
    with ProposalNet(seed=11) as net:
        params = net.params()
        net.train_step(x.data_ptr(), t.data_ptr(), N, H, W, 1e-6, wd, wn)
        g = net.grads()
    ws = [(w.clone().requires_grad_(), b.clone().requires_grad_())
          for w, b in torch_weights(params)]
    o = ref_forward(x.cpu().permute(0, 3, 1, 2).contiguous(), ws)
    loss = ref_loss(o, t.cpu().permute(0, 3, 1, 2), wd, wn)
    loss.backward()
    lay, _ = layout()
    for l, ((w, b), (woff, boff, cin, cout, kpad, cpad)) in \
            enumerate(zip(ws, lay)):
        gw = g[woff:woff + cpad * kpad].reshape(cpad, kpad)
        gw = gw[:cout, :9 * cin].reshape(cout, 3, 3, cin).transpose(0, 3, 1, 2)
        ref = w.grad.numpy()
        if cout == 8:                     # the head: 4 real outputs
            gw, ref = gw[:4], ref[:4]
        err = np.linalg.norm(gw - ref) / max(np.linalg.norm(ref), 1e-12)
        assert err < GRAD_RTOL, (l, err)
        gb = g[boff:boff + cout][:ref.shape[0]]
        rb = b.grad.numpy()[:ref.shape[0]]
        errb = np.linalg.norm(gb - rb) / max(np.linalg.norm(rb), 1e-12)
        assert errb < GRAD_RTOL, (l, "bias", errb)


def test_training_lowers_the_loss():
    import torch
    from cvt_amd import synth_gpu
    N, H, W = 2, 64, 96
    x, t = batch(synth_gpu, N, H, W, seed=13)
    with ProposalNet(seed=17) as net:
        losses = [net.train_step(x.data_ptr(), t.data_ptr(), N, H, W, 2e-3)
                  for _ in range(30)]
    assert np.isfinite(losses).all()
    assert np.mean(losses[-5:]) < 0.7 * np.mean(losses[:3]), losses


@pytest.mark.parametrize("views", [2, 1])
def test_setup_proposal_bitexact(oracle, synth, views):
    """pm_setup_proposal(_views) against the oracle: the proposal (the
    network's output on the padded pair, and on the mirrored pair for the
    right view) competes with the random init plane at every pixel; with
    views = 1 the right view takes, through the view scatter of the seeded
    left view, the nearest left surface at each pixel (viewprop), the
    random init plane where none lands, and no cost (NaN)."""
    import torch
    L, R, _ = synth.make_pair(61, 37, cells=4, dmax=16.0, channels=4, seed=19)
    H, W = L.shape[:2]
    Hp, Wp = 40, 64
    r, dmax = 3, 16.0
    with ProposalNet(seed=23) as net:
        p = net.params()
        lay, _ = layout()
        _, boff, _, _, _, _ = lay[-1]
        p[boff] = 6.0 / 128.0             # proposals around 6 px
        net.set_params(p)

        def run(a, b):
            x = np.concatenate([a, b], -1)
            x = np.pad(x, ((0, Hp - H), (0, Wp - W), (0, 0)), mode="edge")
            xd = torch.from_numpy(np.ascontiguousarray(x[None])).cuda()
            out = torch.empty((1, Hp, Wp, 4), device="cuda")
            net.forward(xd.data_ptr(), 1, Hp, Wp, out.data_ptr())
            torch.cuda.synchronize()
            return out[0].cpu().numpy()

        prop_l = run(L, R)
        prop_r = run(R[:, ::-1], L[:, ::-1])
        with cvt_amd.Context(radius=r, disparity_max=dmax) as ctx:
            ctx.set_images(L, R)
            ctx.setup_proposal(net, views)
            got = (ctx.planes(0), ctx.planes(1))
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    for view, (a, b, lr) in enumerate(((tl, tr, 1), (tr, tl, 0))):
        rnd = oracle.init(a, b, r, dmax, lr)
        if view >= views:
            # from the left view's planes (checked above) by the scatter
            left = got[0]
            key0 = (oracle.scatter(left, 1) & 0xFFFFFFFF).astype(np.int64)
            want = rnd.copy()
            landed = key0 != 0xFFFFFFFF
            for y, x in zip(*np.nonzero(landed)):
                want[y, x] = oracle.viewprop(left[y, W - 1 - key0[y, x]])
            want[..., 3] = np.nan
            assert 0.3 < landed.mean() < 1.0, landed.mean()
            assert same(got[view], want), (view, first_diff(got[view], want))
            continue
        prop = np.zeros_like(rnd)
        have = np.zeros((H, W), bool)
        for y in range(H):
            for x in range(W):
                o = prop_l[y, x] if lr else prop_r[y, W - 1 - x]
                d = np.float32(o[0]) * np.float32(128.0)
                n = np.array([o[1], o[2] if lr else -o[2], o[3]], np.float32)
                if d > 0 and np.isfinite(d) and np.isfinite(n).all():
                    # normalize(): |n| from the built-in dot, a chain of
                    # fused multiply-adds (exact product in float64)
                    l2 = np.float32(np.float64(n[1]) * n[1] +
                                    np.float64(np.float32(n[0] * n[0])))
                    l2 = np.float32(np.float64(n[2]) * n[2] +
                                    np.float64(l2))
                    ln = np.sqrt(l2)
                    n = (n / ln).astype(np.float32) if ln > 0 else n
                    prop[y, x] = oracle.from_disparity_normal(
                        x, y, d if lr else -d, n)
                    have[y, x] = True
        cp = oracle.cost_field(a, b, r, prop)
        prop[..., 3] = cp
        keep = have & ((cp <= rnd[..., 3]) | np.isnan(rnd[..., 3]))
        want = np.where(keep[..., None], prop, rnd)
        assert 0.05 < keep.mean() < 0.95, keep.mean()
        assert same(got[view], want), (view, first_diff(got[view], want))


def test_shipped_weights_match_plain_with_one_iteration(synth):
    """The committed weights (cvt_amd/weights/proposal.bin) load, and on
    the 720p bench pair (24 cells, disparity 0-128, 9x9 patch, the
    colour+gradient cost) the proposal seeding the left view plus one
    refine iteration reaches at least the bad-1px of plain PatchMatch
    after 8 iterations (LR check at 1 px, the angle test off)."""
    import os
    from cvt_amd.net import WEIGHTS
    assert os.path.exists(WEIGHTS)
    L, R, gt = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                               seed=0)
    kw = dict(radius=4, disparity_max=128.0,
              cost_kind=cvt_amd.PM_COST_COLOR_GRAD, cost_weight=25.5)
    bad = {}
    with cvt_amd.Context(**kw) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(8)
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
        bad["plain"] = cvt_amd.disparity.bad_pixel_rate(d, gt, 1.0)
        with ProposalNet(weights=WEIGHTS) as net:
            ctx.setup_proposal(net, 1)
            ctx.iterate(1)
            d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
            bad["learned"] = cvt_amd.disparity.bad_pixel_rate(d, gt, 1.0)
    print(bad, flush=True)
    assert 0.0 < bad["learned"] <= bad["plain"] < 0.3, bad
