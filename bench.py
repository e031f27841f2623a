#!/usr/bin/env python3
"""720p stereo pairs/s of plain PatchMatch on one MI355X.

Workload (one step = one stereo pair, inputs resident in HBM): a seeded
synthetic 1280x720 RGBA float32 pair with 24 slanted planar cells and
disparities in 0-128 px; precompute (bilateral weights, gradients), random
plane init of both views, 8 PatchMatch iterations (left then right view,
one pm_iterate call) with a 9x9 patch and the gamma-10 truncated
colour+gradient cost (patch_eval_color_grad_weighted, support-weight gain
255 / 10 = 25.5 on [0, 1] data), LR check at 1 px (angle test off) and
disparity conversion, ordered on the engine's stream (the engine runs each
view's propagate as two half launches on a second stream of its own and
joins it back).  --cost census runs the reference's default census cost
instead.

Prints one JSON line with the throughput, the roofline of the dominant
kernel (k_propagate: FP32 VALU bound, algorithmic FLOPs from an exact tap
count, per view propagate) and the CPU baseline (the oracle on the host
cores: one full run of the same pair, and the 320x240 gray configuration
as one warm-up plus the median of five full runs).  --pipeline huber runs
the PM-Huber pipeline instead and quotes the ROF kernel it adds.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("720p stereo pairs/s at matched bad-1.0px; learned-init vs plain "
          "PatchMatch")
W, H, CELLS, DMAX, RADIUS, ITERS, SEED = 1280, 720, 24, 128.0, 4, 8, 0
MAXDISP, MAXDEG = 1.0, 180.0
# the small CPU configuration: 320x240 gray, 8 cells, disparity 0-32, 11x11
# patch, 3 iterations (census cost, the reference's)
C1 = dict(W=320, H=240, cells=8, dmax=32.0, radius=5, iters=3, seed=1)
C1_RUNS = 5
# candidate costs: the gamma-10 colour+gradient cost (255 / 10 on [0, 1]
# data) and the reference's census cost (weight CENSUS_COLOR_WEIGHT 12)
COSTS = {"color_grad": (1, 25.5), "census": (0, 16.0)}
# algorithmic FP32 work per valid tap (FMA = 2 FLOP, abs and min count 1):
#  census tap     = transform 4 + colour lerp 10 + gradient lerp 13 +
#                   colour term 13 + sign term 6 + accumulate/exit 4   = 50
#  colour+grad tap = transform 4 + colour lerp 10 + gradient lerp 13 +
#                   colour term (3 sub, 3 abs, 3 min, 2 add) 11 +
#                   gradient term (4 sub, 4 abs, 4 min, 3 add) 15 +
#                   the alpha blend (mul + fma) 3 + accumulate (fma +
#                   add) 3                                             = 59
#  KLT tap        = 3 lerped colour/gradient samples + 3 cost terms +
#                   the outer-product accumulate                      = 150
FLOP_PER_COST_TAP = {"census": 50, "color_grad": 59}
FLOP_PER_KLT_TAP = 150
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0
# HBM bytes per view propagate (its two k_propagate half launches) from
# rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by the same counters read
# for a copy kernel of known size; used only while the kernel sources
# still hash to what was profiled
HBM_PROFILE = os.path.join(ROOT, "profiles",
                           "k_propagate_1280x720_rgba_r4_hbm.json")
KERNEL_SOURCES = ("cvt_amd/csrc/pm_engine.hip", "cvt_amd/csrc/pm_device.h")
# the same for one k_rof_steps launch of the PM-Huber pipeline
ROF_HBM_PROFILE = os.path.join(ROOT, "profiles",
                               "k_rof_steps_1280x720_rgba_huber_hbm.json")
ROF_KERNEL_SOURCES = ("cvt_amd/csrc/pm_smooth.hip", "cvt_amd/csrc/pm_device.h",
                      "cvt_amd/csrc/pm_engine.h")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(path=HBM_PROFILE, sources=KERNEL_SOURCES,
                     key="hbm_bytes_per_propagate"):
    """Calibrated HBM bytes (per view propagate by default) from a
    profile file, or None when the kernel sources changed since the
    counters were collected."""
    import hashlib
    try:
        with open(path) as f:
            prof = json.load(f)
        h = hashlib.sha256()
        for rel in sources:
            with open(os.path.join(ROOT, rel), "rb") as f:
                h.update(f.read())
    except OSError:
        return None
    if h.hexdigest() != prof.get("source_sha256"):
        return None
    return float(prof[key])


def host_cpu():
    """The host the CPU baseline ran on: CPU model, logical CPUs seen by
    the process, and the oracle's thread count."""
    from oracle import oracle as O
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
             else os.cpu_count())
    return {"cpu_model": model, "nproc": nproc, "threads": O.num_threads()}


def schedules_by_iteration(dL, dR, gt, cost_kind, cost_weight, stream):
    """bad-1.0px after each of the 8 iterations, and the time of 8
    iterations as one pm_iterate call (HIP events on the engine's stream),
    for the reference's Jacobi sweep and the red-black schedule."""
    import numpy as np
    import torch

    import cvt_amd
    out = {}
    dD = torch.empty((H, W), dtype=torch.float32, device=dL.device)
    for name, sched in (("jacobi", cvt_amd.PM_SCHEDULE_JACOBI),
                        ("red_black", cvt_amd.PM_SCHEDULE_RED_BLACK)):
        with cvt_amd.Context(radius=RADIUS, disparity_max=DMAX,
                             stream=stream.cuda_stream, cost_kind=cost_kind,
                             cost_weight=cost_weight,
                             schedule=sched) as ctx, \
                torch.cuda.stream(stream):
            bad = []
            ctx.set_images_device(dL.data_ptr(), dR.data_ptr(), W, H, 4)
            ctx.setup()
            for it in range(ITERS):
                ctx.iterate(1)
                ctx.disparity_device(dD.data_ptr(),
                                     cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP,
                                     MAXDEG)
                torch.cuda.synchronize()
                bad.append(cvt_amd.disparity.bad_pixel_rate(
                    dD.cpu().numpy(), gt, 1.0))
            ms = []
            for _ in range(3):
                ctx.setup()
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                ctx.iterate(ITERS)
                e1.record(stream)
                torch.cuda.synchronize()
                ms.append(e0.elapsed_time(e1))
        out[name] = {"bad_1px_by_iteration": bad,
                     "ms_per_8_iterations": float(np.min(ms))}
    return out


def cpu_baseline(L, R, gt, cost, gpu_disp):
    """The oracle (C, OpenMP on all host cores): one full run of the 720p
    pair (setup, all 8 iterations, LR check and disparity, the GPU step's
    sequence), and the 320x240 gray configuration as one warm-up plus the
    median of C1_RUNS full runs.  Also reports how the oracle's disparity
    compares with the GPU's (0.5 px bar, bit equality)."""
    import numpy as np

    from cvt_amd import disparity, synth
    from oracle import oracle as O
    kind, kw = COSTS[cost]
    t0 = time.perf_counter()
    ref = O.run(L, R, RADIUS, DMAX, ITERS, maxdisp=MAXDISP, maxdeg=MAXDEG,
                cost_kind=kind, cost_weight=kw)
    t1 = time.perf_counter()
    per_pair = t1 - t0
    d = ref["disp"]
    within = float((np.abs(d - gpu_disp) <= 0.5).mean())
    identical = bool(np.array_equal(d.view(np.uint32),
                                    gpu_disp.view(np.uint32)))
    cL, cR, cgt = synth.make_pair(C1["W"], C1["H"], cells=C1["cells"],
                                  dmax=C1["dmax"], channels=1,
                                  seed=C1["seed"])
    times = []
    for k in range(C1_RUNS + 1):
        s0 = time.perf_counter()
        c1 = O.run(cL, cR, C1["radius"], C1["dmax"], C1["iters"],
                   maxdisp=MAXDISP, maxdeg=MAXDEG)
        s1 = time.perf_counter()
        if k > 0:                       # the first run is the warm-up
            times.append(s1 - s0)
    c1_s = float(np.median(times))
    return {"value": 1.0 / per_pair, "unit": "pairs/s",
            "cores": O.num_threads(), "kind": "port", "host": host_cpu(),
            "sample": "the same 1280x720 pair, one full run (setup + %d "
                      "iterations + LR check: %.1f s)" % (ITERS, per_pair),
            "ms_per_pair": per_pair * 1e3,
            "bad_1px": disparity.bad_pixel_rate(d, gt, 1.0),
            "gpu_disp_within_0.5px": within,
            "gpu_disp_bit_identical": identical,
            "c1": {"workload": "320x240 gray f32 pair, 8 cells, disparity "
                               "0-32 px, 11x11 patch, 3 iterations, census "
                               "cost, LR check 1 px",
                   "pairs_per_s": 1.0 / c1_s, "ms_per_pair": c1_s * 1e3,
                   "runs": "1 warm-up + median of %d" % C1_RUNS,
                   "bad_1px": disparity.bad_pixel_rate(c1["disp"], cgt,
                                                       1.0)}}


# algorithmic HBM bytes per pixel of one k_rof_steps launch (two ROF
# steps; the state between them stays on chip): read u, q, f, tensor
# (16 B each) and the two-float4 dual p (32 B); write u, q (16 B each) and
# p (32 B)
ROF_BYTES_PER_PIXEL = 96 + 64
ROF_STEPS_PER_LAUNCH = 2


def time_rof_launch(ctx, stream):
    """Average duration of one k_rof_steps launch, from HIP events on the
    engine's stream around pm_rof_device runs of 10 and 50 steps on a field
    in HBM (best of three each): the difference, 20 launches, removes the
    buffer set-up both runs share (one copy and two fills)."""
    import numpy as np
    import torch
    H, Wd = ctx.shape
    gen = torch.Generator(device=stream.device).manual_seed(0)
    f = torch.randn((H, Wd, 4), generator=gen, device=stream.device)
    out = torch.empty_like(f)
    lam = np.float32([1.5, 1.5, 1.5, 1.0])
    best = {}
    with torch.cuda.stream(stream):
        ctx.rof_device(0, f.data_ptr(), lam, 2, out.data_ptr())
        for n in (10, 50):
            best[n] = float("inf")
            for _ in range(3):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                ctx.rof_device(0, f.data_ptr(), lam, n, out.data_ptr())
                e1.record(stream)
                torch.cuda.synchronize()
                best[n] = min(best[n], e0.elapsed_time(e1))
    return (best[50] - best[10]) / (40 / ROF_STEPS_PER_LAUNCH)


def cpu_baseline_huber(L, R):
    """The oracle on a bounded sample of the PM-Huber pipeline: setup
    (texels, diffusion tensors, init) and one iteration (two propagates,
    two smoothings of 40 ROF steps); per pair = setup + 8 x iteration."""
    import numpy as np
    from oracle import oracle as O
    t0 = time.perf_counter()
    tl, tr = O.prepare(L, RADIUS), O.prepare(R, RADIUS)
    dtl, dtr = O.diffusion_tensor(L), O.diffusion_tensor(R)
    sl = O.init(tl, tr, RADIUS, DMAX, 1)
    sr = O.init(tr, tl, RADIUS, DMAX, 0)
    sml = np.zeros(sl.shape, np.float32)
    smr = np.zeros(sr.shape, np.float32)
    t1 = time.perf_counter()
    th = 1e-4
    lam = np.float32([th * 10 + 1] * 3 + [1.0])
    vb2 = O.viewbuf_empty(W, H)
    sl, vb1 = O.propagate(tl, tr, RADIUS, DMAX, 1, 0, sl, vb2, sr, smooth=sml,
                          theta=th * 50)
    sr, vb2 = O.propagate(tr, tl, RADIUS, DMAX, 0, 0, sr, vb1, sl, smooth=smr,
                          theta=th * 50)
    sml = O.rof(O.normal_depth(sl, 1.0 / DMAX, 1), dtl, lam, 40)
    smr = O.rof(O.normal_depth(sr, 1.0 / DMAX, 0), dtr, lam, 40)
    t2 = time.perf_counter()
    per_pair = (t1 - t0) + ITERS * (t2 - t1)
    return {"value": 1.0 / per_pair, "unit": "pairs/s",
            "cores": O.num_threads(), "kind": "port", "host": host_cpu(),
            "sample": "1280x720 pair, PM-Huber: setup + 1 of %d iterations "
                      "timed (%.1f s); per pair = setup %.2f s + %d x "
                      "iteration %.2f s" % (ITERS, t2 - t0, t1 - t0, ITERS,
                                            t2 - t1),
            "ms_per_pair": per_pair * 1e3}


BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 MFMA peak


def learned_section(dL, dR, gt, stream, cost_kind, cost_weight, plain_bad,
                    steps, warmup):
    """Learned-init PatchMatch on the 720p pair: the proposal network seeds
    both views (pm_setup_proposal: one network pass over the pair and the
    mirrored pair, each pixel keeping the proposed plane unless the random
    init plane costs less) or the left view only (pm_setup_proposal_views
    with views = 1: a pass over the pair, the right view from the random
    init), then k refine iterations, in the reference's Jacobi sweep and in
    the red-black schedule.  For each of the four: bad-1.0px after k =
    0..8 refine iterations, and k = the fewest that match the plain
    8-iteration bad-1.0px (8 when none does), timed like the plain step
    (set images from HBM, setup, iterations, LR check, disparity).  The
    line reports the fastest matched one.  None when no trained weights
    file exists."""
    import numpy as np
    import torch

    import cvt_amd
    from cvt_amd.net import WEIGHTS, ProposalNet, forward_flops, padded
    if not os.path.exists(WEIGHTS):
        return None
    dD = torch.empty((H, W), dtype=torch.float32, device=dL.device)
    net = ProposalNet(stream=stream.cuda_stream, weights=WEIGHTS)
    runs = {}
    for name, sched, views in (
            ("jacobi", cvt_amd.PM_SCHEDULE_JACOBI, 2),
            ("red_black", cvt_amd.PM_SCHEDULE_RED_BLACK, 2),
            ("jacobi_left_seeded", cvt_amd.PM_SCHEDULE_JACOBI, 1),
            ("red_black_left_seeded", cvt_amd.PM_SCHEDULE_RED_BLACK, 1)):
        ctx = cvt_amd.Context(radius=RADIUS, disparity_max=DMAX,
                              stream=stream.cuda_stream, cost_kind=cost_kind,
                              cost_weight=cost_weight, schedule=sched)
        bad = []
        with torch.cuda.stream(stream):
            for k in range(ITERS + 1):
                ctx.set_images_device(dL.data_ptr(), dR.data_ptr(), W, H, 4)
                ctx.setup_proposal(net, views)
                if k:
                    ctx.iterate(k)
                ctx.disparity_device(dD.data_ptr(),
                                     cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP,
                                     MAXDEG)
                torch.cuda.synchronize()
                bad.append(cvt_amd.disparity.bad_pixel_rate(
                    dD.cpu().numpy(), gt, 1.0))
            matched = matched_iterations(bad, plain_bad)
            k = matched if matched is not None else ITERS

            def step(ev=None):
                ctx.set_images_device(dL.data_ptr(), dR.data_ptr(), W, H, 4)
                if ev:
                    ev[0].record(stream)
                ctx.setup_proposal(net, views)
                if ev:
                    ev[1].record(stream)
                ctx.iterate(k)
                if ev:
                    ev[2].record(stream)
                ctx.disparity_device(dD.data_ptr(),
                                     cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP,
                                     MAXDEG)

            for _ in range(warmup):
                step()
            ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)]
                  for _ in range(steps)]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for i in range(steps):
                step(ev[i])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            disp = dD.cpu().numpy()
        ctx.close()
        runs[name] = {
            "value": steps / (t1 - t0), "ms_per_step": (t1 - t0) / steps * 1e3,
            "seeded_views": views,
            "refine_iterations": k, "matched": matched is not None,
            "bad_1px": cvt_amd.disparity.bad_pixel_rate(disp, gt, 1.0),
            "bad_1px_by_refine_iterations": bad,
            "setup_ms": float(np.mean([e[0].elapsed_time(e[1]) for e in ev])),
            "iterations_ms": float(np.mean([e[1].elapsed_time(e[2])
                                            for e in ev]))}
    best = fastest_matched(runs, "jacobi")
    views = runs[best]["seeded_views"]
    # the network alone as the chosen setup runs it: one pass over the
    # padded pair (and its mirror, as a batch of two, when both views are
    # seeded); best of five after a warm-up
    Hp, Wp = padded(H, W)
    x = torch.rand((views, Hp, Wp, 8), device=dL.device)
    out = torch.empty((views, Hp, Wp, 4), device=dL.device)
    fwd = []
    with torch.cuda.stream(stream):
        for _ in range(6):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            net.forward(x.data_ptr(), views, Hp, Wp, out.data_ptr())
            e1.record(stream)
            torch.cuda.synchronize()
            fwd.append(e0.elapsed_time(e1))
    fwd_ms = float(np.min(fwd[1:]))
    net.close()
    flops = forward_flops(H, W, N=views)
    tflops = flops / (fwd_ms * 1e-3) / 1e12
    out = dict(runs[best])
    out.update({
        "schedule": best, "schedules": runs,
        "weights": os.path.relpath(WEIGHTS, ROOT),
        "roofline_net": {"bound": "mfma", "kernel": "k_conv (forward pass)",
                         "achieved": tflops, "peak": BF16_PEAK_TFLOPS,
                         "unit": "TFLOP/s",
                         "frac": tflops / BF16_PEAK_TFLOPS,
                         "traffic": None,
                         "forward_ms": fwd_ms, "flop_per_forward": flops,
                         "forward": ("one pass over the pair and its mirror "
                                     "(batch of two)" if views == 2 else
                                     "one pass over the pair")}})
    return out


# the batched configuration: 64 pairs of seeded uneven sizes
BATCH_PAIRS, BATCH_SEED = 64, 3


def batched_section(stream, cost_kind, cost_weight, runs=2):
    """Batched plain PatchMatch (pm_batch): 64 RGBA pairs from the GPU
    generator, width uniform in 320-1920, height in 240-1080, 24 cells,
    disparity range 0.1 x width; 8 iterations, LR check at 1 px and the
    error counts, all pairs in one graph launch per run, after the
    autotuner picked the launch shape.  Beside it, the same pairs one after
    another through one single-pair context; every pair's disparity is
    compared bit for bit."""
    import numpy as np
    import torch

    import cvt_amd
    from cvt_amd import synth_gpu
    from cvt_amd.batch import Batch
    rng = np.random.default_rng(BATCH_SEED)
    sizes = [(int(rng.integers(320, 1921)), int(rng.integers(240, 1081)))
             for _ in range(BATCH_PAIRS)]
    data = []
    for i, (w, h) in enumerate(sizes):
        L, R, g = synth_gpu.make_pairs(1, w, h, cells=24, dmax=0.1 * w,
                                       seed=BATCH_SEED * 1000 + i)
        data.append((L[0], R[0], g[0]))
    torch.cuda.synchronize()
    pairs = [{"width": w, "height": h, "channels": 4, "disparity_max": 0.1 * w,
              "left": L.data_ptr(), "right": R.data_ptr(),
              "gt": g.data_ptr()}
             for (w, h), (L, R, g) in zip(sizes, data)]
    npix = sum(w * h for w, h in sizes)
    with Batch(pairs, radius=RADIUS, cost_kind=cost_kind,
               cost_weight=cost_weight, stream=stream.cuda_stream) as b:
        table = b.autotune(reps=2)
        tuning = b.tuning
        b.run(ITERS, cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP, MAXDEG, 1.0, 0.0)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(runs):
            b.run(ITERS, cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP, MAXDEG, 1.0,
                  0.0)
        b.sync()
        t1 = time.perf_counter()
        counts = b.error_counts().astype(np.float64)
        bdisp = [b.disparity(i) for i in range(len(sizes))]
    # the same pairs one after another, each through a single-pair context
    # of its own (its disparity range is a context parameter), after one
    # untimed pass that sizes every context's buffers and streams
    outs = [torch.empty((h, w), device="cuda") for w, h in sizes]
    ctxs = [cvt_amd.Context(radius=RADIUS, disparity_max=0.1 * w,
                            stream=stream.cuda_stream, cost_kind=cost_kind,
                            cost_weight=cost_weight) for w, h in sizes]

    def one(i):
        (w, h), (L, R, _) = sizes[i], data[i]
        ctxs[i].set_images_device(L.data_ptr(), R.data_ptr(), w, h, 4)
        ctxs[i].setup()
        ctxs[i].iterate(ITERS)
        ctxs[i].disparity_device(outs[i].data_ptr(),
                                 cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP,
                                 MAXDEG)

    with torch.cuda.stream(stream):
        for i in range(len(sizes)):
            one(i)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        for i in range(len(sizes)):
            one(i)
        torch.cuda.synchronize()
        s1 = time.perf_counter()
    for c in ctxs:
        c.close()
    identical = all(np.array_equal(bdisp[i].view(np.uint32),
                                   outs[i].cpu().numpy().view(np.uint32))
                    for i in range(len(sizes)))
    bad = counts[:, 0] / counts[:, 2]
    return {"workload": "batched plain PatchMatch: 64 RGBA f32 pairs, width "
                        "320-1920, height 240-1080 (seeded), 24 cells, "
                        "disparity 0.1 x width, 9x9 patch, 8 iterations, LR "
                        "check 1 px, one graph launch per batch",
            "value": BATCH_PAIRS * runs / (t1 - t0), "unit": "pairs/s",
            "ms_per_batch": (t1 - t0) / runs * 1e3,
            "mpixel_per_s": npix * runs / (t1 - t0) / 1e6,
            "sequential_single_pair": {
                "value": BATCH_PAIRS / (s1 - s0), "unit": "pairs/s",
                "ms_per_batch": (s1 - s0) * 1e3},
            "speedup_vs_sequential": (s1 - s0) / ((t1 - t0) / runs),
            "bit_identical_to_single_pair": identical,
            "bad_1px_mean": float(np.mean(bad)),
            "bad_1px_pooled": float(counts[:, 0].sum() / counts[:, 2].sum()),
            "tuning": tuning, "autotune_candidates": table,
            "pixels": npix, "seed": BATCH_SEED}


def fastest_matched(runs, fallback):
    """the key of the fastest run (by "value") among those that matched
    the plain run's quality ("matched"), or fallback when none did"""
    matched = [n for n in runs if runs[n]["matched"]]
    return max(matched, key=lambda n: runs[n]["value"]) if matched \
        else fallback


def matched_iterations(bad, target):
    """the fewest refine iterations k >= 1 whose bad-1.0px (bad[k]) is at
    most the target, or None"""
    return next((k for k in range(1, len(bad)) if bad[k] <= target), None)


# the batched serving mix: 16 pairs of seeded uneven sizes
SERVE_PAIRS, SERVE_SEED = 16, 5


def serving_section(stream, cost_kind, cost_weight, runs=3):
    """Batched learned serving mix: 16 RGBA pairs from the GPU generator,
    width uniform in 320-1920, height in 240-1080, disparity range 0.1 x
    width (up to 192 px); plain 8-iteration batched PatchMatch on them,
    then the proposal network seeding both views of every pair, or the
    left views only, and k refine iterations, k the fewest whose pooled
    bad-1.0px (from the on-device error counts) is at most plain's; the
    faster matched seeding is reported; each run one graph launch.  None
    without a trained weights file."""
    import numpy as np
    import torch

    import cvt_amd
    from cvt_amd import synth_gpu
    from cvt_amd.batch import Batch
    from cvt_amd.net import WEIGHTS, ProposalNet
    if not os.path.exists(WEIGHTS):
        return None
    rng = np.random.default_rng(SERVE_SEED)
    sizes = [(int(rng.integers(320, 1921)), int(rng.integers(240, 1081)))
             for _ in range(SERVE_PAIRS)]
    data = []
    for i, (w, h) in enumerate(sizes):
        L, R, g = synth_gpu.make_pairs(1, w, h, cells=24, dmax=0.1 * w,
                                       seed=SERVE_SEED * 1000 + i)
        data.append((L[0], R[0], g[0]))
    torch.cuda.synchronize()
    pairs = [{"width": w, "height": h, "channels": 4, "disparity_max": 0.1 * w,
              "left": L.data_ptr(), "right": R.data_ptr(),
              "gt": g.data_ptr()}
             for (w, h), (L, R, g) in zip(sizes, data)]
    out = {}
    with ProposalNet(stream=stream.cuda_stream, weights=WEIGHTS) as net, \
            Batch(pairs, radius=RADIUS, cost_kind=cost_kind,
                  cost_weight=cost_weight, stream=stream.cuda_stream) as b:
        b.autotune(reps=2)
        pp = cvt_amd.DISPARITY_PP_LRCHECK

        def pooled_bad():
            c = b.error_counts().astype(np.float64)
            return (float(c[:, 0].sum() / c[:, 2].sum()),
                    float(np.mean(c[:, 0] / c[:, 2])))

        def timed(iters):
            b.run(iters, pp, MAXDISP, MAXDEG, 1.0, 0.0)
            b.sync()
            t0 = time.perf_counter()
            for _ in range(runs):
                b.run(iters, pp, MAXDISP, MAXDEG, 1.0, 0.0)
            b.sync()
            t1 = time.perf_counter()
            pooled, mean = pooled_bad()
            return {"value": SERVE_PAIRS * runs / (t1 - t0), "unit": "pairs/s",
                    "ms_per_batch": (t1 - t0) / runs * 1e3,
                    "iterations": iters, "bad_1px_mean": mean,
                    "bad_1px_pooled": pooled}

        b.set_proposal(None)
        out["plain"] = timed(ITERS)
        # learned, both views or the left view seeded: bad-1px after k = 0,
        # 1, ... refine iterations until it reaches plain's
        target = out["plain"]["bad_1px_pooled"]
        seeded = {}
        for views in (2, 1):
            b.set_proposal(net, views)
            bad = []
            for k in range(ITERS + 1):
                b.run(k, pp, MAXDISP, MAXDEG, 1.0, 0.0)
                bad.append(pooled_bad()[0])
                if matched_iterations(bad, target):
                    break
            matched = matched_iterations(bad, target)
            r = timed(matched if matched is not None else ITERS)
            r["matched"] = matched is not None
            r["seeded_views"] = views
            r["bad_1px_pooled_by_refine_iterations"] = bad
            seeded[views] = r
        out["learned"] = dict(seeded[fastest_matched(seeded, 2)])
        out["learned"]["other_seeding"] = seeded[
            1 if out["learned"]["seeded_views"] == 2 else 2]
        tuning = b.tuning
    return {"workload": "batched serving mix: 16 RGBA f32 pairs, width "
                        "320-1920, height 240-1080 (seeded), 24 cells, "
                        "disparity 0.1 x width (<= 192 px), 9x9 patch, LR "
                        "check 1 px; learned: proposal network seeding both "
                        "views or the left view (the faster matched one) + "
                        "the fewest refine iterations whose pooled bad-1px "
                        "is at most plain's, plain: 8 iterations; one graph "
                        "launch per batch",
            "learned": out["learned"], "plain": out["plain"],
            "speedup_vs_plain": out["learned"]["value"] /
            out["plain"]["value"],
            "tuning": tuning, "seed": SERVE_SEED}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-schedules", action="store_true",
                    help="skip the bad-1px-by-iteration comparison of the "
                         "Jacobi and red-black schedules")
    ap.add_argument("--pipeline", choices=["plain", "huber"], default="plain",
                    help="plain PatchMatch (the headline workload) or the "
                         "PM-Huber pipeline (theta schedule plus "
                         "diffusion-tensor ROF smoothing after every "
                         "iteration)")
    ap.add_argument("--cost", choices=sorted(COSTS), default="color_grad",
                    help="candidate cost: the gamma-10 truncated "
                         "colour+gradient cost (default) or the reference's "
                         "census cost")
    ap.add_argument("--io-steps", type=int, default=2,
                    help="pairs run from and to host memory after the timed "
                         "region (PCIe-inclusive rate)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-init run (the line is then plain "
                         "PatchMatch)")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched 64-pair run")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the engine's plain copy kernel over 512 "
                         "MiB (calibration run for the HBM byte counters)")
    ap.add_argument("--net-probe", action="store_true",
                    help="only run the proposal network's forward pass over "
                         "the padded 720p pair and its mirror (batch of "
                         "two), 12 passes (profiling run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the PatchMatch engine runs on a "
              "single MI355X and has no multi-GPU path (--gpus %d given)"
              % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import cvt_amd
    from cvt_amd import synth

    if args.copy_probe:
        nbytes = 512 << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        stream = torch.cuda.Stream()
        with cvt_amd.Context(radius=1, disparity_max=1.0,
                             stream=stream.cuda_stream) as ctx:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            reps = 5
            with torch.cuda.stream(stream):
                e0.record(stream)
                for _ in range(reps):
                    ctx.copy_probe(src.data_ptr(), dst.data_ptr(), nbytes)
                e1.record(stream)
                torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        assert bool((dst == 1).all())
        print(json.dumps({"copy_probe_bytes_read": nbytes,
                          "copy_probe_bytes_written": nbytes,
                          "avg_ms": ms,
                          "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9}))
        return 0

    if args.net_probe:
        from cvt_amd.net import WEIGHTS, ProposalNet, forward_flops, padded
        Hp, Wp = padded(H, W)
        x = torch.rand((2, Hp, Wp, 8), device="cuda:0")
        out = torch.empty((2, Hp, Wp, 4), device="cuda:0")
        stream = torch.cuda.Stream()
        with ProposalNet(stream=stream.cuda_stream, weights=WEIGHTS) as net:
            with torch.cuda.stream(stream):
                net.forward(x.data_ptr(), 2, Hp, Wp, out.data_ptr())
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                reps = 11
                e0.record(stream)
                for _ in range(reps):
                    net.forward(x.data_ptr(), 2, Hp, Wp, out.data_ptr())
                e1.record(stream)
                torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        print(json.dumps({"net_forward_ms": ms, "passes": reps,
                          "TFLOP_per_s": forward_flops(H, W, N=2) /
                          (ms * 1e-3) / 1e12}))
        return 0

    log("bench: generating the 1280x720 pair")
    L, R, gt = synth.make_pair(W, H, cells=CELLS, dmax=DMAX, channels=4,
                               seed=SEED)
    dev = torch.device("cuda:0")
    dL = torch.from_numpy(L).to(dev)
    dR = torch.from_numpy(R).to(dev)
    dD = torch.empty((H, W), dtype=torch.float32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    huber = args.pipeline == "huber"
    cost_kind, cost_weight = COSTS[args.cost]
    ctx = cvt_amd.Context(
        radius=RADIUS, disparity_max=DMAX, stream=stream.cuda_stream,
        smoothing=cvt_amd.PM_SMOOTH_HUBER if huber else cvt_amd.PM_SMOOTH_NONE,
        schedule_iterations=ITERS, cost_kind=cost_kind,
        cost_weight=cost_weight)

    def step(events=None):
        ctx.set_images_device(dL.data_ptr(), dR.data_ptr(), W, H, 4)
        ctx.setup()
        # the iterations as one call, as disparityMap runs them: the engine
        # then overlaps consecutive propagates (two half launches each)
        if events is not None:
            events[0].record(stream)
        ctx.iterate(ITERS)
        if events is not None:
            events[1].record(stream)
        ctx.disparity_device(dD.data_ptr(), cvt_amd.DISPARITY_PP_LRCHECK,
                             MAXDISP, MAXDEG)

    with torch.cuda.stream(stream):
        for i in range(args.warmup):
            step()
            log("bench: warmup step %d done" % (i + 1))
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)]
              for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(ev[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        log("bench: %d timed steps done" % args.steps)
        disp = dD.cpu().numpy()
        # per iteration = one propagate per view, each two k_propagate
        # half launches on two streams (+ the 8 B/pixel slot clears, <
        # 0.1 %); a propagate's time is its share of the iterations' span
        iter_ms = [ev[k][0].elapsed_time(ev[k][1]) / ITERS
                   for k in range(args.steps)]
        prop_ms = float(np.mean(iter_ms)) / 2.0
        # the same pair from and to host memory (pm_set_images /
        # pm_disparity): the PCIe-inclusive rate, reported beside `value`
        t2 = time.perf_counter()
        for k in range(args.io_steps):
            ctx.set_images(L, R)
            ctx.setup()
            ctx.iterate(ITERS)
            ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, MAXDISP, MAXDEG)
        t3 = time.perf_counter()
        # untimed counting pass: identical inputs, identical work
        ctx.set_images_device(dL.data_ptr(), dR.data_ptr(), W, H, 4)
        ctx.setup()
        cost_taps, klt_taps = ctx.iterate_counted(ITERS)
        torch.cuda.synchronize()
        if huber:
            rof_ms = time_rof_launch(ctx, stream)
    ctx.close()

    propagates = 2 * ITERS
    flop = (FLOP_PER_COST_TAP[args.cost] * cost_taps +
            FLOP_PER_KLT_TAP * klt_taps) / propagates
    achieved = flop / (prop_ms * 1e-3) / 1e12
    elapsed = t1 - t0
    value = args.steps / elapsed
    bad1 = cvt_amd.disparity.bad_pixel_rate(disp, gt, 1.0)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Voronoi slanted planes, Perlin texture)",
        "config": {"workload": "plain PatchMatch, 1280x720 RGBA f32 pair, 24 "
                               "planar cells, disparity 0-128 px, 9x9 patch, "
                               "%s, 8 iterations, LR check 1 px" %
                               ("gamma-10 truncated colour+gradient cost "
                                "(weight 25.5)" if args.cost == "color_grad"
                                else "census cost (weight 12)"),
                   "variant": "plain", "cost": args.cost,
                   "cost_weight": cost_weight, "width": W, "height": H,
                   "radius": RADIUS, "iterations": ITERS,
                   "disparity_max": DMAX, "seed": SEED},
        "quality": {"bad_1px": bad1},
        "host_io": ({"value": args.io_steps / (t3 - t2), "unit": "pairs/s",
                     "ms_per_pair": (t3 - t2) / args.io_steps * 1e3,
                     "note": "pair uploaded from and disparity returned to "
                             "host memory per step"}
                    if args.io_steps > 0 else None),
        "roofline": {"bound": "valu", "kernel": "k_propagate",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": profiled_traffic(),
                     "traffic_unit": "bytes per view propagate",
                     "avg_propagate_ms": prop_ms,
                     "launches_per_propagate": 2,
                     "cost_taps_per_propagate": cost_taps / propagates,
                     "klt_taps_per_propagate": klt_taps / propagates},
    }
    if huber:
        # the PM-Huber pipeline: same workload, smoothing after every
        # iteration; its added kernel runs two ROF steps of an HBM stencil
        # per launch
        result["metric"] = ("720p stereo pairs/s, PM-Huber pipeline "
                            "(PatchMatch + diffusion-tensor ROF smoothing)")
        result["config"]["variant"] = "huber"
        result["config"]["workload"] = (
            "PM-Huber PatchMatch, 1280x720 RGBA f32 pair, 24 planar cells, "
            "disparity 0-128 px, 9x9 patch, 8 iterations with 2 x 40 ROF "
            "steps each, LR check 1 px")
        gbs = ROF_BYTES_PER_PIXEL * W * H / (rof_ms * 1e-3) / 1e9
        result["roofline"] = {
            "bound": "hbm", "kernel": "k_rof_steps", "achieved": gbs,
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
            "traffic": profiled_traffic(ROF_HBM_PROFILE, ROF_KERNEL_SOURCES,
                                        "hbm_bytes_per_launch"),
            "traffic_unit": "bytes per k_rof_steps launch",
            "avg_launch_ms": rof_ms,
            "steps_per_launch": ROF_STEPS_PER_LAUNCH,
            "iteration_ms": float(np.mean(iter_ms))}
    if not huber and not args.no_learned:
        log("bench: learned-init PatchMatch (proposal network + refine)")
        learned = learned_section(dL, dR, gt, stream, cost_kind, cost_weight,
                                  bad1, args.steps, args.warmup)
        if learned is None:
            result["learned"] = {"note": "no proposal weights file (%s); "
                                         "the line is plain PatchMatch" %
                                         os.path.relpath(
                                             cvt_amd.net.WEIGHTS, ROOT)}
        else:
            # the headline: learned-init pairs/s at the plain run's bad-1px
            plain = {"value": value, "ms_per_step": result["ms_per_step"],
                     "bad_1px": bad1, "iterations": ITERS}
            k = learned["refine_iterations"]
            result["value"] = learned["value"]
            result["ms_per_step"] = learned["ms_per_step"]
            result["config"]["variant"] = "learned"
            result["config"]["refine_iterations"] = k
            result["config"]["refine_schedule"] = (
                "red_black" if learned["schedule"].startswith("red")
                else "jacobi")
            result["config"]["seeded_views"] = learned["seeded_views"]
            seeding = ("on the pair and its mirror) seeding both views"
                       if learned["seeded_views"] == 2 else
                       "on the pair) seeding the left view, the right view "
                       "from the random init,")
            sched = ("red-black" if learned["schedule"].startswith("red")
                     else "jacobi")
            result["config"]["workload"] = (
                "learned-init PatchMatch: bf16 proposal network (4-level "
                "32-64-96-128 U-Net %s with per-pixel fallback to random "
                "init, then %d refine iteration%s (%s schedule); 1280x720 "
                "RGBA f32 pair, 24 planar cells, disparity 0-128 px, 9x9 "
                "patch, %s, LR check 1 px; against plain PatchMatch with 8 "
                "iterations"
                % (seeding, k, "" if k == 1 else "s", sched,
                   "gamma-10 truncated colour+gradient cost (weight 25.5)"
                   if args.cost == "color_grad" else "census cost (weight 12)"))
            result["quality"] = {"bad_1px": learned["bad_1px"],
                                 "plain_bad_1px": bad1,
                                 "matched": learned["matched"]}
            result["plain"] = plain
            result["learned"] = learned
            result["speedup_vs_plain"] = learned["value"] / value
            result["dtype"] = "f32 (PatchMatch), bf16/f32-acc (network)"
    if not huber and not args.no_batched:
        log("bench: batched PatchMatch, 64 pairs of uneven sizes")
        result["batched"] = batched_section(stream, cost_kind, cost_weight)
        log("bench: batched serving mix, 16 pairs, learned vs plain")
        serving = serving_section(stream, cost_kind, cost_weight)
        if serving is not None:
            result["serving"] = serving
    if not huber and not args.no_schedules:
        log("bench: bad-1px by iteration, Jacobi and red-black schedules")
        result["schedules"] = schedules_by_iteration(dL, dR, gt, cost_kind,
                                                     cost_weight, stream)
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle) on a bounded sample")
        result["cpu_baseline"] = (cpu_baseline_huber(L, R) if huber
                                  else cpu_baseline(L, R, gt, args.cost,
                                                    disp))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
