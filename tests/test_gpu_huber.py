"""Parity of the PM-Huber smoothing coupling (PM_SMOOTH_HUBER) with the
oracle, through the C ABI: diffusion tensor, ROF smoothing, the coupled
iterations with the theta schedule, and the prior-seeded smooth field.
Floats are compared bit for bit."""
import numpy as np
import pytest

import cvt_amd

pytestmark = pytest.mark.gpu

F = np.float32


def bits(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


def assert_same(got, want, what):
    d = np.argwhere(bits(got) != bits(want))
    assert d.size == 0, "%s: %d differ, first %s: %r vs %r" % (
        what, len(d), tuple(d[0]), got[tuple(d[0])], want[tuple(d[0])])


def huber_ctx(r, dmax, **kw):
    return cvt_amd.Context(radius=r, disparity_max=dmax,
                           smoothing=cvt_amd.PM_SMOOTH_HUBER, **kw)


@pytest.fixture(scope="module")
def pair(synth):
    return synth.make_pair(131, 77, cells=5, dmax=20.0, channels=4, seed=12)


@pytest.fixture(scope="module")
def pair_gray(synth):
    return synth.make_pair(96, 64, cells=4, dmax=16.0, channels=1, seed=13)


@pytest.mark.parametrize("which", ["rgba", "gray"])
def test_tensor_bitexact(oracle, pair, pair_gray, which):
    L, R, _ = pair if which == "rgba" else pair_gray
    print("tensor", which, L.shape, flush=True)
    with huber_ctx(3, 16.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, img in ((0, L), (1, R)):
            assert_same(ctx.tensor(view), oracle.diffusion_tensor(img),
                        "tensor view %d" % view)


@pytest.mark.parametrize("iters", [40, 3])
def test_rof_bitexact(oracle, pair, iters):
    L, R, _ = pair
    H, W = L.shape[:2]
    rng = np.random.default_rng(iters)
    f = rng.normal(0, 1, (H, W, 4)).astype(np.float32)
    lam = F([1.3, 1.3, 1.3, 1.0])
    print("rof", iters, flush=True)
    with huber_ctx(3, 20.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        got = ctx.rof(1, f, lam, iters)
    want = oracle.rof(f, oracle.diffusion_tensor(R), lam, iters)
    assert_same(got, want, "rof")


@pytest.mark.parametrize("which,r,dmax,n,tmin", [
    ("rgba", 3, 20.0, 3, 1e-4),
    ("gray", 5, 16.0, 3, 1e-4),     # 11 x 11 patch, CL_INTENSITY reads
    ("rgba", 3, 20.0, 2, 0.05),     # a larger theta floor
])
def test_huber_iterations_bitexact(oracle, pair, pair_gray, which, r, dmax,
                                   n, tmin):
    L, R, _ = pair if which == "rgba" else pair_gray
    print("huber pipeline", which, r, n, tmin, flush=True)
    with huber_ctx(r, dmax, theta_min=tmin) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(n)
        sl, sr = ctx.planes(0), ctx.planes(1)
        ml, mr = ctx.smooth(0), ctx.smooth(1)
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
    ref = oracle.run_huber(L, R, r, dmax, n, tmin, 1.0, 5.0)
    assert_same(sl, ref["stateL"], "left planes")
    assert_same(sr, ref["stateR"], "right planes")
    assert_same(ml, ref["smoothL"], "left smooth field")
    assert_same(mr, ref["smoothR"], "right smooth field")
    assert_same(d, ref["disp"], "disparity")


@pytest.mark.parametrize("smoothing,n,split", [
    (cvt_amd.PM_SMOOTH_HUBER, 3, (2, 1)),
    (cvt_amd.PM_SMOOTH_DEPTHMAP, 6, (4, 2)),   # the cubic part after 3
])
def test_schedule_over_split_calls(pair, smoothing, n, split):
    """With schedule_iterations = N, iterating in two calls gives what one
    call of N gives (PM-Huber: theta = (i / N)^4 + theta_min after
    iteration i; depthMap: ((i - 3) / (N - 4))^3 + 1e-3 from i = 3)."""
    L, R, _ = pair
    out = []
    for calls in ((n,), split):
        with cvt_amd.Context(radius=3, disparity_max=20.0,
                             smoothing=smoothing,
                             schedule_iterations=n) as ctx:
            ctx.set_images(L, R)
            ctx.setup()
            for k in calls:
                ctx.iterate(k)
            out.append((ctx.planes(0), ctx.smooth(1)))
    assert_same(out[1][0], out[0][0], "planes")
    assert_same(out[1][1], out[0][1], "smooth")


@pytest.mark.parametrize("smoothing", [cvt_amd.PM_SMOOTH_HUBER,
                                       cvt_amd.PM_SMOOTH_DEPTHMAP])
def test_split_calls_past_the_schedule(pair, smoothing):
    """Six calls of one iteration with schedule_iterations = 0 (each call's
    schedule ends with it) run past the schedule's end, where the
    reference never uses theta: theta holds at the end value instead of
    dividing by zero (depthMap's N - 4 at N = 4).  The fields stay finite;
    for depthMap they equal one call of six with schedule_iterations = 4,
    whose thetas after iterations 3 and 4 are the same held end value."""
    L, R, _ = pair
    out = []
    for sched, calls in ((0, (1,) * 6), (4, (6,))):
        with cvt_amd.Context(radius=3, disparity_max=20.0,
                             smoothing=smoothing,
                             schedule_iterations=sched) as ctx:
            ctx.set_images(L, R)
            ctx.setup()
            for k in calls:
                ctx.iterate(k)
            out.append((ctx.planes(0), ctx.planes(1), ctx.smooth(0),
                        ctx.smooth(1)))
    for a in out[0]:
        assert np.isfinite(a[..., :3]).all()
    if smoothing == cvt_amd.PM_SMOOTH_DEPTHMAP:
        for a, b in zip(out[0], out[1]):
            assert_same(a, b, "split calls past the schedule")


def test_prior_seeds_the_smooth_field(oracle, pair):
    L, R, gt = pair
    r, dmax = 3, 20.0
    H, W = gt.shape
    prior = gt.copy()
    prior[::3, ::5] = 0.0                 # some pixels without a prior
    print("huber prior", flush=True)
    with huber_ctx(r, dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup_prior(prior)
        sm = ctx.smooth(0)
        sl = ctx.planes(0)
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    want_sl = oracle.init_prior(tl, tr, r, dmax, 1, prior,
                                oracle.disparity_to_normal(prior))
    assert_same(sl, want_sl, "seeded planes")
    lam_v = F(F(1e-4) * F(10) + F(1))
    want = oracle.rof(oracle.normal_depth(want_sl, F(1) / F(dmax), 1),
                      oracle.diffusion_tensor(L),
                      F([lam_v, lam_v, lam_v, 1.0]), 40)
    assert_same(sm, want, "seeded smooth field")


def test_plain_context_has_no_smooth_field(pair):
    L, R, _ = pair
    with cvt_amd.Context(radius=2, disparity_max=8.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        with pytest.raises(cvt_amd.PMError):
            ctx.smooth(0)


def test_rof_device_matches_host(pair):
    """pm_rof_device on HBM buffers gives what pm_rof gives; overlapping
    buffers are refused."""
    import torch
    L, R, _ = pair
    H, W = L.shape[:2]
    f = np.random.default_rng(5).normal(0, 1, (H, W, 4)).astype(np.float32)
    lam = F([1.2, 1.2, 1.2, 1.0])
    print("rof device", flush=True)
    with huber_ctx(3, 20.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        want = ctx.rof(0, f, lam, 7)
        df = torch.from_numpy(f).to("cuda:0")
        dout = torch.empty_like(df)
        torch.cuda.synchronize()
        ctx.rof_device(0, df.data_ptr(), lam, 7, dout.data_ptr())
        ctx.sync()
        got = dout.cpu().numpy()
        with pytest.raises(cvt_amd.PMError):
            ctx.rof_device(0, df.data_ptr(), lam, 7, df.data_ptr())
    assert_same(got, want, "rof_device")


def test_huber_full_size_bitexact(oracle, synth):
    """The PM-Huber pipeline on the 720p bench pair: two coupled iterations
    (propagates, both smoothings beside them, theta schedule), the smooth
    fields and the disparity, bit for bit."""
    L, R, _ = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                              seed=0)
    r, dmax, n = 4, 128.0, 2
    print("huber full size", flush=True)
    ref = oracle.run_huber(L, R, r, dmax, n, 1e-4, 1.0, 5.0)
    with huber_ctx(r, dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(n)
        got = {"stateL": ctx.planes(0), "stateR": ctx.planes(1),
               "smoothL": ctx.smooth(0), "smoothR": ctx.smooth(1),
               "disp": ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)}
    for k, v in got.items():
        assert_same(v, ref[k], k)


def _random_huber_configs(n, seed=77):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        W = int(rng.integers(12, 140))
        H = int(rng.integers(6, 90))
        C = int(rng.choice([1, 4]))
        r = int(rng.integers(1, 7))
        dmax = float(rng.choice([6.0, 16.0, 40.0]))
        n_it = int(rng.integers(1, 4))
        tmin = float(rng.choice([1e-4, 1e-2]))
        out.append((W, H, C, r, dmax, n_it, tmin, int(rng.integers(0, 1000))))
    return out


@pytest.mark.parametrize("W,H,C,r,dmax,n,tmin,seed", _random_huber_configs(6))
def test_random_huber_configurations_bitexact(oracle, synth, W, H, C, r, dmax,
                                              n, tmin, seed):
    """A seeded sweep of the coupled pipeline over sizes, formats, radii,
    disparity ranges, iteration counts and theta floors: planes, smooth
    fields and the disparity against oracle.run_huber, bit for bit."""
    L, R, _ = synth.make_pair(W, H, cells=3, dmax=min(dmax, W / 3.0),
                              channels=C, seed=seed)
    print("random huber config", W, H, C, r, dmax, n, tmin, seed, flush=True)
    ref = oracle.run_huber(L, R, r, dmax, n, tmin, 1.0, 5.0)
    with huber_ctx(r, dmax, theta_min=tmin) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(n)
        got = {"stateL": ctx.planes(0), "stateR": ctx.planes(1),
               "smoothL": ctx.smooth(0), "smoothR": ctx.smooth(1),
               "disp": ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)}
    for k, v in got.items():
        assert_same(v, ref[k], k)


# ---- depthMap's coupling (PM_SMOOTH_DEPTHMAP) ------------------------------

def depth_ctx(r, dmax, **kw):
    return cvt_amd.Context(radius=r, disparity_max=dmax,
                           smoothing=cvt_amd.PM_SMOOTH_DEPTHMAP, **kw)


@pytest.mark.parametrize("which,r,dmax,n", [
    ("rgba", 3, 20.0, 1),
    ("rgba", 3, 20.0, 3),
    ("rgba", 2, 20.0, 5),     # iterations 3 and 4 take the cubic schedule
    ("gray", 4, 16.0, 3),
])
def test_depthmap_iterations_bitexact(oracle, pair, pair_gray, which, r, dmax,
                                      n):
    """depthMap's pipeline (PMHuberStereo.cpp:147-347): planes, the
    Huber-smoothed filled fields and the disparity against
    oracle.run_depthmap, bit for bit."""
    L, R, _ = pair if which == "rgba" else pair_gray
    print("depthmap pipeline", which, r, n, flush=True)
    with depth_ctx(r, dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(n)
        got = {"stateL": ctx.planes(0), "stateR": ctx.planes(1),
               "smoothL": ctx.smooth(0), "smoothR": ctx.smooth(1),
               "disp": ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)}
    ref = oracle.run_depthmap(L, R, r, dmax, n, 1.0, 5.0)
    for k, v in got.items():
        assert_same(v, ref[k], k)


def test_depthmap_full_size_bitexact(oracle, synth):
    """depthMap's pipeline on the 720p bench pair, two iterations."""
    L, R, _ = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                              seed=0)
    r, dmax, n = 4, 128.0, 2
    print("depthmap full size", flush=True)
    ref = oracle.run_depthmap(L, R, r, dmax, n, 1.0, 5.0)
    with depth_ctx(r, dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(n)
        got = {"stateL": ctx.planes(0), "stateR": ctx.planes(1),
               "smoothL": ctx.smooth(0), "smoothR": ctx.smooth(1),
               "disp": ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)}
    for k, v in got.items():
        assert_same(v, ref[k], k)


def test_depthmap_entry_scales_the_disparity(oracle, pair):
    """cvt_amd.PMHuberStereo.depthMap: depthMap's pipeline for exactly the
    given iteration count (no `| 1`), the disparity times dscale, and the
    default dscale 1 / depthmax."""
    L, R, _ = pair
    ref = oracle.run_depthmap(L, R, 3, 20.0, 2, 1.0, 5.0)
    pm = cvt_amd.PMHuberStereo()
    try:
        d = pm.depthMap(L, R, 3, 20.0, 2, dscale=0.5)
        assert_same(d, ref["disp"] * np.float32(0.5), "dscale 0.5")
        d = pm.depthMap(L, R, 3, 20.0, 2)
        assert_same(d, ref["disp"] * np.float32(1.0 / 20.0), "default")
        # the object's own pipeline is back for disparityMap
        assert pm._cfg[2] == cvt_amd.PM_SMOOTH_DEPTHMAP
        pm.disparityMap(L, R, 3, 20.0, 1)
        assert pm._cfg[2] == cvt_amd.PM_SMOOTH_HUBER
    finally:
        pm.close()
