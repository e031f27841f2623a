"""The colour+gradient cost (PM_COST_COLOR_GRAD, patch_eval_color_grad_
weighted, pmhstereo.cl:510-565) through the C ABI, bit for bit against the
oracle: the cost field (including planes that miss the other image, which
cost NaN), the initial fields, iterations, the whole pipeline, and the
720p workload with the gamma-10 weight (25.5) at full size."""
import numpy as np
import pytest

import cvt_amd
from test_gpu_parity import cost_close, first_diff, same

pytestmark = pytest.mark.gpu

CG = cvt_amd.PM_COST_COLOR_GRAD
KW = 25.5          # gamma 10 on [0, 1] data


@pytest.fixture(scope="module")
def rgba_small(synth):
    return synth.make_pair(157, 83, cells=6, dmax=24.0, channels=4, seed=7)


def ctx_cg(r, dmax, kw=KW):
    return cvt_amd.Context(radius=r, disparity_max=dmax, cost_kind=CG,
                           cost_weight=kw)


def test_cg_cost_field_bitexact(oracle, rgba_small):
    L, R, _ = rgba_small
    H, W = L.shape[:2]
    r = 4
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    rng = np.random.default_rng(14)
    planes = np.zeros((H, W, 4), np.float32)
    planes[..., 0] = 1 + rng.normal(0, 0.05, (H, W))
    planes[..., 1] = rng.normal(0, 0.05, (H, W))
    planes[..., 2] = -rng.uniform(0, 24, (H, W))
    planes[::17, ::13, 0] = np.nan         # non-finite planes: NaN
    planes[5::19, 3::11, 2] = 1e4          # outside the image: NaN
    planes[2::23, 1::7, 1] = 2e5           # steep in y: only its own row
    with ctx_cg(r, 24.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, (a, b) in enumerate(((tl, tr), (tr, tl))):
            got = ctx.cost(view, planes)
            want = oracle.cost_field(a, b, r, planes, cost_kind=1,
                                     cost_weight=KW)
            assert np.isnan(want).any()
            fin = np.isfinite(want)
            assert cost_close(got[fin], want[fin])
            assert same(got, want), first_diff(got, want)


@pytest.mark.parametrize("channels,r,dmax,iters", [(4, 4, 24.0, 3),
                                                   (1, 5, 32.0, 2)])
def test_cg_iterations_bitexact(oracle, synth, channels, r, dmax, iters):
    L, R, _ = synth.make_pair(143, 61, cells=5, dmax=dmax, channels=channels,
                              seed=17 + channels)
    cost = dict(cost_kind=1, cost_weight=KW)
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    sl = oracle.init(tl, tr, r, dmax, 1, **cost)
    sr = oracle.init(tr, tl, r, dmax, 0, **cost)
    vb2 = oracle.viewbuf_empty(L.shape[1], L.shape[0])
    with ctx_cg(r, dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, want in ((0, sl), (1, sr)):
            got = ctx.planes(view)
            assert same(got, want), ("init", view, first_diff(got, want))
        for it in range(iters):
            sl, vb1 = oracle.propagate(tl, tr, r, dmax, 1, it, sl, vb2, sr,
                                       **cost)
            sr, vb2 = oracle.propagate(tr, tl, r, dmax, 0, it, sr, vb1, sl,
                                       **cost)
            ctx.iterate(1)
            for view, want in ((0, sl), (1, sr)):
                got = ctx.planes(view)
                assert same(got, want), (it, view, first_diff(got, want))


def test_cg_pipeline_matches_oracle_run(oracle, synth):
    L, R, _ = synth.make_pair(211, 97, cells=6, dmax=20.0, channels=4,
                              seed=23)
    ref = oracle.run(L, R, 3, 20.0, 3, maxdisp=1.0, maxdeg=180.0,
                     cost_kind=1, cost_weight=16.0)
    with ctx_cg(3, 20.0, kw=16.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(3)
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
        for view, key in ((0, "stateL"), (1, "stateR")):
            got = ctx.planes(view)
            assert same(got, ref[key]), (view, first_diff(got, ref[key]))
    assert same(d, ref["disp"]), first_diff(d, ref["disp"])


def test_cg_counted_iteration_counts_every_tap(rgba_small):
    """No early exit: the counted cost taps of one iteration are the valid
    taps of every evaluated candidate, and the fields are unchanged by
    counting."""
    L, R, _ = rgba_small
    out = []
    for counted in (False, True):
        with ctx_cg(4, 24.0) as ctx:
            ctx.set_images(L, R)
            ctx.setup()
            if counted:
                taps, klt = ctx.iterate_counted(1)
                H, W = L.shape[:2]
                assert 0 < taps <= 2 * H * W * 10 * 81
                assert 0 < klt <= 2 * H * W * 6 * 81
            else:
                ctx.iterate(1)
            out.append((ctx.planes(0), ctx.planes(1)))
    assert same(out[0][0], out[1][0]) and same(out[0][1], out[1][1])


def full_size_eight_iterations(oracle, synth, cost_kind, kw):
    """The 720p workload over all 8 iterations against oracle.run: each
    iteration's per-pixel cost of both views (within 1e-4 relative, then
    bit for bit), the final plane fields and the LR-checked disparity (0.5
    px on >= 99.5 % of pixels, then bit for bit)."""
    L, R, _ = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                              seed=0)
    print("full size, 8 iterations, cost kind", cost_kind, flush=True)
    ref = oracle.run(L, R, 4, 128.0, 8, maxdisp=1.0, maxdeg=180.0,
                     trace=True, cost_kind=cost_kind, cost_weight=kw)
    print("oracle done", flush=True)
    with cvt_amd.Context(radius=4, disparity_max=128.0, cost_kind=cost_kind,
                         cost_weight=kw) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for it in range(8):
            ctx.iterate(1)
            for v in (0, 1):
                got = ctx.planes(v)[..., 3]
                want = ref["cost_trace"][it, v]
                fin = np.isfinite(want)
                assert np.array_equal(fin, np.isfinite(got)), (it, v)
                assert cost_close(got[fin], want[fin]), (it, v)
                assert same(got, want), (it, v, first_diff(got, want))
        for view, key in ((0, "stateL"), (1, "stateR")):
            got = ctx.planes(view)
            assert same(got, ref[key]), (view, first_diff(got, ref[key]))
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
    assert (np.abs(d - ref["disp"]) <= 0.5).mean() >= 0.995
    assert same(d, ref["disp"]), first_diff(d, ref["disp"])


def test_full_size_eight_iterations_census(oracle, synth):
    full_size_eight_iterations(oracle, synth, cvt_amd.PM_COST_CENSUS, 16.0)


def test_full_size_eight_iterations_color_grad(oracle, synth):
    full_size_eight_iterations(oracle, synth, CG, KW)


@pytest.mark.parametrize("cost_kind,kw", [(cvt_amd.PM_COST_CENSUS, 16.0),
                                          (CG, KW)])
@pytest.mark.parametrize("W,H,C,r", [(157, 83, 4, 4), (300, 61, 1, 5)])
def test_red_black_schedule_bitexact(oracle, synth, cost_kind, kw, W, H, C,
                                     r):
    """PM_SCHEDULE_RED_BLACK (not in the reference) against the oracle's
    red-black run: per-iteration costs, the final fields and the
    LR-checked disparity, bit for bit.  The widths leave a partial block of
    128 columns (a pass's block covers every other column of 128)."""
    L, R, _ = synth.make_pair(W, H, cells=6, dmax=24.0, channels=C, seed=W)
    iters = 3
    ref = oracle.run(L, R, r, 24.0, iters, maxdisp=1.0, maxdeg=180.0,
                     trace=True, cost_kind=cost_kind, cost_weight=kw,
                     schedule=1)
    with cvt_amd.Context(radius=r, disparity_max=24.0, cost_kind=cost_kind,
                         cost_weight=kw,
                         schedule=cvt_amd.PM_SCHEDULE_RED_BLACK) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for it in range(iters):
            ctx.iterate(1)
            for v in (0, 1):
                got = ctx.planes(v)[..., 3]
                want = ref["cost_trace"][it, v]
                assert same(got, want), (it, v, first_diff(got, want))
        for view, key in ((0, "stateL"), (1, "stateR")):
            got = ctx.planes(view)
            assert same(got, ref[key]), (view, first_diff(got, ref[key]))
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
    assert same(d, ref["disp"]), first_diff(d, ref["disp"])
