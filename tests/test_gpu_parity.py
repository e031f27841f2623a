"""Parity of the HIP engine with the oracle, through the C ABI.

The engine and the oracle evaluate the same operation sequence (see the
arithmetic contract in cvt_amd/csrc/pm_device.h and oracle/pm_oracle.h),
so integer and index results must match exactly and float results are
compared bit for bit; the looser tolerance bar (per-pixel cost within
1e-4 relative, final disparity within 0.5 px on >= 99.5% of pixels) is
asserted as well so a deviation is reported against it.
"""
import numpy as np
import pytest

import cvt_amd

pytestmark = pytest.mark.gpu

COST_RTOL = 1e-4


def bits(a):
    """float32 bit patterns with every NaN mapped to one pattern (x86 and
    gfx950 produce NaNs with different sign/payload bits)"""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


def same(a, b):
    """bitwise equality of float arrays, any NaN == any NaN"""
    return np.shape(a) == np.shape(b) and np.array_equal(bits(a), bits(b))


def first_diff(a, b):
    idx = np.argwhere(bits(a) != bits(b))
    return None if idx.size == 0 else (tuple(idx[0]), a[tuple(idx[0])],
                                       b[tuple(idx[0])], len(idx))


def cost_close(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.all(np.abs(a - b) <= COST_RTOL * np.maximum(np.abs(b), 1e-6))


@pytest.fixture(scope="module")
def c1(synth):
    L, R, gt = synth.make_pair(320, 240, cells=8, dmax=32.0, channels=1,
                               seed=1)
    return L, R, gt


@pytest.fixture(scope="module")
def rgba_small(synth):
    L, R, gt = synth.make_pair(157, 83, cells=6, dmax=24.0, channels=4,
                               seed=7)
    return L, R, gt


def test_math_bitexact(oracle):
    with cvt_amd.Context(radius=1, disparity_max=8) as ctx:
        xs = np.concatenate([np.linspace(-86.5, 0.0, 3001),
                             np.linspace(0.0, 6.2832, 3001),
                             np.linspace(-1.0, 1.0, 3001),
                             np.random.default_rng(0).uniform(0, 50, 3001)])
        xs = xs.astype(np.float32)
        out = ctx.math_selftest(xs)
    for i, x in enumerate(xs):
        s, c = oracle.sincosf(x)
        want = [oracle.expf(x), s, c]
        got = out[i, :3]
        assert same(got, np.array(want, np.float32)), (x, got, want)
        if -1.0 <= x <= 1.0:
            assert same(out[i, 3:4], np.float32([oracle.acospif(x)])), x
        if x >= 0:
            assert out[i, 4] == np.sqrt(np.float32(x)), x


@pytest.mark.parametrize("k", [0, 14 * 921607 - 100, 2 ** 40 + 123])
def test_draws_bitexact(oracle, k):
    with cvt_amd.Context(radius=1, disparity_max=8) as ctx:
        got = ctx.draws(k, 5003)
    assert same(got, oracle.draws(k, 5003))


@pytest.mark.parametrize("which,radius", [("c1", 5), ("rgba", 4)])
def test_texels_bitexact(oracle, c1, rgba_small, which, radius):
    L, R, _ = c1 if which == "c1" else rgba_small
    with cvt_amd.Context(radius=radius, disparity_max=32) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, img in ((0, L), (1, R)):
            got = ctx.texels(view)
            want = oracle.prepare(img, radius)
            assert same(got, want), first_diff(got, want)


@pytest.mark.parametrize("which,radius,dmax", [("c1", 5, 32.0),
                                               ("rgba", 4, 24.0)])
def test_init_bitexact(oracle, c1, rgba_small, which, radius, dmax):
    L, R, _ = c1 if which == "c1" else rgba_small
    tl, tr = oracle.prepare(L, radius), oracle.prepare(R, radius)
    with cvt_amd.Context(radius=radius, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, (a, b, lr) in enumerate(((tl, tr, 1), (tr, tl, 0))):
            want = oracle.init(a, b, radius, dmax, lr)
            got = ctx.planes(view)
            assert cost_close(got[..., 3], want[..., 3])
            assert same(got, want), first_diff(got, want)


def test_cost_field_bitexact(oracle, rgba_small):
    L, R, _ = rgba_small
    H, W = L.shape[:2]
    r = 4
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    rng = np.random.default_rng(4)
    planes = np.zeros((H, W, 4), np.float32)
    planes[..., 0] = 1 + rng.normal(0, 0.05, (H, W))
    planes[..., 1] = rng.normal(0, 0.05, (H, W))
    planes[..., 2] = -rng.uniform(0, 24, (H, W))
    planes[::17, ::13, 0] = np.nan         # invalid planes
    planes[5::19, 3::11, 2] = 1e4          # outside the image
    with cvt_amd.Context(radius=r, disparity_max=24) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, (a, b) in enumerate(((tl, tr), (tr, tl))):
            got = ctx.cost(view, planes)
            want = oracle.cost_field(a, b, r, planes)
            assert cost_close(got, want)
            assert same(got, want), first_diff(got, want)


def test_klt_bitexact(oracle, synth):
    L, R, _ = synth.make_pair(48, 32, cells=3, dmax=10.0, channels=4, seed=3)
    r = 3
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    rng = np.random.default_rng(5)
    H, W = 32, 48
    planes = np.zeros((H, W, 4), np.float32)
    planes[..., 0] = 1 + rng.normal(0, 0.02, (H, W))
    planes[..., 1] = rng.normal(0, 0.02, (H, W))
    planes[..., 2] = -rng.uniform(0, 10, (H, W))
    with cvt_amd.Context(radius=r, disparity_max=10) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for passes in (1, 6):
            got = ctx.klt(0, planes, passes)
            want = np.zeros_like(planes)
            for y in range(H):
                for x in range(W):
                    s = planes[y, x]
                    for _ in range(passes):
                        s = oracle.klt_at(tl, tr, r, x, y, s)
                    want[y, x] = s
            assert same(got, want), first_diff(got, want)


def run_oracle_iterations(oracle, L, R, radius, dmax, iters):
    tl, tr = oracle.prepare(L, radius), oracle.prepare(R, radius)
    sl = oracle.init(tl, tr, radius, dmax, 1)
    sr = oracle.init(tr, tl, radius, dmax, 0)
    vb2 = oracle.viewbuf_empty(L.shape[1], L.shape[0])
    fields = []
    for it in range(iters):
        sl, vb1 = oracle.propagate(tl, tr, radius, dmax, 1, it, sl, vb2, sr)
        sr, vb2 = oracle.propagate(tr, tl, radius, dmax, 0, it, sr, vb1, sl)
        fields.append((sl, sr))
    return fields


@pytest.mark.parametrize("which,radius,dmax,iters",
                         [("c1", 5, 32.0, 3), ("rgba", 4, 24.0, 4)])
def test_iterations_bitexact(oracle, c1, rgba_small, which, radius, dmax,
                             iters):
    L, R, _ = c1 if which == "c1" else rgba_small
    want = run_oracle_iterations(oracle, L, R, radius, dmax, iters)
    with cvt_amd.Context(radius=radius, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for it in range(iters):
            ctx.iterate(1)
            for view in (0, 1):
                got = ctx.planes(view)
                w = want[it][view]
                # the tolerance bar first, then bit equality
                assert cost_close(got[..., 3], w[..., 3]), (it, view)
                assert same(got, w), (it, view, first_diff(got, w))


@pytest.mark.parametrize("cw", [0.0, 5.0, 30.0])
def test_color_weight_bitexact(oracle, rgba_small, cw):
    """The support weights' colour factor (CENSUS_COLOR_WEIGHT, 12 in the
    reference) as a parameter: 0 makes every weight exactly 1, 30 drives
    most towards underflow.  Texels (their weight sums), the initial fields
    and two iterations, bit for bit."""
    L, R, _ = rgba_small
    r, dmax = 4, 24.0
    tl = oracle.prepare(L, r, color_weight=cw)
    tr = oracle.prepare(R, r, color_weight=cw)
    sl = oracle.init(tl, tr, r, dmax, 1, color_weight=cw)
    sr = oracle.init(tr, tl, r, dmax, 0, color_weight=cw)
    with cvt_amd.Context(radius=r, disparity_max=dmax,
                         color_weight=cw) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for view, want in ((0, tl), (1, tr)):
            got = ctx.texels(view)
            assert same(got, want), (view, first_diff(got, want))
        for view, want in ((0, sl), (1, sr)):
            got = ctx.planes(view)
            assert same(got, want), (view, first_diff(got, want))
        vb2 = oracle.viewbuf_empty(L.shape[1], L.shape[0])
        for it in range(2):
            sl, vb1 = oracle.propagate(tl, tr, r, dmax, 1, it, sl, vb2, sr,
                                       color_weight=cw)
            sr, vb2 = oracle.propagate(tr, tl, r, dmax, 0, it, sr, vb1, sl,
                                       color_weight=cw)
            ctx.iterate(1)
            for view, want in ((0, sl), (1, sr)):
                got = ctx.planes(view)
                assert same(got, want), (it, view, first_diff(got, want))


def test_pipeline_matches_oracle_run(oracle, c1):
    """C1 configuration end to end: 320x240 gray, 11x11 patch, 3
    iterations, LR check at 1 px (angle test off)."""
    L, R, gt = c1
    ref = oracle.run(L, R, 5, 32.0, 3, maxdisp=1.0, maxdeg=180.0)
    with cvt_amd.Context(radius=5, disparity_max=32.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(3)
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
        assert same(ctx.planes(0), ref["stateL"])
        assert same(ctx.planes(1), ref["stateR"])
    close = np.abs(d - ref["disp"]) <= 0.5
    assert close.mean() >= 0.995
    assert same(d, ref["disp"]), first_diff(d, ref["disp"])


@pytest.mark.parametrize("pp", [0, 1, 2])
def test_postprocessing_bitexact(oracle, rgba_small, pp):
    L, R, _ = rgba_small
    r, dmax = 4, 24.0
    with cvt_amd.Context(radius=r, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(3)
        sl, sr = ctx.planes(0), ctx.planes(1)
        d, chk = ctx.disparity(pp, 1.0, 5.0, checked=True)
        nrm = ctx.normals(pp)
    if pp == 0:
        want_chk = sl
        want_d = oracle.to_disparity(sl)
    else:
        want_chk = oracle.consistency(sl, sr, 1.0, 5.0)
        want_d = (oracle.to_disparity(want_chk) if pp == 1
                  else oracle.fill_depthmap(want_chk))
    assert same(chk, want_chk), first_diff(chk, want_chk)
    assert same(d, want_d), first_diff(d, want_d)
    want_n = oracle.fill_normalmap(want_chk)
    assert same(nrm, want_n), first_diff(nrm, want_n)


def test_prior_init_bitexact(oracle, rgba_small):
    L, R, gt = rgba_small
    r, dmax = 4, 24.0
    prior = gt.copy()
    prior[::7, :] = 0.0                      # unknown rows -> random init
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    normals = oracle.disparity_to_normal(prior)
    with cvt_amd.Context(radius=r, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup_prior(prior)                 # normals from the disparity
        want = oracle.init_prior(tl, tr, r, dmax, 1, prior, normals)
        assert same(ctx.planes(0), want), first_diff(ctx.planes(0), want)
        assert same(ctx.planes(1), oracle.init(tr, tl, r, dmax, 0))
        ctx.setup_prior(prior, normals)        # explicit normals
        assert same(ctx.planes(0), want)


def test_resume_from_fields(rgba_small):
    L, R, _ = rgba_small
    with cvt_amd.Context(radius=4, disparity_max=24.0) as a:
        a.set_images(L, R)
        a.setup()
        a.iterate(2)
        sl, sr = a.planes(0), a.planes(1)
        a.iterate(2)
        end = (a.planes(0), a.planes(1))
    with cvt_amd.Context(radius=4, disparity_max=24.0) as b:
        b.set_images(L, R)
        b.setup()
        b.set_planes(0, sl, 2)
        b.set_planes(1, sr, 2)
        b.iterate(2)
        assert same(b.planes(0), end[0]) and same(b.planes(1), end[1])


def test_counted_iteration_is_identical(rgba_small):
    L, R, _ = rgba_small
    out = []
    for counted in (False, True):
        with cvt_amd.Context(radius=4, disparity_max=24.0) as ctx:
            ctx.set_images(L, R)
            ctx.setup()
            if counted:
                taps, klt = ctx.iterate_counted(2)
                H, W = L.shape[:2]
                assert 0 < taps < 2 * 2 * H * W * 13 * 81
                assert klt <= 2 * 2 * H * W * 6 * 81
            else:
                ctx.iterate(2)
            out.append((ctx.planes(0), ctx.planes(1)))
    assert same(out[0][0], out[1][0]) and same(out[0][1], out[1][1])


def test_error_paths():
    with pytest.raises(cvt_amd.PMError) as e:
        cvt_amd.Context(radius=0, disparity_max=8)
    assert e.value.code == 1
    with cvt_amd.Context(radius=2, disparity_max=8) as ctx:
        with pytest.raises(cvt_amd.PMError) as e:
            ctx.setup()
        assert e.value.code == 2
        with pytest.raises(ValueError):
            ctx.set_images(np.zeros((4, 5), np.float32),
                           np.zeros((4, 6), np.float32))
        ctx.set_images(np.zeros((4, 5), np.float32),
                       np.zeros((4, 5), np.float32))
        with pytest.raises(cvt_amd.PMError) as e:
            ctx.iterate(1)
        assert e.value.code == 2
        ctx.setup()
        with pytest.raises(cvt_amd.PMError) as e:
            ctx.planes(2)
        assert e.value.code == 1
        with pytest.raises(cvt_amd.PMError):
            ctx.disparity(pp=7)


def _random_pair(W, H, C, seed):
    """A textured pair whose right view is the left shifted by one column
    (edge column repeated); random enough that every candidate kind wins
    somewhere."""
    rng = np.random.default_rng(seed)
    shape = (H, W) if C == 1 else (H, W, 4)
    L = rng.random(shape).astype(np.float32)
    R = np.concatenate([L[:, 1:], L[:, -1:]], axis=1) if W > 1 else L.copy()
    R = np.ascontiguousarray(R + np.float32(0.01) *
                             rng.standard_normal(shape).astype(np.float32))
    return L, R


@pytest.mark.parametrize("W,H,C,radius", [
    (1, 1, 4, 2),     # one pixel: no KLT taps, singular KLT system
    (1, 9, 1, 3),     # one column
    (11, 1, 4, 2),    # one row, narrower than a wavefront
    (5, 3, 1, 4),     # image smaller than the patch
    (70, 6, 4, 1),    # smallest patch, a partial second block column
    (40, 21, 1, 16),  # the largest patch the engine accepts
])
def test_small_and_odd_shapes_bitexact(oracle, W, H, C, radius):
    """Edges everywhere: patches, KLT windows, texel pads and tiles reach
    past the image on every side; planes of every iteration, then the
    LR-checked disparity, bit for bit."""
    L, R = _random_pair(W, H, C, seed=W * 100 + H)
    dmax, iters = float(max(2, W // 2)), 2
    print("shape", W, H, C, radius, flush=True)
    want = run_oracle_iterations(oracle, L, R, radius, dmax, iters)
    ref = oracle.run(L, R, radius, dmax, iters, maxdisp=1.0, maxdeg=180.0)
    with cvt_amd.Context(radius=radius, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for it in range(iters):
            ctx.iterate(1)
            for view in (0, 1):
                got = ctx.planes(view)
                assert same(got, want[it][view]), (
                    it, view, first_diff(got, want[it][view]))
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
    assert same(d, ref["disp"]), first_diff(d, ref["disp"])


@pytest.mark.parametrize("kind", ["flat", "saturated"])
def test_flat_and_saturated_images_bitexact(oracle, kind):
    """A flat image (every support weight 1, every plane costs the same)
    and one whose colour steps make nearly every off-centre support weight
    underflow to 0 (most evaluations end at the wsum <= 1 rule)."""
    W, H = 90, 13
    if kind == "flat":
        L = np.full((H, W, 4), 0.5, np.float32)
        R = L.copy()
    else:
        rng = np.random.default_rng(3)
        L = (rng.random((H, W, 4)) * 400.0).astype(np.float32)
        R = np.ascontiguousarray(np.roll(L, 2, axis=1))
    print("image", kind, flush=True)
    want = run_oracle_iterations(oracle, L, R, 3, 12.0, 2)
    with cvt_amd.Context(radius=3, disparity_max=12.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(2)
        for view in (0, 1):
            got = ctx.planes(view)
            assert same(got, want[1][view]), (view,
                                               first_diff(got, want[1][view]))


@pytest.mark.parametrize("W,H,dmax,C,r", [(1280, 720, 128.0, 4, 4),
                                          (1920, 1080, 192.0, 4, 4),
                                          (1242, 375, 128.0, 4, 4),
                                          (1280, 720, 128.0, 1, 5)])
def test_full_size_bitexact(oracle, synth, W, H, dmax, C, r):
    """The 720p bench configuration, a 1080p pair, a KITTI-sized pair
    (1242 x 375: partial work-groups at the right edge and a partial last
    block row) and a gray 720p pair with the 11 x 11 patch at full size:
    the initial fields, one iteration and the LR-checked disparity, bit
    for bit (the oracle takes seconds per iteration on the host cores)."""
    L, R, _ = synth.make_pair(W, H, cells=24, dmax=dmax, channels=C, seed=0)
    print("full size", W, H, C, r, flush=True)
    ref = oracle.run(L, R, r, dmax, 1, maxdisp=1.0, maxdeg=180.0)
    with cvt_amd.Context(radius=r, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(1)
        for view, key in ((0, "stateL"), (1, "stateR")):
            got = ctx.planes(view)
            assert same(got, ref[key]), (view, first_diff(got, ref[key]))
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 180.0)
    assert same(d, ref["disp"]), first_diff(d, ref["disp"])


def test_two_contexts_in_two_threads(synth):
    """Contexts are independent: two threads running their own contexts at
    once (each with its own streams and events) get what each gets alone."""
    import threading
    pairs = [synth.make_pair(211, 97, cells=6, dmax=20.0, channels=4,
                             seed=s)[:2] for s in (21, 22)]

    def run(L, R, out, k):
        with cvt_amd.Context(radius=3, disparity_max=20.0) as ctx:
            ctx.set_images(L, R)
            ctx.setup()
            ctx.iterate(3)
            out[k] = (ctx.planes(0), ctx.planes(1),
                      ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0))

    alone = {}
    for k, (L, R) in enumerate(pairs):
        run(L, R, alone, k)
    together = {}
    threads = [threading.Thread(target=run, args=(L, R, together, k))
               for k, (L, R) in enumerate(pairs)]
    print("two threads", flush=True)
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for k in range(2):
        for a, b in zip(alone[k], together[k]):
            assert same(a, b), (k, first_diff(a, b))


def test_strided_inputs_match_packed(synth):
    """Rows with padding (stride > width x channels), from host and from
    device memory, give the fields packed rows give."""
    import torch
    from cvt_amd.stereo import _ptr
    L, R, _ = synth.make_pair(150, 40, cells=4, dmax=16.0, channels=4,
                              seed=31)
    H, W, C, pad = L.shape[0], L.shape[1], 4, 7

    def padded(img):
        buf = np.full((H, W * C + pad), np.nan, np.float32)
        buf[:, :W * C] = img.reshape(H, W * C)
        return buf

    def fields(ctx):
        ctx.setup()
        ctx.iterate(2)
        return ctx.planes(0), ctx.planes(1)

    pl, pr = padded(L), padded(R)
    print("strided inputs", flush=True)
    with cvt_amd.Context(radius=2, disparity_max=16.0) as ctx:
        ctx.set_images(L, R)
        want = fields(ctx)
        ctx._check(ctx._lib.pm_set_images(ctx._h, _ptr(pl), _ptr(pr), W, H,
                                          C, W * C + pad))
        from_host = fields(ctx)
        dl = torch.from_numpy(pl).to("cuda:0")
        dr = torch.from_numpy(pr).to("cuda:0")
        torch.cuda.synchronize()
        ctx.set_images_device(dl.data_ptr(), dr.data_ptr(), W, H, C,
                              stride=W * C + pad)
        from_device = fields(ctx)
    for got in (from_host, from_device):
        for a, b in zip(got, want):
            assert same(a, b), first_diff(a, b)


def test_full_size_eight_iteration_invariants(synth):
    """The 720p bench configuration over all 8 iterations, by properties
    that hold at any size (plain PatchMatch, theta = 0): a pixel's stored
    cost never rises from one iteration to the next, and it equals a fresh
    evaluation of its stored plane (pm_cost), bit for bit."""
    L, R, _ = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                              seed=0)
    print("full size invariants", flush=True)
    with cvt_amd.Context(radius=4, disparity_max=128.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        prev = [ctx.planes(v)[..., 3] for v in (0, 1)]
        for it in range(8):
            ctx.iterate(1)
            for v in (0, 1):
                planes = ctx.planes(v)
                w = planes[..., 3]
                rose = np.argwhere(w > prev[v])
                assert rose.size == 0, (it, v, len(rose), tuple(rose[0]))
                fresh = ctx.cost(v, planes)
                assert same(fresh, w), (it, v, first_diff(fresh, w))
                prev[v] = w


@pytest.mark.parametrize("pp", [cvt_amd.DISPARITY_PP_LRCHECK,
                                cvt_amd.DISPARITY_PP_LRCHECK_FILL])
def test_full_size_postprocessing_bitexact(oracle, synth, pp):
    """LR check, hole filling and normals on the 720p bench pair's fields
    after all 8 iterations, the fields the bench turns into its output:
    each is a pure function of the two plane fields, so the oracle applied
    to the GPU's own fields must give the GPU's outputs bit for bit."""
    L, R, _ = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                              seed=0)
    print("full size post-processing", pp, flush=True)
    with cvt_amd.Context(radius=4, disparity_max=128.0) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        ctx.iterate(8)
        sl, sr = ctx.planes(0), ctx.planes(1)
        d, chk = ctx.disparity(pp, 1.0, 5.0, checked=True)
        nrm = ctx.normals(pp)
    want_chk = oracle.consistency(sl, sr, 1.0, 5.0)
    rejected = (want_chk[..., 3] == -1.0).mean()     # the check's holes
    assert 0.001 < rejected < 0.95, rejected
    assert same(chk, want_chk), first_diff(chk, want_chk)
    want_d = (oracle.to_disparity(want_chk)
              if pp == cvt_amd.DISPARITY_PP_LRCHECK
              else oracle.fill_depthmap(want_chk))
    assert same(d, want_d), first_diff(d, want_d)
    want_n = oracle.fill_normalmap(want_chk)
    assert same(nrm, want_n), first_diff(nrm, want_n)


def test_prior_init_full_size_bitexact(oracle, synth):
    """The prior-seeded start (the hook a plane proposal feeds) on the 720p
    bench pair: ground-truth disparity with holes seeds the left view, then
    two iterations refine it; planes after each step bit for bit."""
    L, R, gt = synth.make_pair(1280, 720, cells=24, dmax=128.0, channels=4,
                               seed=0)
    r, dmax = 4, 128.0
    prior = gt.copy()
    prior[::5, ::3] = 0.0                    # pixels without a prior
    print("prior full size", flush=True)
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    sl = oracle.init_prior(tl, tr, r, dmax, 1, prior,
                           oracle.disparity_to_normal(prior))
    sr = oracle.init(tr, tl, r, dmax, 0)
    vb2 = oracle.viewbuf_empty(1280, 720)
    want = []
    for it in range(2):
        sl, vb1 = oracle.propagate(tl, tr, r, dmax, 1, it, sl, vb2, sr)
        sr, vb2 = oracle.propagate(tr, tl, r, dmax, 0, it, sr, vb1, sl)
        want.append((sl, sr))
    with cvt_amd.Context(radius=r, disparity_max=dmax) as ctx:
        ctx.set_images(L, R)
        ctx.setup_prior(prior)
        for it in range(2):
            ctx.iterate(1)
            for v in (0, 1):
                got = ctx.planes(v)
                assert same(got, want[it][v]), (it, v,
                                                 first_diff(got, want[it][v]))


def _random_configs(n, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        W = int(rng.integers(8, 200))
        H = int(rng.integers(4, 120))
        C = int(rng.choice([1, 4]))
        r = int(rng.integers(1, 9))
        dmax = float(rng.choice([4.0, 12.5, 31.0, 64.0]))
        cw = float(rng.choice([0.0, 6.0, 12.0, 20.0]))
        iters = int(rng.integers(1, 4))
        out.append((W, H, C, r, dmax, cw, iters, int(rng.integers(0, 1000))))
    return out


@pytest.mark.parametrize("W,H,C,r,dmax,cw,iters,seed", _random_configs(12))
def test_random_configurations_bitexact(oracle, synth, W, H, C, r, dmax, cw,
                                        iters, seed):
    """A seeded sweep over sizes (one or many block rows and columns,
    partial work-groups), formats, radii, disparity ranges, colour weights
    and iteration counts: the planes after every iteration, the LR-checked
    disparity, its hole-filled variant and the normals, bit for bit."""
    L, R, _ = synth.make_pair(W, H, cells=3, dmax=min(dmax, W / 3.0),
                              channels=C, seed=seed)
    print("random config", W, H, C, r, dmax, cw, iters, seed, flush=True)
    tl = oracle.prepare(L, r, color_weight=cw)
    tr = oracle.prepare(R, r, color_weight=cw)
    sl = oracle.init(tl, tr, r, dmax, 1, color_weight=cw)
    sr = oracle.init(tr, tl, r, dmax, 0, color_weight=cw)
    vb2 = oracle.viewbuf_empty(W, H)
    with cvt_amd.Context(radius=r, disparity_max=dmax,
                         color_weight=cw) as ctx:
        ctx.set_images(L, R)
        ctx.setup()
        for it in range(iters):
            sl, vb1 = oracle.propagate(tl, tr, r, dmax, 1, it, sl, vb2, sr,
                                       color_weight=cw)
            sr, vb2 = oracle.propagate(tr, tl, r, dmax, 0, it, sr, vb1, sl,
                                       color_weight=cw)
            ctx.iterate(1)
            for view, want in ((0, sl), (1, sr)):
                got = ctx.planes(view)
                assert same(got, want), (it, view, first_diff(got, want))
        d = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK, 1.0, 5.0)
        df = ctx.disparity(cvt_amd.DISPARITY_PP_LRCHECK_FILL, 1.0, 5.0)
        nrm = ctx.normals(cvt_amd.DISPARITY_PP_LRCHECK_FILL)
    chk = oracle.consistency(sl, sr, 1.0, 5.0)
    want_d = oracle.to_disparity(chk)
    assert same(d, want_d), first_diff(d, want_d)
    want_df = oracle.fill_depthmap(chk)
    assert same(df, want_df), first_diff(df, want_df)
    want_n = oracle.fill_normalmap(chk)
    assert same(nrm, want_n), first_diff(nrm, want_n)


def test_eval_error_device_matches_host(synth):
    """pm_eval_error_device against the host pm_eval_error
    (Disparity::evaluateError): the same integer counts, with NaN
    estimates, rejected pixels (0), pixels without ground truth (gt <= 0)
    and a relative threshold, on a strided estimate."""
    import torch
    rng = np.random.default_rng(41)
    H, W = 203, 317
    gt = rng.uniform(-2, 60, (H, W)).astype(np.float32)
    est = gt + rng.normal(0, 1.5, (H, W)).astype(np.float32)
    est[::7, ::5] = 0.0
    est[3::11, 2::13] = np.nan
    pad = np.zeros((H, W + 9), np.float32)
    pad[:, :W] = est
    d_est = torch.from_numpy(pad).cuda()
    d_gt = torch.from_numpy(gt).cuda()
    counts = torch.zeros(3, dtype=torch.int64, device="cuda")
    for abs_t, rel_t in ((1.0, 0.0), (2.0, 0.05), (0.5, 0.0)):
        counts.fill_(123)           # the call zeroes before counting
        torch.cuda.synchronize()
        cvt_amd.disparity.evaluate_error_device(
            d_est.data_ptr(), d_gt.data_ptr(), W, H, counts.data_ptr(),
            abs_t, rel_t, stride_d=W + 9)
        torch.cuda.synchronize()
        got = tuple(int(v) for v in counts.cpu())
        want = cvt_amd.disparity.evaluate_error(est, gt, abs_t, rel_t)
        assert got == want, (abs_t, rel_t, got, want)
