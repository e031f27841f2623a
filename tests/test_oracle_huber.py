"""Pins of the oracle's PM-Huber smoothing coupling (CPU only): the
diffusion tensor (PDDiffTensor.cl:25-116) and normal-depth conversion
(pmhstereo.cl:1384-1402) against closed forms, the diffusion-tensor ROF
step (PDDROFDiffTensor.cl:43-125) against an independent numpy statement
of the kernel text, and the PM-Huber pipeline (PMHuberStereo.cpp:406-581)
against its documented sequence."""
import numpy as np
import pytest

F = np.float32


def bits(a):
    a = np.ascontiguousarray(a, np.float32).copy()
    a[np.isnan(a)] = np.nan
    return a.view(np.uint32)


def test_tensor_of_flat_image_is_identity(oracle):
    img = np.full((6, 7, 4), 0.3, np.float32)
    dt = oracle.diffusion_tensor(img)
    # interior: no gradient -> (1, 0, 0, 1); the last row and column see
    # the zero outside the image
    np.testing.assert_array_equal(dt[:-1, :-1],
                                  np.broadcast_to(F([1, 0, 0, 1]), (5, 6, 4)))
    assert (dt[-1, :, :] != F([1, 0, 0, 1])).any(axis=-1).all()


@pytest.mark.parametrize("channels", [1, 4])
def test_tensor_of_ramp(oracle, channels):
    H, W, s = 5, 9, 0.01
    x = np.arange(W, dtype=np.float32) * F(s)
    if channels == 1:
        img = np.broadcast_to(x, (H, W)).copy()
        g = img
    else:
        img = np.zeros((H, W, 4), np.float32)
        img[..., 0] = img[..., 1] = img[..., 2] = x
        img[..., 3] = 1.0
        g = ((img[..., 0] * F(0.2126) + img[..., 1] * F(0.7152)) +
             img[..., 2] * F(0.0722)) + img[..., 3] * F(0.0)
    dt = oracle.diffusion_tensor(img)
    # interior, not the last row/column: the gradient is (g(x+1) - g(x), 0)
    for yy in range(H - 1):
        for xx in range(W - 1):
            dx = F(g[yy, xx + 1] - g[yy, xx])
            w = F(np.exp(-24.0 * float(abs(dx))))
            # unit gradient direction (1, 0): tensor (w, 0, 0, 1)
            assert abs(dt[yy, xx, 0] - w) < 3e-7 * max(1, w)
            assert dt[yy, xx, 1] == 0 and dt[yy, xx, 2] == 0
            assert dt[yy, xx, 3] == 1.0


def test_tensor_is_symmetric_with_eigenvalues_w_and_1(oracle):
    rng = np.random.default_rng(1)
    img = rng.random((12, 14, 4)).astype(np.float32)
    dt = oracle.diffusion_tensor(img).astype(np.float64)
    m = np.stack([np.stack([dt[..., 0], dt[..., 1]], -1),
                  np.stack([dt[..., 2], dt[..., 3]], -1)], -2)
    assert np.allclose(m[..., 0, 1], m[..., 1, 0], atol=1e-7)
    ev = np.linalg.eigvalsh(m)
    assert np.allclose(ev[..., 1], 1.0, atol=1e-6)
    # the other eigenvalue is w = exp(-24 |grad gray|) (zero outside)
    g = np.pad(((img[..., 0] * F(0.2126) + img[..., 1] * F(0.7152)) +
                img[..., 2] * F(0.0722)).astype(np.float64), ((0, 1), (0, 1)))
    gx, gy = g[:-1, 1:] - g[:-1, :-1], g[1:, :-1] - g[:-1, :-1]
    w = np.exp(-24.0 * np.hypot(gx, gy))
    assert np.allclose(ev[..., 0], w, atol=2e-6)


def test_normal_depth_of_fronto_planes(oracle):
    W, H, d = 8, 3, 5.0
    s = np.zeros((H, W, 4), np.float32)
    s[..., 0] = 1.0
    s[..., 2] = -d                       # left view: x' = x - d
    nd = oracle.normal_depth(s, 1.0 / 64.0, 1)
    np.testing.assert_array_equal(
        nd, np.broadcast_to(F([0, 0, 1, d / 64.0]), (H, W, 4)))
    s[..., 2] = d                        # right view: x' = x + d
    nd = oracle.normal_depth(s, 0.5, 0)
    np.testing.assert_array_equal(
        nd, np.broadcast_to(F([0, 0, 1, d * 0.5]), (H, W, 4)))


SIGMA, TAU, ETA, THETA = F(4.0), F(0.05), F(8.0), F(0.5)


def fma(a, b, c):
    """elementwise float32 a * b + c with one rounding (the product of two
    float32 is exact in float64)"""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) +
            np.asarray(c, np.float64)).astype(np.float32)


def rof_step_numpy(u, p, q, f, dt, lam):
    """PDDROFDiffTensor.cl:43-125 restated with numpy float32 array
    operations: the dual of every position (-1..W-1, -1..H-1) first, then
    the divergence, primal and q updates.  The stored dual is the
    tensor-applied one, as the kernel writes it.  A product that is an
    operand of a sum in the kernel's source expression is fused into it
    (the left one first), as an OpenCL compiler contracts."""
    H, W = u.shape[:2]
    jj, ii = np.mgrid[-1:H, -1:W]
    inside = ((ii >= 0) & (jj >= 0))[..., None]
    cj, ci = np.clip(jj, 0, H - 1), np.clip(ii, 0, W - 1)
    u0 = u[cj, ci]
    ux = u[cj, np.clip(ii + 1, 0, W - 1)]
    uy = u[np.clip(jj + 1, 0, H - 1), ci]
    D = np.where(inside, dt[cj, ci], F(0))
    plo = np.where(inside, p[cj, ci, :4], F(0))
    phi = np.where(inside, p[cj, ci, 4:], F(0))
    dx, dy = u0 - ux, u0 - uy
    d0, d1, d2, d3 = (D[..., k:k + 1] for k in range(4))
    DTdx = fma(d0, dx, d2 * dy)
    DTdy = fma(d1, dx, d3 * dy)
    plo = fma(SIGMA, DTdx, plo)
    phi = fma(SIGMA, DTdy, phi)
    n = fma(plo, plo, phi * phi)
    nxyz = np.sqrt((n[..., 0] + n[..., 1]) + n[..., 2])
    proj = np.stack([np.fmax(F(1), nxyz)] * 3 +
                    [np.fmax(F(1), np.sqrt(n[..., 3]))], -1)
    plo, phi = plo / proj, phi / proj
    lo = fma(d0, plo, d2 * phi)
    hi = fma(d1, plo, d3 * phi)
    Plo, Phi = lo[1:, 1:], hi[1:, 1:]
    Llo, Uhi = lo[1:, :-1], hi[:-1, 1:]
    div = ((Plo - Llo) + Phi) - Uhi
    lam = np.asarray(lam, np.float32)
    un = fma(TAU, fma(lam, f, -div) - q, u) / fma(TAU, lam, F(1))
    un = fma(THETA, un - u, un)
    qn = fma(ETA, un - f, q) / (F(1) + F(4) * ETA / lam)
    qn = fma(THETA, qn - q, qn)
    return un, np.concatenate([Plo, Phi], -1), qn


@pytest.mark.parametrize("seed", [0, 1])
def test_rof_step_matches_numpy_statement(oracle, seed):
    rng = np.random.default_rng(seed)
    H, W = 7, 11
    f = rng.normal(0, 1, (H, W, 4)).astype(np.float32)
    dt = oracle.diffusion_tensor(rng.random((H, W, 4)).astype(np.float32))
    lam = F([1.7, 1.7, 1.7, 1.0])
    u = f.copy()
    p = np.zeros((H, W, 8), np.float32)
    q = np.zeros((H, W, 4), np.float32)
    for _ in range(4):
        got = oracle.rof_step(u, p, q, f, dt, lam)
        want = rof_step_numpy(u, p, q, f, dt, lam)
        for g_, w_ in zip(got, want):
            np.testing.assert_array_equal(bits(g_), bits(w_))
        u, p, q = got
    # the 40-step solver is 40 such steps from u = f, zero duals
    u, p, q = f.copy(), np.zeros_like(p), np.zeros_like(q)
    for _ in range(40):
        u, p, q = rof_step_numpy(u, p, q, f, dt, lam)
    np.testing.assert_array_equal(bits(oracle.rof(f, dt, lam, 40)), bits(u))


def test_rof_flat_input_is_a_fixed_point(oracle):
    f = np.full((6, 8, 4), 0.25, np.float32)
    dt = oracle.diffusion_tensor(np.full((6, 8), 0.5, np.float32))
    u = oracle.rof(f, dt, F([2.0, 2.0, 2.0, 1.0]), 40)
    assert np.abs(u - f).max() < 1e-6


def test_rof_smooths(oracle):
    rng = np.random.default_rng(3)
    H, W = 24, 32
    clean = np.zeros((H, W, 4), np.float32)
    clean[:, W // 2:] = 1.0
    f = (clean + rng.normal(0, 0.2, clean.shape)).astype(np.float32)
    dt = oracle.diffusion_tensor(np.zeros((H, W), np.float32))

    def tv(a):
        return np.abs(np.diff(a, axis=0)).sum() + \
            np.abs(np.diff(a, axis=1)).sum()

    u = oracle.rof(f, dt, F([1.0, 1.0, 1.0, 1.0]), 40)
    assert tv(u) < 0.7 * tv(f)
    assert np.abs(u - clean).mean() < np.abs(f - clean).mean()


def test_run_huber_is_the_documented_sequence(oracle, synth):
    O = oracle
    R, dmax, iters, tmin = 2, 8.0, 3, 1e-4
    L, Rt, _ = synth.make_pair(20, 12, cells=3, dmax=dmax, channels=4,
                               seed=9)
    res = O.run_huber(L, Rt, R, dmax, iters, tmin, 1.0, 5.0)
    tl, tr = O.prepare(L, R), O.prepare(Rt, R)
    dtl, dtr = O.diffusion_tensor(L), O.diffusion_tensor(Rt)
    sl, sr = O.init(tl, tr, R, dmax, 1), O.init(tr, tl, R, dmax, 0)
    H, W = sl.shape[:2]
    sml = np.zeros((H, W, 4), np.float32)
    smr = np.zeros((H, W, 4), np.float32)
    vb2 = O.viewbuf_empty(W, H)
    theta = F(tmin)
    for it in range(iters):
        lam_v = F(theta * F(10) + F(1))
        lam = F([lam_v, lam_v, lam_v, 1.0])
        th = float(F(theta * F(50)))
        nl, vb1 = O.propagate(tl, tr, R, dmax, 1, it, sl, vb2, sr, smooth=sml,
                              theta=th)
        nr, vb2 = O.propagate(tr, tl, R, dmax, 0, it, sr, vb1, nl,
                              smooth=smr, theta=th)
        sml = O.rof(O.normal_depth(nl, F(1) / F(dmax), 1), dtl, lam, 40)
        smr = O.rof(O.normal_depth(nr, F(1) / F(dmax), 0), dtr, lam, 40)
        theta = F(np.float32(it) / np.float32(iters)) ** F(4) + F(tmin)
        sl, sr = nl, nr
    np.testing.assert_array_equal(bits(res["stateL"]), bits(sl))
    np.testing.assert_array_equal(bits(res["stateR"]), bits(sr))
    np.testing.assert_array_equal(bits(res["smoothL"]), bits(sml))
    np.testing.assert_array_equal(bits(res["smoothR"]), bits(smr))
    d = O.to_disparity(O.consistency(sl, sr, 1.0, 5.0))
    np.testing.assert_array_equal(bits(res["disp"]), bits(d))


def test_huber_coupling_changes_the_result(oracle, synth):
    """theta > 0 and the smooth field are live: the PM-Huber fields differ
    from plain PatchMatch's, and the smooth field tracks the planes."""
    O = oracle
    L, Rt, gt = synth.make_pair(48, 32, cells=3, dmax=12.0, channels=4,
                                seed=4)
    hub = O.run_huber(L, Rt, 3, 12.0, 5)
    pla = O.run(L, Rt, 3, 12.0, 5, 1.0, 5.0)
    assert (bits(hub["stateL"]) != bits(pla["stateL"])).any()
    nd = O.normal_depth(hub["stateL"], 1.0 / 12.0, 1)
    ok = np.isfinite(nd).all(-1)
    assert np.median(np.abs(hub["smoothL"][..., 3] - nd[..., 3])[ok]) < 0.05



# ---- the depthMap coupling: Huber dual, fill_state, pipeline wiring ------

def _bits(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).copy()
    u[np.isnan(a)] = 0x7FC00000
    return u


def test_huber_step_with_eps_0_is_the_rof_step_when_unprojected(oracle):
    # tiny variations: every dual stays inside the unit ball, so neither
    # projection acts and the divisions by 1 + 4 * 0 are exact
    rng = np.random.default_rng(5)
    H, W = 9, 13
    u = (0.5 + 1e-3 * rng.standard_normal((H, W, 4))).astype(np.float32)
    f = (0.5 + 1e-3 * rng.standard_normal((H, W, 4))).astype(np.float32)
    p = (1e-3 * rng.standard_normal((H, W, 8))).astype(np.float32)
    q = (1e-3 * rng.standard_normal((H, W, 4))).astype(np.float32)
    dt = np.zeros((H, W, 4), np.float32)
    dt[..., 0] = 1.0
    dt[..., 3] = 1.0
    lam = [2.5, 2.5, 2.5, 2.5]
    a = oracle.rof_step(u, p, q, f, dt, lam)
    b = oracle.huber_step(u, p, q, f, dt, lam, 0.0)
    for x, y in zip(a, b):
        assert np.array_equal(_bits(x), _bits(y))


def test_huber_dual_projects_all_four_channels_jointly(oracle):
    # identity tensor: the stored dual is the projected one.  A step edge
    # in every channel drives the ascent far outside the unit ball: the
    # Huber dual ends on the 4-channel unit sphere, the ROF dual only on
    # the xyz and w spheres separately (joint norm up to sqrt 2)
    H, W = 6, 8
    u = np.zeros((H, W, 4), np.float32)
    u[:, W // 2:, :] = 1.0
    dt = np.zeros((H, W, 4), np.float32)
    dt[..., 0] = 1.0
    dt[..., 3] = 1.0
    z4, z8 = np.zeros((H, W, 4), np.float32), np.zeros((H, W, 8), np.float32)
    lam = [2.0] * 4
    _, ph, _ = oracle.huber_step(u, z8, z4, u, dt, lam, 1e-4)
    _, pr, _ = oracle.rof_step(u, z8, z4, u, dt, lam)
    joint = lambda p: np.sqrt((p.astype(np.float64) ** 2).sum(axis=-1))
    assert joint(ph).max() <= 1.0 + 1e-6
    assert abs(joint(ph).max() - 1.0) < 1e-6
    assert joint(pr).max() > 1.3


def test_fill_state_keeps_valid_planes(oracle):
    # disparities inside [0, W] and costs >= 0: fill_state is
    # normal_depth
    rng = np.random.default_rng(2)
    H, W = 5, 17
    s = np.zeros((H, W, 4), np.float32)
    s[..., 0] = 1.0 + 0.01 * rng.standard_normal((H, W))
    s[..., 1] = 0.01 * rng.standard_normal((H, W))
    s[..., 2] = -rng.uniform(1, 6, (H, W))
    s[..., 3] = rng.uniform(0, 1, (H, W))
    for lr, st in ((1, s), (0, s * np.float32([1, 1, -1, 1]))):
        got = oracle.fill_state(st, 0.25, lr)
        want = oracle.normal_depth(st, 0.25, lr)
        assert np.array_equal(_bits(got), _bits(want))


def _fronto(d):
    # left-view plane of constant disparity d: x' = x - d
    return np.float32([1.0, 0.0, -d, 0.5])


def test_fill_state_takes_the_neighbour_with_the_smaller_disparity(oracle):
    W = 6
    s = np.stack([_fronto(3.0)] * W)[None].copy()
    s[0, 2] = _fronto(-5.0)          # negative disparity: filled
    s[0, 1] = _fronto(4.0)           # left neighbour, disparity 4
    s[0, 3] = _fronto(2.0)           # right neighbour, disparity 2
    out = oracle.fill_state(s, 1.0, 1)
    want = oracle.normal_depth(s[:, 3:4], 1.0, 1)[0, 0]
    # the right neighbour's plane evaluated at x = 2
    nd = oracle.normal_depth(np.stack([s[0, 3]] * W)[None], 1.0, 1)
    assert np.array_equal(_bits(out[0, 2]), _bits(nd[0, 2]))
    s[0, 1] = _fronto(1.0)           # now the left one is nearer
    out = oracle.fill_state(s, 1.0, 1)
    nd = oracle.normal_depth(np.stack([s[0, 1]] * W)[None], 1.0, 1)
    assert np.array_equal(_bits(out[0, 2]), _bits(nd[0, 2]))
    # the left edge has no left neighbour: the right one is taken
    s[0, 0] = _fronto(-2.0)
    out = oracle.fill_state(s, 1.0, 1)
    nd = oracle.normal_depth(np.stack([s[0, 1]] * W)[None], 1.0, 1)
    assert np.array_equal(_bits(out[0, 0]), _bits(nd[0, 0]))
    assert want.shape == (4,)


def test_fill_state_zeroes_non_finite_slopes(oracle):
    s = np.stack([_fronto(2.0)] * 4)[None].copy()
    s[0, 1, 0] = np.nan                # a NaN plane keeps its place ...
    out = oracle.fill_state(s, 1.0, 1)
    t = s.copy()
    t[0, 1, 0] = 0.0                   # ... with a, b set to 0
    t[0, 1, 1] = 0.0
    want = oracle.normal_depth(t, 1.0, 1)
    assert np.array_equal(_bits(out[0, 1]), _bits(want[0, 1]))


def test_run_depthmap_composes_its_pieces(oracle, synth):
    """Two iterations of the pipeline against the same sequence built from
    the oracle's pinned pieces: tensor alpha 12, theta 1e-3 (accept rule
    theta * 10, lambda theta * 15 + 2), fill_state, 40 Huber steps with
    eps 1e-4."""
    L, R, _ = synth.make_pair(40, 24, cells=3, dmax=8.0, channels=4, seed=4)
    r, dmax, n = 2, 8.0, 2
    ref = oracle.run_depthmap(L, R, r, dmax, n, 1.0, 5.0)
    tl, tr = oracle.prepare(L, r), oracle.prepare(R, r)
    dtl = oracle.diffusion_tensor(L, 12.0)
    dtr = oracle.diffusion_tensor(R, 12.0)
    sl = oracle.init(tl, tr, r, dmax, 1)
    sr = oracle.init(tr, tl, r, dmax, 0)
    sml = np.zeros_like(sl)
    smr = np.zeros_like(sr)
    vb2 = oracle.viewbuf_empty(40, 24)
    theta = np.float32(1e-3)
    for it in range(n):
        lam = np.float32(theta * np.float32(15.0) + np.float32(2.0))
        th = np.float32(theta * np.float32(10.0))
        sl, vb1 = oracle.propagate(tl, tr, r, dmax, 1, it, sl, vb2, sr,
                                   smooth=sml, theta=th)
        sr, vb2 = oracle.propagate(tr, tl, r, dmax, 0, it, sr, vb1, sl,
                                   smooth=smr, theta=th)
        sml = oracle.huber(oracle.fill_state(sl, np.float32(1.0) / np.float32(dmax), 1),
                           dtl, [lam] * 4, 1e-4)
        smr = oracle.huber(oracle.fill_state(sr, np.float32(1.0) / np.float32(dmax), 0),
                           dtr, [lam] * 4, 1e-4)
    for k, v in (("stateL", sl), ("stateR", sr), ("smoothL", sml),
                 ("smoothR", smr)):
        assert np.array_equal(_bits(ref[k]), _bits(v)), k
