"""Closed-form pins of the oracle's plane algebra (pmhstereo.cl:77-224)."""
import math

import numpy as np
import pytest


def test_viewprop_involution(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        s = np.array([rng.uniform(0.5, 1.5), rng.uniform(-0.3, 0.3),
                      rng.uniform(-50, 50), 0.3], np.float32)
        r = oracle.viewprop(oracle.viewprop(s))
        assert np.allclose(r[:3], s[:3], rtol=1e-5, atol=1e-4)
        assert r[3] == 0.0


def test_viewprop_inverse_map(oracle):
    # viewprop maps x' = a x + b y + c to its inverse on the same row
    s = np.array([0.9, 0.05, -12.0, 0.0], np.float32)
    v = oracle.viewprop(s)
    x, y = 123.0, 45.0
    xp = oracle.transform(s, x, y)
    assert abs(oracle.transform(v, xp, y) - x) < 1e-3


def test_viewprop_degenerate(oracle):
    s = np.array([1e-9, 2.0, 3.0, 4.0], np.float32)
    assert np.array_equal(oracle.viewprop(s),
                          np.array([0.0, -2.0, -3.0, 0.0], np.float32))


def test_from_disparity_normal_maps_to_x_minus_d(oracle):
    rng = np.random.default_rng(2)
    for _ in range(100):
        n = rng.normal(size=3)
        n[2] = abs(n[2]) + 0.3
        n /= np.linalg.norm(n)
        x0, y0, d = rng.uniform(0, 600), rng.uniform(0, 400), rng.uniform(0, 64)
        s = oracle.from_disparity_normal(x0, y0, d, n.astype(np.float32))
        assert abs(oracle.transform(s, x0, y0) - (x0 - d)) < 2e-3
        # recovered normal equals n (nd_state_to_normal)
        assert np.allclose(oracle.to_normal(s), n, atol=2e-4)


def test_transform_definition(oracle):
    s = np.array([0.75, -0.125, 3.5, 0.0], np.float32)
    assert oracle.transform(s, 10.0, 4.0) == np.float32(0.75 * 10 - 0.125 * 4
                                                         + 3.5)
    # (a x + b y) + c with a x fused into the first sum: b y rounds, the
    # fused sum rounds once, then c is added.  Over random planes this
    # differs from the unfused sum and from b y + c first.
    rng = np.random.default_rng(3)
    f64 = np.float64
    diff_unfused = diff_other = 0
    for _ in range(400):
        a, b = np.float32(rng.normal(1, 0.3)), np.float32(rng.normal(0, 0.3))
        c = np.float32(rng.normal(0, 50))
        x, y = np.float32(rng.integers(0, 1920)), np.float32(rng.integers(0, 1080))
        by = np.float32(b * y)
        want = np.float32(np.float32(f64(a) * f64(x) + f64(by)) + c)
        s = np.array([a, b, c, 0.0], np.float32)
        assert oracle.transform(s, x, y) == want
        unfused = np.float32(np.float32(np.float32(a * x) + by) + c)
        other = np.float32(f64(a) * f64(x) +
                           f64(np.float32(f64(b) * f64(y) + f64(c))))
        diff_unfused += want != unfused
        diff_other += want != other
    assert diff_unfused > 0 and diff_other > 0


def test_fronto_parallel_zero_normal(oracle):
    # the cleared smooth field gives normalize(0) = 0 -> n = (0, 0, 1e-4)
    # -> (0, 0, 1): the d = 0 fronto-parallel plane x' = x
    s = oracle.from_disparity_normal(5.0, 7.0, -0.0, np.zeros(3, np.float32))
    assert np.array_equal(s, np.array([1, 0, 0, 0], np.float32))


@pytest.mark.parametrize("lr", [1, 0])
def test_state_init_properties(oracle, lr):
    rng = np.random.default_rng(3)
    dmax = 64.0
    for _ in range(200):
        u = rng.uniform(0, 1, 3).astype(np.float32)
        x, y = rng.integers(0, 500), rng.integers(0, 300)
        s = oracle.state_init(u, x, y, lr, dmax)
        left = s if lr else oracle.viewprop(s)
        # disparity at the pixel is u0 * dmax (x' = x - z on the left)
        d = x - oracle.transform(left, x, y)
        assert abs(d - u[0] * dmax) < 2e-3 * max(1.0, dmax)
        # the normal's z component is max(u2 * 0.95, 1e-4) before
        # normalisation; the slopes are -nx/nz, -ny/nz
        n = oracle.to_normal(left)
        nz = max(u[2] * 0.95, 1e-4)
        w = math.sqrt(1 - (u[2] * 0.95) ** 2)
        az = 2 * math.pi * u[1]
        ref = np.array([w * math.cos(az), w * math.sin(az), nz])
        ref /= np.linalg.norm(ref)
        if nz > 0.05:
            assert np.allclose(n, ref, atol=5e-3), (n, ref)


def test_state_refine_right_view_quirk(oracle):
    # pmhstereo.cl:169-170: the right view returns viewprop of the
    # converted, unperturbed state: viewprop((1,0,0,0) - viewprop(s))
    s = np.array([0.95, 0.02, 4.0, 0.1], np.float32)
    r = oracle.state_refine([0.3, 0.6, 0.9], s, 40, 20, 64.0, 0)
    om = np.array([1, 0, 0, 0], np.float32) - oracle.viewprop(s)
    assert np.array_equal(r, oracle.viewprop(om))


def test_state_refine_left_view(oracle):
    # left view: depth moves by (u0 - 0.5) * 4 and stays in [0, dmax]
    s = oracle.from_disparity_normal(100, 50, 20.0,
                                     np.array([0, 0, 1], np.float32))
    for u0 in (0.0, 0.25, 0.5, 0.75, 1.0):
        r = oracle.state_refine([u0, 0.0, 0.0], s, 100, 50, 64.0, 1)
        d = 100 - oracle.transform(r, 100, 50)
        assert abs(d - (20.0 + (u0 - 0.5) * 4.0)) < 1e-3
    r = oracle.state_refine([0.0, 0.0, 0.0], oracle.from_disparity_normal(
        1, 1, 0.5, np.array([0, 0, 1], np.float32)), 1, 1, 64.0, 1)
    assert abs(1 - oracle.transform(r, 1, 1)) < 1e-6  # clamped to 0


def test_to_ref_normal_depth(oracle):
    n = np.array([0.2, -0.1, 0.9], np.float32)
    n = n / np.linalg.norm(n)
    s = oracle.from_disparity_normal(30, 40, 12.0, n.astype(np.float32))
    r = oracle.to_ref_normal_depth(s, 30, 40, 1)
    assert abs(r[3] - 12.0) < 1e-3
    assert np.allclose(r[:3], n, atol=1e-5)
    bad = np.array([np.nan, 0, 0, 0], np.float32)
    assert np.array_equal(oracle.to_ref_normal_depth(bad, 1, 1, 1),
                          np.array([0, 0, 1, 0], np.float32))


def test_right_view_plane_from_the_mirrored_pass(oracle):
    """The proposal's right-view planes: a network run on the mirrored pair
    (both images flipped, swapped) sees the right view as a left view and
    predicts, at column W - 1 - x, its disparity d' and normal n' in the
    left view's convention (disparity slopes -n'x/n'z along its x, -n'y/n'z
    along y).  The engine builds the right plane as
    from_disparity_normal(x, y, -d', (n'x, -n'y, n'z)), the
    pmhstereo_init_disparity_normal form for the right view
    (pmhstereo.cl:818-819).  For a slanted surface seen from both views
    that plane must map every right pixel to its left match."""
    O = oracle
    W = 200
    d0, gx, gy, x0, y0 = 30.0, 0.08, -0.05, 90.0, 40.0

    def d_left(x, y):
        return d0 + gx * (x - x0) + gy * (y - y0)

    def d_right(xr, y):     # the same surface: x_l = x_r + d_R, d_L(x_l) = d_R
        return (d0 + gx * (xr - x0) + gy * (y - y0)) / (1.0 - gx)

    # the left view's own plane from its normal: slopes -nx/nz = gx, ...
    nl = np.array([-gx, -gy, 1.0], np.float64)
    nl /= np.linalg.norm(nl)
    sl = O.from_disparity_normal(x0, y0, d_left(x0, y0), nl.astype(np.float32))
    for (x, y) in [(x0, y0), (120.0, 10.0), (60.0, 70.0)]:
        assert abs((x - O.transform(sl, x, y)) - d_left(x, y)) < 1e-3
    # the mirrored pass at column xm = W - 1 - xr: d'(xm, y) = d_R(xr, y),
    # so d' slopes -gx/(1-gx) along xm and gy/(1-gx) along y
    xr, y = 70.0, 30.0
    sxm, sy = -gx / (1.0 - gx), gy / (1.0 - gx)
    nm = np.array([-sxm, -sy, 1.0], np.float64)   # -n'x/n'z = sxm, ...
    nm /= np.linalg.norm(nm)
    nr = np.array([nm[0], -nm[1], nm[2]], np.float32)
    sr = O.from_disparity_normal(xr, y, -d_right(xr, y), nr)
    for (x2, y2) in [(xr, y), (xr + 25.0, y - 12.0), (xr - 30.0, y + 20.0)]:
        xl = O.transform(sr, x2, y2)
        assert abs((xl - x2) - d_right(x2, y2)) < 1e-3, (x2, y2)
        # and the left view sees the same disparity at the match
        assert abs(d_left(xl, y2) - (xl - x2)) < 1e-3
    # the unflipped normal gives the wrong slope
    bad = O.from_disparity_normal(xr, y, -d_right(xr, y), nm.astype(np.float32))
    assert abs((O.transform(bad, xr, y + 20.0) - xr) -
               d_right(xr, y + 20.0)) > 0.5
