"""Pins of the oracle's precompute, costs (patch_eval_census,
pmhstereo.cl:661-736, and patch_eval_color_grad_weighted, :510-565) and KLT
step (patch_census_klt2, :326-410): closed forms and special cases first,
then a literal brute force on tiny images (tests/refimpl.py)."""
import numpy as np
import pytest

import refimpl

FRONTO = np.array([1.0, 0.0, 0.0, 0.0], np.float32)   # x' = x


def rand_img(rng, H, W, C):
    if C == 1:
        return rng.uniform(0, 1, (H, W)).astype(np.float32)
    img = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    img[..., 3] = 1.0
    return img


def shift_plane(d):
    # left-view plane with constant disparity d: x' = x - d
    return np.array([1.0, 0.0, -float(d), 0.0], np.float32)


# ---------------------------------------------------------------- precompute
def test_gradient_closed_form(oracle):
    img = np.arange(20, dtype=np.float32).reshape(4, 5) / 20
    tex = oracle.prepare(img, 1)
    g = img * np.float32(np.float32(0.2126) + np.float32(0.7152)
                         + np.float32(0.0722))
    # interior: forward and central differences of the gray image
    assert np.allclose(tex[1, 2, 4], g[1, 3] - g[1, 2], atol=1e-6)
    assert np.allclose(tex[1, 2, 5], g[2, 2] - g[1, 2], atol=1e-6)
    assert np.allclose(tex[1, 2, 6], g[1, 3] - g[1, 1], atol=1e-6)
    assert np.allclose(tex[1, 2, 7], g[2, 2] - g[0, 2], atol=1e-6)
    # zero border (CLK_ADDRESS_CLAMP): right edge sees gray 0
    assert np.allclose(tex[1, 4, 4], -g[1, 4], atol=1e-6)
    assert np.allclose(tex[3, 0, 5], -g[3, 0], atol=1e-6)
    assert np.allclose(tex[0, 0, 6], g[0, 1], atol=1e-6)


def test_bilateral_weight_closed_form(oracle):
    # uniform image: every in-image tap weighs exp(0) = 1
    img = np.full((9, 11), 0.3, np.float32)
    tex = oracle.prepare(img, 2)
    assert tex[4, 5, 3] == 25.0
    assert tex[0, 0, 3] == 9.0       # corner: 3 x 3 taps inside
    assert tex[0, 5, 3] == 15.0
    # two-level image: weight exp(-12 |dot(delta, gray)|)
    img = np.zeros((1, 3), np.float32)
    img[0, 2] = 0.1
    tex = oracle.prepare(img, 1)
    gsum = np.float32(0.2126) + np.float32(0.7152) + np.float32(0.0722)
    assert abs(tex[0, 0, 3] - 2.0) < 1e-6
    assert abs(tex[0, 1, 3] - (2.0 + np.exp(-12 * 0.1 * gsum))) < 1e-6
    assert np.array_equal(tex[0, 1, :3], img[0, 1].repeat(3))


def test_bilateral_weight_follows_color_weight(oracle):
    # colour factor 0: every tap weighs exp(-0) = 1 whatever the image, so
    # the sum is the count of taps inside the image
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 1, (9, 11, 4)).astype(np.float32)
    tex = oracle.prepare(img, 2, color_weight=0.0)
    ys, xs = np.mgrid[0:9, 0:11]
    count = ((np.minimum(xs + 2, 10) - np.maximum(xs - 2, 0) + 1) *
             (np.minimum(ys + 2, 8) - np.maximum(ys - 2, 0) + 1))
    assert np.array_equal(tex[..., 3], count.astype(np.float32))
    # factor 5 on the two-level image: exp(-5 |dot(delta, gray)|)
    img = np.zeros((1, 3), np.float32)
    img[0, 2] = 0.1
    tex = oracle.prepare(img, 1, color_weight=5.0)
    gsum = np.float32(0.2126) + np.float32(0.7152) + np.float32(0.0722)
    assert abs(tex[0, 1, 3] - (2.0 + np.exp(-5 * 0.1 * gsum))) < 1e-6


@pytest.mark.parametrize("C", [1, 4])
def test_prepare_brute_force(oracle, C):
    rng = np.random.default_rng(10 + C)
    img = rand_img(rng, 7, 9, C)
    got = oracle.prepare(img, 2)
    want = refimpl.prepare(img, 2)
    assert np.array_equal(got[..., [0, 1, 2, 4, 5, 6, 7]],
                          want[..., [0, 1, 2, 4, 5, 6, 7]])
    assert np.allclose(got[..., 3], want[..., 3], rtol=2e-6)


# ---------------------------------------------------------------- cost
def test_identical_images_cost_zero(oracle):
    rng = np.random.default_rng(0)
    img = rand_img(rng, 12, 16, 4)
    tex = oracle.prepare(img, 2)
    for (x, y) in [(0, 0), (5, 6), (15, 11)]:
        assert oracle.cost_at(tex, tex, 2, x, y, FRONTO) == 0.0


def test_integer_shift_cost_zero(oracle):
    rng = np.random.default_rng(1)
    H, W, d, r = 10, 30, 4, 2
    left = rand_img(rng, H, W, 1)
    right = rng.uniform(0, 1, (H, W)).astype(np.float32)
    right[:, :W - d] = left[:, d:]          # right(x) = left(x + d)
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    for x in range(d + r + 1, W - r - 2):
        for y in range(1, H - 1):
            assert oracle.cost_at(tl, tr, r, x, y, shift_plane(d)) == 0.0
            assert oracle.cost_at(tl, tr, r, x, y, shift_plane(d + 1)) > 0.02


def test_alternating_pattern_closed_form(oracle):
    # L(x) = 0.5 + 0.01(-1)^x, R(x) = 0.5 - 0.01(-1)^x: every tap at x' = x
    # has |R - L| = 0.02 per channel and exactly one of the four gradient
    # signs differs (gx), so cost = 0.125*0.33*3*(0.02*20) + 0.875*0.25
    H, W, r = 9, 13, 3
    alt = np.where(np.arange(W) % 2 == 0, 0.01, -0.01).astype(np.float32)
    left = np.tile(np.float32(0.5) + alt, (H, 1))
    right = np.tile(np.float32(0.5) - alt, (H, 1))
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    want = 0.125 * 0.33 * 3 * (0.02 * 20) + 0.875 * 0.25
    got = oracle.cost_at(tl, tr, r, 6, 4, FRONTO)
    assert abs(got - want) < 1e-5, (got, want)


def test_uniform_images_mean_and_truncation(oracle):
    H, W, r = 9, 9, 2
    left = np.zeros((H, W, 4), np.float32)
    right = np.zeros((H, W, 4), np.float32)
    left[..., :3] = [0.2, 0.5, 0.7]
    right[..., :3] = [0.21, 0.52, 0.8]     # diffs 0.01, 0.02, 0.1 (> 0.05)
    left[..., 3] = right[..., 3] = 1.0
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    want = 0.125 * 0.33 * (0.2 + 0.4 + 1.0)
    assert abs(oracle.cost_at(tl, tr, r, 4, 4, FRONTO) - want) < 1e-5


def test_center_only_tap_gives_1_1(oracle):
    rng = np.random.default_rng(2)
    img = rand_img(rng, 11, 11, 1)
    tex = oracle.prepare(img, 2)
    # steep plane: only the centre tap maps into [0, W)
    # x'(5+dx, 5+dy) = 1000 dx + 20 dy + 5.5
    s = np.array([1000.0, 20.0, 0.0, 0.0], np.float32)
    s[2] = 5.5 - 1000.0 * 5 - 20.0 * 5
    assert oracle.cost_at(tex, tex, 2, 5, 5, s) == np.float32(1.1)


def test_invalid_planes_cost_1e5(oracle):
    rng = np.random.default_rng(3)
    img = rand_img(rng, 8, 8, 1)
    tex = oracle.prepare(img, 1)
    for s in ([np.nan, 0, 0, 0], [1, np.inf, 0, 0],
              [1, 0, -20, 0],            # x' outside the image
              [1, 0, 100, 0],
              [2e5, 0, -2e5 * 3, 0]):    # normal z below 1e-4
        assert oracle.cost_at(tex, tex, 1, 3, 3,
                              np.array(s, np.float32)) == np.float32(1e5)


def test_early_exit_iff_total_reaches_mincost(oracle):
    # uniform weights make bilw == wsum, so census/bilw is the full cost
    H, W, r = 9, 9, 2
    left = np.zeros((H, W), np.float32) + 0.3
    right = np.zeros((H, W), np.float32) + 0.32
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    full = oracle.cost_at(tl, tr, r, 4, 4, FRONTO)
    assert 0 < full < 1
    assert oracle.cost_at(tl, tr, r, 4, 4, FRONTO, mincost=full) == 1.0
    assert oracle.cost_at(tl, tr, r, 4, 4, FRONTO,
                          mincost=full * 1.001) == full


@pytest.mark.parametrize("C", [1, 4])
def test_cost_brute_force(oracle, C):
    rng = np.random.default_rng(20 + C)
    H, W, r = 7, 10, 2
    left, right = rand_img(rng, H, W, C), rand_img(rng, H, W, C)
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    for _ in range(60):
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H))
        d = rng.uniform(0, 4)
        s = np.array([1 + rng.normal(0, 0.05), rng.normal(0, 0.05),
                      -d + rng.normal(0, 0.5), 0], np.float32)
        mc = float(rng.choice([1e10, rng.uniform(0, 0.5)]))
        got = oracle.cost_at(tl, tr, r, x, y, s, mincost=mc)
        want = refimpl.cost(tl, tr, r, x, y, s, mincost=mc)
        assert abs(got - want) <= 1e-6 * max(1.0, abs(want)), (x, y, s, mc)


# ------------------------------------------- colour + gradient cost
CG = dict(cost_kind=1)


def test_cg_identical_images_cost_zero(oracle):
    rng = np.random.default_rng(30)
    img = rand_img(rng, 12, 16, 4)
    tex = oracle.prepare(img, 2)
    for (x, y) in [(0, 0), (5, 6), (15, 11)]:
        assert oracle.cost_at(tex, tex, 2, x, y, FRONTO, **CG) == 0.0


def test_cg_integer_shift_cost_zero(oracle):
    # right(x) = left(x + d) on [0, W - d); the gradients of both agree at
    # x' = x - d when x' - 1 >= 0 and x' + 1 < W - d (central differences),
    # so every tap of the pixels x in [d + r + 1, W - r - 1) matches
    rng = np.random.default_rng(31)
    H, W, d, r = 10, 30, 4, 2
    left = rand_img(rng, H, W, 4)
    right = rand_img(rng, H, W, 4)
    right[:, :W - d] = left[:, d:]
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    for x in range(d + r + 1, W - r - 1):
        for y in range(0, H):
            assert oracle.cost_at(tl, tr, r, x, y, shift_plane(d), **CG) == 0
            assert oracle.cost_at(tl, tr, r, x, y, shift_plane(d + 1),
                                  **CG) > 0.01


def test_cg_uniform_images_colour_truncation(oracle):
    # uniform images: unit weights, zero gradients in the interior, so the
    # cost is 0.05 x the truncated colour sum
    H, W, r = 9, 9, 2
    left = np.zeros((H, W, 4), np.float32)
    right = np.zeros((H, W, 4), np.float32)
    left[..., :3] = [0.2, 0.5, 0.7]
    right[..., :3] = [0.21, 0.52, 0.8]     # diffs 0.01, 0.02, 0.1 (> 0.05)
    left[..., 3] = right[..., 3] = 1.0
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    want = 0.05 * (0.01 + 0.02 + 0.05)
    assert abs(oracle.cost_at(tl, tr, r, 4, 4, FRONTO, **CG) - want) < 1e-7


def test_cg_gradient_term_and_truncation(oracle):
    # L: horizontal gray ramp of slope s per px; R: constant far from L.
    # Every colour difference truncates (0.15); the gradient differences
    # are (s, 0, 2s, 0) * gsum, truncated at 0.05 each; the weights vary,
    # but C is the same at every interior tap, so the mean is C.
    H, W, r = 9, 15, 2
    gsum = float(np.float32(0.2126) + np.float32(0.7152) + np.float32(0.0722))
    for slope, trunc in ((0.01, False), (0.04, True)):
        left = np.tile((0.5 + slope * np.arange(W)).astype(np.float32),
                       (H, 1))
        right = np.zeros((H, W), np.float32) - 1.0
        tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
        g = [slope * gsum, 0.0, 2 * slope * gsum, 0.0]
        sg = sum(min(abs(v), 0.05) for v in g)
        want = 0.05 * 0.15 + 0.95 * sg
        got = oracle.cost_at(tl, tr, r, 7, 4, FRONTO, **CG)
        assert abs(got - want) < 2e-6, (slope, got, want)
        if trunc:
            assert abs(sg - (min(0.04 * gsum, 0.05) + 0.05)) < 1e-9


def test_cg_no_tap_in_range_is_nan(oracle):
    rng = np.random.default_rng(32)
    tex = oracle.prepare(rand_img(rng, 8, 8, 4), 1)
    for s in ([1, 0, -20, 0], [1, 0, 100, 0], [np.nan, 0, 0, 0],
              [1, np.inf, 0, 0]):
        assert np.isnan(oracle.cost_at(tex, tex, 1, 3, 3,
                                       np.array(s, np.float32), **CG))


def test_cg_has_no_invalid_plane_test(oracle):
    # a plane steep in y (normal z < 1e-4) that maps its own row into the
    # image: the census cost rejects it (1e5), this cost is finite, and
    # the mincost argument has no effect (no early exit)
    rng = np.random.default_rng(33)
    tex = oracle.prepare(rand_img(rng, 11, 11, 4), 2)
    s = np.array([1.0, 2e5, -2e5 * 5, 0.0], np.float32)
    assert oracle.cost_at(tex, tex, 2, 5, 5, s) == np.float32(1e5)
    c = oracle.cost_at(tex, tex, 2, 5, 5, s, **CG)
    assert c == 0.0      # only row 5 taps count, at x' = x
    s = shift_plane(0.5)
    full = oracle.cost_at(tex, tex, 2, 5, 5, s, **CG)
    assert 0 < full and oracle.cost_at(tex, tex, 2, 5, 5, s, mincost=0.0,
                                       **CG) == full


def test_cg_weight_factor(oracle):
    # cost_weight 0 gives unit weights: the plain mean of the tap costs;
    # other factors weigh the taps differently
    rng = np.random.default_rng(34)
    H, W, r = 7, 10, 2
    left, right = rand_img(rng, H, W, 4), rand_img(rng, H, W, 4)
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    s = shift_plane(1.0)
    taps = []
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            # cost of a single tap: the patch of radius 0 at that pixel
            taps.append(oracle.cost_at(tl, tr, 0, 5 + dx, 3 + dy, s, **CG))
    mean0 = oracle.cost_at(tl, tr, r, 5, 3, s, cost_kind=1, cost_weight=0.0)
    assert abs(mean0 - np.mean(taps)) < 1e-6
    c16 = oracle.cost_at(tl, tr, r, 5, 3, s, **CG)
    c25 = oracle.cost_at(tl, tr, r, 5, 3, s, cost_kind=1, cost_weight=25.5)
    assert c16 != mean0 and c25 != c16
    assert abs(c25 - refimpl.cost_cg(tl, tr, r, 5, 3, s, kw=25.5)) < 1e-6


@pytest.mark.parametrize("C", [1, 4])
def test_cg_cost_brute_force(oracle, C):
    rng = np.random.default_rng(40 + C)
    H, W, r = 7, 10, 2
    left, right = rand_img(rng, H, W, C), rand_img(rng, H, W, C)
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    nan = 0
    for _ in range(80):
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H))
        d = rng.uniform(0, 12)
        s = np.array([1 + rng.normal(0, 0.1), rng.normal(0, 0.1),
                      -d + rng.normal(0, 0.5), 0], np.float32)
        kw = float(rng.choice([16.0, 25.5]))
        got = oracle.cost_at(tl, tr, r, x, y, s, cost_kind=1, cost_weight=kw)
        want = refimpl.cost_cg(tl, tr, r, x, y, s, kw=kw)
        if np.isnan(want):
            nan += 1
            assert np.isnan(got), (x, y, s)
            continue
        assert abs(got - want) <= 1e-6 * max(1.0, abs(want)), (x, y, s, kw)
    assert nan > 0      # some planes map the whole patch out of the image


# ---------------------------------------------------------------- KLT
def test_klt_singular_system_is_rejected(oracle):
    # identical uniform images: every curvature is 0, H = 0, the solve
    # gives NaN and the candidate then costs 1e5
    img = np.full((9, 9), 0.4, np.float32)
    tex = oracle.prepare(img, 2)
    k = oracle.klt_at(tex, tex, 2, 4, 4, FRONTO)
    assert not np.isfinite(k[:3]).all()
    assert oracle.cost_at(tex, tex, 2, 4, 4, k) == np.float32(1e5)


def test_klt_brute_force(oracle):
    rng = np.random.default_rng(5)
    H, W, r = 8, 12, 2
    left = rand_img(rng, H, W, 4)
    right = np.roll(left, -2, axis=1) + rng.normal(0, 0.02, left.shape)
    right = right.astype(np.float32)
    tl, tr = oracle.prepare(left, r), oracle.prepare(right, r)
    for _ in range(40):
        x, y = int(rng.integers(0, W)), int(rng.integers(0, H))
        s = np.array([1 + rng.normal(0, 0.02), rng.normal(0, 0.02),
                      -2 + rng.normal(0, 0.3), 0.5], np.float32)
        got = oracle.klt_at(tl, tr, r, x, y, s)
        want = refimpl.klt(tl, tr, r, x, y, s)
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.allclose(got[fin], want[fin], rtol=2e-5, atol=1e-4), \
            (x, y, got, want)
