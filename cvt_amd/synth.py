"""Seeded synthetic stereo pairs with ground truth, and the bad-pixel metric.

The reference has no stereo data or generator; this module produces the
scenes the benchmark configurations describe:

* a piecewise-planar scene of ``cells`` Voronoi cells (seeded uniform
  sites); cell ``c`` carries the slanted left-view disparity plane
  ``d = d0 + gx (x - sx) + gy (y - sy)`` kept inside ``(0, dmax]``;
* a left image of seeded multi-octave Perlin gradient noise (one field per
  colour channel for RGBA, alpha = 1) plus Gaussian noise of standard
  deviation ``noise``;
* a right image that samples the clean left texture by horizontal linear
  interpolation at ``xL = xr + d(xL, y)``, the nearest surface (largest
  disparity) winning; right pixels no cell maps to (occluded or outside)
  are uniform noise; then its own Gaussian noise.

The error metrics of the generated pairs are cvt's Disparity ones, in
``cvt_amd.disparity``.
"""
import numpy as np


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def perlin(H, W, rng, scale=24.0, octaves=4, persistence=0.55):
    """Multi-octave 2-D gradient noise in [0, 1], shape (H, W), float64."""
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    total = np.zeros((H, W))
    amp, freq, norm = 1.0, 1.0 / scale, 0.0
    for _ in range(octaves):
        gw = int(W * freq) + 2
        gh = int(H * freq) + 2
        ang = rng.uniform(0.0, 2.0 * np.pi, size=(gh + 1, gw + 1))
        gx, gy = np.cos(ang), np.sin(ang)
        u, v = xs * freq, ys * freq
        i0, j0 = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
        fu, fv = u - i0, v - j0

        def corner(di, dj):
            g_x = gx[j0 + dj, i0 + di]
            g_y = gy[j0 + dj, i0 + di]
            return g_x * (fu - di) + g_y * (fv - dj)

        su, sv = _fade(fu), _fade(fv)
        n0 = corner(0, 0) * (1 - su) + corner(1, 0) * su
        n1 = corner(0, 1) * (1 - su) + corner(1, 1) * su
        total += amp * (n0 * (1 - sv) + n1 * sv)
        norm += amp
        amp *= persistence
        freq *= 2.0
    total /= norm
    lo, hi = total.min(), total.max()
    return (total - lo) / max(hi - lo, 1e-12)


def _planes(W, H, cells, dmax, rng):
    sites = np.stack([rng.uniform(0, W, cells), rng.uniform(0, H, cells)], 1)
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    d2 = (xs[None] - sites[:, 0, None, None]) ** 2 + \
         (ys[None] - sites[:, 1, None, None]) ** 2
    label = np.argmin(d2, axis=0)
    planes = np.zeros((cells, 5))   # d0, gx, gy, sx, sy
    lo_lim, hi_lim = 0.05 * dmax, dmax
    for c in range(cells):
        sx, sy = sites[c]
        d0 = rng.uniform(0.15, 0.85) * dmax
        gx, gy = rng.uniform(-0.12, 0.12, 2)
        m = label == c
        if m.any():
            dd = gx * (xs[m] - sx) + gy * (ys[m] - sy)
            lo, hi = d0 + dd.min(), d0 + dd.max()
            s = 1.0
            if lo < lo_lim:
                s = min(s, (d0 - lo_lim) / max(d0 - lo, 1e-9))
            if hi > hi_lim:
                s = min(s, (hi_lim - d0) / max(hi - d0, 1e-9))
            gx, gy = gx * s, gy * s
        planes[c] = (d0, gx, gy, sx, sy)
    return sites, label, planes


def make_pair(W, H, cells=8, dmax=32.0, channels=1, seed=0, noise=0.01):
    """Returns (left, right, gt) as float32; images are (H, W) for
    channels == 1 or (H, W, 4) RGBA, gt is the left-view disparity."""
    rng = np.random.default_rng(seed)
    sites, label, planes = _planes(W, H, cells, dmax, rng)
    nch = 1 if channels == 1 else 3
    tex = np.stack([perlin(H, W, rng) for _ in range(nch)], -1)  # H,W,nch
    ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
    d0, gx, gy, sx, sy = (planes[label, k] for k in range(5))
    gt = d0 + gx * (xs - sx) + gy * (ys - sy)

    # right view: xr = xL - d(xL); solve per cell, nearest surface wins
    best_d = np.full((H, W), -1.0)
    best_x = np.zeros((H, W))
    for c in range(cells):
        cd0, cgx, cgy, csx, csy = planes[c]
        xl = (xs + cd0 - cgx * csx + cgy * (ys - csy)) / (1.0 - cgx)
        inb = (xl >= 0) & (xl <= W - 1)
        dist = (xl - sites[c, 0]) ** 2 + (ys - sites[c, 1]) ** 2
        own = inb.copy()
        for c2 in range(cells):
            if c2 != c:
                own &= dist <= (xl - sites[c2, 0]) ** 2 + \
                    (ys - sites[c2, 1]) ** 2
        dv = cd0 + cgx * (xl - csx) + cgy * (ys - csy)
        take = own & (dv > best_d)
        best_d = np.where(take, dv, best_d)
        best_x = np.where(take, xl, best_x)
    hit = best_d > 0
    i0 = np.clip(np.floor(best_x).astype(np.int64), 0, W - 1)
    i1 = np.clip(i0 + 1, 0, W - 1)
    fr = best_x - np.floor(best_x)
    rows = np.arange(H)[:, None]
    right_clean = tex[rows, i0] * (1 - fr)[..., None] + \
        tex[rows, i1] * fr[..., None]
    occl = rng.uniform(0.0, 1.0, size=(H, W, nch))
    right_clean = np.where(hit[..., None], right_clean, occl)

    left = tex + rng.normal(0.0, noise, size=tex.shape)
    right = right_clean + rng.normal(0.0, noise, size=tex.shape)
    if channels == 1:
        left, right = left[..., 0], right[..., 0]
    else:
        one = np.ones((H, W, 1))
        left = np.concatenate([left, one], -1)
        right = np.concatenate([right, one], -1)
    return (left.astype(np.float32), right.astype(np.float32),
            gt.astype(np.float32))
