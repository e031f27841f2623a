"""Slow, literal float32 transcription of three reference functions, for
brute-force comparison with the oracle on tiny images:

* gradxy (pmhstereo.cl:1549-1602) and bilateral_weight_to_alpha
  (pmhstereo.cl:1462-1498), with math.exp;
* patch_eval_census (pmhstereo.cl:661-736), per tap with the early exit;
* patch_eval_color_grad_weighted (pmhstereo.cl:510-565);
* patch_census_klt2 (pmhstereo.cl:326-410) + mat3f_lusolve
  (Matrix3.cl:286-311).

Written against the reference text with plain Python loops and numpy
float32 scalars (an fma is formed exactly in float64 and rounded once), so
a wrong index, sign or dropped term in the C oracle shows up as a mismatch.
"""
import math

import numpy as np

f32 = np.float32
GW = (f32(0.2126), f32(0.7152), f32(0.0722))


def fma(a, b, c):
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def rgba(img, x, y):
    if img.ndim == 2:
        v = f32(img[y, x])
        return (v, v, v, v)
    return tuple(f32(v) for v in img[y, x])


def gray_at(img, x, y):
    H, W = img.shape[:2]
    if x < 0 or x >= W or y < 0 or y >= H:
        return f32(0.0)
    v = rgba(img, x, y)
    # the built-in dot as a chain of fused multiply-adds
    return fma(v[3], f32(0.0),
               fma(v[2], GW[2], fma(v[1], GW[1], f32(v[0] * GW[0]))))


def weight(d, cw=12.0):
    g = fma(d[2], GW[2], fma(d[1], GW[1], f32(d[0] * GW[0])))
    return f32(math.exp(float(f32(-abs(g) * f32(cw)))))


def prepare(img, r, cw=12.0):
    H, W = img.shape[:2]
    tex = np.zeros((H, W, 8), np.float32)
    for y in range(H):
        for x in range(W):
            c = rgba(img, x, y)
            ws = f32(0.0)
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    qx, qy = x + dx, y + dy
                    if 0 <= qx < W and 0 <= qy < H:
                        q = rgba(img, qx, qy)
                        ws = f32(ws + weight([q[k] - c[k] for k in range(3)],
                                             cw))
            g = gray_at(img, x, y)
            tex[y, x] = [c[0], c[1], c[2], ws,
                         gray_at(img, x + 1, y) - g, gray_at(img, x, y + 1) - g,
                         gray_at(img, x + 1, y) - gray_at(img, x - 1, y),
                         gray_at(img, x, y + 1) - gray_at(img, x, y - 1)]
    return tex


def transform(s, x, y):
    return f32(fma(s[0], f32(x), f32(s[1] * f32(y))) + s[2])


def sample(row, u):
    W = row.shape[0]
    u = max(min(f32(u), f32(W)), f32(-1.0))
    fl = f32(math.floor(u))
    a = f32(u - fl)
    i0 = min(max(int(fl), 0), W - 1)
    i1 = min(max(int(fl) + 1, 0), W - 1)
    t0, t1 = row[i0], row[i1]
    return [fma(a, f32(t1[k] - t0[k]), t0[k]) for k in range(8)]


def signs(g):
    return [1 if np.signbit(v) else 0 for v in g]


def to_normal_z(s):
    sx, sy = f32(1.0 - s[0]), f32(0.0 - s[1])
    nz = f32(f32(1.0) / f32(math.sqrt(f32(fma(sx, sx, f32(sy * sy)) +
                                          f32(1.0)))))
    nx, ny = f32(-sx * nz), f32(-sy * nz)
    # (1 - nx nx) - ny ny, each product fused into its difference
    t = fma(-ny, ny, fma(-nx, nx, f32(1.0)))
    return f32(math.sqrt(t)) if t >= 0 else f32("nan")


def cost(tex1, tex2, r, x, y, s, mincost=1e10, cw=12.0):
    H, W = tex1.shape[:2]
    s = [f32(v) for v in s]
    if not all(math.isfinite(v) for v in s[:3]):
        return f32(1e5)
    xp = transform(s, x, y)
    if to_normal_z(s) < f32(1e-4) or xp < 0 or xp >= W:
        return f32(1e5)
    c = tex1[y, x]
    bilw = c[3]
    census, wsum = f32(0.0), f32(0.0)
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            qx, qy = x + dx, y + dy
            xq = transform(s, qx, qy)
            if xq < 0 or xq >= W or qx < 0 or qx >= W or qy < 0 or qy >= H:
                continue
            t1 = tex1[qy, qx]
            w = weight([t1[k] - c[k] for k in range(3)], cw)
            v2 = sample(tex2[qy], xq)
            cc = [f32(min(abs(f32(v2[k] - t1[k])), f32(0.05)) * f32(20.0))
                  for k in range(3)]
            cad = f32(f32(0.33) * f32(f32(cc[0] + cc[1]) + cc[2]))
            cost_ = f32(f32(0.125) * cad)
            nd = sum(a != b for a, b in zip(signs(t1[4:8]), signs(v2[4:8])))
            cost_ = f32(cost_ + f32(f32(0.875) * f32(f32(0.25) * f32(nd))))
            census = fma(w, cost_, census)
            wsum = f32(wsum + w)
            if f32(census / bilw) >= mincost:
                return f32(1.0)
    if wsum <= 1.0:
        return f32(1.1)
    return f32(census / wsum)


def cost_cg(tex1, tex2, r, x, y, s, kw=16.0):
    """patch_eval_color_grad_weighted: w1 = exp(-|dot(valcenter - val1,
    gray)| * kw); C = 0.05 dot(fmin(|val1 - val2|, 0.05), 1) + (1 - 0.05)
    dot(fmin(|gval1 - gval2|, 0.05), 1); ret1 / wsum1.  The tap loop's
    tests are the kernel's; a NaN x' skips the tap (the oracle's stated
    choice where the reference samples at a NaN coordinate)."""
    H, W = tex1.shape[:2]
    s = [f32(v) for v in s]
    c = tex1[y, x]
    ret, wsum = f32(0.0), f32(0.0)
    alpha = f32(0.05)
    with np.errstate(all="ignore"):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                qx, qy = x + dx, y + dy
                if qx < 0 or qx >= W or qy < 0 or qy >= H:
                    continue
                t1 = tex1[qy, qx]
                d = [f32(c[k] - t1[k]) for k in range(3)]
                g = fma(d[2], GW[2], fma(d[1], GW[1], f32(d[0] * GW[0])))
                w1 = f32(math.exp(float(f32(-abs(g) * f32(kw)))))
                xq = transform(s, qx, qy)
                if not (xq >= 0 and xq < W):
                    continue
                wsum = f32(wsum + w1)
                v2 = sample(tex2[qy], xq)
                dc = [min(abs(f32(t1[k] - v2[k])), f32(0.05)) for k in range(3)]
                dg = [min(abs(f32(t1[k] - v2[k])), f32(0.05))
                      for k in range(4, 8)]
                sc = f32(f32(dc[0] + dc[1]) + dc[2])
                sg = f32(f32(f32(dg[0] + dg[1]) + dg[2]) + dg[3])
                C = fma(alpha, sc, f32(f32(f32(1.0) - alpha) * sg))
                ret = fma(w1, C, ret)
        return f32(ret / wsum) if wsum != 0 else f32("nan")


def klt_term(row2, u, val1, ga, w1):
    v2 = sample(row2, u)
    cc = [f32(min(abs(f32(v2[k] - val1[k])), f32(0.05)) * f32(20.0))
          for k in range(3)]
    C = f32(f32(f32(w1 * f32(0.125)) * f32(0.33)) * f32(f32(cc[0] + cc[1]) +
                                                        cc[2]))
    nd = sum(a != b for a, b in zip(ga, signs(v2[4:8])))
    # (double)C + (double)(w1 * 0.875f) * 0.25 * n, rounded to float
    return f32(np.float64(C) + np.float64(f32(w1 * f32(0.875))) * 0.25 * nd)


def lusolve(m, b):
    m = [[f32(v) for v in row] for row in m]
    b = [f32(v) for v in b]
    m[1][0] = f32(m[1][0] / m[0][0])
    m[2][0] = f32(m[2][0] / m[0][0])
    # x -= a * b as fma(-a, b, x); dot(u, v) = fma(u1, v1, u0 * v0)
    m[1][1] = fma(-m[1][0], m[0][1], m[1][1])
    m[1][2] = fma(-m[1][0], m[0][2], m[1][2])
    m[2][1] = fma(-m[2][0], m[0][1], m[2][1])
    m[2][2] = fma(-m[2][0], m[0][2], m[2][2])
    m[2][1] = f32(m[2][1] / m[1][1])
    m[2][2] = fma(-m[2][1], m[1][2], m[2][2])
    b[1] = fma(-m[1][0], b[0], b[1])
    b[2] = f32(b[2] - fma(m[2][1], b[1], f32(m[2][0] * b[0])))
    z = f32(b[2] / m[2][2])
    yv = f32(fma(-m[1][2], z, b[1]) / m[1][1])
    xv = f32(f32(b[0] - fma(m[0][2], z, f32(m[0][1] * yv))) / m[0][0])
    return [xv, yv, z]


def klt(tex1, tex2, r, x, y, s, cw=12.0):
    H, W = tex1.shape[:2]
    s = [f32(v) for v in s]
    c = tex1[y, x]
    J = [f32(0.0)] * 3
    Hm = [[f32(0.0)] * 3 for _ in range(3)]
    with np.errstate(all="ignore"):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                qx, qy = x + dx, y + dy
                if qx < 0 or qx >= W or qy < 0 or qy >= H:
                    continue
                t1 = tex1[qy, qx]
                val1 = [f32(t1[k] - c[k]) for k in range(3)]
                w1 = weight(val1, cw)
                if qx < 1 or qx >= W - 1:
                    continue
                xq = transform(s, qx, qy)
                ga = signs(t1[4:8])
                C = klt_term(tex2[qy], xq, val1, ga, w1)
                Cp = klt_term(tex2[qy], f32(xq + f32(1.0)), val1, ga, w1)
                Cm = klt_term(tex2[qy], f32(xq - f32(1.0)), val1, ga, w1)
                curv = f32(f32(Cp + Cm) - f32(2.0 * C))
                if curv > 0:
                    p3 = [f32(qx), f32(qy), f32(1.0)]
                    g = f32(f32(w1 * f32(Cm - Cp)) * f32(0.5))
                    J = [fma(g, p3[i], J[i]) for i in range(3)]
                    h = f32(w1 * curv)
                    for i in range(3):
                        for j in range(3):
                            Hm[i][j] = f32(Hm[i][j] + f32(h * f32(p3[i] * p3[j])))
        d = lusolve(Hm, J)
        return np.array([s[0] - d[0], s[1] - d[1], s[2] - d[2],
                         s[3] - f32(1e5)], np.float32)
