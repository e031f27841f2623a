"""Batches of seeded synthetic stereo pairs generated on the GPU, for
training the plane proposal.

The scene model is cvt_amd.synth's (Voronoi cells carrying slanted
disparity planes kept inside (0.05 dmax, dmax], a multi-octave Perlin
texture per colour channel plus Gaussian noise, a right view sampled from
the clean left texture with the nearest surface winning and uniform noise
where no cell maps), vectorised over a batch with torch's generator; the
random streams differ from the numpy generator's, so the pairs are not the
same as synth.make_pair's for equal seeds.  Returns float32 device tensors
left, right (B, H, W, 4, alpha 1) and the left-view ground truth (B, H, W).
"""
import math

import torch


def _fade(t):
    return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)


def _perlin(B, H, W, g, device, scale=24.0, octaves=4, persistence=0.55):
    ys = torch.arange(H, device=device, dtype=torch.float32)[:, None]
    xs = torch.arange(W, device=device, dtype=torch.float32)[None, :]
    total = torch.zeros((B, H, W), device=device)
    amp, freq, norm = 1.0, 1.0 / scale, 0.0
    bidx = torch.arange(B, device=device)[:, None, None]
    for _ in range(octaves):
        gw, gh = int(W * freq) + 2, int(H * freq) + 2
        ang = torch.rand((B, gh + 1, gw + 1), generator=g,
                         device=device) * (2.0 * math.pi)
        gxg, gyg = torch.cos(ang), torch.sin(ang)
        u, v = xs * freq, ys * freq
        i0, j0 = torch.floor(u).long(), torch.floor(v).long()
        fu, fv = (u - i0).expand(H, W), (v - j0).expand(H, W)
        i0, j0 = i0.expand(H, W), j0.expand(H, W)

        def corner(di, dj):
            jj, ii = (j0 + dj)[None], (i0 + di)[None]
            return gxg[bidx, jj, ii] * (fu - di) + gyg[bidx, jj, ii] * (fv - dj)

        su, sv = _fade(fu), _fade(fv)
        n0 = corner(0, 0) * (1 - su) + corner(1, 0) * su
        n1 = corner(0, 1) * (1 - su) + corner(1, 1) * su
        total += amp * (n0 * (1 - sv) + n1 * sv)
        norm += amp
        amp *= persistence
        freq *= 2.0
    total /= norm
    lo = total.amin(dim=(1, 2), keepdim=True)
    hi = total.amax(dim=(1, 2), keepdim=True)
    return (total - lo) / torch.clamp(hi - lo, min=1e-12)


def make_pairs(B, W, H, cells=24, dmax=128.0, seed=0, device="cuda",
               noise=0.01):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f = dict(device=device, dtype=torch.float32)
    ys = torch.arange(H, **f)[None, :, None].expand(B, H, W)
    xs = torch.arange(W, **f)[None, None, :].expand(B, H, W)
    sx = torch.rand((B, cells), generator=g, device=device) * W
    sy = torch.rand((B, cells), generator=g, device=device) * H
    d2 = (xs[:, None] - sx[:, :, None, None]) ** 2 + \
         (ys[:, None] - sy[:, :, None, None]) ** 2
    label = d2.argmin(1)                                     # B, H, W
    del d2
    d0 = (0.15 + 0.7 * torch.rand((B, cells), generator=g,
                                  device=device)) * dmax
    gx = (torch.rand((B, cells), generator=g, device=device) * 2 - 1) * 0.12
    gy = (torch.rand((B, cells), generator=g, device=device) * 2 - 1) * 0.12

    def at(t):
        return torch.gather(t, 1, label.view(B, -1)).view(B, H, W)

    # keep each cell's plane inside [0.05 dmax, dmax] over the cell
    dd = at(gx) * (xs - at(sx)) + at(gy) * (ys - at(sy))
    lab = label.view(B, -1)
    lo = torch.full((B, cells), float("inf"), **f).scatter_reduce(
        1, lab, dd.view(B, -1), "amin")
    hi = torch.full((B, cells), float("-inf"), **f).scatter_reduce(
        1, lab, dd.view(B, -1), "amax")
    present = torch.isfinite(lo)
    lo_lim, hi_lim = 0.05 * dmax, dmax
    s = torch.ones((B, cells), **f)
    lo_abs, hi_abs = d0 + lo, d0 + hi
    s = torch.where(present & (lo_abs < lo_lim),
                    torch.minimum(s, (d0 - lo_lim) /
                                  torch.clamp(d0 - lo_abs, min=1e-9)), s)
    s = torch.where(present & (hi_abs > hi_lim),
                    torch.minimum(s, (hi_lim - d0) /
                                  torch.clamp(hi_abs - d0, min=1e-9)), s)
    gx, gy = gx * s, gy * s
    gt = at(d0) + at(gx) * (xs - at(sx)) + at(gy) * (ys - at(sy))

    tex = torch.stack([_perlin(B, H, W, g, device) for _ in range(3)], -1)

    # right view: xr = xL - d(xL); solve per cell, nearest surface wins
    best_d = torch.full((B, H, W), -1.0, **f)
    best_x = torch.zeros((B, H, W), **f)
    for c in range(cells):
        cd0, cgx, cgy = d0[:, c, None, None], gx[:, c, None, None], \
            gy[:, c, None, None]
        csx, csy = sx[:, c, None, None], sy[:, c, None, None]
        xl = (xs + cd0 - cgx * csx + cgy * (ys - csy)) / (1.0 - cgx)
        inb = (xl >= 0) & (xl <= W - 1)
        dall = (xl[:, None] - sx[:, :, None, None]) ** 2 + \
               (ys[:, None] - sy[:, :, None, None]) ** 2
        dist = dall[:, c].clone()
        dall[:, c] = float("inf")
        own = inb & (dist <= dall.amin(1))
        del dall
        dv = cd0 + cgx * (xl - csx) + cgy * (ys - csy)
        take = own & (dv > best_d)
        best_d = torch.where(take, dv, best_d)
        best_x = torch.where(take, xl, best_x)
    hit = best_d > 0
    fl = torch.floor(best_x)
    i0 = torch.clamp(fl.long(), 0, W - 1)
    i1 = torch.clamp(i0 + 1, 0, W - 1)
    fr = (best_x - fl)[..., None]

    def col(i):
        return torch.gather(tex, 2, i[..., None].expand(B, H, W, 3))

    right_clean = col(i0) * (1 - fr) + col(i1) * fr
    occl = torch.rand((B, H, W, 3), generator=g, device=device)
    right_clean = torch.where(hit[..., None], right_clean, occl)
    left = tex + torch.randn((B, H, W, 3), generator=g, device=device) * noise
    right = right_clean + torch.randn((B, H, W, 3), generator=g,
                                      device=device) * noise
    one = torch.ones((B, H, W, 1), **f)
    return (torch.cat([left, one], -1).contiguous(),
            torch.cat([right, one], -1).contiguous(), gt.contiguous())
