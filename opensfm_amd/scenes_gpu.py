"""The scene renderer of ``scenes.render`` on the GPU, for the large synthetic
workloads (2048x1536 ring views, 128 street views, 4000x3000 nadir views).

Test and benchmark scaffolding only: it produces the input images and the
ground-truth depth; the depthmap engine never calls it.  PyTorch supplies
the device arithmetic.  It follows the host renderer step by step (FP64
rays, FP32 heightfield marching refined by FP64 bisection, the same value-
noise hash and shading), so both render the same scene; the images agree
with the host's up to rare 1-level differences where transcendental
functions round differently.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = (1 << 64) - 1


def _i64(c):
    """A 64-bit unsigned constant as the int64 with the same bits."""
    c &= _M64
    return c - (1 << 64) if c >= (1 << 63) else c


_C1, _C2, _C3 = _i64(0x9E3779B97F4A7C15), _i64(0xC2B2AE3D27D4EB4F), _i64(0x165667B19E3779F9)
_F1, _F2 = _i64(0xFF51AFD7ED558CCD), _i64(0xC4CEB9FE1A85EC53)


def _shr(h, k):
    """Logical right shift of int64 bits."""
    import torch
    return (h >> k) & torch.tensor((1 << (64 - k)) - 1, dtype=torch.int64, device=h.device)


def _hash3(ix, iy, iz, seed):
    """scenes._hash3 in int64 arithmetic (wrapping products and xors give the
    uint64 bits)."""
    h = (ix * _C1) ^ (iy * _C2) ^ (iz * _C3) ^ _i64(seed)
    h = h ^ _shr(h, 33)
    h = h * _F1
    h = h ^ _shr(h, 33)
    h = h * _F2
    h = h ^ _shr(h, 33)
    import torch
    return _shr(h, 40).to(torch.float32) * (1.0 / (1 << 24))


def value_noise3(p, spacing, seed):
    import torch
    q = p / spacing
    i0 = torch.floor(q)
    f = (q - i0).to(torch.float32)
    f = f * f * (3.0 - 2.0 * f)
    i0 = i0.to(torch.int64)
    out = torch.zeros(p.shape[0], dtype=torch.float32, device=p.device)
    for dz in (0, 1):
        wz = f[:, 2] if dz else 1.0 - f[:, 2]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1.0 - f[:, 1]
            for dx in (0, 1):
                wx = f[:, 0] if dx else 1.0 - f[:, 0]
                out += wx * wy * wz * _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed)
    return out


class _Terrain:
    """Heights and gradients of a SceneGeometry's surface on the device."""

    def __init__(self, geom, device):
        import torch
        self.kind = None
        if geom.heightfield is not None:
            hf = geom.heightfield
            self.kind = "grid"
            self.amp = hf.amp
            self.extent, self.res, self.n = hf.extent, hf.res, hf.n
            self.grid = torch.as_tensor(hf.grid, device=device)
            self.steps = getattr(hf, "steps", 48)
            self.bisect = getattr(hf, "bisect", 12)
        elif geom.terrain is not None:
            self.kind = "sin"
            self.amp, per = geom.terrain
            self.k = 2 * math.pi / per
            self.steps, self.bisect = 48, 12

    def _cell(self, x, y):
        import torch
        gx = torch.clamp((x.double() + self.extent) / self.res, 0, self.n - 1.001)
        gy = torch.clamp((y.double() + self.extent) / self.res, 0, self.n - 1.001)
        ix, iy = gx.to(torch.int64), gy.to(torch.int64)
        return ix, iy, gx - ix, gy - iy

    def height(self, x, y):
        import torch
        if self.kind == "sin":
            return self.amp * torch.sin(self.k * x) * torch.sin(self.k * y)
        ix, iy, fx, fy = self._cell(x, y)
        g = self.grid
        g00, g01 = g[iy, ix].double(), g[iy, ix + 1].double()
        g10, g11 = g[iy + 1, ix].double(), g[iy + 1, ix + 1].double()
        top = g00 + fx * (g01 - g00)
        bot = g10 + fx * (g11 - g10)
        return (top + fy * (bot - top)).to(x.dtype)

    def grad(self, x, y):
        import torch
        if self.kind == "sin":
            k = self.k
            return (self.amp * k * torch.cos(k * x) * torch.sin(k * y),
                    self.amp * k * torch.sin(k * x) * torch.cos(k * y))
        ix, iy, fx, fy = self._cell(x, y)
        g = self.grid
        g00, g01 = g[iy, ix].double(), g[iy, ix + 1].double()
        g10, g11 = g[iy + 1, ix].double(), g[iy + 1, ix + 1].double()
        dx = (1 - fy) * (g01 - g00) + fy * (g11 - g10)
        dy = (1 - fx) * (g10 - g00) + fx * (g11 - g01)
        return dx / self.res, dy / self.res


def _hit_heightfield(o, d, terr):
    import torch
    n_rays = d.shape[1]
    dev = d.device
    t_hit = torch.full((n_rays,), math.inf, dtype=torch.float64, device=dev)
    normal = torch.zeros((3, n_rays), dtype=torch.float64, device=dev)
    zlo, zhi = -terr.amp, terr.amp
    ta = (zhi - o[2]) / d[2]
    tb = (zlo - o[2]) / d[2]
    idx = torch.nonzero((d[2] < 0) & (tb > 0)).flatten()
    if idx.numel() == 0:
        return t_hit, normal
    dx, dy, dz = (d[k, idx].float() for k in range(3))
    a = torch.clamp(ta[idx], min=1e-3).float()
    step = ((tb[idx] - a.double()) / terr.steps).float()
    ox, oy, oz = (float(np.float32(v)) for v in o)

    def f(t):
        return (oz + t * dz) - terr.height(ox + t * dx, oy + t * dy)

    prev_t = a
    prev_f = f(prev_t)
    found = torch.zeros(idx.numel(), dtype=torch.bool, device=dev)
    lo_t = torch.zeros(idx.numel(), dtype=torch.float32, device=dev)
    for s in range(1, terr.steps + 1):
        tc = a + float(s) * step
        fc = f(tc)
        cross = (~found) & (prev_f > 0) & (fc <= 0)
        lo_t = torch.where(cross, prev_t, lo_t)
        found |= cross
        prev_t, prev_f = tc, fc
    sel = torch.nonzero(found).flatten()
    lo = lo_t[sel].double()
    hi = lo + step[sel].double()
    gi = idx[sel]
    ddx, ddy, ddz = d[0, gi], d[1, gi], d[2, gi]
    for _ in range(terr.bisect):
        mid = 0.5 * (lo + hi)
        fm = (o[2] + mid * ddz) - terr.height(o[0] + mid * ddx, o[1] + mid * ddy)
        above = fm > 0
        lo = torch.where(above, mid, lo)
        hi = torch.where(above, hi, mid)
    th = 0.5 * (lo + hi)
    t_hit[gi] = th
    gx, gy = terr.grad(o[0] + th * ddx, o[1] + th * ddy)
    nrm = torch.sqrt(gx * gx + gy * gy + 1.0)
    normal[0, gi] = -gx / nrm
    normal[1, gi] = -gy / nrm
    normal[2, gi] = 1.0 / nrm
    return t_hit, normal


def _hit_box(o, d, lo, hi):
    import torch
    n = d.shape[1]
    dev = d.device
    tmin = torch.full((n,), -math.inf, dtype=torch.float64, device=dev)
    tmax = torch.full((n,), math.inf, dtype=torch.float64, device=dev)
    axis = torch.zeros(n, dtype=torch.int64, device=dev)
    for ax in range(3):
        inv = 1.0 / d[ax]
        t1 = (lo[ax] - o[ax]) * inv
        t2 = (hi[ax] - o[ax]) * inv
        lo_a = torch.fmin(t1, t2)
        hi_a = torch.fmax(t1, t2)
        upd = lo_a > tmin
        axis = torch.where(upd, ax, axis)
        tmin = torch.where(upd, lo_a, tmin)
        tmax = torch.fmin(tmax, hi_a)
    hit = (tmin <= tmax) & (tmin > 0)
    t = torch.where(hit, tmin, torch.full_like(tmin, math.inf))
    nrm = torch.zeros((3, n), dtype=torch.float64, device=dev)
    cols = torch.arange(n, device=dev)
    nrm[axis, cols] = -torch.sign(d[axis, cols])
    return t, nrm


def render(geom, K, R, t, width, height, device="cuda"):
    """Ray-cast one view on the GPU: (image u8, z-depth f32) as numpy."""
    import torch
    from . import scenes
    dev = torch.device(device)
    jj, ii = torch.meshgrid(torch.arange(width, dtype=torch.float64, device=dev),
                            torch.arange(height, dtype=torch.float64, device=dev), indexing="xy")
    Kinv = torch.as_tensor(np.linalg.inv(K), device=dev)
    Rt = torch.as_tensor(np.asarray(R).T.copy(), device=dev)
    q = torch.stack([jj.reshape(-1), ii.reshape(-1), torch.ones(jj.numel(), dtype=torch.float64,
                                                                  device=dev)])
    d = Rt @ (Kinv @ q)
    o = (-np.asarray(R).T @ np.asarray(t)).astype(np.float64)
    best_t = torch.full((d.shape[1],), math.inf, dtype=torch.float64, device=dev)
    best_n = torch.zeros_like(d)

    def take(th, nh, cols=None):
        nonlocal best_t, best_n
        if cols is None:
            better = th < best_t
            best_t = torch.where(better, th, best_t)
            best_n = torch.where(better[None, :], nh, best_n)
        else:
            better = th < best_t[cols]
            c = cols[better]
            best_t[c] = th[better]
            best_n[:, c] = nh[:, better]

    terr = _Terrain(geom, dev)
    if terr.kind is not None:
        take(*_hit_heightfield(o, d, terr))
    elif geom.ground:
        tz = (0.0 - o[2]) / d[2]
        tz = torch.where((d[2] < 0) & (tz > 0), tz, torch.full_like(tz, math.inf))
        nz = torch.zeros_like(d)
        nz[2] = 1.0
        take(tz, nz)
    for lo, hi in geom.boxes:
        lo = np.asarray(lo, np.float64)
        hi = np.asarray(hi, np.float64)
        bb = scenes._box_pixel_bbox(lo, hi, K, R, t, width, height)
        if bb is None:
            continue
        j0, j1, i0, i1 = bb
        cols = (torch.arange(i0, i1, device=dev)[:, None] * width +
                torch.arange(j0, j1, device=dev)[None, :]).reshape(-1)
        th, nh = _hit_box(o, d[:, cols], lo, hi)
        take(th, nh, cols)

    hit = torch.isfinite(best_t)
    p = torch.as_tensor(o, device=dev)[None, :] + torch.where(hit, best_t, 0.0)[:, None] * d.T
    tex = value_noise3(p, geom.tex_spacing, geom.tex_seed)
    L = np.asarray(geom.light, np.float64)
    L = torch.as_tensor(L / np.linalg.norm(L), device=dev)
    shade = 0.55 + 0.45 * torch.clamp(L @ best_n, 0.0, 1.0)
    val = torch.where(hit, 255.0 * tex.double() * shade, torch.zeros_like(shade))
    image = torch.clamp(torch.round(val), 0, 255).to(torch.uint8).reshape(height, width)
    depth = torch.where(hit, best_t, 0.0).float().reshape(height, width)
    return image.cpu().numpy(), depth.cpu().numpy()
