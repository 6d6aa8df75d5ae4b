"""Seeded synthetic multi-view scenes with ground-truth depth.

The reference ships no image renderer (opensfm/synthetic_data/ produces
cameras and 3D points only), so this module ray-casts piecewise-planar and
heightfield scenes into uint8 grayscale views with exact z-depth.

Conventions follow the dense path's inputs:
  * cameras are world->camera, x_c = R X + t (opensfm/dense.py:681-682);
  * K is a pinhole with focal f = 0.85 W and the principal point at the image
    centre in the engine's pixel convention, where pixel (j, i) is the image
    point (j, i) (depthmap.cc:448-450), so c = ((W-1)/2, (H-1)/2);
  * ground-truth depth is the camera-frame z of the first surface hit,
    0 where a ray misses everything.

Texture is 3D value noise (a hashed lattice with smoothstep-trilinear
interpolation) evaluated at the world-space hit point, so every view sees the
same surface pattern.  The lattice spacing is chosen so that one cell spans
`corr_px` pixels at the reference camera's distance to the scene centre.
A fixed directional light adds Lambertian shading so that faces differ.
"""
from __future__ import annotations

import concurrent.futures
import dataclasses
import math
import os

import numpy as np


@dataclasses.dataclass
class View:
    K: np.ndarray          # (3, 3) float64
    R: np.ndarray          # (3, 3) float64, world -> camera
    t: np.ndarray          # (3,) float64
    image: np.ndarray      # (H, W) uint8
    mask: np.ndarray       # (H, W) uint8, 255 everywhere
    depth: np.ndarray      # (H, W) float32 ground-truth z-depth (0 = miss)


@dataclasses.dataclass
class Scene:
    views: list            # View, index 0 is the reference view
    min_depth: float
    max_depth: float
    name: str = ""

    @property
    def ref(self) -> View:
        return self.views[0]


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


# --------------------------------------------------------------------------
# cameras


def look_at(center, target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0)):
    """World->camera (R, t) for a camera at `center` looking at `target`,
    image x to the right and y down."""
    c = np.asarray(center, np.float64)
    f = np.asarray(target, np.float64) - c
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    if np.linalg.norm(r) < 1e-9:
        r = np.cross(f, np.array([0.0, 1.0, 0.0]))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    t = -R @ c
    return R, t


def intrinsics(width, height, focal_factor=0.85):
    f = focal_factor * width
    return np.array([[f, 0.0, (width - 1) / 2.0],
                     [0.0, f, (height - 1) / 2.0],
                     [0.0, 0.0, 1.0]])


def arc_cameras(n_neighbours, radius=10.0, spacing_deg=8.0, elevation_deg=45.0):
    """Reference camera at azimuth 0 and neighbours at +-spacing, +-2 spacing,
    ... on a horizontal arc of the given radius, all looking at the origin."""
    az = [0.0]
    k = 1
    while len(az) < n_neighbours + 1:
        az.append(k * spacing_deg)
        if len(az) < n_neighbours + 1:
            az.append(-k * spacing_deg)
        k += 1
    el = math.radians(elevation_deg)
    poses = []
    for a in az:
        a = math.radians(a)
        c = radius * np.array([math.cos(a) * math.cos(el),
                               math.sin(a) * math.cos(el),
                               math.sin(el)])
        poses.append(look_at(c))
    return poses


# --------------------------------------------------------------------------
# value noise


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
         ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
         ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)
         ^ np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / (1 << 24))


def value_noise3(p, spacing, seed):
    """Uniform-random lattice values in [0, 1), smoothstep-trilinear."""
    q = p / spacing
    i0 = np.floor(q)
    f = (q - i0).astype(np.float32)
    f = f * f * (3.0 - 2.0 * f)
    i0 = i0.astype(np.int64)
    out = np.zeros(p.shape[0], np.float32)
    with np.errstate(over="ignore"):
        for dz in (0, 1):
            wz = f[:, 2] if dz else 1.0 - f[:, 2]
            for dy in (0, 1):
                wy = f[:, 1] if dy else 1.0 - f[:, 1]
                for dx in (0, 1):
                    wx = f[:, 0] if dx else 1.0 - f[:, 0]
                    out += wx * wy * wz * _hash3(i0[:, 0] + dx, i0[:, 1] + dy,
                                                 i0[:, 2] + dz, seed)
    return out


# --------------------------------------------------------------------------
# surfaces: rays are o + t d with d = R^T K^-1 (j, i, 1), so t is camera z


def _hit_plane_z(o, d, z0=0.0):
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (z0 - o[2]) / d[2]
    t = np.where((d[2] < 0) & (t > 0), t, np.inf)
    n = np.zeros((3, t.size))
    n[2] = 1.0
    return t, n


def _hit_box(o, d, lo, hi):
    with np.errstate(divide="ignore", invalid="ignore"):
        tmin = np.full(d.shape[1], -np.inf)
        tmax = np.full(d.shape[1], np.inf)
        axis = np.zeros(d.shape[1], np.int64)
        for a in range(3):
            inv = 1.0 / d[a]
            t1 = (lo[a] - o[a]) * inv
            t2 = (hi[a] - o[a]) * inv
            lo_a = np.fmin(t1, t2)
            hi_a = np.fmax(t1, t2)
            upd = lo_a > tmin
            axis = np.where(upd, a, axis)
            tmin = np.where(upd, lo_a, tmin)
            tmax = np.fmin(tmax, hi_a)
    hit = (tmin <= tmax) & (tmin > 0)
    t = np.where(hit, tmin, np.inf)
    n = np.zeros((3, t.size))
    cols = np.arange(t.size)
    n[axis, cols] = -np.sign(d[axis, cols])
    return t, n


def _hit_heightfield(o, d, height_fn, grad_fn, zlo, zhi, steps=48, bisect=12):
    """First crossing of z = height(x, y), marching the slab [zlo, zhi] in
    float32 and refining by bisection in float64."""
    n_rays = d.shape[1]
    t_hit = np.full(n_rays, np.inf)
    normal = np.zeros((3, n_rays))
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = (zhi - o[2]) / d[2]
        tb = (zlo - o[2]) / d[2]
    idx = np.nonzero((d[2] < 0) & (tb > 0))[0]
    if idx.size == 0:
        return t_hit, normal
    dx, dy, dz = (d[k, idx].astype(np.float32) for k in range(3))
    a = np.maximum(ta[idx], 1e-3).astype(np.float32)
    step = ((tb[idx] - a) / steps).astype(np.float32)
    ox, oy, oz = (np.float32(v) for v in o)

    def f(t):
        return (oz + t * dz) - height_fn(ox + t * dx, oy + t * dy)

    prev_t = a
    prev_f = f(prev_t)
    found = np.zeros(idx.size, bool)
    lo_t = np.zeros(idx.size, np.float32)
    for s in range(1, steps + 1):
        tc = a + np.float32(s) * step
        fc = f(tc)
        cross = (~found) & (prev_f > 0) & (fc <= 0)
        lo_t[cross] = prev_t[cross]
        found |= cross
        prev_t, prev_f = tc, fc
    sel = np.nonzero(found)[0]
    lo = lo_t[sel].astype(np.float64)
    hi = lo + step[sel].astype(np.float64)
    gi = idx[sel]
    ddx, ddy, ddz = d[0, gi], d[1, gi], d[2, gi]
    for _ in range(bisect):
        mid = 0.5 * (lo + hi)
        fm = (o[2] + mid * ddz) - height_fn(o[0] + mid * ddx, o[1] + mid * ddy)
        above = fm > 0
        lo = np.where(above, mid, lo)
        hi = np.where(above, hi, mid)
    th = 0.5 * (lo + hi)
    t_hit[gi] = th
    gx, gy = grad_fn(o[0] + th * ddx, o[1] + th * ddy)
    nrm = np.sqrt(gx * gx + gy * gy + 1.0)
    normal[0, gi] = -gx / nrm
    normal[1, gi] = -gy / nrm
    normal[2, gi] = 1.0 / nrm
    return t_hit, normal


# --------------------------------------------------------------------------
# scene description and rendering


@dataclasses.dataclass
class SceneGeometry:
    boxes: list                  # list of (lo[3], hi[3])
    ground: bool = True
    terrain: tuple | None = None  # (amplitude, period) sinusoidal heightfield
    heightfield: object = None    # or any object with amp, height(x, y), grad(x, y)
    tex_spacing: float = 0.15
    tex_seed: int = 0
    light: tuple = (0.4, 0.3, 0.866)

    def height(self, x, y):
        amp, per = self.terrain
        k = 2 * math.pi / per
        if x.dtype == np.float32:
            amp, k = np.float32(amp), np.float32(k)
        return amp * np.sin(k * x) * np.sin(k * y)

    def grad(self, x, y):
        amp, per = self.terrain
        k = 2 * math.pi / per
        return (amp * k * np.cos(k * x) * np.sin(k * y),
                amp * k * np.sin(k * x) * np.cos(k * y))


def _box_pixel_bbox(lo, hi, K, R, t, width, height):
    """Pixel rectangle covering the projected box, or None when a corner is
    behind the camera (then every ray is tested)."""
    corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1])
                        for z in (lo[2], hi[2])])
    pc = corners @ R.T + t
    if np.any(pc[:, 2] <= 1e-6):
        return 0, width, 0, height
    uv = (pc @ K.T)
    u = uv[:, 0] / uv[:, 2]
    v = uv[:, 1] / uv[:, 2]
    j0 = max(0, int(math.floor(u.min())) - 1)
    j1 = min(width, int(math.ceil(u.max())) + 2)
    i0 = max(0, int(math.floor(v.min())) - 1)
    i1 = min(height, int(math.ceil(v.max())) + 2)
    if j0 >= j1 or i0 >= i1:
        return None
    return j0, j1, i0, i1


def render(geom: SceneGeometry, K, R, t, width, height):
    """Ray-cast one view: returns (image u8, z-depth f32)."""
    jj, ii = np.meshgrid(np.arange(width, dtype=np.float64),
                         np.arange(height, dtype=np.float64))
    Kinv = np.linalg.inv(K)
    rc = Kinv @ np.stack([jj.ravel(), ii.ravel(), np.ones(jj.size)])
    d = R.T @ rc                             # world-frame rays, camera z = 1
    o = -R.T @ t                             # camera centre
    best_t = np.full(d.shape[1], np.inf)
    best_n = np.zeros_like(d)

    def take(th, nh, cols=None):
        if cols is None:
            better = th < best_t
            best_t[better] = th[better]
            best_n[:, better] = nh[:, better]
        else:
            better = th < best_t[cols]
            c = cols[better]
            best_t[c] = th[better]
            best_n[:, c] = nh[:, better]

    if geom.heightfield is not None:
        hf = geom.heightfield
        take(*_hit_heightfield(o, d, hf.height, hf.grad, -hf.amp, hf.amp,
                               steps=getattr(hf, "steps", 48), bisect=getattr(hf, "bisect", 12)))
    elif geom.terrain is not None:
        amp = geom.terrain[0]
        take(*_hit_heightfield(o, d, geom.height, geom.grad, -amp, amp))
    elif geom.ground:
        take(*_hit_plane_z(o, d))
    for lo, hi in geom.boxes:
        lo = np.asarray(lo, np.float64)
        hi = np.asarray(hi, np.float64)
        bb = _box_pixel_bbox(lo, hi, K, R, t, width, height)
        if bb is None:
            continue
        j0, j1, i0, i1 = bb
        cols = (np.arange(i0, i1)[:, None] * width + np.arange(j0, j1)[None, :]).ravel()
        th, nh = _hit_box(o, d[:, cols], lo, hi)
        take(th, nh, cols)

    hit = np.isfinite(best_t)
    p = o[None, :] + np.where(hit, best_t, 0)[:, None] * d.T
    tex = value_noise3(p, geom.tex_spacing, geom.tex_seed)
    L = np.asarray(geom.light, np.float64)
    L = L / np.linalg.norm(L)
    shade = 0.55 + 0.45 * np.clip(L @ best_n, 0.0, 1.0)
    val = np.where(hit, 255.0 * tex * shade, 0.0)
    image = np.clip(np.rint(val), 0, 255).astype(np.uint8).reshape(height, width)
    depth = np.where(hit, best_t, 0.0).astype(np.float32).reshape(height, width)
    return image, depth


def _render_all(geom, cams, device=None):
    """Renders (K, R, t, width, height) cameras: on the host with a thread
    pool, or one after another on a torch device (scenes_gpu)."""
    if device is not None:
        from . import scenes_gpu
        return [scenes_gpu.render(geom, K, R, t, w, h, device=device) for K, R, t, w, h in cams]
    with concurrent.futures.ThreadPoolExecutor(max(1, min(len(cams), _threads()))) as ex:
        return list(ex.map(lambda c: render(geom, *c), cams))


def random_boxes(rng, n, extent=4.0, base_fn=None):
    boxes = []
    for _ in range(n):
        cx, cy = rng.uniform(-extent, extent, size=2)
        sx, sy = rng.uniform(0.4, 1.6, size=2)
        h = rng.uniform(0.5, 2.5)
        z0 = -2.0 if base_fn is not None else 0.0
        boxes.append((np.array([cx - sx / 2, cy - sy / 2, z0]),
                      np.array([cx + sx / 2, cy + sy / 2, h])))
    return boxes


def make_scene(width, height, n_neighbours, n_boxes, seed, min_depth, max_depth,
               terrain=None, corr_px=4.0, radius=10.0, spacing_deg=8.0,
               elevation_deg=45.0, name="", device=None):
    """Arc-camera scene of a ground plane (or sinusoidal terrain) and random
    axis-aligned boxes, textured with value noise of `corr_px` correlation."""
    rng = np.random.default_rng(seed)
    K = intrinsics(width, height)
    spacing = corr_px * radius / K[0, 0]
    geom = SceneGeometry(
        boxes=random_boxes(rng, n_boxes, base_fn=terrain),
        ground=True, terrain=terrain, tex_spacing=spacing,
        tex_seed=int(rng.integers(1, 2 ** 62)))
    poses = arc_cameras(n_neighbours, radius, spacing_deg, elevation_deg)
    rendered = _render_all(geom, [(K, R, t, width, height) for R, t in poses], device)
    views = [View(K=K.copy(), R=R, t=t, image=img, mask=np.full_like(img, 255),
                  depth=dep) for (R, t), (img, dep) in zip(poses, rendered)]
    return Scene(views=views, min_depth=min_depth, max_depth=max_depth, name=name)


def scene_config1(seed=1000):
    """320x240 reference + 4 neighbours, ground plane + 6 boxes, depth [2, 20]."""
    return make_scene(320, 240, 4, 6, seed, 2.0, 20.0, name="cfg1")


def scene_config2(seed=2000, width=1600, height=1200):
    """1600x1200 reference + 8 neighbours, 20 boxes on a sinusoidal terrain of
    amplitude 1.5 and period 6, depth [1, 40]."""
    return make_scene(width, height, 8, 20, seed, 1.0, 40.0, terrain=(1.5, 6.0),
                      name="cfg2")


def ring_scene(n_refs, n_neighbours, width, height, n_boxes, seed, min_depth, max_depth,
               terrain=None, corr_px=4.0, radius=10.0, spacing_deg=8.0, elevation_deg=45.0,
               name="", device=None):
    """The make_scene family as a batch: cameras on an arc at `spacing_deg`
    steps, n_refs reference views in the middle of the arc, each with its
    n_neighbours nearest cameras along the arc (half on each side) as
    neighbours; every camera is rendered once."""
    rng = np.random.default_rng(seed)
    K = intrinsics(width, height)
    spacing = corr_px * radius / K[0, 0]
    geom = SceneGeometry(
        boxes=random_boxes(rng, n_boxes, base_fn=terrain),
        ground=True, terrain=terrain, tex_spacing=spacing,
        tex_seed=int(rng.integers(1, 2 ** 62)))
    half = (n_neighbours + 1) // 2
    n_cams = n_refs + 2 * half
    el = math.radians(elevation_deg)
    poses = []
    for k in range(n_cams):
        a = math.radians((k - (n_cams - 1) / 2.0) * spacing_deg)
        c = radius * np.array([math.cos(a) * math.cos(el), math.sin(a) * math.cos(el),
                               math.sin(el)])
        poses.append(look_at(c))
    jobs = []
    for r in range(half, half + n_refs):
        order = sorted((v for v in range(n_cams) if v != r), key=lambda v: (abs(v - r), v))
        jobs.append((r, order[:n_neighbours]))
    rendered = _render_all(geom, [(K, R, t, width, height) for R, t in poses], device)
    views = [View(K=K.copy(), R=R, t=t, image=img, mask=np.full_like(img, 255), depth=dep)
             for (R, t), (img, dep) in zip(poses, rendered)]
    return BatchScene(views=views, jobs=jobs, min_depth=min_depth, max_depth=max_depth,
                      name=name)


def scene_config3(seed=3000, n_refs=16, width=2048, height=1536, neighbours=10, device=None):
    """16 reference views of 2048x1536, 10 neighbours each, from the 1600x1200
    scene's family (20 boxes on a sinusoidal terrain of amplitude 1.5 and
    period 6, depth [1, 40])."""
    return ring_scene(n_refs, neighbours, width, height, 20, seed, 1.0, 40.0,
                      terrain=(1.5, 6.0), name="cfg3", device=device)


def normals_from_depth(K, depth):
    """Camera-frame normals (facing the camera, unnormalised) of a depthmap:
    the cross product of the back-projected point's horizontal and vertical
    central differences; (0, 0, -1) where that vanishes."""
    h, w = depth.shape
    jj, ii = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    P = (np.linalg.inv(K) @ np.stack([jj.ravel(), ii.ravel(), np.ones(jj.size)])).T
    P = (P * depth.reshape(-1, 1).astype(np.float64)).reshape(h, w, 3)
    n = np.cross(np.gradient(P, axis=1), np.gradient(P, axis=0))
    n[n[..., 2] > 0] *= -1
    n[np.linalg.norm(n, axis=-1) == 0] = (0, 0, -1)
    return n.astype(np.float32)


def noisy_depth(view, rng, noise=0.002, holes=0.02):
    """The view's ground-truth depth with relative Gaussian noise and holes
    (depth 0), as float32."""
    d = view.depth.astype(np.float32) * (1 + rng.normal(0, noise, view.depth.shape)).astype(np.float32)
    d[rng.random(d.shape) < holes] = 0
    return d.astype(np.float32)


def depth_accuracy(depth, gt, min_depth=None, max_depth=None, rel=0.01):
    """Fraction of ground-truth-valid pixels whose estimate lies within `rel`
    relative error (depth 0 counts as a miss)."""
    valid = gt > 0
    if min_depth is not None:
        valid &= gt >= min_depth
    if max_depth is not None:
        valid &= gt <= max_depth
    good = valid & (depth > 0) & (np.abs(depth - gt) < rel * gt)
    return float(good.sum()) / max(1, int(valid.sum()))


# --------------------------------------------------------------------------
# street scenes (batched workloads)

STREET_SIZES = ((640, 480), (1024, 768), (1600, 1200), (2048, 1536))


@dataclasses.dataclass
class BatchScene:
    views: list            # View, pool of every camera
    jobs: list             # (reference view index, [neighbour view indices])
    min_depth: float
    max_depth: float
    name: str = ""


def street_geometry(rng, length=60.0, street_width=8.0, tex_spacing=0.06):
    """Ground plane plus facade blocks along both sides of a street that
    runs along world x: each block is set back from the kerb by a random
    offset and has a random length and height."""
    boxes = []
    for side in (1.0, -1.0):
        x = -length / 2
        while x < length / 2:
            seg = rng.uniform(3.0, 9.0)
            gap = rng.uniform(0.0, 1.5)
            y0 = street_width / 2 + rng.uniform(0.0, 2.5)
            h = rng.uniform(4.0, 12.0)
            lo = np.array([x, side * y0 if side > 0 else -(y0 + 3.0), 0.0])
            hi = np.array([x + seg, y0 + 3.0 if side > 0 else -y0, h])
            boxes.append((lo, hi))
            x += seg + gap
    return SceneGeometry(boxes=boxes, ground=True, terrain=None,
                         tex_spacing=tex_spacing,
                         tex_seed=int(rng.integers(1, 2 ** 62)))


def street_scene(n_views, seed=4000, sizes=STREET_SIZES, neighbours=(4, 12),
                 min_depth=1.0, max_depth=60.0, name="street", device=None):
    """Cameras along a street (1 m apart, height 1.6-2.6 m, yaw within 40
    degrees of the street axis, pitched 3-12 degrees down), each with an
    image size drawn uniformly from `sizes`.  Every camera is a job whose
    neighbours are its nearest cameras along the track, their count drawn
    uniformly from the inclusive range `neighbours`."""
    rng = np.random.default_rng(seed)
    length = max(60.0, n_views * 1.0 + 30.0)
    geom = street_geometry(rng, length=length)
    xs = -n_views / 2.0 + np.arange(n_views) * 1.0
    poses, dims = [], []
    for x in xs:
        c = np.array([x, rng.uniform(-1.5, 1.5), rng.uniform(1.6, 2.6)])
        yaw = math.radians(rng.uniform(-40.0, 40.0))
        pitch = math.radians(rng.uniform(3.0, 12.0))
        f = np.array([math.cos(yaw) * math.cos(pitch), math.sin(yaw) * math.cos(pitch),
                      -math.sin(pitch)])
        poses.append(look_at(c, c + f))
        dims.append(sizes[int(rng.integers(0, len(sizes)))])
    jobs = []
    for k in range(n_views):
        nn = int(rng.integers(neighbours[0], neighbours[1] + 1))
        order = np.argsort(np.abs(xs - xs[k]), kind="stable")
        jobs.append((k, [int(v) for v in order if v != k][:nn]))

    cams = [(intrinsics(*dims[k]), poses[k][0], poses[k][1], dims[k][0], dims[k][1])
            for k in range(n_views)]
    views = [View(K=c[0], R=c[1], t=c[2], image=img, mask=np.full_like(img, 255), depth=dep)
             for c, (img, dep) in zip(cams, _render_all(geom, cams, device))]
    return BatchScene(views=views, jobs=jobs, min_depth=min_depth,
                      max_depth=max_depth, name=name)


# --------------------------------------------------------------------------
# nadir aerial views over fractal-noise terrain


def _value_noise2(x, y, seed):
    """2-D uniform lattice noise in [0, 1), smoothstep-bilinear."""
    ix, iy = np.floor(x), np.floor(y)
    fx, fy = x - ix, y - iy
    fx = fx * fx * (3 - 2 * fx)
    fy = fy * fy * (3 - 2 * fy)
    ix, iy = ix.astype(np.int64), iy.astype(np.int64)
    z = np.zeros_like(ix)
    with np.errstate(over="ignore"):
        v00 = _hash3(ix, iy, z, seed)
        v10 = _hash3(ix + 1, iy, z, seed)
        v01 = _hash3(ix, iy + 1, z, seed)
        v11 = _hash3(ix + 1, iy + 1, z, seed)
    top = v00 + fx * (v10 - v00)
    bot = v01 + fx * (v11 - v01)
    return (top + fy * (bot - top)).astype(x.dtype)


class FractalTerrain:
    """Heightfield amp * sum_o 0.5^o (2 noise(2^o p / period) - 1) / norm
    (within [-amp, amp]), tabulated on a `res`-metre grid over
    [-extent, extent]^2 and read bilinearly (constant beyond the grid)."""

    steps, bisect = 12, 10  # marching for near-vertical (nadir) rays

    def __init__(self, seed, amp=6.0, period=24.0, octaves=5, extent=80.0, res=0.1):
        self.amp, self.extent, self.res = float(amp), float(extent), float(res)
        n = int(round(2 * extent / res)) + 1
        g = np.linspace(-extent, extent, n)
        X, Y = np.meshgrid(g, g)
        h = np.zeros_like(X)
        norm = sum(0.5 ** o for o in range(octaves))
        for o in range(octaves):
            f = (2.0 ** o) / period
            h += 0.5 ** o * (2.0 * _value_noise2(X * f, Y * f, int(seed) + o) - 1.0)
        self.grid = (h * (amp / norm)).astype(np.float32)
        self.n = n

    def _cell(self, x, y):
        gx = np.clip((np.asarray(x, np.float64) + self.extent) / self.res, 0, self.n - 1.001)
        gy = np.clip((np.asarray(y, np.float64) + self.extent) / self.res, 0, self.n - 1.001)
        ix, iy = gx.astype(np.int64), gy.astype(np.int64)
        return ix, iy, gx - ix, gy - iy

    def height(self, x, y):
        ix, iy, fx, fy = self._cell(x, y)
        g = self.grid
        top = g[iy, ix] + fx * (g[iy, ix + 1] - g[iy, ix])
        bot = g[iy + 1, ix] + fx * (g[iy + 1, ix + 1] - g[iy + 1, ix])
        return (top + fy * (bot - top)).astype(np.asarray(x).dtype)

    def grad(self, x, y):
        ix, iy, fx, fy = self._cell(x, y)
        g = self.grid
        dx = ((1 - fy) * (g[iy, ix + 1] - g[iy, ix]) + fy * (g[iy + 1, ix + 1] - g[iy + 1, ix]))
        dy = ((1 - fx) * (g[iy + 1, ix] - g[iy, ix]) + fx * (g[iy + 1, ix + 1] - g[iy, ix + 1]))
        return dx / self.res, dy / self.res


def nadir_scene(rows, cols, width, height, seed=5000, neighbours=10, altitude=40.0,
                spacing=6.0, amp=6.0, min_depth=20.0, max_depth=60.0, name="nadir",
                device=None):
    """rows x cols cameras looking straight down from `altitude` over fractal
    terrain, on a grid of `spacing` metres; every camera is a job whose
    neighbours are its `neighbours` nearest cameras."""
    rng = np.random.default_rng(seed)
    K = intrinsics(width, height)
    geom = SceneGeometry(boxes=[], ground=False,
                         heightfield=FractalTerrain(
                             int(rng.integers(1, 2 ** 40)), amp=amp,
                             extent=max(rows, cols) * spacing / 2 + 0.75 * altitude + 5.0),
                         tex_spacing=4.0 * altitude / K[0, 0],
                         tex_seed=int(rng.integers(1, 2 ** 62)))
    centres = []
    for r in range(rows):
        for c in range(cols):
            centres.append(np.array([(c - (cols - 1) / 2) * spacing + rng.uniform(-1, 1),
                                     (r - (rows - 1) / 2) * spacing + rng.uniform(-1, 1),
                                     altitude + rng.uniform(-2, 2)]))
    poses = [look_at(c, c - np.array([0.0, 0.0, 1.0]), up=(0.0, 1.0, 0.0)) for c in centres]
    n = len(poses)
    jobs = []
    for k in range(n):
        dist = [np.linalg.norm(centres[k][:2] - centres[v][:2]) for v in range(n)]
        order = [int(v) for v in np.argsort(dist, kind="stable") if v != k]
        jobs.append((k, order[:neighbours]))

    cams = [(K, R, t, width, height) for R, t in poses]
    views = [View(K=K.copy(), R=R, t=t, image=img, mask=np.full_like(img, 255), depth=dep)
             for (R, t), (img, dep) in zip(poses, _render_all(geom, cams, device))]
    return BatchScene(views=views, jobs=jobs, min_depth=min_depth, max_depth=max_depth,
                      name=name)


def sfm_landmarks(nx, ny, n_points, seed=6000, spacing=10.0, altitude=30.0, radius=25.0,
                  dup=0.05):
    """A synthetic sparse reconstruction for neighbour selection: nx x ny
    camera centres on a grid at `altitude`, n_points landmarks on a wavy
    ground (a fraction `dup` repeated exactly, as re-triangulated tracks),
    each seen by the cameras within `radius` of it horizontally (at least
    one).  Returns (origins (N, 3), coords (n, 3), observers: per landmark
    the camera indices)."""
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid(np.arange(nx) * spacing, np.arange(ny) * spacing)
    origins = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, altitude)], 1)
    origins += rng.normal(0, 0.5, origins.shape)
    xy = rng.uniform([-spacing, -spacing], [nx * spacing, ny * spacing], (n_points, 2))
    z = 2.0 * np.sin(xy[:, 0] / 17.0) * np.cos(xy[:, 1] / 23.0)
    coords = np.concatenate([xy, z[:, None]], 1)
    nd = int(dup * n_points)
    if nd:
        coords[-nd:] = coords[:nd]
    cx = np.clip(np.round(xy[:, 0] / spacing).astype(int), 0, nx - 1)
    cy = np.clip(np.round(xy[:, 1] / spacing).astype(int), 0, ny - 1)
    reach = int(np.ceil(radius / spacing))
    obs = [[] for _ in range(n_points)]
    for dy in range(-reach, reach + 1):
        for dx in range(-reach, reach + 1):
            ox, oy = cx + dx, cy + dy
            ok = (ox >= 0) & (ox < nx) & (oy >= 0) & (oy < ny)
            cam = np.where(ok, oy * nx + ox, 0)
            d = np.hypot(origins[cam, 0] - coords[:, 0], origins[cam, 1] - coords[:, 1])
            for i in np.flatnonzero(ok & (d < radius)):
                obs[i].append(int(cam[i]))
    for i in range(n_points):
        if not obs[i]:
            obs[i] = [int(cy[i] * nx + cx[i])]
    return origins, coords, obs
