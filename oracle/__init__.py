"""CPU oracle for the dense depthmap path -- TEST INFRASTRUCTURE ONLY.

A ctypes wrapper over ``oracle/pm_oracle.c``, a plain-C restatement of
OpenSfM's legacy ``dense::DepthmapEstimator`` / ``dense::DepthmapCleaner``
(opensfm/src/lib/dense/src/depthmap.cc:8-555).  Only ``tests/``,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this package; the engine in ``opensfm_amd`` never does.

Parity pins (each checked in tests/test_oracle.py):
  inv3, homography_baked, plane_from_depth_and_normal,
  depth_of_plane_backprojection  -- closed forms of doc/dense_matching.md
  linear_interp_u8/f32          -- textbook bilinear + zero-outside rule
  variance                      -- numpy.var
  ncc / bilateral_weight        -- independent FP64 weighted Pearson correlation
  score / score_f64             -- identity-camera and affine-intensity cases,
                                   FP64 twin, brute-force numpy re-statement
  compute_brute_force           -- textured fronto-parallel plane at a sweep depth
  median5                       -- scipy.ndimage.median_filter(mode="nearest")
  clean                         -- FP64 numpy vote count, identical/perturbed views
  compute_patch_match           -- synthetic ground truth and invariants
Parity unpinned at the last-ulp level (no reference binary can be built here:
depthmap.cc needs OpenCV, which the image lacks): the exact rounding of
cv::Matx33d::inv and the expf/exp overload choice in BilateralWeight.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpm_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    L.orc_create.restype = C.c_void_p
    L.orc_destroy.argtypes = [C.c_void_p]
    L.orc_add_view.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _u8p, _u8p,
                               C.c_int, C.c_int]
    L.orc_set_depth_range.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
    L.orc_set_patchmatch_iterations.argtypes = [C.c_void_p, C.c_int]
    L.orc_set_patch_size.argtypes = [C.c_void_p, C.c_int]
    L.orc_set_min_patch_sd.argtypes = [C.c_void_p, C.c_float]
    L.orc_set_seed.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_last_scorings.argtypes = [C.c_void_p]
    L.orc_last_scorings.restype = C.c_int64
    L.orc_compute.argtypes = [C.c_void_p, C.c_int, _f32p, _f32p, _f32p, _i32p,
                              C.c_int]
    L.orc_score.argtypes = [C.c_void_p, C.c_int, C.c_int, _f32p, C.c_int]
    L.orc_score.restype = C.c_float
    L.orc_score_f64.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p, C.c_int]
    L.orc_score_f64.restype = C.c_double
    L.orc_inv3.argtypes = [_f64p, _f64p]
    L.orc_inv3.restype = C.c_int
    L.orc_linear_interp_u8.argtypes = [_u8p, C.c_int, C.c_int, C.c_float, C.c_float]
    L.orc_linear_interp_u8.restype = C.c_float
    L.orc_linear_interp_f32.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, C.c_float]
    L.orc_linear_interp_f32.restype = C.c_float
    L.orc_variance.argtypes = [_f32p, C.c_int]
    L.orc_variance.restype = C.c_float
    L.orc_ncc.argtypes = [_f32p, _f32p, _f32p, C.c_int]
    L.orc_ncc.restype = C.c_float
    L.orc_bilateral_weight.argtypes = [C.c_float, C.c_float, C.c_float]
    L.orc_bilateral_weight.restype = C.c_float
    L.orc_homography_baked.argtypes = [_f64p, _f64p, _f64p, _f64p, _f32p, _f32p]
    L.orc_plane_from_depth_and_normal.argtypes = [C.c_float, C.c_float, _f64p,
                                                  C.c_float, _f32p, _f32p]
    L.orc_depth_of_plane_backprojection.argtypes = [C.c_double, C.c_double,
                                                    _f64p, _f32p]
    L.orc_depth_of_plane_backprojection.restype = C.c_float
    L.orc_median5.argtypes = [_f32p, _f32p, C.c_int, C.c_int]
    L.orc_clean.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.POINTER(C.c_void_p),
                            _i32p, _i32p, C.c_float, C.c_int, _f32p]
    L.orc_clean.restype = C.c_int
    L.orc_prune.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.POINTER(C.c_void_p),
                            C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                            C.POINTER(C.c_void_p), _i32p, _i32p, C.c_float, C.c_long,
                            _f32p, _f32p, _u8p, _u8p]
    L.orc_prune.restype = C.c_long
    L.orc_fuse.argtypes = [C.c_int, _f64p, _f64p, _f64p] + [C.POINTER(C.c_void_p)] * 4 + [
        _i32p, _i32p, _i32p, _i32p, _i32p, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int,
        C.c_int, C.c_float, C.c_float, C.c_long, _f32p, _f32p, _u8p, C.POINTER(C.c_long)]
    L.orc_fuse.restype = C.c_long
    L.orc_sor.argtypes = [C.c_long, _f32p, _f32p, _u8p, C.c_int, C.c_float, _f32p]
    L.orc_sor.restype = C.c_long
    _i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    L.orc_covis.argtypes = [C.c_long, _f64p, _i64p, _i32p, C.c_int, C.c_int, C.c_double, _i32p,
                            _f64p, _i64p, _i32p, _i64p, _i32p, C.c_long, _i32p, _i32p, C.c_long,
                            C.POINTER(C.c_long), C.POINTER(C.c_float)]
    L.orc_covis.restype = C.c_long
    L.orc_select_neighbors.argtypes = [C.c_int, _f64p, C.c_long, _f64p, _i64p, _i32p, _i32p,
                                       C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       _i64p, _i32p, C.c_long, _i64p, _i32p, C.c_long]
    L.orc_select_neighbors.restype = C.c_long
    L.orc_fuse_inverse3f.argtypes = [_f32p, _f32p]
    L.orc_fuse_inverse3f.restype = None
    _lib = L
    return L


def _f64(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    assert a.size == n, a.shape
    return a


def inv3(K):
    out = np.zeros(9)
    ok = lib().orc_inv3(_f64(K, 9), out)
    if not ok:
        raise ValueError("singular matrix")
    return out.reshape(3, 3)


def linear_interp_u8(img, y, x):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    return lib().orc_linear_interp_u8(img, img.shape[1], img.shape[0], y, x)


def linear_interp_f32(img, y, x):
    img = np.ascontiguousarray(img, dtype=np.float32)
    return lib().orc_linear_interp_f32(img, img.shape[1], img.shape[0], y, x)


def variance(x):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    return lib().orc_variance(x, x.size)


def ncc(x, y, w):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
    y = np.ascontiguousarray(y, dtype=np.float32).reshape(-1)
    w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
    return lib().orc_ncc(x, y, w, x.size)


def bilateral_weight(dcolor, dx, dy):
    return lib().orc_bilateral_weight(dcolor, dx, dy)


def homography_baked(K1inv, Q2, a2, K2, v):
    H = np.zeros(9, np.float32)
    lib().orc_homography_baked(_f64(K1inv, 9), _f64(Q2, 9), _f64(a2, 3),
                               _f64(K2, 9),
                               np.ascontiguousarray(v, dtype=np.float32), H)
    return H.reshape(3, 3)


def plane_from_depth_and_normal(x, y, K, depth, normal):
    out = np.zeros(3, np.float32)
    lib().orc_plane_from_depth_and_normal(
        x, y, _f64(K, 9), depth, np.ascontiguousarray(normal, dtype=np.float32),
        out)
    return out


def depth_of_plane_backprojection(x, y, K, plane):
    return lib().orc_depth_of_plane_backprojection(
        x, y, _f64(K, 9), np.ascontiguousarray(plane, dtype=np.float32))


def median5(img):
    img = np.ascontiguousarray(img, dtype=np.float32)
    out = np.zeros_like(img)
    lib().orc_median5(img, out, img.shape[1], img.shape[0])
    return out


def clean(views, same_depth_threshold=0.01, min_consistent_views=2):
    """views: list of (K, R, t, depth[H,W] float32); view 0 is cleaned."""
    n = len(views)
    Ks = np.ascontiguousarray(np.stack([np.asarray(v[0], np.float64).reshape(9) for v in views]))
    Rs = np.ascontiguousarray(np.stack([np.asarray(v[1], np.float64).reshape(9) for v in views]))
    ts = np.ascontiguousarray(np.stack([np.asarray(v[2], np.float64).reshape(3) for v in views]))
    depths = [np.ascontiguousarray(v[3], dtype=np.float32) for v in views]
    ptrs = (C.c_void_p * n)(*[d.ctypes.data for d in depths])
    ws = np.array([d.shape[1] for d in depths], np.int32)
    hs = np.array([d.shape[0] for d in depths], np.int32)
    out = np.zeros_like(depths[0])
    rc = lib().orc_clean(n, Ks.reshape(-1), Rs.reshape(-1), ts.reshape(-1),
                         ptrs, ws, hs, same_depth_threshold,
                         min_consistent_views, out)
    if rc != 0:
        raise ValueError("clean needs at least one view")
    return out


def prune(views, same_depth_threshold=0.01):
    """DepthmapPruner::Prune (depthmap.cc:574-636).  views: list of (K, R, t,
    depth (H, W) f32, plane (H, W, 3) f32, color (H, W, 3) u8, label (H, W)
    u8); view 0 is pruned.  Returns (points (N, 3), normals (N, 3), colors
    (N, 3) u8, labels (N,) u8)."""
    n = len(views)
    Ks = np.ascontiguousarray(np.stack([np.asarray(v[0], np.float64).reshape(9) for v in views]))
    Rs = np.ascontiguousarray(np.stack([np.asarray(v[1], np.float64).reshape(9) for v in views]))
    ts = np.ascontiguousarray(np.stack([np.asarray(v[2], np.float64).reshape(3) for v in views]))
    keep = []
    depths = [np.ascontiguousarray(v[3], np.float32) for v in views]
    planes = [np.ascontiguousarray(v[4], np.float32) for v in views]
    colors = [np.ascontiguousarray(v[5], np.uint8) for v in views]
    labels = [np.ascontiguousarray(v[6], np.uint8) for v in views]
    keep += depths + planes + colors + labels
    arr = lambda xs: (C.c_void_p * n)(*[x.ctypes.data for x in xs])  # noqa: E731
    ws = np.array([d.shape[1] for d in depths], np.int32)
    hs = np.array([d.shape[0] for d in depths], np.int32)
    cap = depths[0].size
    pts = np.zeros((cap, 3), np.float32)
    nrm = np.zeros((cap, 3), np.float32)
    col = np.zeros((cap, 3), np.uint8)
    lab = np.zeros(cap, np.uint8)
    cnt = lib().orc_prune(n, Ks.reshape(-1), Rs.reshape(-1), ts.reshape(-1), arr(depths),
                          arr(planes), arr(colors), arr(labels), ws, hs,
                          same_depth_threshold, cap, pts.reshape(-1), nrm.reshape(-1),
                          col.reshape(-1), lab)
    if cnt < 0:
        raise ValueError("prune needs at least one view")
    return pts[:cnt], nrm[:cnt], col[:cnt], lab[:cnt]


FUSER_DEFAULTS = dict(min_num_consistent=3, max_reproj_error=2.0, max_depth_error=0.01,
                      max_normal_error=10.0, border_margin=5, sor_knn=12,
                      sor_stddev_factor=2.5, behind_depth_factor=0.3)


def fuse(views, **params):
    """DepthmapFuser::Fuse (fuser.cc:113-480) with one thread.  views: list of
    (K, R, t, depth (H, W) f32, normal (H, W, 3) f32, color (H, W, 3) u8,
    mask (H, W) u8 or None, neighbor_ids, primary=True); params: the
    setters' arguments (FUSER_DEFAULTS).  Returns (points, normals, colors)
    and the fused count before outlier removal."""
    p = dict(FUSER_DEFAULTS)
    p.update(params)
    n = len(views)
    if n == 0:
        return (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32),
                np.zeros((0, 3), np.uint8)), 0
    Ks = np.stack([np.asarray(v[0], np.float64).reshape(9) for v in views])
    Rs = np.stack([np.asarray(v[1], np.float64).reshape(9) for v in views])
    ts = np.stack([np.asarray(v[2], np.float64).reshape(3) for v in views])
    depths = [np.ascontiguousarray(v[3], np.float32) for v in views]
    normals = [np.ascontiguousarray(v[4], np.float32) for v in views]
    colors = [np.ascontiguousarray(v[5], np.uint8) for v in views]
    masks = [None if v[6] is None or np.asarray(v[6]).size == 0
             else np.ascontiguousarray(v[6], np.uint8) for v in views]
    nbs = [np.asarray(v[7], np.int32).reshape(-1) for v in views]
    prim = np.array([int(bool(v[8])) if len(v) > 8 else 1 for v in views], np.int32)
    off = np.zeros(n + 1, np.int32)
    off[1:] = np.cumsum([len(x) for x in nbs])
    ids = np.ascontiguousarray(np.concatenate(nbs + [np.zeros(1, np.int32)]), np.int32)
    arr = lambda xs: (C.c_void_p * n)(*[0 if x is None else x.ctypes.data for x in xs])  # noqa: E731
    ws = np.array([d.shape[1] for d in depths], np.int32)
    hs = np.array([d.shape[0] for d in depths], np.int32)
    cap = int(sum(d.size for d, pr in zip(depths, prim) if pr)) + 1
    pts = np.zeros((cap, 3), np.float32)
    nrm = np.zeros((cap, 3), np.float32)
    col = np.zeros((cap, 3), np.uint8)
    nf = C.c_long(0)
    cnt = lib().orc_fuse(
        n, np.ascontiguousarray(Ks).reshape(-1), np.ascontiguousarray(Rs).reshape(-1),
        np.ascontiguousarray(ts).reshape(-1), arr(depths), arr(normals), arr(colors),
        arr(masks), ws, hs, off, ids, prim, int(p["min_num_consistent"]),
        float(p["max_reproj_error"]), float(p["max_depth_error"]), float(p["max_normal_error"]),
        int(p["border_margin"]), int(p["sor_knn"]), float(p["sor_stddev_factor"]),
        float(p["behind_depth_factor"]), cap, pts.reshape(-1), nrm.reshape(-1),
        col.reshape(-1), C.byref(nf))
    assert cnt >= 0
    return (pts[:cnt].copy(), nrm[:cnt].copy(), col[:cnt].copy()), int(nf.value)


def sor(points, normals, colors, knn=12, factor=2.5):
    """Statistical outlier removal of the fuser (fuser.cc:369-478) by brute
    force: returns the kept (points, normals, colors) and every point's mean
    distance to its knn nearest others."""
    pts = np.ascontiguousarray(points, np.float32).copy()
    nrm = np.ascontiguousarray(normals, np.float32).copy()
    col = np.ascontiguousarray(colors, np.uint8).copy()
    mk = np.zeros(max(len(pts), 1), np.float32)
    kept = lib().orc_sor(len(pts), pts.reshape(-1), nrm.reshape(-1), col.reshape(-1),
                         int(knn), float(factor), mk)
    return (pts[:kept], nrm[:kept], col[:kept]), mk[:len(pts)]


def inverse3f(m):
    """Eigen's fixed-size 3x3 float inverse (cofactor form) as the fuser
    uses it (Kf.inverse(), fuser.cc:162)."""
    a = np.ascontiguousarray(np.asarray(m, np.float32).reshape(9))
    out = np.zeros(9, np.float32)
    lib().orc_fuse_inverse3f(a, out)
    return out.reshape(3, 3)


def _csr(lists):
    off = np.zeros(len(lists) + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    idx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in lists] +
                                              [np.zeros(0, np.int32)]), np.int32)
    return off, idx


def covis(coords, observers, n_shots, fuse_knn=15, fuse_radius_factor=0.5):
    """BuildCovisibilityGraph (dense_helpers.cc:30-326) on plain arrays:
    coords (n, 3) float64, observers: per point a list of shot indices.
    Returns dict(point_sp, sp_coord, sp_members (list of arrays), sp_vis
    (list of arrays), edges (E, 2), weights (E,), scale)."""
    coords = np.ascontiguousarray(np.asarray(coords, np.float64).reshape(-1, 3))
    n = len(coords)
    off, shots = _csr(observers)
    point_sp = np.zeros(max(n, 1), np.int32)
    sp_coord = np.zeros((max(n, 1), 3))
    moff = np.zeros(n + 2, np.int64)
    members = np.zeros(max(n, 1), np.int32)
    voff = np.zeros(n + 2, np.int64)
    cap_v = max(int(off[-1]), 1)
    vis = np.zeros(cap_v, np.int32)
    cap_e = max(n_shots * (n_shots - 1) // 2, 1)
    edges = np.zeros((cap_e, 2), np.int32)
    weights = np.zeros(cap_e, np.int32)
    ne = C.c_long(0)
    scale = C.c_float(0)
    nsp = lib().orc_covis(n, coords.reshape(-1), off, shots, int(n_shots), int(fuse_knn),
                          float(fuse_radius_factor), point_sp, sp_coord.reshape(-1), moff,
                          members, voff, vis, cap_v, edges.reshape(-1), weights, cap_e,
                          C.byref(ne), C.byref(scale))
    assert nsp >= 0
    return dict(point_sp=point_sp[:n].copy(), sp_coord=sp_coord[:nsp].copy(),
                sp_members=[members[moff[s]:moff[s + 1]].copy() for s in range(nsp)],
                sp_vis=[vis[voff[s]:voff[s + 1]].copy() for s in range(nsp)],
                edges=edges[:ne.value].copy(), weights=weights[:ne.value].copy(),
                scale=scale.value)


def select_neighbors(origins, sp_coord, sp_vis, sp_ntracks, num_neighbors, min_point_best=20,
                     min_point_all=40, theta_min_deg=3.0, theta_max_deg=60.0):
    """SelectNeighbors (dense_helpers.cc:331-512) on plain arrays: origins
    (N, 3) camera centres, super-point centroids, visibility lists and track
    counts.  Returns (best, all): per shot index a list of shot indices."""
    origins = np.ascontiguousarray(np.asarray(origins, np.float64).reshape(-1, 3))
    N = len(origins)
    sp_coord = np.ascontiguousarray(np.asarray(sp_coord, np.float64).reshape(-1, 3))
    voff, vis = _csr(sp_vis)
    nt = np.ascontiguousarray(np.asarray(sp_ntracks, np.int32))
    boff = np.zeros(N + 1, np.int64)
    aoff = np.zeros(N + 1, np.int64)
    cap = max(N * N, 1)
    best = np.zeros(cap, np.int32)
    allv = np.zeros(cap, np.int32)
    r = lib().orc_select_neighbors(N, origins.reshape(-1), len(sp_coord), sp_coord.reshape(-1),
                                   voff, vis, nt, int(num_neighbors), int(min_point_best),
                                   int(min_point_all), float(theta_min_deg),
                                   float(theta_max_deg), boff, best, cap, aoff, allv, cap)
    assert r >= 0
    return ([best[boff[i]:boff[i + 1]].tolist() for i in range(N)],
            [allv[aoff[i]:aoff[i + 1]].tolist() for i in range(N)])


class OracleEstimator:
    """Mirror of pydense.DepthmapEstimator (pydense.pyi:43-60) over the oracle."""

    def __init__(self):
        self._L = lib()
        self._h = self._L.orc_create()
        self._shape = None
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.orc_destroy(self._h)
            self._h = None

    def add_view(self, K, R, t, image, mask):
        image = np.ascontiguousarray(image, dtype=np.uint8)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        if self._shape is None:
            self._shape = image.shape
        self._L.orc_add_view(self._h, _f64(K, 9), _f64(R, 9), _f64(t, 3), image,
                             mask, image.shape[1], image.shape[0])

    def set_depth_range(self, mn, mx, n):
        self._L.orc_set_depth_range(self._h, mn, mx, n)

    def set_patchmatch_iterations(self, n):
        self._L.orc_set_patchmatch_iterations(self._h, n)

    def set_patch_size(self, s):
        self._L.orc_set_patch_size(self._h, s)

    def set_min_patch_sd(self, sd):
        self._L.orc_set_min_patch_sd(self._h, sd)

    def set_seed(self, seed):
        self._L.orc_set_seed(self._h, seed)

    @property
    def last_scorings(self):
        return self._L.orc_last_scorings(self._h)

    def _compute(self, mode, stop_after=-1):
        h, w = self._shape
        depth = np.zeros((h, w), np.float32)
        plane = np.zeros((h, w, 3), np.float32)
        score = np.zeros((h, w), np.float32)
        nghbr = np.zeros((h, w), np.int32)
        self._L.orc_compute(self._h, mode, depth, plane.reshape(-1), score, nghbr,
                            stop_after)
        return [depth, plane, score, nghbr]

    def compute_brute_force(self):
        return self._compute(0)

    compute_plane_sweep = compute_brute_force

    def compute_patch_match(self, stop_after_iterations=-1):
        return self._compute(1, stop_after_iterations)

    def compute_patch_match_sample(self):
        return self._compute(2)

    def score(self, i, j, plane, other):
        return self._L.orc_score(self._h, i, j,
                                 np.ascontiguousarray(plane, dtype=np.float32),
                                 other)

    def score_f64(self, i, j, plane, other):
        return self._L.orc_score_f64(self._h, i, j,
                                     np.ascontiguousarray(plane, dtype=np.float64),
                                     other)


# ---------------------------------------------------------------------------
# GPUDepthmapCleaner (opensfm/src/lib/dense/src/cleaner.cc:56-179 and its
# kernel acmmp_clean_depthmap, opencl_kernels.h:2155-2273), restated in
# numpy float32 with the operation order written there.

def resize_linear_f32(depth, w, h):
    """cv::resize(INTER_LINEAR) of a float32 map to (h, w): source
    coordinate (x + 0.5) * sw / w - 0.5, floor and fraction, clamped at
    the borders (cleaner.cc:114-117)."""
    src = np.asarray(depth, np.float32)
    sh, sw = src.shape
    if (sh, sw) == (h, w):
        return src.copy()

    def axis(n_dst, n_src):
        scale = np.float64(n_src) / n_dst
        f = ((np.arange(n_dst) + 0.5) * scale - 0.5).astype(np.float32)
        i = np.floor(f).astype(np.int64)
        a = (f - i.astype(np.float32)).astype(np.float32)
        low = i < 0
        a[low], i[low] = 0, 0
        high = i >= n_src - 1
        a[high], i[high] = 0, n_src - 1
        return i, np.minimum(i + 1, n_src - 1), a

    x0, x1, ax = axis(w, sw)
    y0, y1, ay = axis(h, sh)
    one = np.float32(1)
    r0 = src[y0][:, x0] * (one - ax) + src[y0][:, x1] * ax
    r1 = src[y1][:, x0] * (one - ax) + src[y1][:, x1] * ax
    return (r0 * (one - ay)[:, None] + r1 * ay[:, None]).astype(np.float32)


def gpu_clean(views, ref_idx, neighbor_ids, same_depth_threshold=0.01,
              min_consistent_views=3, carving_threshold=0.2, max_carved_views=1):
    """views: list of (K, R, t, depth).  Returns the cleaned depth of
    views[ref_idx] checked against views[neighbor_ids] (at most 16)."""
    f = np.float32
    Kr, Rr, tr, dr = views[ref_idx]
    dr = np.asarray(dr, f)
    h, w = dr.shape
    Kr, Rr, tr = (np.asarray(a, np.float64).astype(f) for a in (Kr, Rr, tr))
    nb = list(neighbor_ids)[:16]
    num_views = 1 + len(neighbor_ids)
    eff_min = min(min_consistent_views, num_views)
    carving = f(max(0.0, carving_threshold))
    thr = f(same_depth_threshold)
    yy, xx = np.meshgrid(np.arange(h, dtype=f), np.arange(w, dtype=f), indexing="ij")
    d = dr
    cam = [(xx - Kr[0, 2]) / Kr[0, 0] * d, (yy - Kr[1, 2]) / Kr[1, 1] * d, d]
    pc = [cam[k] - tr[k] for k in range(3)]
    wp = [Rr[0, k] * pc[0] + Rr[1, k] * pc[1] + Rr[2, k] * pc[2] for k in range(3)]
    consistent = np.ones((h, w), np.int32)
    carved = np.zeros((h, w), np.int32)
    for v in nb:
        Ks, Rs, ts, ds = views[v]
        Ks, Rs, ts = (np.asarray(a, np.float64).astype(f) for a in (Ks, Rs, ts))
        src = resize_linear_f32(ds, w, h)
        s = [Rs[k, 0] * wp[0] + Rs[k, 1] * wp[1] + Rs[k, 2] * wp[2] + ts[k] for k in range(3)]
        with np.errstate(divide="ignore", invalid="ignore"):
            su = Ks[0, 0] * s[0] / s[2] + Ks[0, 2]
            sv = Ks[1, 1] * s[1] / s[2] + Ks[1, 2]
        ok = (d > 0) & (s[2] > f(1e-6))
        ok &= ~((su < 0) | (sv < 0) | (su >= f(w - 1)) | (sv >= f(h - 1)))
        ok &= np.isfinite(su) & np.isfinite(sv)
        sus, svs = np.where(ok, su, 0).astype(f), np.where(ok, sv, 0).astype(f)
        x0 = np.floor(sus).astype(np.int64)
        y0 = np.floor(svs).astype(np.int64)
        a = sus - x0.astype(f)
        b = svs - y0.astype(f)
        x1, y1 = np.minimum(x0 + 1, w - 1), np.minimum(y0 + 1, h - 1)
        one = f(1)
        r0 = src[y0, x0] * (one - a) + src[y0, x1] * a
        r1 = src[y1, x0] * (one - a) + src[y1, x1] * a
        sd = (r0 * (one - b) + r1 * b).astype(f)
        ok &= sd > 0
        carve = ok & (sd > s[2] * (one + carving))
        carved += carve
        ok &= ~carve
        ok &= np.abs(sd - s[2]) < s[2] * thr
        sc = [(sus - Ks[0, 2]) / Ks[0, 0] * sd, (svs - Ks[1, 2]) / Ks[1, 1] * sd, sd]
        spc = [sc[k] - ts[k] for k in range(3)]
        w2 = [Rs[0, k] * spc[0] + Rs[1, k] * spc[1] + Rs[2, k] * spc[2] for k in range(3)]
        r = [Rr[k, 0] * w2[0] + Rr[k, 1] * w2[1] + Rr[k, 2] * w2[2] + tr[k] for k in range(3)]
        with np.errstate(divide="ignore", invalid="ignore"):
            ru = Kr[0, 0] * r[0] / r[2] + Kr[0, 2]
            rv = Kr[1, 1] * r[1] / r[2] + Kr[1, 2]
        ok &= r[2] > f(1e-6)
        ok &= ~(np.abs(ru - xx) + np.abs(rv - yy) > f(2))
        ok &= np.abs(r[2] - d) < d * thr
        consistent += ok
    keep = (d > 0) & (consistent >= eff_min) & (carved < max_carved_views)
    return np.where(keep, d, f(0)).astype(f)


# ---------------------------------------------------------------------------
# The cluster estimator's readback (opensfm/src/lib/dense/src/estimator.cc:
# 960-993 and its confidence, :1022-1040), restated in numpy float32 over a
# PatchMatch result (depth, plane, score) instead of the OpenCL plane/cost
# buffers: cost = 1 - score.

def cluster_readback(depth, plane, score, depth_min, depth_max):
    f = np.float32
    depth = np.asarray(depth, f)
    plane = np.asarray(plane, f)
    score = np.asarray(score, f)
    finite = np.isfinite(plane).all(axis=-1)
    d = np.where(np.isfinite(depth), depth, f(0))
    keep = finite & (d > 0) & (d >= f(depth_min)) & (d <= f(depth_max))
    out_depth = np.where(keep, d, f(0)).astype(f)
    p64 = plane.astype(np.float64)
    nrm = np.sqrt((p64 * p64).sum(-1)).astype(f)
    unit = keep & (nrm > f(1e-6))
    normal = np.where(unit[..., None], plane / np.where(unit, nrm, f(1))[..., None], f(0)).astype(f)
    with np.errstate(invalid="ignore"):
        c = (f(1) - score).astype(f)
    cost = np.where(finite & np.isfinite(c), c, f(0)).astype(f)
    with np.errstate(over="ignore"):
        conf = np.where(out_depth > 0, f(1) / (f(1) + np.exp((cost - f(0.5)) * f(5))), f(0))
    return out_depth, normal, cost, conf.astype(f)
