"""``pydense``-compatible surface of the MI355X depthmap engine.

Mirrors ``pydense.DepthmapEstimator``, ``DepthmapCleaner``,
``DepthmapPruner`` and ``GPUDepthmapCleaner``
(opensfm/src/lib/dense/pydense.pyi:28-77, 129-148; C++ classes in
opensfm/src/lib/dense/depthmap.h:57-153 and cleaner.h:16-63): same method
names, argument order,
defaults and return layout, so code written against the reference reads the
same.  Every compute runs on the GPU through the C ABI in
include/opensfm_amd.h; the calls release the GIL (ctypes does) like the
reference's bindings do (depthmap_bind.h:57-60).

Additive API: ``compute_plane_sweep`` (alias of ``compute_brute_force``),
``set_seed``, ``set_far_probes``, ``set_refinement_steps``,
``set_exploration``, ``set_consistency``, ``set_precision``, ``set_post_process``,
``score_hypotheses``, ``run``/``download``, ``stats`` and the ``device``
constructor argument.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi


def _mat(a, shape, name):
    a = np.asarray(a, dtype=np.float64)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"{name} must have {int(np.prod(shape))} elements, got shape {a.shape}")
    return np.ascontiguousarray(a.reshape(shape))


class DepthmapEstimator:
    """PatchMatch / plane-sweep depthmap estimator (depthmap.h:57-111)."""

    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self._lib.pm_estimator_create(int(device), C.byref(h)))
        self._h = h
        self._shape = None
        self._init_map = None  # the device tensor the init map lives in

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.pm_estimator_destroy(h)
            self._h = None

    # -- DepthmapEstimator::AddView (depthmap.cc:146-162)
    def add_view(self, K, R, t, image, mask) -> None:
        K = _mat(K, (3, 3), "K")
        R = _mat(R, (3, 3), "R")
        t = _mat(t, (3,), "t")
        image = np.asarray(image)
        mask = np.asarray(mask)
        if image.ndim != 2 or image.dtype != np.uint8:
            raise ValueError("image must be a 2-D uint8 array")
        if mask.shape != image.shape or mask.dtype != np.uint8:
            raise ValueError("mask must be a uint8 array of the image's shape")
        image = np.ascontiguousarray(image)
        mask = np.ascontiguousarray(mask)
        h, w = image.shape
        _capi.check(self._lib.pm_estimator_add_view(
            self._h, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double),
            _capi.ptr(t, C.c_double), _capi.ptr(image, C.c_uint8),
            _capi.ptr(mask, C.c_uint8), w, h, w))
        if self._shape is None:
            self._shape = (h, w)

    def set_depth_range(self, min_depth: float, max_depth: float,
                        num_depth_planes: int) -> None:
        _capi.check(self._lib.pm_estimator_set_depth_range(
            self._h, float(min_depth), float(max_depth), int(num_depth_planes)))

    def set_patchmatch_iterations(self, n: int) -> None:
        _capi.check(self._lib.pm_estimator_set_patchmatch_iterations(self._h, int(n)))

    def set_patch_size(self, size: int) -> None:
        _capi.check(self._lib.pm_estimator_set_patch_size(self._h, int(size)))

    def set_min_patch_sd(self, sd: float) -> None:
        _capi.check(self._lib.pm_estimator_set_min_patch_sd(self._h, float(sd)))

    def set_seed(self, seed: int) -> None:
        _capi.check(self._lib.pm_estimator_set_seed(self._h, int(seed) & (2 ** 64 - 1)))

    def set_pyramid_consistency(self, enable: bool) -> None:
        """The median-consistent re-selection at the pyramid's coarser
        levels (default on, as set_consistency; off: at the full level
        only)."""
        _capi.check(self._lib.pm_estimator_set_pyramid_consistency(self._h, int(bool(enable))))

    def set_pyramid(self, half_iterations, levels: int = 2) -> None:
        """Plain PatchMatch seeded by half_iterations iterations at half
        resolution (0: off; not in the reference's legacy estimator); with
        levels > 2 each coarser level is seeded by the next, down to
        1 / 2^(levels - 1) of the full size.  A sequence gives each coarser
        level its own iterations, coarsest first (levels = its length + 1)."""
        if isinstance(half_iterations, (int, np.integer)):
            _capi.check(self._lib.pm_estimator_set_pyramid_levels(self._h, int(levels)))
            _capi.check(self._lib.pm_estimator_set_pyramid(self._h, int(half_iterations)))
            return
        it = np.ascontiguousarray(half_iterations, np.int32).ravel()
        _capi.check(self._lib.pm_estimator_set_pyramid_schedule(
            self._h, it.ctypes.data_as(C.POINTER(C.c_int32)), int(it.size)))

    def set_lanes_per_pixel(self, lanes: int) -> None:
        """Lanes per pixel of the red/black launches: 0 (default) 8 for
        launches with few pixels (FP32 samples), else 1; 1 or 8 forces it.
        The maps are the same for every setting."""
        _capi.check(self._lib.pm_estimator_set_lanes_per_pixel(self._h, int(lanes)))

    def set_far_probes(self, n: int) -> None:
        _capi.check(self._lib.pm_estimator_set_far_probes(self._h, int(n)))

    def set_refinement_steps(self, n: int) -> None:
        _capi.check(self._lib.pm_estimator_set_refinement_steps(self._h, int(n)))

    def set_exploration(self, n: int) -> None:
        _capi.check(self._lib.pm_estimator_set_exploration(self._h, int(n)))

    def set_precision(self, precision) -> None:
        """PatchMatch sample precision: "fp32" (default), "fp16" or "bf16"."""
        _capi.check(self._lib.pm_estimator_set_precision(self._h, _capi.precision_code(precision)))

    def set_consistency(self, enable: bool) -> None:
        """Median-consistent re-selection between iterations (default on;
        off = the reference's candidate search)."""
        _capi.check(self._lib.pm_estimator_set_consistency(self._h, int(bool(enable))))

    def set_post_process(self, enable: bool) -> None:
        _capi.check(self._lib.pm_estimator_set_post_process(self._h, int(bool(enable))))

    def _alloc(self):
        if self._shape is None:
            raise ValueError("need a reference view and at least one neighbour view (got 0)")
        h, w = self._shape
        return (np.zeros((h, w), np.float32), np.zeros((h, w, 3), np.float32),
                np.zeros((h, w), np.float32), np.zeros((h, w), np.int32))

    def _compute(self, mode):
        depth, plane, score, nghbr = self._alloc()
        _capi.check(self._lib.pm_estimator_compute(
            self._h, mode, _capi.ptr(depth, C.c_float), _capi.ptr(plane, C.c_float),
            _capi.ptr(score, C.c_float), _capi.ptr(nghbr, C.c_int32)))
        return [depth, plane, score, nghbr]

    # -- ComputeBruteForce / ComputePatchMatch / ComputePatchMatchSample
    def compute_brute_force(self) -> list:
        return self._compute(_capi.PM_MODE_PLANE_SWEEP)

    compute_plane_sweep = compute_brute_force

    def compute_patch_match(self) -> list:
        return self._compute(_capi.PM_MODE_PATCHMATCH)

    def compute_patch_match_sample(self) -> list:
        return self._compute(_capi.PM_MODE_PATCHMATCH_SAMPLE)

    # -- device-resident variants (timing)
    def run(self, mode: int) -> None:
        _capi.check(self._lib.pm_estimator_run(self._h, int(mode)))

    def download(self) -> list:
        depth, plane, score, nghbr = self._alloc()
        _capi.check(self._lib.pm_estimator_download(
            self._h, _capi.ptr(depth, C.c_float), _capi.ptr(plane, C.c_float),
            _capi.ptr(score, C.c_float), _capi.ptr(nghbr, C.c_int32)))
        return [depth, plane, score, nghbr]

    def score_hypotheses(self, ij, planes, views, fast: bool = False) -> np.ndarray:
        """ComputePlaneImageScore (depthmap.cc:441-478) of fixed hypotheses:
        ij (n, 2) = (row, col), planes (n, 3), views (n,) in 1..N-1.  With
        fast=True, the PatchMatch scorer in the set sample precision."""
        ij = np.ascontiguousarray(np.asarray(ij, np.int32).reshape(-1, 2))
        n = ij.shape[0]
        planes = np.ascontiguousarray(np.asarray(planes, np.float32).reshape(n, 3))
        views = np.ascontiguousarray(np.asarray(views, np.int32).reshape(n))
        out = np.zeros(n, np.float32)
        fn = (self._lib.pm_estimator_score_hypotheses_fast if fast
              else self._lib.pm_estimator_score_hypotheses)
        _capi.check(fn(self._h, n, _capi.ptr(ij, C.c_int32), _capi.ptr(planes, C.c_float),
                       _capi.ptr(views, C.c_int32), _capi.ptr(out, C.c_float)))
        return out

    def initializer_input(self, out, out_w: int, out_h: int, scale: int = 4,
                          nplanes: int = 32) -> None:
        """Writes the learned initialiser's input (BF16 NHWC, 40 channels,
        out_h x out_w cells) into the device tensor `out`."""
        ptr = _capi._device_tensor(out, out_w * out_h * 40, "initializer input", "bfloat16")
        _capi.check(self._lib.pm_estimator_initializer_input(
            self._h, int(scale), int(nplanes), C.c_void_p(ptr), int(out_w), int(out_h)))

    def set_init_map(self, init_map, w: int = 0, h: int = 0, stride: int = 8,
                     scale: int = 4) -> None:
        """Seeds PatchMatch's initialisation from a coarse (t, sx, sy, c) map:
        a float32 device tensor of h x w cells of `stride` values (None:
        random initialisation).  The estimator keeps a reference to the
        tensor until the map is replaced or cleared."""
        ptr = None
        if init_map is not None:
            ptr = _capi._device_tensor(init_map, w * h * stride, "init map", "float32")
        _capi.check(self._lib.pm_estimator_set_init_map(
            self._h, C.c_void_p(ptr) if ptr else None, int(w), int(h), int(stride), int(scale)))
        self._init_map = init_map

    def set_init_search(self, n: int) -> None:
        """Depth hypotheses tried around the init map's depth (1..15)."""
        _capi.check(self._lib.pm_estimator_set_init_search(self._h, int(n)))

    def set_init_spread(self, sigma_log_depth: float, sigma_slope: float) -> None:
        """Per-pixel spread of the seeded planes around the init map's."""
        _capi.check(self._lib.pm_estimator_set_init_spread(self._h, float(sigma_log_depth),
                                                           float(sigma_slope)))

    def stats(self) -> dict:
        s = _capi.pm_stats()
        _capi.check(self._lib.pm_estimator_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


class DepthmapCleaner:
    """Cross-view depth consistency check (depthmap.h:113-129)."""

    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self._lib.pm_cleaner_create(int(device), C.byref(h)))
        self._h = h
        self._shape = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.pm_cleaner_destroy(h)
            self._h = None

    def set_same_depth_threshold(self, t: float) -> None:
        _capi.check(self._lib.pm_cleaner_set_same_depth_threshold(self._h, float(t)))

    def set_min_consistent_views(self, n: int) -> None:
        _capi.check(self._lib.pm_cleaner_set_min_consistent_views(self._h, int(n)))

    def add_view(self, K, R, t, depth) -> None:
        K = _mat(K, (3, 3), "K")
        R = _mat(R, (3, 3), "R")
        t = _mat(t, (3,), "t")
        depth = np.asarray(depth)
        if depth.ndim != 2:
            raise ValueError("depth must be a 2-D array")
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        h, w = depth.shape
        _capi.check(self._lib.pm_cleaner_add_view(
            self._h, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double),
            _capi.ptr(t, C.c_double), _capi.ptr(depth, C.c_float), w, h))
        if self._shape is None:
            self._shape = (h, w)

    def clean(self) -> np.ndarray:
        if self._shape is None:
            raise ValueError("no views added")
        out = np.zeros(self._shape, np.float32)
        _capi.check(self._lib.pm_cleaner_clean(self._h, _capi.ptr(out, C.c_float)))
        return out

    def last_ms(self) -> float:
        ms = C.c_float(0)
        _capi.check(self._lib.pm_cleaner_last_ms(self._h, C.byref(ms)))
        return ms.value


class GPUDepthmapCleaner:
    """The live pipeline's cleaner with space carving and a backward
    reprojection check (GPUDepthmapCleaner, cleaner.h:16-63; Python surface
    pydense.pyi:129-148)."""

    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        self._device = int(device)
        self._create()
        self._shapes = []

    def _create(self):
        h = C.c_void_p()
        _capi.check(self._lib.pm_gpu_cleaner_create(self._device, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.pm_gpu_cleaner_destroy(h)
            self._h = None

    def set_same_depth_threshold(self, t: float) -> None:
        _capi.check(self._lib.pm_gpu_cleaner_set_same_depth_threshold(self._h, float(t)))

    def set_min_consistent_views(self, n: int) -> None:
        _capi.check(self._lib.pm_gpu_cleaner_set_min_consistent_views(self._h, int(n)))

    def set_carving_threshold(self, t: float) -> None:
        _capi.check(self._lib.pm_gpu_cleaner_set_carving_threshold(self._h, float(t)))

    def set_max_carved_views(self, n: int) -> None:
        _capi.check(self._lib.pm_gpu_cleaner_set_max_carved_views(self._h, int(n)))

    def set_device(self, idx: int) -> None:
        """Selects the device; only before the first add_view (the views
        live in that device's memory)."""
        if self._shapes:
            raise RuntimeError("set_device must come before add_view")
        self._lib.pm_gpu_cleaner_destroy(self._h)
        self._device = int(idx)
        self._create()

    def add_view(self, K, R, t, depth) -> None:
        K = _mat(K, (3, 3), "K")
        R = _mat(R, (3, 3), "R")
        t = _mat(t, (3,), "t")
        depth = np.asarray(depth)
        if depth.ndim != 2:
            raise ValueError("depth must be a 2-D array")
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        h, w = depth.shape
        idx = C.c_int32(-1)
        _capi.check(self._lib.pm_gpu_cleaner_add_view(
            self._h, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double),
            _capi.ptr(t, C.c_double), _capi.ptr(depth, C.c_float), w, h, C.byref(idx)))
        self._shapes.append((h, w))

    def clean(self, ref_idx: int, neighbor_ids) -> np.ndarray:
        nb = np.ascontiguousarray(np.asarray(neighbor_ids, np.int32).reshape(-1))
        if not (0 <= int(ref_idx) < len(self._shapes)):
            shape = (1, 1)
        else:
            shape = self._shapes[int(ref_idx)]
        out = np.zeros(shape, np.float32)
        _capi.check(self._lib.pm_gpu_cleaner_clean(
            self._h, int(ref_idx), _capi.ptr(nb, C.c_int32), int(nb.size),
            _capi.ptr(out, C.c_float)))
        return out

    def clear(self) -> None:
        _capi.check(self._lib.pm_gpu_cleaner_clear(self._h))
        self._shapes = []

    def last_ms(self) -> float:
        ms = C.c_float(0)
        _capi.check(self._lib.pm_gpu_cleaner_last_ms(self._h, C.byref(ms)))
        return ms.value

    @staticmethod
    def is_available() -> bool:
        try:
            return _capi.device_count() > 0
        except RuntimeError:
            return False



class DepthmapPruner:
    """Merges view 0's depthmap into a point cloud, dropping the points that
    another view samples more finely (DepthmapPruner, depthmap.h:131-153;
    Python surface pydense.pyi:63-77)."""

    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self._lib.pm_pruner_create(int(device), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.pm_pruner_destroy(h)
            self._h = None

    def set_same_depth_threshold(self, t: float) -> None:
        _capi.check(self._lib.pm_pruner_set_same_depth_threshold(self._h, float(t)))

    def add_view(self, K, R, t, depth, plane, color, label) -> None:
        """depth (H, W), plane (H, W, 3), color (H, W, 3) uint8, label (H, W)
        uint8, as DepthmapPruner::AddView (depthmap.cc:563-575)."""
        K = _mat(K, (3, 3), "K")
        R = _mat(R, (3, 3), "R")
        t = _mat(t, (3,), "t")
        depth = np.asarray(depth)
        if depth.ndim != 2:
            raise ValueError("depth must be a 2-D array")
        h, w = depth.shape
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        plane = np.ascontiguousarray(np.asarray(plane, dtype=np.float32))
        color = np.ascontiguousarray(np.asarray(color, dtype=np.uint8))
        label = np.ascontiguousarray(np.asarray(label, dtype=np.uint8))
        if plane.shape != (h, w, 3) or color.shape != (h, w, 3) or label.shape != (h, w):
            raise ValueError("plane and color must be (H, W, 3) and label (H, W) like depth")
        _capi.check(self._lib.pm_pruner_add_view(
            self._h, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double),
            _capi.ptr(t, C.c_double), _capi.ptr(depth, C.c_float),
            _capi.ptr(plane, C.c_float), _capi.ptr(color, C.c_uint8),
            _capi.ptr(label, C.c_uint8), w, h))

    def prune(self) -> list:
        """[points (N, 3) f32, normals (N, 3) f32, colors (N, 3) u8,
        labels (N,) u8], kept pixels in row-major order."""
        n = C.c_int64(0)
        _capi.check(self._lib.pm_pruner_prune(self._h, C.byref(n)))
        m = int(n.value)
        points = np.zeros((m, 3), np.float32)
        normals = np.zeros((m, 3), np.float32)
        colors = np.zeros((m, 3), np.uint8)
        labels = np.zeros(m, np.uint8)
        _capi.check(self._lib.pm_pruner_download(
            self._h, _capi.ptr(points, C.c_float), _capi.ptr(normals, C.c_float),
            _capi.ptr(colors, C.c_uint8), _capi.ptr(labels, C.c_uint8)))
        return [points, normals, colors, labels]

    def last_ms(self) -> float:
        ms = C.c_float(0)
        _capi.check(self._lib.pm_pruner_last_ms(self._h, C.byref(ms)))
        return ms.value


class DepthmapFuser:
    """Fuses depthmaps into a point cloud by multi-view consistency with
    outlier removal (DepthmapFuser, fuser.h:28-75; Python surface
    pydense.pyi:78-100).  Views fuse in index order and pixels in row-major
    order -- the reference's result with one thread; ``set_num_threads`` is
    kept for the surface and does not change the result."""

    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self._lib.pm_fuser_create(int(device), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.pm_fuser_destroy(h)
            self._h = None

    def set_min_num_consistent(self, n: int) -> None:
        _capi.check(self._lib.pm_fuser_set_min_num_consistent(self._h, int(n)))

    def set_max_reproj_error(self, px: float) -> None:
        _capi.check(self._lib.pm_fuser_set_max_reproj_error(self._h, float(px)))

    def set_max_depth_error(self, ratio: float) -> None:
        _capi.check(self._lib.pm_fuser_set_max_depth_error(self._h, float(ratio)))

    def set_max_normal_error(self, degrees: float) -> None:
        _capi.check(self._lib.pm_fuser_set_max_normal_error(self._h, float(degrees)))

    def set_border_margin(self, px: int) -> None:
        _capi.check(self._lib.pm_fuser_set_border_margin(self._h, int(px)))

    def set_num_threads(self, n: int) -> None:
        _capi.check(self._lib.pm_fuser_set_num_threads(self._h, int(n)))

    def set_sor_params(self, knn: int, stddev_factor: float) -> None:
        _capi.check(self._lib.pm_fuser_set_sor_params(self._h, int(knn), float(stddev_factor)))

    def set_behind_depth_factor(self, f: float) -> None:
        _capi.check(self._lib.pm_fuser_set_behind_depth_factor(self._h, float(f)))

    def add_view(self, K, R, t, depth, normal, color, mask, neighbor_ids,
                 primary: bool = True) -> None:
        """As DepthmapFuserWrapper::AddView (depthmap_bind.h:244-264): depth
        (H, W), normal and color (H, W, 3); mask (H, W) uint8, or empty /
        None for no mask."""
        K = _mat(K, (3, 3), "K")
        R = _mat(R, (3, 3), "R")
        t = _mat(t, (3,), "t")
        depth = np.asarray(depth)
        if depth.ndim != 2:
            raise ValueError("depth must be a 2-D array")
        h, w = depth.shape
        depth = np.ascontiguousarray(depth, dtype=np.float32)
        normal = np.asarray(normal)
        if normal.ndim != 3 or normal.shape[0] != h or normal.shape[1] != w:
            raise ValueError("depth and normal must have matching shapes.")
        color = np.asarray(color)
        if color.ndim != 3 or color.shape[0] != h or color.shape[1] != w:
            raise ValueError("depth and color must have matching shapes.")
        if normal.shape[2] != 3 or color.shape[2] != 3:
            raise ValueError("normal and color must have 3 channels")
        normal = np.ascontiguousarray(normal, dtype=np.float32)
        color = np.ascontiguousarray(color, dtype=np.uint8)
        mptr = None
        if mask is not None and np.asarray(mask).size > 0:
            mask = np.ascontiguousarray(np.asarray(mask), dtype=np.uint8)
            if mask.shape != (h, w):
                raise ValueError("mask must be empty or match the depth map's shape")
            mptr = _capi.ptr(mask, C.c_uint8)
        nb = np.ascontiguousarray(np.asarray(neighbor_ids, np.int32).reshape(-1))
        _capi.check(self._lib.pm_fuser_add_view(
            self._h, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double),
            _capi.ptr(t, C.c_double), _capi.ptr(depth, C.c_float),
            _capi.ptr(normal, C.c_float), _capi.ptr(color, C.c_uint8), mptr, w, h,
            _capi.ptr(nb, C.c_int32), int(nb.size), int(bool(primary))))

    def fuse(self) -> list:
        """[points (N, 3) f32, normals (N, 3) f32, colors (N, 3) u8]."""
        n = C.c_int64(0)
        _capi.check(self._lib.pm_fuser_fuse(self._h, C.byref(n)))
        m = int(n.value)
        points = np.zeros((m, 3), np.float32)
        normals = np.zeros((m, 3), np.float32)
        colors = np.zeros((m, 3), np.uint8)
        _capi.check(self._lib.pm_fuser_download(
            self._h, _capi.ptr(points, C.c_float), _capi.ptr(normals, C.c_float),
            _capi.ptr(colors, C.c_uint8)))
        return [points, normals, colors]

    def stats(self) -> dict:
        nf, rounds = C.c_int64(0), C.c_int64(0)
        fms, sms = C.c_float(0), C.c_float(0)
        _capi.check(self._lib.pm_fuser_stats(self._h, C.byref(nf), C.byref(rounds),
                                             C.byref(fms), C.byref(sms)))
        return {"fused_before_sor": nf.value, "rounds": rounds.value,
                "fuse_ms": fms.value, "sor_ms": sms.value}
