"""Batched depthmap estimation: many reference views in one launch.

``DepthmapBatch`` holds a pool of views (camera + image, uploaded once) and
a list of jobs, each a reference view with its neighbour views and depth
range.  Job k computes what ``DepthmapEstimator`` computes with
``add_view(ref)``, ``add_view(n)`` for each neighbour, the same settings and
``set_seed(seed + k)`` -- bit for bit -- but every job runs in the same
persistent launches, and a run is replayed from one captured hipGraph.  The
reference runs one ``DepthmapEstimator`` per reference view in a thread
pool (opensfm/dense.py:187-241); the batch replaces that loop, then cleans
every depthmap against its neighbours' (DepthmapCleaner::Clean,
opensfm/src/lib/dense/src/depthmap.cc:524-555) on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .pydense import _mat


class DepthmapBatch:
    def __init__(self, device: int = 0):
        self._lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self._lib.pm_batch_create(int(device), C.byref(h)))
        self._h = h
        self._shapes = []   # per pool view (h, w)
        self._jobs = []     # per job: ref view index
        self._init_maps = {}  # job -> the device tensor its init map lives in

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.pm_batch_destroy(h)
            self._h = None

    def add_view(self, K, R, t, image, mask=None) -> int:
        """Adds a camera and its image to the pool; returns its index."""
        K = _mat(K, (3, 3), "K")
        R = _mat(R, (3, 3), "R")
        t = _mat(t, (3,), "t")
        image = np.asarray(image)
        if image.ndim != 2 or image.dtype != np.uint8:
            raise ValueError("image must be a 2-D uint8 array")
        mask = np.full_like(image, 255) if mask is None else np.asarray(mask)
        if mask.shape != image.shape or mask.dtype != np.uint8:
            raise ValueError("mask must be a uint8 array of the image's shape")
        image = np.ascontiguousarray(image)
        mask = np.ascontiguousarray(mask)
        h, w = image.shape
        vid = C.c_int32(-1)
        _capi.check(self._lib.pm_batch_add_view(
            self._h, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double),
            _capi.ptr(t, C.c_double), _capi.ptr(image, C.c_uint8),
            _capi.ptr(mask, C.c_uint8), w, h, w, C.byref(vid)))
        self._shapes.append((h, w))
        return vid.value

    def add_job(self, ref: int, neighbours, min_depth: float, max_depth: float) -> int:
        """Adds a reference view to estimate against `neighbours` (pool
        indices, in the order a DepthmapEstimator would add them)."""
        nb = np.ascontiguousarray(np.asarray(neighbours, np.int32).reshape(-1))
        jid = C.c_int32(-1)
        _capi.check(self._lib.pm_batch_add_job(
            self._h, int(ref), _capi.ptr(nb, C.c_int32), int(nb.size),
            float(min_depth), float(max_depth), C.byref(jid)))
        self._jobs.append(int(ref))
        return jid.value

    @property
    def num_jobs(self) -> int:
        return len(self._jobs)

    def job_shape(self, job: int):
        return self._shapes[self._jobs[job]]

    def set_patchmatch_iterations(self, n: int) -> None:
        _capi.check(self._lib.pm_batch_set_patchmatch_iterations(self._h, int(n)))

    def set_patch_size(self, size: int) -> None:
        _capi.check(self._lib.pm_batch_set_patch_size(self._h, int(size)))

    def set_min_patch_sd(self, sd: float) -> None:
        _capi.check(self._lib.pm_batch_set_min_patch_sd(self._h, float(sd)))

    def set_seed(self, seed: int) -> None:
        _capi.check(self._lib.pm_batch_set_seed(self._h, int(seed) & (2 ** 64 - 1)))

    def set_pyramid_consistency(self, enable: bool) -> None:
        """The median-consistent re-selection at the pyramid's coarser
        levels (default on, as set_consistency; off: at the full level
        only)."""
        _capi.check(self._lib.pm_batch_set_pyramid_consistency(self._h, int(bool(enable))))

    def set_pyramid(self, half_iterations, levels: int = 2) -> None:
        """Plain PatchMatch of every job seeded by half_iterations iterations
        at half resolution (0: off), with levels - 1 coarser levels, or a
        sequence of iterations per coarser level (coarsest first), as
        DepthmapEstimator.set_pyramid."""
        if isinstance(half_iterations, (int, np.integer)):
            _capi.check(self._lib.pm_batch_set_pyramid_levels(self._h, int(levels)))
            _capi.check(self._lib.pm_batch_set_pyramid(self._h, int(half_iterations)))
            return
        it = np.ascontiguousarray(half_iterations, np.int32).ravel()
        _capi.check(self._lib.pm_batch_set_pyramid_schedule(
            self._h, it.ctypes.data_as(C.POINTER(C.c_int32)), int(it.size)))

    def set_exploration(self, n: int) -> None:
        _capi.check(self._lib.pm_batch_set_exploration(self._h, int(n)))

    def set_init_map(self, job: int, init_map, w: int = 0, h: int = 0, stride: int = 8,
                     scale: int = 4) -> None:
        """Seeds job `job`'s initialisation from a coarse (t, sx, sy, c) map,
        a float32 device tensor, as DepthmapEstimator.set_init_map (None:
        random).  The batch keeps a reference to each job's tensor until it is
        replaced or cleared."""
        ptr = None
        if init_map is not None:
            ptr = _capi._device_tensor(init_map, w * h * stride, "init map", "float32")
        _capi.check(self._lib.pm_batch_set_init_map(
            self._h, int(job), C.c_void_p(ptr) if ptr else None, int(w), int(h),
            int(stride), int(scale)))
        if init_map is None:
            self._init_maps.pop(int(job), None)
        else:
            self._init_maps[int(job)] = init_map

    def initializer_input(self, job: int, out, out_w: int, out_h: int, scale: int = 4,
                          nplanes: int = 32) -> None:
        """The learned initialiser's input of job `job` into the device
        tensor `out` (BF16, out_h x out_w x 40), from the pool's views at
        1/scale (as DepthmapEstimator.initializer_input)."""
        ptr = _capi._device_tensor(out, out_w * out_h * 40, "initializer input", "bfloat16")
        _capi.check(self._lib.pm_batch_initializer_input(
            self._h, int(job), int(scale), int(nplanes), C.c_void_p(ptr), int(out_w), int(out_h)))

    def set_init_search(self, n: int) -> None:
        _capi.check(self._lib.pm_batch_set_init_search(self._h, int(n)))

    def set_init_spread(self, sigma_log_depth: float, sigma_slope: float) -> None:
        _capi.check(self._lib.pm_batch_set_init_spread(self._h, float(sigma_log_depth),
                                                       float(sigma_slope)))

    def set_consistency(self, enable: bool) -> None:
        """Median-consistent re-selection for every job (default on)."""
        _capi.check(self._lib.pm_batch_set_consistency(self._h, int(bool(enable))))

    def set_post_process(self, enable: bool) -> None:
        _capi.check(self._lib.pm_batch_set_post_process(self._h, int(bool(enable))))

    def set_precision(self, precision) -> None:
        """PatchMatch sample precision: "fp32" (default), "fp16" or "bf16"."""
        _capi.check(self._lib.pm_batch_set_precision(self._h, _capi.precision_code(precision)))

    def set_tile(self, tile: int) -> None:
        _capi.check(self._lib.pm_batch_set_tile(self._h, int(tile)))

    def set_graph(self, enable: bool) -> None:
        _capi.check(self._lib.pm_batch_set_graph(self._h, int(bool(enable))))

    def autotune(self, patch_sizes=(), sample: bool = False, accuracy=None,
                 tolerance: float = 0.002):
        """Chooses the lane layout by speed and, among `patch_sizes`, the
        patch size; returns (patch, tile).

        Without `accuracy`, the fastest (patch, tile) on the first jobs'
        converged iteration wins.  Patch size changes the maps, so a caller
        that can measure accuracy (ground truth, or a reference run) passes
        `accuracy(batch) -> float`: each patch size then gets its fastest
        tile and one full run, and the fastest patch size whose accuracy is
        within `tolerance` of the most accurate one wins.  The per-candidate
        figures are kept in `self.tune_report`."""
        mode = _capi.PM_MODE_PATCHMATCH_SAMPLE if sample else _capi.PM_MODE_PATCHMATCH

        def tune(sizes):
            ps = np.ascontiguousarray(np.asarray(list(sizes), np.int32).reshape(-1))
            p, t = C.c_int32(0), C.c_int32(0)
            _capi.check(self._lib.pm_batch_autotune(
                self._h, mode, _capi.ptr(ps, C.c_int32) if ps.size else None, int(ps.size),
                C.byref(p), C.byref(t)))
            return p.value, t.value

        if accuracy is None or len(patch_sizes) < 2:
            self.tune_report = None
            return tune(patch_sizes)
        report = []
        for p in patch_sizes:
            _, t = tune([p])
            self.run(sample)            # first run of this configuration (graph capture)
            self.run(sample)
            report.append({"patch": int(p), "tile": int(t),
                           "device_ms": float(self.stats()["total_ms"]),
                           "accuracy": float(accuracy(self))})
        best = max(r["accuracy"] for r in report)
        ok = [r for r in report if r["accuracy"] >= best - tolerance]
        pick = min(ok, key=lambda r: r["device_ms"])
        self.set_patch_size(pick["patch"])
        self.set_tile(pick["tile"])
        self.tune_report = {"candidates": report, "tolerance": tolerance,
                            "patch": pick["patch"], "tile": pick["tile"]}
        return pick["patch"], pick["tile"]

    def run(self, sample: bool = False) -> None:
        """Runs every job; the maps stay on the device."""
        mode = _capi.PM_MODE_PATCHMATCH_SAMPLE if sample else _capi.PM_MODE_PATCHMATCH
        _capi.check(self._lib.pm_batch_run(self._h, mode))

    def result(self, job: int) -> list:
        """[depth, plane, score, nghbr] of a job, as compute_patch_match."""
        h, w = self.job_shape(job)
        depth = np.zeros((h, w), np.float32)
        plane = np.zeros((h, w, 3), np.float32)
        score = np.zeros((h, w), np.float32)
        nghbr = np.zeros((h, w), np.int32)
        _capi.check(self._lib.pm_batch_download(
            self._h, int(job), _capi.ptr(depth, C.c_float), _capi.ptr(plane, C.c_float),
            _capi.ptr(score, C.c_float), _capi.ptr(nghbr, C.c_int32)))
        return [depth, plane, score, nghbr]

    def cluster_result(self, job: int) -> list:
        """[depth, normal, cost, confidence] of a job in the cluster
        estimator's layout (DepthmapResult, estimator.h:33-38)."""
        h, w = self.job_shape(job)
        depth = np.zeros((h, w), np.float32)
        normal = np.zeros((h, w, 3), np.float32)
        cost = np.zeros((h, w), np.float32)
        conf = np.zeros((h, w), np.float32)
        _capi.check(self._lib.pm_batch_download_cluster(
            self._h, int(job), _capi.ptr(depth, C.c_float), _capi.ptr(normal, C.c_float),
            _capi.ptr(cost, C.c_float), _capi.ptr(conf, C.c_float)))
        return [depth, normal, cost, conf]

    def clean(self, same_depth_threshold: float = 0.01, min_consistent_views: int = 2) -> None:
        """DepthmapCleaner::Clean of every job against the jobs of its
        neighbour views (defaults of depthmap.cc:504-505)."""
        _capi.check(self._lib.pm_batch_clean(self._h, float(same_depth_threshold),
                                             int(min_consistent_views)))

    def clean_result(self, job: int) -> np.ndarray:
        h, w = self.job_shape(job)
        out = np.zeros((h, w), np.float32)
        _capi.check(self._lib.pm_batch_download_clean(self._h, int(job),
                                                      _capi.ptr(out, C.c_float)))
        return out

    def stats(self) -> dict:
        s = _capi.pm_stats()
        _capi.check(self._lib.pm_batch_stats(self._h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in s._fields_}
        ms = C.c_float(0)
        _capi.check(self._lib.pm_batch_clean_ms(self._h, C.byref(ms)))
        out["clean_ms"] = ms.value
        return out
