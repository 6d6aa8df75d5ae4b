"""``pydense.DepthmapClusterEstimator``-compatible surface over the engine.

The live pipeline (opensfm/dense.py:657-731) builds a cluster of reference
views, each with its source views, runs it, and saves per reference view
``[depth, normal, cost, confidence]`` (depthmap_bind.h:19-114,
dense/cluster.h:42-78).  This class keeps that call sequence and result
layout and computes every reference view with the engine's batched
PatchMatch (one ``DepthmapBatch`` job per reference view, all jobs in one
set of launches): depth, the unit camera-frame normal of the winning
plane, cost = 1 - NCC score, and the reference's logistic confidence
(estimator.cc:1022-1040).

The reference's estimator is a different PatchMatch (top-k view costs
with census, SfM planar prior, cross-view geometric consistency).  Settings
with a counterpart here are applied: ``max_iterations``, ``patch_size`` and
``hierarchy_levels`` -- as in the reference (estimator.cc:140, 168-172),
N levels at scales 1 / 2^(N - 1), ..., 1/2, 1, each running max_iterations
iterations and seeding the next (the engine's pyramid; levels whose
smallest view would be narrower than one patch are dropped).  The others
are accepted and kept (``max_image_size`` -- the caller already scales
images --, ``sigma_spatial``, ``sigma_color``, ``top_k``,
``census_weight``, ``smooth_weight``, the geometric-consistency weight and
links, the SfM points) without changing the result.
"""
from __future__ import annotations

import ctypes as C

from typing import List

import numpy as np

try:
    import xxhash

    def _digest(image: np.ndarray) -> int:
        return xxhash.xxh3_64_intdigest(memoryview(image).cast("B"))
except ImportError:  # the built-in hash of a copy: ten times slower
    def _digest(image: np.ndarray) -> int:
        return hash(image.tobytes())

from . import _capi
from .batch import DepthmapBatch


class DepthmapClusterEstimator:
    """Cluster-of-reference-views estimator (pydense.pyi DepthmapClusterEstimator)."""

    def __init__(self):
        self._device = 0
        self._params = dict(max_iterations=3, patch_size=11, max_image_size=3200,
                            hierarchy_levels=1, sigma_spatial=5.0, sigma_color=3.0 / 255.0,
                            top_k=4, census_weight=0.3, smooth_weight=0.0,
                            geom_consistency_weight=0.0)
        self.clear()

    # ---- parameters (depthmap_bind.h:21-33) ----
    def set_max_iterations(self, n: int) -> None:
        self._params["max_iterations"] = int(n)

    def set_patch_size(self, size: int) -> None:
        self._params["patch_size"] = int(size)

    def set_max_image_size(self, size: int) -> None:
        self._params["max_image_size"] = int(size)

    def set_hierarchy_levels(self, levels: int) -> None:
        self._params["hierarchy_levels"] = int(levels)

    def set_sigma_spatial(self, v: float) -> None:
        self._params["sigma_spatial"] = float(v)

    def set_sigma_color(self, v: float) -> None:
        self._params["sigma_color"] = float(v)

    def set_top_k(self, k: int) -> None:
        self._params["top_k"] = int(k)

    def set_census_weight(self, w: float) -> None:
        self._params["census_weight"] = float(w)

    def set_smooth_weight(self, w: float) -> None:
        self._params["smooth_weight"] = float(w)

    def set_geom_consistency_weight(self, w: float) -> None:
        self._params["geom_consistency_weight"] = float(w)

    def set_device(self, idx: int) -> None:
        self._device = int(idx)

    # ---- cluster contents (depthmap_bind.h:36-53) ----
    def begin_ref_view(self, K, R, t, image, depth_min: float, depth_max: float) -> int:
        self._refs.append(dict(view=self._view(K, R, t, image), sources=[],
                               depth_range=(float(depth_min), float(depth_max)),
                               sfm_points=None))
        return len(self._refs) - 1

    def add_source_view(self, K, R, t, image) -> None:
        if not self._refs:
            raise RuntimeError("add_source_view before begin_ref_view")
        self._refs[-1]["sources"].append(self._view(K, R, t, image))

    def set_sfm_points(self, points) -> None:
        if not self._refs:
            raise RuntimeError("set_sfm_points before begin_ref_view")
        self._refs[-1]["sfm_points"] = np.asarray(points, np.float64).reshape(-1, 3)

    def add_geom_link(self, ref_idx: int, source_view_idx: int, from_ref_idx: int) -> None:
        self._links.append((int(ref_idx), int(source_view_idx), int(from_ref_idx)))

    def run(self) -> List[list]:
        """Per reference view, in begin_ref_view order: [depth (H, W),
        normal (H, W, 3), cost (H, W), confidence (H, W)], float32."""
        if not self._refs:
            return []
        b = DepthmapBatch(self._device)
        pool = {}

        def view_index(v):
            # the same shot arrives once per reference view that uses it:
            # one pool view per distinct (camera, image)
            key = (v["image"].shape, v["K"].tobytes(), v["R"].tobytes(), v["t"].tobytes(),
                   _digest(v["image"]))
            if key not in pool:
                pool[key] = b.add_view(v["K"], v["R"], v["t"], v["image"])
            return pool[key]

        for ref in self._refs:
            r = view_index(ref["view"])
            nb = [view_index(s) for s in ref["sources"]]
            lo, hi = ref["depth_range"]
            b.add_job(r, nb, lo, hi)
        iters = max(1, self._params["max_iterations"])
        patch = self._params["patch_size"]
        b.set_patchmatch_iterations(iters)
        b.set_patch_size(patch)
        levels = self.pyramid_levels(
            min(min(v["image"].shape) for ref in self._refs
                for v in [ref["view"]] + ref["sources"]), patch)
        if levels >= 2:
            b.set_pyramid(iters, levels=levels)
        b.run()
        self.last_stats = b.stats()
        return [b.cluster_result(k) for k in range(len(self._refs))]

    def pyramid_levels(self, min_side: int, patch: int) -> int:
        """hierarchy_levels, less the levels at which a view of min_side
        pixels would be narrower than one patch (at most 8)."""
        levels = min(8, max(1, self._params["hierarchy_levels"]))
        while levels > 1 and (min_side >> (levels - 1)) < patch:
            levels -= 1
        return levels

    def clear(self) -> None:
        self._refs = []
        self._links = []
        self.last_stats = None

    def num_ref_views(self) -> int:
        return len(self._refs)

    @staticmethod
    def _view(K, R, t, image):
        image = np.ascontiguousarray(np.asarray(image), np.uint8)
        if image.ndim != 2:
            raise ValueError("image must be a 2-D uint8 array")
        return dict(K=np.asarray(K, np.float64), R=np.asarray(R, np.float64),
                    t=np.asarray(t, np.float64).reshape(3), image=image)

    # ---- devices (depthmap_bind.h:77-107) ----
    @staticmethod
    def is_available() -> bool:
        try:
            return _capi.device_count() > 0
        except RuntimeError:
            return False

    @staticmethod
    def num_devices() -> int:
        return _capi.device_count()

    @staticmethod
    def device_name(idx: int) -> str:
        return _capi.device_info(idx)[0]

    @staticmethod
    def device_memory_bytes(idx: int) -> int:
        return int(_capi.device_info(idx)[1])

    @staticmethod
    def device_is_gpu(idx: int) -> bool:
        _capi.device_info(idx)  # range check: every visible HIP device is a GPU
        return True

    @staticmethod
    def device_available_memory(idx: int) -> int:
        """Global memory minus the bytes reserved with reserve_device_memory
        (opencl_utils.h:68-72)."""
        n = C.c_uint64(0)
        _capi.check(_capi.load().pm_device_available(int(idx), C.byref(n)))
        return int(n.value)

    @staticmethod
    def gpu_memory_bytes() -> int:
        return DepthmapClusterEstimator.device_memory_bytes(0)

    @staticmethod
    def reserve_device_memory(idx: int, nbytes: int) -> None:
        """Bookkeeping of the dense scheduler, as the reference's OpenCL
        context keeps it (opencl_utils.h:73): device memory itself is
        allocated on use."""
        _capi.check(_capi.load().pm_device_reserve(int(idx), int(nbytes)))

    @staticmethod
    def release_device_memory(idx: int, nbytes: int) -> None:
        """Releases reserved bytes; more than is reserved leaves zero
        (opencl_utils.h:74-76)."""
        _capi.check(_capi.load().pm_device_reserve(int(idx), -int(nbytes)))
