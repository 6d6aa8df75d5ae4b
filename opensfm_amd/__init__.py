"""MI355X-native PatchMatch multi-view-stereo depthmap engine.

A drop-in for OpenSfM's dense depthmap path (pydense.DepthmapEstimator /
pydense.DepthmapCleaner, opensfm/src/lib/dense/pydense.pyi:28-60), computed
by hand-written HIP kernels for gfx950 behind the C ABI in
include/opensfm_amd.h.
"""
from .batch import DepthmapBatch  # noqa: F401
from .cluster import DepthmapClusterEstimator  # noqa: F401
from .pydense import (  # noqa: F401
    DepthmapCleaner, DepthmapEstimator, DepthmapFuser, DepthmapPruner, GPUDepthmapCleaner)

__all__ = ["DepthmapEstimator", "DepthmapCleaner", "GPUDepthmapCleaner", "DepthmapPruner",
           "DepthmapFuser", "DepthmapBatch", "DepthmapClusterEstimator"]
