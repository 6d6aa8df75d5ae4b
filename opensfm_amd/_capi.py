"""ctypes binding of the engine's C ABI (include/opensfm_amd.h).

The shared library is built in-tree by ``opensfm_amd._build`` (hipcc,
gfx950).  There is no fallback: if the library is missing or no HIP device is
present, the calls fail loudly with the library's own error message.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _build

PM_OK = 0
PM_ERR_INVALID = -1
PM_ERR_RUNTIME = -2
PM_ERR_RANGE = -3

PM_NOALLOC_MAX_VIEWS = 64  # include/opensfm_amd.h

PM_MODE_PLANE_SWEEP = 0
PM_MODE_PATCHMATCH = 1
PM_MODE_PATCHMATCH_SAMPLE = 2

PM_PRECISION_FP32 = 0
PM_PRECISION_FP16 = 1
PM_PRECISION_BF16 = 2
PRECISIONS = {"fp32": PM_PRECISION_FP32, "fp16": PM_PRECISION_FP16, "bf16": PM_PRECISION_BF16}

PM_TILE_8x8 = 0
PM_TILE_16x16 = 1
PM_TILE_32x8 = 2
TILE_NAMES = {PM_TILE_8x8: "8x8", PM_TILE_16x16: "16x16", PM_TILE_32x8: "32x8"}


class pm_stats(C.Structure):
    _fields_ = [
        ("scorings", C.c_int64),
        ("dominant_scorings", C.c_int64),
        ("patch_size", C.c_int32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("launches", C.c_int32),
        ("dominant_ms", C.c_float),
        ("total_ms", C.c_float),
    ]


# Every symbol declared in include/opensfm_amd.h, with (restype, argtypes).
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)
SIGNATURES = {
    "pm_last_error": (C.c_char_p, []),
    "pm_device_count": (C.c_int, [_ip]),
    "pm_device_info": (C.c_int, [C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]),
    "pm_device_reserve": (C.c_int, [C.c_int, C.c_int64]),
    "pm_device_available": (C.c_int, [C.c_int, C.POINTER(C.c_uint64)]),
    "pm_estimator_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_estimator_destroy": (C.c_int, [_vp]),
    "pm_estimator_add_view": (C.c_int, [_vp, _dp, _dp, _dp, _u8p, _u8p, C.c_int,
                                        C.c_int, C.c_int]),
    "pm_estimator_set_depth_range": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int]),
    "pm_estimator_set_patchmatch_iterations": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_patch_size": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_min_patch_sd": (C.c_int, [_vp, C.c_float]),
    "pm_estimator_set_seed": (C.c_int, [_vp, C.c_uint64]),
    "pm_estimator_set_far_probes": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_pyramid": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_pyramid_levels": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_pyramid_schedule": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_int]),
    "pm_estimator_set_pyramid_consistency": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_lanes_per_pixel": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_refinement_steps": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_exploration": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_precision": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_consistency": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_post_process": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_compute": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _ip]),
    "pm_estimator_run": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_download": (C.c_int, [_vp, _fp, _fp, _fp, _ip]),
    "pm_estimator_score_hypotheses": (C.c_int, [_vp, C.c_int, _ip, _fp, _ip, _fp]),
    "pm_estimator_score_hypotheses_fast": (C.c_int, [_vp, C.c_int, _ip, _fp, _ip, _fp]),
    "pm_estimator_initializer_input": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int]),
    "pm_estimator_set_init_map": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pm_estimator_set_init_search": (C.c_int, [_vp, C.c_int]),
    "pm_estimator_set_init_spread": (C.c_int, [_vp, C.c_float, C.c_float]),
    "pm_estimator_stats": (C.c_int, [_vp, C.POINTER(pm_stats)]),
    "pm_cleaner_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_cleaner_destroy": (C.c_int, [_vp]),
    "pm_cleaner_set_same_depth_threshold": (C.c_int, [_vp, C.c_float]),
    "pm_cleaner_set_min_consistent_views": (C.c_int, [_vp, C.c_int]),
    "pm_cleaner_add_view": (C.c_int, [_vp, _dp, _dp, _dp, _fp, C.c_int, C.c_int]),
    "pm_cleaner_clean": (C.c_int, [_vp, _fp]),
    "pm_cleaner_last_ms": (C.c_int, [_vp, _fp]),
    "pm_cleaner_clean_noalloc": (C.c_int, [C.c_int, C.c_int, _dp, _dp, _dp, C.POINTER(_vp),
                                           _ip, _ip, C.c_float, C.c_int, _vp, _vp]),
    "pm_gpu_cleaner_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_gpu_cleaner_destroy": (C.c_int, [_vp]),
    "pm_gpu_cleaner_set_same_depth_threshold": (C.c_int, [_vp, C.c_float]),
    "pm_gpu_cleaner_set_min_consistent_views": (C.c_int, [_vp, C.c_int]),
    "pm_gpu_cleaner_set_carving_threshold": (C.c_int, [_vp, C.c_float]),
    "pm_gpu_cleaner_set_max_carved_views": (C.c_int, [_vp, C.c_int]),
    "pm_gpu_cleaner_add_view": (C.c_int, [_vp, _dp, _dp, _dp, _fp, C.c_int, C.c_int, _ip]),
    "pm_gpu_cleaner_clean": (C.c_int, [_vp, C.c_int, _ip, C.c_int, _fp]),
    "pm_gpu_cleaner_clear": (C.c_int, [_vp]),
    "pm_gpu_cleaner_last_ms": (C.c_int, [_vp, _fp]),
    "pm_pruner_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_pruner_destroy": (C.c_int, [_vp]),
    "pm_pruner_set_same_depth_threshold": (C.c_int, [_vp, C.c_float]),
    "pm_pruner_add_view": (C.c_int, [_vp, _dp, _dp, _dp, _fp, _fp, _u8p, _u8p, C.c_int,
                                     C.c_int]),
    "pm_pruner_prune": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "pm_pruner_download": (C.c_int, [_vp, _fp, _fp, _u8p, _u8p]),
    "pm_pruner_last_ms": (C.c_int, [_vp, _fp]),
    "pm_fuser_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_fuser_destroy": (C.c_int, [_vp]),
    "pm_fuser_set_min_num_consistent": (C.c_int, [_vp, C.c_int]),
    "pm_fuser_set_max_reproj_error": (C.c_int, [_vp, C.c_float]),
    "pm_fuser_set_max_depth_error": (C.c_int, [_vp, C.c_float]),
    "pm_fuser_set_max_normal_error": (C.c_int, [_vp, C.c_float]),
    "pm_fuser_set_border_margin": (C.c_int, [_vp, C.c_int]),
    "pm_fuser_set_num_threads": (C.c_int, [_vp, C.c_int]),
    "pm_fuser_set_sor_params": (C.c_int, [_vp, C.c_int, C.c_float]),
    "pm_fuser_set_behind_depth_factor": (C.c_int, [_vp, C.c_float]),
    "pm_fuser_add_view": (C.c_int, [_vp, _dp, _dp, _dp, _fp, _fp, _u8p, _u8p, C.c_int,
                                    C.c_int, _ip, C.c_int, C.c_int]),
    "pm_fuser_fuse": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "pm_fuser_download": (C.c_int, [_vp, _fp, _fp, _u8p]),
    "pm_fuser_stats": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _fp, _fp]),
    "pm_covis_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_covis_destroy": (C.c_int, [_vp]),
    "pm_covis_build": (C.c_int, [_vp, C.c_int64, _dp, _i64p, _ip, C.c_int, C.c_int, C.c_double,
                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                 C.POINTER(C.c_int64)]),
    "pm_covis_download": (C.c_int, [_vp, _ip, _dp, _i64p, _ip, _i64p, _ip, _ip, _ip, _fp]),
    "pm_covis_select_neighbors": (C.c_int, [_vp, C.c_int, _dp, C.c_int64, _dp, _i64p, _ip, _ip,
                                            C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "pm_covis_download_neighbors": (C.c_int, [_vp, _i64p, _ip, _i64p, _ip]),
    "pm_covis_last_ms": (C.c_int, [_vp, _fp]),
    "pm_lz4_block_decompress": (C.c_int, [_u8p, C.c_int64, _u8p, C.c_int64, C.c_int64,
                                          C.POINTER(C.c_int64)]),
    "pm_lz4_block_compress": (C.c_int, [_u8p, C.c_int64, _u8p, C.c_int64,
                                        C.POINTER(C.c_int64)]),
    "pm_batch_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "pm_batch_destroy": (C.c_int, [_vp]),
    "pm_batch_add_view": (C.c_int, [_vp, _dp, _dp, _dp, _u8p, _u8p, C.c_int, C.c_int,
                                    C.c_int, _ip]),
    "pm_batch_add_job": (C.c_int, [_vp, C.c_int, _ip, C.c_int, C.c_double, C.c_double,
                                   _ip]),
    "pm_batch_set_patchmatch_iterations": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_patch_size": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_min_patch_sd": (C.c_int, [_vp, C.c_float]),
    "pm_batch_set_seed": (C.c_int, [_vp, C.c_uint64]),
    "pm_batch_initializer_input": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                                             C.c_int]),
    "pm_batch_set_init_map": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "pm_batch_set_init_search": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_init_spread": (C.c_int, [_vp, C.c_float, C.c_float]),
    "pm_batch_set_exploration": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_pyramid": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_pyramid_levels": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_pyramid_schedule": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_int]),
    "pm_batch_set_pyramid_consistency": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_consistency": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_post_process": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_tile": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_precision": (C.c_int, [_vp, C.c_int]),
    "pm_batch_set_graph": (C.c_int, [_vp, C.c_int]),
    "pm_batch_autotune": (C.c_int, [_vp, C.c_int, _ip, C.c_int, _ip, _ip]),
    "pm_batch_run": (C.c_int, [_vp, C.c_int]),
    "pm_batch_download": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _ip]),
    "pm_batch_clean": (C.c_int, [_vp, C.c_float, C.c_int]),
    "pm_batch_download_clean": (C.c_int, [_vp, C.c_int, _fp]),
    "pm_batch_download_cluster": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _fp]),
    "pm_batch_stats": (C.c_int, [_vp, C.POINTER(pm_stats)]),
    "pm_batch_clean_ms": (C.c_int, [_vp, _fp]),
    "pm_nn_conv_kpad": (C.c_int, [C.c_int]),
    "pm_nn_conv3x3": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                _vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                C.c_int, _vp, C.c_int, C.c_int, _vp]),
    "pm_nn_conv3x3_wgrad": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                                      C.c_int, C.c_int, _vp, C.c_int, _vp, _vp]),
    "pm_nn_weight_flip": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp]),
    "pm_nn_lrelu_backward": (C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp,
                                       C.c_int, C.c_int, _vp]),
    "pm_nn_maxpool2": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 _vp, _vp, _vp]),
    "pm_nn_maxpool2_backward": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                                          C.c_int, C.c_int, _vp]),
    "pm_nn_upsample2": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                                  C.c_int, _vp]),
    "pm_nn_upsample2_backward": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, _vp, _vp]),
    "pm_nn_loss": (C.c_int, [_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_float, C.c_float, _vp,
                             _vp, _vp]),
    "pm_nn_adam": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_float, C.c_float,
                             C.c_float, C.c_float, C.c_int, _vp]),
    "pm_nn_to_bf16": (C.c_int, [_vp, _vp, C.c_int, _vp]),
}

_lock = threading.Lock()
_lib = None


def library_path() -> str:
    """The in-tree build, or the library named by OPENSFM_AMD_LIB."""
    return os.environ.get("OPENSFM_AMD_LIB") or _build.LIB


def load():
    """Load the engine library (building it first if it is absent)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        # PyTorch (which the learned initialiser uses for device memory)
        # ships its own HIP runtime; loading it first makes the engine bind
        # to that same runtime instead of bringing up a second one in the
        # process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        path = library_path()
        if not os.path.exists(path):
            _build.build()
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def last_error() -> str:
    msg = load().pm_last_error()
    return msg.decode() if msg else ""


def check(rc: int) -> None:
    """Map a C status to the exception pybind11 would raise for the
    reference's C++ exception (depthmap_bind.h:253-259)."""
    if rc == PM_OK:
        return
    msg = last_error() or f"status {rc}"
    if rc == PM_ERR_INVALID:
        raise ValueError(msg)
    if rc == PM_ERR_RANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    n = C.c_int32(0)
    check(load().pm_device_count(C.byref(n)))
    return n.value


def device_info(idx: int):
    """(name, global memory bytes, free bytes) of HIP device idx."""
    name = C.create_string_buffer(256)
    total, free = C.c_uint64(0), C.c_uint64(0)
    check(load().pm_device_info(int(idx), name, 256, C.byref(total), C.byref(free)))
    return name.value.decode(errors="replace"), total.value, free.value


def precision_code(p) -> int:
    """PM_PRECISION_* from a code or one of "fp32", "fp16", "bf16"."""
    if isinstance(p, str):
        if p.lower() not in PRECISIONS:
            raise ValueError(f"unknown precision {p!r}")
        return PRECISIONS[p.lower()]
    return int(p)


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _device_tensor(t, n, what, dtype_name):
    """Checks that t is a torch-like device tensor (data_ptr, numel, device,
    dtype) of at least n elements of the given dtype; returns its address."""
    if not (hasattr(t, "data_ptr") and hasattr(t, "numel")):
        raise TypeError(f"{what} must be a device tensor (an object with data_ptr() and numel())")
    dev = getattr(t, "device", None)
    if dev is not None and getattr(dev, "type", "cuda") != "cuda":
        raise ValueError(f"{what} must be in device memory, got {dev}")
    dt = getattr(t, "dtype", None)
    if dt is not None and str(dt).split(".")[-1] != dtype_name:
        raise ValueError(f"{what} must be {dtype_name}, got {dt}")
    if t.numel() < n:
        raise ValueError(f"{what} holds {t.numel()} elements, needs at least {n}")
    return t.data_ptr()
