"""CPU-side checks of the C ABI boundary (no GPU needed).

* the shared library loads and exports every function include/*.h declares;
* the Python binding declares exactly those symbols;
* with no HIP device, constructors fail with a clean RuntimeError (the
  reference's no-device stubs raise RuntimeError, depthmap_bind.h:178-190)
  rather than crashing, and argument validation happens before any device
  work.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from opensfm_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pm_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_abi():
    names = declared_functions()
    assert {"pm_estimator_create", "pm_estimator_add_view", "pm_estimator_compute",
            "pm_cleaner_clean", "pm_last_error"} <= names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_capi.library_path())
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert set(_capi.SIGNATURES) == declared_functions()


def test_library_has_gfx950_code():
    data = open(_capi.library_path(), "rb").read()
    assert b"gfx950" in data


def test_no_device_errors_are_clean():
    from opensfm_amd import DepthmapCleaner, DepthmapEstimator
    if _capi.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(RuntimeError, match="no HIP device"):
        DepthmapEstimator()
    with pytest.raises(RuntimeError, match="no HIP device"):
        DepthmapCleaner()
    # device queries name the missing device (IndexError, like the
    # reference's out-of-range device index)
    with pytest.raises(IndexError, match="out of range"):
        _capi.device_info(0)
    with pytest.raises(IndexError, match="out of range"):
        _capi.check(_capi.load().pm_device_reserve(0, 1))


def test_null_handle_and_argument_validation():
    lib = _capi.load()
    assert lib.pm_estimator_set_patch_size(None, 7) == _capi.PM_ERR_INVALID
    assert "null" in _capi.last_error()
    assert lib.pm_estimator_run(None, 1) == _capi.PM_ERR_INVALID
    assert lib.pm_cleaner_clean(None, None) == _capi.PM_ERR_INVALID
    assert lib.pm_estimator_destroy(None) == _capi.PM_OK
    assert lib.pm_cleaner_destroy(None) == _capi.PM_OK
    out = C.c_void_p()
    assert lib.pm_estimator_create(0, None) == _capi.PM_ERR_INVALID
    n = C.c_int32(-1)
    assert lib.pm_device_count(C.byref(n)) == _capi.PM_OK and n.value >= 0
    assert out.value is None


def test_status_mapping():
    lib = _capi.load()
    lib.pm_estimator_set_patch_size(None, 4)
    with pytest.raises(ValueError):
        _capi.check(_capi.PM_ERR_INVALID)
    with pytest.raises(IndexError):
        _capi.check(_capi.PM_ERR_RANGE)
    with pytest.raises(RuntimeError):
        _capi.check(_capi.PM_ERR_RUNTIME)
    _capi.check(_capi.PM_OK)


def test_stats_struct_layout():
    # int64, int64, 4 x int32, 2 x float
    assert C.sizeof(_capi.pm_stats) == 8 + 8 + 4 * 4 + 2 * 4
    assert np.dtype(np.int64).itemsize == 8
