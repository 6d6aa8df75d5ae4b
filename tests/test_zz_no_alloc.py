"""pm_cleaner_clean_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against pm_cleaner_clean, the
allocating DepthmapCleaner entry point (depthmap.cc:524-555), through the C
ABI: bit-equal maps on two calls in a row, the second with fewer and
smaller views, and a clear range error beyond the view limit."""
import ctypes as C

import numpy as np
import pytest

from opensfm_amd import DepthmapCleaner, _capi, scenes

pytestmark = pytest.mark.gpu


def _views(width, height, n_neighbours, seed):
    scene = scenes.make_scene(width, height, n_neighbours, 5, seed=seed, min_depth=2.0,
                              max_depth=20.0)
    rng = np.random.default_rng(seed)
    out = []
    for v in scene.views:
        d = v.depth.astype(np.float32) * (1 + rng.normal(0, 0.004, v.depth.shape)).astype(np.float32)
        d[rng.random(d.shape) < 0.03] = 0
        out.append((v.K, v.R, v.t, np.ascontiguousarray(d, np.float32)))
    return out


def _allocating(views, thr, min_views):
    c = DepthmapCleaner()
    c.set_same_depth_threshold(thr)
    c.set_min_consistent_views(min_views)
    for K, R, t, d in views:
        c.add_view(K, R, t, d)
    return c.clean()


def _noalloc(lib, views, thr, min_views, dev_depths, out):
    n = len(views)
    K = np.ascontiguousarray(np.stack([np.asarray(v[0], np.float64).reshape(9) for v in views]))
    R = np.ascontiguousarray(np.stack([np.asarray(v[1], np.float64).reshape(9) for v in views]))
    t = np.ascontiguousarray(np.stack([np.asarray(v[2], np.float64).reshape(3) for v in views]))
    ptrs = (C.c_void_p * n)(*[d.data_ptr() for d in dev_depths])
    w = np.array([v[3].shape[1] for v in views], np.int32)
    h = np.array([v[3].shape[0] for v in views], np.int32)
    return lib.pm_cleaner_clean_noalloc(
        0, n, _capi.ptr(K, C.c_double), _capi.ptr(R, C.c_double), _capi.ptr(t, C.c_double),
        ptrs, _capi.ptr(w, C.c_int32), _capi.ptr(h, C.c_int32), float(thr), int(min_views),
        C.c_void_p(out.data_ptr()), None)


def test_noalloc_clean_equals_allocating_clean():
    import torch
    lib = _capi.load()
    for size, nb, seed, thr, mv in (((320, 240), 6, 41, 0.01, 2), ((160, 120), 3, 42, 0.02, 3)):
        print(f"no-alloc clean {size} with {nb} neighbours", flush=True)
        views = _views(size[0], size[1], nb, seed)
        want = _allocating(views, thr, mv)
        dev = [torch.from_numpy(v[3]).to("cuda") for v in views]
        out = torch.full((size[1], size[0]), -7.0, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _capi.check(_noalloc(lib, views, thr, mv, dev, out))
        got = out.cpu().numpy()
        assert np.array_equal(got, want), np.abs(got - want).max()
        assert (got > 0).mean() > 0.3


def test_noalloc_clean_rejects_too_many_views():
    import torch
    lib = _capi.load()
    views = _views(64, 48, 1, 43)
    views = views * (_capi.PM_NOALLOC_MAX_VIEWS // 2 + 1)
    dev = [torch.from_numpy(views[0][3]).to("cuda")] * len(views)
    out = torch.zeros((48, 64), dtype=torch.float32, device="cuda")
    rc = _noalloc(lib, views, 0.01, 2, dev, out)
    assert rc == _capi.PM_ERR_RANGE
    assert "exceed the maximum" in _capi.last_error()
