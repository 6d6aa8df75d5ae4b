"""GPU tests of the DepthmapClusterEstimator surface (depthmap_bind.h:19-114,
the call sequence of opensfm/dense.py:657-731): results in begin_ref_view
order with the reference's layout; depth, normal, cost and confidence equal
to oracle.cluster_readback (the readback rules of estimator.cc:960-993 and
the logistic of :1022-1040) applied to the batched engine's maps for the
same jobs; accurate depths on a rendered scene."""
import numpy as np
import pytest

import oracle
from opensfm_amd import DepthmapBatch, DepthmapClusterEstimator, scenes

pytestmark = pytest.mark.gpu


def test_cluster_surface(cfg1_scene):
    print("DepthmapClusterEstimator over the batched engine", flush=True)
    sc = cfg1_scene
    v = sc.views
    refs = [(0, [1, 2, 3, 4]), (2, [0, 1, 3]), (4, [3, 2, 0])]
    ce = DepthmapClusterEstimator()
    ce.set_max_iterations(4)
    ce.set_patch_size(7)
    ce.set_top_k(3)                        # kept, no effect
    for r, nb in refs:
        idx = ce.begin_ref_view(v[r].K, v[r].R, v[r].t, v[r].image.copy(), sc.min_depth,
                                sc.max_depth)
        ce.set_sfm_points(np.zeros((5, 3)))
        for s in nb:                       # fresh copies, as dense.py loads them
            ce.add_source_view(v[s].K, v[s].R, v[s].t, v[s].image.copy())
    ce.add_geom_link(0, 1, 1)
    res = ce.run()
    assert len(res) == 3
    b = DepthmapBatch()
    ids = [b.add_view(x.K, x.R, x.t, x.image) for x in v]
    for r, nb in refs:
        b.add_job(ids[r], [ids[s] for s in nb], sc.min_depth, sc.max_depth)
    b.set_patchmatch_iterations(4)
    b.set_patch_size(7)
    b.run()
    for k, (r, _) in enumerate(refs):
        depth, normal, cost, conf = res[k]
        h, w = v[r].image.shape
        assert depth.shape == (h, w) and normal.shape == (h, w, 3)
        assert cost.shape == (h, w) and conf.shape == (h, w)
        assert all(a.dtype == np.float32 for a in res[k])
        d, plane, score, _ = b.result(k)
        wd, wn, wc, wk = oracle.cluster_readback(d, plane, score, sc.min_depth, sc.max_depth)
        assert np.array_equal(depth, wd)
        ok = depth > 0
        np.testing.assert_allclose(normal, wn, rtol=1e-6, atol=1e-6)
        assert (normal[~ok] == 0).all()
        assert np.array_equal(cost, wc)
        np.testing.assert_allclose(conf, wk, rtol=3e-7, atol=1e-7)
        assert (conf[~ok] == 0).all()
        assert (normal[ok][:, 2] < 0).mean() > 0.99           # facing the camera
        acc = scenes.depth_accuracy(depth, v[r].depth, sc.min_depth, sc.max_depth)
        print(k, round(acc, 4), flush=True)
        assert acc > 0.8
    ce.clear()
    assert ce.run() == [] and ce.num_ref_views() == 0
    assert DepthmapClusterEstimator.is_available() and DepthmapClusterEstimator.num_devices() >= 1


def test_cluster_hierarchy_levels(cfg1_scene):
    """hierarchy_levels = 3 (estimator.cc:140, 168-172: scales 1/4, 1/2, 1,
    max_iterations at each): the depths equal the batched engine's with a
    3-level pyramid of max_iterations per level; levels are dropped where
    the smallest view would be narrower than one patch."""
    print("DepthmapClusterEstimator with hierarchy levels", flush=True)
    sc = cfg1_scene
    v = sc.views
    refs = [(0, [1, 2, 3, 4]), (3, [2, 4])]
    ce = DepthmapClusterEstimator()
    ce.set_max_iterations(2)
    ce.set_patch_size(7)
    ce.set_hierarchy_levels(3)
    for r, nb in refs:
        ce.begin_ref_view(v[r].K, v[r].R, v[r].t, v[r].image, sc.min_depth, sc.max_depth)
        for s in nb:
            ce.add_source_view(v[s].K, v[s].R, v[s].t, v[s].image)
    res = ce.run()
    assert ce.last_stats["launches"] == 3 * 2 * 2
    b = DepthmapBatch()
    ids = [b.add_view(x.K, x.R, x.t, x.image) for x in v]
    for r, nb in refs:
        b.add_job(ids[r], [ids[s] for s in nb], sc.min_depth, sc.max_depth)
    b.set_patchmatch_iterations(2)
    b.set_patch_size(7)
    b.set_pyramid(2, levels=3)
    b.run()
    for k, (r, _) in enumerate(refs):
        assert np.array_equal(res[k][0], oracle.cluster_readback(*b.result(k)[:3], sc.min_depth,
                                                                 sc.max_depth)[0])
        acc = scenes.depth_accuracy(res[k][0], v[r].depth, sc.min_depth, sc.max_depth)
        print(k, round(acc, 4), flush=True)
        assert acc > 0.7  # view 3 has two sources only: 0.77
    assert ce.pyramid_levels(240, 7) == 3 and ce.pyramid_levels(20, 7) == 2
    assert ce.pyramid_levels(6, 7) == 1


@pytest.mark.gpu
def test_device_queries_and_reservations():
    """Device name and memory (depthmap_bind.h:85-100) through the C ABI,
    and the reservation bookkeeping of opencl_utils.h:68-76: available =
    global memory - reserved, releasing more than is reserved leaves zero."""
    from opensfm_amd import DepthmapClusterEstimator as CE
    n = CE.num_devices()
    assert n >= 1 and CE.is_available()
    name = CE.device_name(0)
    total = CE.device_memory_bytes(0)
    assert isinstance(name, str) and name
    assert total > 1 << 30 and CE.device_is_gpu(0)
    assert CE.device_available_memory(0) == total
    CE.reserve_device_memory(0, 5 << 30)
    CE.reserve_device_memory(0, 1 << 30)
    assert CE.device_available_memory(0) == total - (6 << 30)
    CE.release_device_memory(0, 2 << 30)
    assert CE.device_available_memory(0) == total - (4 << 30)
    CE.release_device_memory(0, 100 << 30)
    assert CE.device_available_memory(0) == total
    CE.reserve_device_memory(0, total + 1)
    assert CE.device_available_memory(0) == 0
    CE.release_device_memory(0, total + 1)
    assert CE.device_available_memory(0) == total
    with pytest.raises(IndexError):
        CE.device_name(n)
