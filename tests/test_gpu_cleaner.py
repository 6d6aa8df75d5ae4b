"""GPU tests of GPUDepthmapCleaner (space carving + backward reprojection,
cleaner.cc:56-179): bit-exact against the oracle's float32 restatement
(oracle.gpu_clean) for same-size views, neighbours of another size (the
cv::resize path), more than 16 neighbours (truncated), and the error rules."""
import numpy as np
import pytest

import oracle
from opensfm_amd import GPUDepthmapCleaner, scenes

pytestmark = pytest.mark.gpu


def _noisy_views(scene, rng):
    out = []
    for v in scene.views:
        d = v.depth.astype(np.float32) * (1 + rng.normal(0, 0.004, v.depth.shape)).astype(np.float32)
        d[rng.random(d.shape) < 0.03] = 0
        out.append((v.K, v.R, v.t, d.astype(np.float32)))
    return out


def _gpu(views, ref, nb, **kw):
    c = GPUDepthmapCleaner()
    for k, v in kw.items():
        getattr(c, "set_" + k)(v)
    for K, R, t, d in views:
        c.add_view(K, R, t, d)
    return c.clean(ref, nb)


def test_gpu_cleaner_matches_oracle(cfg1_scene):
    print("GPUDepthmapCleaner vs oracle, config 1 views with noise", flush=True)
    rng = np.random.default_rng(3)
    views = _noisy_views(cfg1_scene, rng)
    for ref in (0, 2):
        nb = [v for v in range(len(views)) if v != ref]
        for kw in ({}, {"min_consistent_views": 2, "carving_threshold": 0.1},
                   {"max_carved_views": 3, "same_depth_threshold": 0.02}):
            got = _gpu(views, ref, nb, **kw)
            args = dict(same_depth_threshold=kw.get("same_depth_threshold", 0.01),
                        min_consistent_views=kw.get("min_consistent_views", 3),
                        carving_threshold=kw.get("carving_threshold", 0.2),
                        max_carved_views=kw.get("max_carved_views", 1))
            want = oracle.gpu_clean(views, ref, nb, **args)
            assert np.array_equal(got, want), (ref, kw)
            assert (got > 0).mean() > 0.3


def test_gpu_cleaner_resized_and_many_neighbours():
    print("GPUDepthmapCleaner: resized neighbours, 18 neighbours", flush=True)
    s = scenes.make_scene(160, 120, 18, 4, seed=21, min_depth=2.0, max_depth=20.0,
                          spacing_deg=2.0)
    rng = np.random.default_rng(4)
    views = _noisy_views(s, rng)
    # neighbour 3 given at another size (its K kept, as the reference does)
    K, R, t, d = views[3]
    views[3] = (K, R, t, oracle.resize_linear_f32(d, 200, 150))
    nb = list(range(1, 19))
    got = _gpu(views, 0, nb, min_consistent_views=4)
    want = oracle.gpu_clean(views, 0, nb, min_consistent_views=4)
    assert np.array_equal(got, want)
    c = GPUDepthmapCleaner()
    for K, R, t, d in views[:3]:
        c.add_view(K, R, t, d)
    with pytest.raises(IndexError, match="ref_idx"):
        c.clean(5, [1])
    with pytest.raises(IndexError, match="neighbor_ids"):
        c.clean(0, [1, 7])
    c.clear()
    with pytest.raises(IndexError):
        c.clean(0, [])
    assert GPUDepthmapCleaner.is_available()
