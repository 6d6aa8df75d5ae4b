"""GPU tests of DepthmapFuser (fuser.cc:113-480): bit-exact against the
oracle's one-thread restatement (oracle.fuse) -- the round-based parallel
greedy must reproduce the sequential fused-flag order exactly -- on noisy
multi-view depthmaps with the pipeline's neighbour lists (every view of the
cluster, itself included), with masks, source-only views, views of another
size, invalid and repeated neighbour ids, and the outlier removal on
clouds with far outliers and duplicates; plus the surface's rules."""
import numpy as np
import pytest

import oracle
from opensfm_amd import DepthmapFuser, scenes

pytestmark = pytest.mark.gpu


def _normals(K, depth):
    h, w = depth.shape
    jj, ii = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    P = (np.linalg.inv(K) @ np.stack([jj.ravel(), ii.ravel(), np.ones(jj.size)])).T
    P = (P * depth.reshape(-1, 1)).reshape(h, w, 3)
    n = np.cross(np.gradient(P, axis=1), np.gradient(P, axis=0))
    n[n[..., 2] > 0] *= -1
    n[np.linalg.norm(n, axis=-1) == 0] = (0, 0, -1)
    return n.astype(np.float32)


def _views(scene, rng, noise=0.002, holes=0.02, masks=False):
    out = []
    n = len(scene.views)
    for v in scene.views:
        d = v.depth.astype(np.float32) * (1 + rng.normal(0, noise, v.depth.shape)).astype(np.float32)
        d[rng.random(d.shape) < holes] = 0
        d = d.astype(np.float32)
        col = np.stack([v.image, 255 - v.image, v.image // 3], -1).astype(np.uint8)
        m = (rng.random(d.shape) > 0.05).astype(np.uint8) * 255 if masks else None
        out.append([v.K, v.R, v.t, d, _normals(v.K, d), col, m, list(range(n)), True])
    return out


def _gpu(views, **params):
    f = DepthmapFuser()
    for k, v in params.items():
        if k == "sor_knn":
            f.set_sor_params(v, params.get("sor_stddev_factor", 2.5))
        elif k != "sor_stddev_factor":
            getattr(f, "set_" + k)(v)
    for v in views:
        f.add_view(*v)
    return f.fuse(), f


def _check(views, **params):
    (gp, gn, gc), f = _gpu(views, **params)
    (op, on, oc), nf = oracle.fuse(views, **params)
    st = f.stats()
    assert st["fused_before_sor"] == nf
    assert gp.shape == op.shape, (gp.shape, op.shape)
    assert np.array_equal(gp, op)
    assert np.array_equal(gn, on)
    assert np.array_equal(gc, oc)
    return gp, st


def test_fuser_matches_oracle_cluster_neighbours():
    print("DepthmapFuser vs oracle: 6 views 160x120, every view a neighbour", flush=True)
    s = scenes.make_scene(160, 120, 5, 4, seed=31, min_depth=2.0, max_depth=20.0,
                          spacing_deg=4.0)
    views = _views(s, np.random.default_rng(1))
    for params in ({"sor_knn": 0}, {"sor_knn": 0, "min_num_consistent": 4},
                   {"sor_knn": 0, "border_margin": 0, "behind_depth_factor": 1.0},
                   {"sor_knn": 12}):
        pts, st = _check(views, **params)
        assert len(pts) > 1000, params
        print(params, len(pts), st, flush=True)


def test_fuser_masks_sources_sizes_and_ids():
    print("DepthmapFuser vs oracle: masks, source-only views, another size, odd ids", flush=True)
    s = scenes.make_scene(128, 96, 5, 4, seed=41, min_depth=2.0, max_depth=20.0,
                          spacing_deg=5.0)
    rng = np.random.default_rng(2)
    views = _views(s, rng, masks=True)
    views[2][8] = False                      # source only
    views[4][7] = [4, 0, -1, 99, 2, 2]       # self, invalid and repeated ids
    # view 3 at half size, K scaled
    K, R, t, d, n, c, m, nb, pr = views[3]
    Kh = K.copy()
    Kh[:2] *= 0.5
    Kh[:2, 2] = (K[:2, 2] + 0.5) * 0.5 - 0.5
    views[3] = [Kh, R, t, np.ascontiguousarray(d[::2, ::2]), np.ascontiguousarray(n[::2, ::2]),
                np.ascontiguousarray(c[::2, ::2]), np.ascontiguousarray(m[::2, ::2]), nb, pr]
    _check(views, sor_knn=0, min_num_consistent=2)
    _check(views, sor_knn=0, max_reproj_error=0.5, max_normal_error=20.0,
           max_depth_error=0.02)


def test_fuser_many_observations():
    print("DepthmapFuser vs oracle: 18 views, more than 12 observations per pixel", flush=True)
    s = scenes.make_scene(96, 72, 17, 3, seed=61, min_depth=4.0, max_depth=20.0,
                          spacing_deg=1.0)
    views = _views(s, np.random.default_rng(5), noise=0.0005, holes=0.0)
    for v, sv in zip(views, s.views):
        v[4] = scenes.normals_from_depth(v[0], sv.depth)
    pts, st = _check(views, sor_knn=0, min_num_consistent=13, border_margin=2)
    assert len(pts) > 100
    pts, st = _check(views, sor_knn=0, min_num_consistent=2, border_margin=2)


def test_fuser_larger_views():
    print("DepthmapFuser vs oracle: 8 views 480x360", flush=True)
    s = scenes.make_scene(480, 360, 7, 6, seed=51, min_depth=2.0, max_depth=20.0,
                          spacing_deg=3.0)
    views = _views(s, np.random.default_rng(3))
    pts, st = _check(views, sor_knn=0, min_num_consistent=4, border_margin=10)
    print(len(pts), st, flush=True)
    assert len(pts) > 10000


def test_outlier_removal_matches_brute_force():
    print("DepthmapFuser outlier removal vs brute force", flush=True)
    rng = np.random.default_rng(4)
    h, w = 90, 120
    K = np.array([[100.0, 0, w / 2], [0, 100.0, h / 2], [0, 0, 1]])
    # a surface with far outliers, exact duplicates and a dense cluster
    d = (5 + np.sin(np.arange(w) / 9.0)[None, :] + 0.3 * np.cos(np.arange(h) / 7.0)[:, None])
    d = d.astype(np.float32)
    far = rng.random((h, w)) < 0.01
    d[far] = rng.uniform(30, 300, far.sum()).astype(np.float32)
    d[10, 10:20] = d[10, 10]                 # a run of equal depths
    d[40:60, 40:60] = np.float32(5.0)        # flat patch
    n = np.zeros((h, w, 3), np.float32)
    n[..., 2] = -1
    c = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    views = [[K, np.eye(3), np.zeros(3), d, n, c, None, [], True]]
    for knn, fac in ((12, 2.5), (5, 1.0), (32, 0.5), (1, 3.0)):
        pts, st = _check(views, min_num_consistent=1, border_margin=0, sor_knn=knn,
                         sor_stddev_factor=fac)
        assert 0 < len(pts) < st["fused_before_sor"], (knn, fac)
    # duplicates: two views on the same pixels with neighbours, min 1
    views2 = [views[0], [K, np.eye(3), np.zeros(3), d, n, c, None, [], True]]
    _check(views2, min_num_consistent=1, border_margin=0, sor_knn=12)


def test_fuser_rules():
    f = DepthmapFuser()
    assert f.fuse()[0].shape == (0, 3)                  # no views
    d = np.full((20, 30), 4.0, np.float32)
    n = np.zeros((20, 30, 3), np.float32)
    n[..., 2] = -1
    c = np.zeros((20, 30, 3), np.uint8)
    K = np.array([[30.0, 0, 15], [0, 30.0, 10], [0, 0, 1]])
    with pytest.raises(ValueError, match="depth and normal"):
        f.add_view(K, np.eye(3), np.zeros(3), d, n[:10], c, None, [])
    with pytest.raises(ValueError, match="depth and color"):
        f.add_view(K, np.eye(3), np.zeros(3), d, n, c[:, :5], None, [])
    f.add_view(K, np.eye(3), np.zeros(3), d, n, c, np.zeros((0, 0), np.uint8), [1], True)
    f.add_view(K, np.eye(3), np.zeros(3), d, n, c, None, [0], False)
    f.set_sor_params(0, 2.5)
    f.set_min_num_consistent(2)
    f.set_border_margin(2)
    first = f.fuse()[0]
    assert len(first) == 16 * 26
    assert len(f.fuse()[0]) == 0                        # flags persist, as in the reference
    g = DepthmapFuser()
    for k in range(66):
        g.add_view(K, np.eye(3), np.zeros(3), d, n, c, None, list(range(66)), k == 0)
    with pytest.raises(ValueError, match="neighbours"):
        g.fuse()
