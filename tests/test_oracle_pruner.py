"""Pins of the oracle's DepthmapPruner::Prune restatement
(depthmap.cc:577-636) by closed forms: the single-view merge is the plain
back-projection of every pixel with depth not <= 0; identical views drop a
pixel exactly when the FP64 area of the other view exceeds the FP32-rounded
area of the pruned one; and occluders in front, samples behind, zero depths,
cameras facing away and projections off the image follow the reference's
comparisons, each case built so the other sign or threshold convention gives
the opposite outcome."""
import numpy as np

import oracle

W, H, F = 40, 30, 100.0


def _K(f=F, w=W, h=H):
    return np.array([[f, 0, w / 2], [0, f, h / 2], [0, 0, 1.0]])


def _rot(axis, ang):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    X = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * X + (1 - np.cos(ang)) * X @ X


def _view(rng, depth, plane=None, K=None, R=None, t=None):
    h, w = depth.shape
    if plane is None:
        plane = np.empty((h, w, 3), np.float32)
        plane[...] = (0.0, 0.0, -1.0)
    return ((_K() if K is None else K), (np.eye(3) if R is None else R),
            (np.zeros(3) if t is None else t), depth.astype(np.float32),
            np.ascontiguousarray(plane, np.float32),
            rng.integers(0, 256, (h, w, 3), dtype=np.uint8),
            rng.integers(0, 256, (h, w), dtype=np.uint8))


def _normalize(plane):
    v = plane.astype(np.float64)
    nv = np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])
    a = np.where(nv != 0, 1.0 / np.where(nv != 0, nv, 1.0), 0.0)
    return (v * a[..., None]).astype(np.float32)


def test_prune_single_view_is_backprojection():
    rng = np.random.default_rng(0)
    depth = rng.uniform(1, 5, (H, W)).astype(np.float32)
    depth[rng.random((H, W)) < 0.1] = 0
    depth[rng.random((H, W)) < 0.05] = -1
    depth[3, 7] = np.nan                          # not <= 0: merged, as NaN
    plane = rng.normal(0, 1, (H, W, 3)).astype(np.float32) * rng.uniform(0.1, 10, (H, W, 1)).astype(np.float32)
    R = _rot((1, 2, 3), 0.4)
    t = np.array([0.5, -1.0, 2.0])
    v = _view(rng, depth, plane, R=R, t=t)
    pts, nrm, col, lab = oracle.prune([v])
    idx = np.flatnonzero(~(depth.ravel() <= 0))
    assert len(pts) == idx.size
    assert np.array_equal(col, v[5].reshape(-1, 3)[idx])
    assert np.array_equal(lab, v[6].ravel()[idx])
    ii, jj = np.divmod(idx, W)
    d = depth.ravel()[idx].astype(np.float64)
    X = (R.T @ (d * (np.linalg.inv(_K()) @ np.stack([jj, ii, np.ones(idx.size)])) - t[:, None])).T
    fin = np.isfinite(d)
    assert not fin.all() and np.isnan(pts[~fin]).all()
    np.testing.assert_allclose(pts[fin], X[fin], rtol=1e-6, atol=1e-6)
    n = _normalize(plane).reshape(-1, 3)[idx].astype(np.float64)
    np.testing.assert_allclose(nrm, (R.T @ n.T).T, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1.0, rtol=1e-5)


def test_prune_identical_views_compare_fp64_area_with_fp32_area():
    rng = np.random.default_rng(1)
    depth = rng.uniform(1, 5, (H, W)).astype(np.float32)
    plane = np.stack([rng.normal(0, 0.3, (H, W)), rng.normal(0, 0.3, (H, W)),
                      -rng.uniform(0.5, 1.5, (H, W))], -1).astype(np.float32)
    v = _view(rng, depth, plane)
    pts, _, _, lab = oracle.prune([v, v])
    n2 = _normalize(plane)[..., 2]
    q = (-n2) / depth                                  # float32 divide
    assert q.dtype == np.float32
    dd = q.astype(np.float64) * F
    drop = dd > dd.astype(np.float32).astype(np.float64)
    assert 0.2 < drop.mean() < 0.8
    assert np.array_equal(lab, v[6].ravel()[~drop.ravel()])


def test_prune_occlusion_rules_closed_form():
    rng = np.random.default_rng(2)
    depth = rng.uniform(2, 4, (H, W)).astype(np.float32)
    tilted = np.empty((H, W, 3), np.float32)
    tilted[...] = (0.8, 0.0, -0.6)
    flat = np.empty((H, W, 3), np.float32)
    flat[...] = (0.0, 0.0, -2.0)
    n = H * W

    def kept(v0, v1, **kw):
        return len(oracle.prune([v0, v1], **kw)[0])

    v_tilted = _view(rng, depth, tilted)
    v_flat = _view(rng, depth, flat)
    # the other view's surface in front (0.5 z < 0.99 z): kept
    assert kept(v_tilted, _view(rng, depth * 0.5, flat)) == n
    # behind, finer (1 / 1.5z > 0.6 / z): dropped; coarser (0.6 / 1.5z < 1 / z): kept
    assert kept(v_tilted, _view(rng, depth * 1.5, flat)) == 0
    assert kept(v_flat, _view(rng, depth * 1.5, tilted)) == n
    # within the threshold band from below: 0.995 z > 0.99 z counts, 0.985 z not
    assert kept(v_tilted, _view(rng, depth * np.float32(0.995), flat)) == 0
    assert kept(v_tilted, _view(rng, depth * np.float32(0.985), flat)) == n
    assert kept(v_tilted, _view(rng, depth * np.float32(0.985), flat),
                same_depth_threshold=0.02) == 0
    # zero depth: below (1 - 0.01) z, kept; above (1 - 1.5) z < 0, dropped
    zero = np.zeros_like(depth)
    assert kept(v_flat, _view(rng, zero, flat)) == n
    assert kept(v_flat, _view(rng, zero, flat), same_depth_threshold=1.5) == 0
    # a camera facing away (z < 0) or looking elsewhere never drops
    away = _view(rng, depth * 1.5, flat, R=_rot((0, 1, 0), np.pi))
    assert kept(v_tilted, away) == n
    aside = _view(rng, depth * 1.5, flat, t=np.array([50.0, 0, 0]))
    assert kept(v_tilted, aside) == n
    # the first view that drops decides; order of the others does not matter
    assert kept(v_tilted, _view(rng, depth * 0.5, flat)) == n
    assert len(oracle.prune([v_tilted, aside, _view(rng, depth * 1.5, flat)])[0]) == 0


def test_prune_rounds_projection_to_nearest_pixel():
    """A second camera shifted so view-0 pixel j lands at j + s: x / z + 0.5
    truncates, so s = 0.4 meets pixel j and s = 0.6 pixel j + 1.  Only the
    other view's last column holds a (finer) sample, so the dropped column
    of view 0 is W - 1, then W - 2."""
    rng = np.random.default_rng(3)
    depth = np.full((H, W), 3.0, np.float32)
    tilted = np.empty((H, W, 3), np.float32)
    tilted[...] = (0.8, 0.0, -0.6)
    flat = np.empty((H, W, 3), np.float32)
    flat[...] = (0.0, 0.0, -1.0)
    v0 = _view(rng, depth, tilted)
    d1 = np.zeros((H, W), np.float32)
    d1[:, W - 1] = 4.5                      # area 1 / 4.5 > 0.6 / 3
    for s, dropped in ((0.4, W - 1), (0.6, W - 2)):
        v1 = _view(rng, d1, flat, t=np.array([s * 3 / F, 0, 0]))
        pts = oracle.prune([v0, v1])[0]
        cols = np.rint(pts[:, 0] * F / 3 + W / 2).astype(int)
        assert len(pts) == H * (W - 1)
        assert set(range(W)) - set(cols.tolist()) == {dropped}
        assert np.array_equal(np.bincount(cols, minlength=W)[np.arange(W) != dropped],
                              np.full(W - 1, H))
