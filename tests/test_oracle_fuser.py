"""Pins of the oracle's DepthmapFuser::Fuse restatement (fuser.cc:113-480,
one thread) by closed forms: Eigen's cofactor inverse, single-view
back-projection with the margin / mask / depth / normal rules, fused flags
(a view fused into another emits nothing), the asymmetric depth test, the
normal-angle test, the upper median, and the outlier removal against
scipy's k-d tree and a lattice's known neighbour distances."""
import numpy as np
import pytest

import oracle

W, H, F = 48, 36, 60.0


def _K():
    return np.array([[F, 0, W / 2 - 0.5], [0, F, H / 2 - 0.5], [0, 0, 1.0]])


def _rot(axis, ang):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    X = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * X + (1 - np.cos(ang)) * X @ X


def _flat(depth=4.0, normal=(0.0, 0.0, -1.0), color=100):
    d = np.full((H, W), depth, np.float32)
    n = np.empty((H, W, 3), np.float32)
    n[...] = normal
    c = np.full((H, W, 3), color, np.uint8)
    return d, n, c


def _view(d, n, c, nb, mask=None, primary=True, K=None, R=None, t=None):
    return (_K() if K is None else K, np.eye(3) if R is None else R,
            np.zeros(3) if t is None else t, d, n, c, mask, nb, primary)


def test_inverse_is_eigen_cofactor_form():
    differ = 0
    for f in np.linspace(500, 1500, 41).astype(np.float32):
        K = np.array([[f, 0, 311.7], [0, f, 240.2], [0, 0, 1]], np.float32)
        inv = oracle.inverse3f(K)
        assert inv[0, 0] == f * (np.float32(1) / (f * f))   # cofactor * (1 / det)
        assert inv[0, 2] == (-np.float32(311.7) * f) * (np.float32(1) / (f * f))
        differ += inv[0, 0] != np.float32(1) / f
    assert differ > 0                                       # so 1 / f would be caught
    rng = np.random.default_rng(0)
    for _ in range(20):
        M = rng.normal(0, 1, (3, 3)).astype(np.float32) + 3 * np.eye(3, dtype=np.float32)
        np.testing.assert_allclose(oracle.inverse3f(M), np.linalg.inv(M.astype(np.float64)),
                                   rtol=1e-5, atol=1e-6)


def test_single_view_emits_every_valid_pixel_in_the_margin():
    rng = np.random.default_rng(1)
    d = rng.uniform(2, 6, (H, W)).astype(np.float32)
    d[rng.random((H, W)) < 0.1] = 0
    n = rng.normal(0, 1, (H, W, 3)).astype(np.float32)
    n[rng.random((H, W)) < 0.05] = 0                        # |n| < 1e-6: skipped
    c = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    mask = (rng.random((H, W)) > 0.1).astype(np.uint8) * 255
    R = _rot((1, -2, 0.5), 0.3)
    t = np.array([0.3, 0.1, -2.0])
    for bm in (0, 5):
        (pts, nrm, col), nf = oracle.fuse([_view(d, n, c, [], mask, R=R, t=t)],
                                          min_num_consistent=1, border_margin=bm, sor_knn=0)
        box = np.zeros((H, W), bool)
        box[bm:H - bm, bm:W - bm] = True
        ok = box & (mask > 0) & (d > 0) & (np.linalg.norm(n, axis=-1) >= 1e-6)
        idx = np.flatnonzero(ok.ravel())
        assert len(pts) == nf == idx.size
        ii, jj = np.divmod(idx, W)
        X = (R.T @ (d.ravel()[idx] * (np.linalg.inv(_K()) @ np.stack([jj, ii, np.ones(idx.size)]))
                    - t[:, None])).T
        np.testing.assert_allclose(pts, X, rtol=2e-5, atol=2e-5)
        nn = n.reshape(-1, 3)[idx].astype(np.float64)
        nn /= np.linalg.norm(nn, axis=1, keepdims=True)
        np.testing.assert_allclose(nrm, (R.T @ nn.T).T, rtol=1e-4, atol=1e-5)
        assert np.array_equal(col, c.reshape(-1, 3)[idx])


def test_fused_pixels_are_not_emitted_again():
    d, n, c = _flat(4.0)
    views = [_view(d, n, c, [1]), _view(d, n, c, [0])]
    (pts, _, _), nf = oracle.fuse(views, min_num_consistent=2, sor_knn=0, border_margin=2)
    assert nf == (H - 4) * (W - 4)                            # view 1 fully fused by view 0
    (pts, _, _), nf = oracle.fuse(views, min_num_consistent=3, sor_knn=0, border_margin=2)
    assert nf == 0                                            # 2 observations < 3
    (pts, _, _), nf = oracle.fuse([views[0], views[1][:8] + (False,)], min_num_consistent=2,
                                  sor_knn=0, border_margin=2)
    assert nf == (H - 4) * (W - 4)                            # non-primary: source only


def test_fused_flags_of_reference_and_source_pixels():
    """View 0 fuses with source view 2, marking its own pixels and view 2's.
    View 1 (same camera) then finds its only candidates, view 0's pixels,
    already fused, so it emits nothing; with view 2 as its neighbour, the
    used source pixels are fused likewise."""
    d, n, c = _flat(4.0)
    nbox = (H - 4) * (W - 4)
    for nb1 in ([0], [2]):
        views = [_view(d, n, c, [2]), _view(d, n, c, nb1), _view(d, n, c, [], primary=False)]
        _, nf = oracle.fuse(views, min_num_consistent=2, sor_knn=0, border_margin=2)
        assert nf == nbox, nb1
    # view 1 after view 0 but with a fresh source: both emit
    views = [_view(d, n, c, [2]), _view(d, n, c, [3]), _view(d, n, c, [], primary=False),
             _view(d, n, c, [], primary=False)]
    _, nf = oracle.fuse(views, min_num_consistent=2, sor_knn=0, border_margin=2)
    assert nf == 2 * nbox


@pytest.mark.parametrize("delta,behind,ok", [(0.009, 0.3, True), (0.011, 0.3, False),
                                             (-0.002, 0.3, True), (-0.004, 0.3, False),
                                             (-0.009, 2.0, True), (-0.004, 0.0, False)])
def test_depth_test_is_asymmetric(delta, behind, ok):
    d0, n, c = _flat(4.0)
    d1, _, _ = _flat(np.float32(4.0 * (1 + delta)))
    views = [_view(d0, n, c, [1]), _view(d1, n, c, [], primary=False)]
    _, nf = oracle.fuse(views, min_num_consistent=2, sor_knn=0, border_margin=2,
                        behind_depth_factor=behind)
    assert nf == ((H - 4) * (W - 4) if ok else 0)


@pytest.mark.parametrize("deg,ok", [(9.0, True), (11.0, False)])
def test_normal_angle_test(deg, ok):
    d, n0, c = _flat(4.0)
    a = np.deg2rad(deg)
    _, n1, _ = _flat(4.0, normal=(np.sin(a), 0.0, -np.cos(a)))
    views = [_view(d, n0, c, [1]), _view(d, n1, c, [], primary=False)]
    _, nf = oracle.fuse(views, min_num_consistent=2, sor_knn=0, border_margin=2)
    assert nf == ((H - 4) * (W - 4) if ok else 0)
    _, nf = oracle.fuse(views, min_num_consistent=2, sor_knn=0, border_margin=2,
                        max_normal_error=deg + 1)
    assert nf == (H - 4) * (W - 4)


def test_medians_take_the_upper_middle():
    d, n, _ = _flat(4.0)
    cols = (10, 200, 50, 90)
    scale = (1.0, 1.002, 1.004, 0.999)
    views = []
    for k, (cv, sc) in enumerate(zip(cols, scale)):
        dd = (d * np.float32(sc)).astype(np.float32)
        views.append(_view(dd, n, np.full((H, W, 3), cv, np.uint8),
                           list(range(1, 4)) if k == 0 else [], primary=(k == 0)))
    (pts, _, col), nf = oracle.fuse(views[:3], min_num_consistent=3, sor_knn=0, border_margin=2)
    assert nf == (H - 4) * (W - 4)
    assert (col == 50).all()                                  # sorted 10, 50, 200
    # z of the median point: the middle depth 4 * 1.002
    np.testing.assert_allclose(pts[:, 2], 4 * 1.002, rtol=1e-6)
    v0 = views[0][:7] + ([1, 2, 3],) + (True,)
    (pts, _, col), nf = oracle.fuse([v0] + views[1:], min_num_consistent=4, sor_knn=0,
                                    border_margin=2)
    assert (col == 90).all()                                  # sorted 10, 50, 90, 200: index 2
    np.testing.assert_allclose(pts[:, 2], 4 * 1.002, rtol=1e-6)   # depths .999, 1, 1.002, 1.004


def test_sor_against_kdtree_and_lattice():
    scipy_spatial = pytest.importorskip("scipy.spatial")
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(np.arange(40.0), np.arange(30.0)), -1).reshape(-1, 2)
    pts = np.concatenate([g, np.zeros((len(g), 1))], 1).astype(np.float32) * np.float32(0.25)
    out = rng.uniform(-50, 50, (20, 3)).astype(np.float32)
    allp = np.concatenate([pts, out])
    nrm = np.zeros_like(allp)
    col = np.zeros(allp.shape, np.uint8)
    (kp, _, _), mk = oracle.sor(allp, nrm, col, knn=12, factor=2.5)
    # scipy's exact k nearest (the point itself first)
    dist, _ = scipy_spatial.cKDTree(allp.astype(np.float64)).query(allp.astype(np.float64), 13)
    np.testing.assert_allclose(mk, dist[:, 1:].mean(1), rtol=1e-5)
    # interior lattice points: 4 at h, 4 at h sqrt 2, 4 at 2h
    interior = ((g[:, 0] >= 2) & (g[:, 0] <= 37) & (g[:, 1] >= 2) & (g[:, 1] <= 27))
    np.testing.assert_allclose(mk[:len(g)][interior], 0.25 * (4 + 4 * np.sqrt(2) + 8) / 12,
                               rtol=1e-6)
    assert len(kp) == len(g)                                  # the 20 far points go
    assert np.array_equal(kp, pts)
    # too few points: nothing removed
    (kp, _, _), _ = oracle.sor(allp[:13], nrm[:13], col[:13], knn=12)
    assert len(kp) == 13
