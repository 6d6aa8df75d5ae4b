"""Pins of the oracle's neighbour selection (dense_helpers.cc:30-512):
the covisibility graph against an independent numpy/scipy construction
(float32 vlfeat distances, stable (distance, index) k nearest, connected
components), exact duplicates merging into super-points with the union of
their shots, the >= 2 processable observers rule; the neighbour scores on
cameras placed at known angles (the [3, 60] degree band, the > 20 / > 40
count rules, score order and index ties)."""
import numpy as np
import pytest

import oracle


def _cloud(rng, n, n_shots):
    pts = rng.normal(0, 1, (n, 3))
    obs = [sorted(rng.choice(n_shots, rng.integers(1, 6), replace=False).tolist())
           for _ in range(n)]
    return pts, obs


def test_covis_against_numpy_and_scipy():
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    sparse = pytest.importorskip("scipy.sparse")
    rng = np.random.default_rng(0)
    n_shots = 14
    pts, obs = _cloud(rng, 600, n_shots)
    obs[3] = [2, 99, -1]                   # one processable observer: dropped
    obs[4] = [5, 5 + n_shots]              # likewise
    for knn, fac in ((15, 0.5), (4, 1.0), (30, 0.25)):
        g = oracle.covis(pts, obs, n_shots, fuse_knn=knn, fuse_radius_factor=fac)
        keep = np.array([sum(0 <= s < n_shots for s in o) >= 2 for o in obs])
        assert (g["point_sp"][~keep] == -1).all() and (g["point_sp"][keep] >= 0).all()
        idx = np.flatnonzero(keep)
        f = pts[idx].astype(np.float32)
        m = len(f)
        k = max(1, min(knn, m - 1))
        d = f[:, None, :] - f[None, :, :]
        d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        np.fill_diagonal(d2, np.inf)
        order = np.argsort(d2, axis=1, kind="stable")[:, :k]
        nd2 = np.take_along_axis(d2, order, 1)
        kdist = np.sqrt(nd2[:, k // 2])
        scale = np.partition(kdist, m // 2)[m // 2]
        assert g["scale"] == scale
        r2 = np.float32((fac * float(scale)) ** 2)
        rows, cols = [], []
        for i in range(m):
            for s in range(k):
                j = order[i, s]
                if j > i and nd2[i, s] <= r2:
                    rows.append(i)
                    cols.append(j)
        A = sparse.coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(m, m))
        ncomp, lab = csgraph.connected_components(A, directed=False)
        assert len(g["sp_coord"]) == ncomp
        # same partition: the oracle numbers components by smallest member
        sp = g["point_sp"][idx]
        first = {}
        for i in range(m):
            first.setdefault(lab[i], len(first))
        assert np.array_equal(sp, np.array([first[lab[i]] for i in range(m)]))
        for s in range(ncomp):
            mem = g["sp_members"][s]
            assert np.array_equal(mem, idx[sp == s])
            np.testing.assert_allclose(g["sp_coord"][s], pts[mem].mean(0), rtol=1e-12, atol=1e-12)
            shots = sorted({o for i in mem for o in obs[i] if 0 <= o < n_shots})
            assert g["sp_vis"][s].tolist() == shots
        W = {}
        for v in g["sp_vis"]:
            for a in range(len(v)):
                for b in range(a + 1, len(v)):
                    W[(v[a], v[b])] = W.get((v[a], v[b]), 0) + 1
        want = sorted(W)
        assert [tuple(e) for e in g["edges"].tolist()] == want
        assert g["weights"].tolist() == [W[e] for e in want]


def test_covis_merges_exact_duplicates():
    rng = np.random.default_rng(1)
    n_shots = 9
    base, obs = _cloud(rng, 200, n_shots)
    obs = [o + [(o[0] + 1) % n_shots] for o in obs]       # >= 2 observers each
    pts = np.concatenate([base, base])
    obs2 = obs + [[(o[0] + 3) % n_shots, (o[0] + 4) % n_shots] for o in obs]
    g = oracle.covis(pts, obs2, n_shots, fuse_knn=15, fuse_radius_factor=1e-6)
    assert len(g["sp_coord"]) == 200
    for s in range(200):
        assert g["sp_members"][s].tolist() == [s, s + 200]
        assert np.array_equal(g["sp_coord"][s], (pts[s] + pts[s + 200]) / 2)
        assert g["sp_vis"][s].tolist() == sorted(set(obs2[s]) | set(obs2[s + 200]))


def _ring(angles_deg, dist=1.0):
    a = np.deg2rad(np.asarray(angles_deg, np.float64))
    return np.stack([dist * np.cos(a), dist * np.sin(a), np.zeros_like(a)], 1)


def test_select_neighbors_angles_and_counts():
    org = _ring([0, 10, 90, 2, 45])
    sp = np.zeros((1, 3))
    best, allv = oracle.select_neighbors(org, sp, [[0, 1, 2, 3, 4]], [50], 3,
                                         min_point_best=20, min_point_all=40)
    # shot 0: angles 10 (good), 90 (beyond 60), 2 (below 3), 45 (good)
    assert best[0] == [0, 1, 4]           # equal scores: lower index first
    assert allv[0] == [0, 1, 4, 2, 3]     # then every shot with count 50 > 40
    # shot 2 at 90: 80 to shot 1, 88 to shot 3, 45 to shot 4
    assert best[2] == [2, 4]
    # count rules: exactly 20 is not > 20, exactly 40 is not > 40
    b, a = oracle.select_neighbors(org, sp, [[0, 1, 2, 3, 4]], [20], 3)
    assert b[0] == [0] and a[0] == [0]
    b, a = oracle.select_neighbors(org, sp, [[0, 1, 2, 3, 4]], [40], 3)
    assert b[0] == [0, 1, 4] and a[0] == [0, 1, 4]
    # scores decide the order; num_neighbors truncates
    sp2 = np.zeros((3, 3))
    b, a = oracle.select_neighbors(org, sp2, [[0, 4], [0, 1], [0, 1]], [30, 30, 30], 1,
                                   min_point_best=20, min_point_all=40)
    assert b[0] == [0, 1]                  # 60 tracks with shot 1 beat 30 with shot 4
    assert a[0] == [0, 1]                  # shot 4: 30 tracks, not > 40
    # shots outside [0, N) are ignored, a single camera scores nothing
    b, a = oracle.select_neighbors(org, sp, [[0, 7, -1]], [50], 3)
    assert b == [[i] for i in range(5)]
