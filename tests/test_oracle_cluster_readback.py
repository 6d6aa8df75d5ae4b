"""CPU pins of oracle.cluster_readback, the cluster estimator's readback
rules (estimator.cc:960-993, confidence :1022-1040), on hand-built pixels:
a non-finite plane gives depth 0, normal 0 and cost 0; depths outside
[min, max] are zeroed with their normals but keep their cost; a plane of
norm <= 1e-6 keeps its depth with a zero normal; a non-finite score costs 0;
confidence is the logistic of the cost where depth > 0."""
import numpy as np

import oracle


def test_cluster_readback_rules():
    f = np.float32
    nan, inf = np.nan, np.inf
    depth = np.array([5.0, 0.5, 50.0, 5.0, 5.0, 5.0, 0.0, 10.0], f)
    plane = np.array([[0, 0, -0.2], [0, 0, -2.0], [0, 0, -0.02], [nan, 0, -0.2],
                      [0, inf, -0.2], [0, 0, 1e-7], [0, 0, 0], [0.03, 0.04, 0.0]], f)
    score = np.array([0.9, 0.8, 0.7, 0.9, 0.9, 0.6, 0.0, nan], f)
    d, n, c, k = oracle.cluster_readback(depth, plane, score, 1.0, 40.0)
    np.testing.assert_array_equal(d, f([5, 0, 0, 0, 0, 5, 0, 10]))
    np.testing.assert_array_equal(n[0], f([0, 0, -1]))
    assert (n[1:5] == 0).all() and (n[5] == 0).all() and (n[6] == 0).all()
    np.testing.assert_allclose(n[7], [0.6, 0.8, 0.0], rtol=1e-6)
    np.testing.assert_allclose(c, [0.1, 0.2, 0.3, 0.0, 0.0, 0.4, 1.0, 0.0], rtol=1e-6)
    want = 1 / (1 + np.exp((c.astype(np.float64) - 0.5) * 5))
    np.testing.assert_allclose(k[[0, 5, 7]], want[[0, 5, 7]], rtol=1e-6)
    assert (k[[1, 2, 3, 4, 6]] == 0).all()
    # a planted mistake: keeping out-of-range depths changes pixel 1 and 2
    bad = np.where(np.isfinite(plane).all(-1) & (depth > 0), depth, 0)
    assert not np.array_equal(bad, d)
