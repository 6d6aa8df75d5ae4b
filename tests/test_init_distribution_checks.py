"""CPU pins of the RandomInitialization distribution checks
(tests/init_distribution.py): samples drawn exactly as depthmap.cc:237-256
describes pass, and each planted mistake (linear instead of log-uniform
depth, the wrong depth range, slopes over [0, 1] or [-2, 2], correlated
slopes, a view range missing one view) fails."""
import numpy as np

from init_distribution import chi2_views, ks_bound, ks_log_depth, ks_slopes

N = 300_000
MN, MX = 1.0, 40.0


def _maps(rng, depth_fn, slope_lo=-1.0, slope_hi=1.0, corr=False):
    depth = depth_fn(rng, N).astype(np.float32)
    s0 = rng.uniform(slope_lo, slope_hi, N)
    s1 = s0.copy() if corr else rng.uniform(slope_lo, slope_hi, N)
    n = np.stack([s0, s1, -np.ones(N)], axis=1)
    # plane v = n / (-n . X), X = depth * ray; a ray along the axis is enough
    # here: the slopes are recovered from ratios of v
    X = np.stack([np.zeros(N), np.zeros(N), depth], axis=1)
    v = n / (-(n * X).sum(1))[:, None]
    return depth.reshape(1, -1), v.astype(np.float32).reshape(1, -1, 3)


def _log_uniform(mn, mx):
    return lambda rng, n: np.exp(rng.uniform(np.log(mn), np.log(mx), n))


def test_reference_law_passes():
    rng = np.random.default_rng(0)
    depth, plane = _maps(rng, _log_uniform(MN, MX))
    d, n = ks_log_depth(depth, MN, MX)
    assert d < ks_bound(n)
    d0, d1, c = ks_slopes(depth, plane)
    assert d0 < ks_bound(n) and d1 < ks_bound(n) and abs(c) < 0.01


def test_linear_depth_fails():
    rng = np.random.default_rng(1)
    depth, _ = _maps(rng, lambda r, n: r.uniform(MN, MX, n))
    d, n = ks_log_depth(depth, MN, MX)
    assert d > 10 * ks_bound(n)


def test_wrong_depth_range_fails():
    rng = np.random.default_rng(2)
    depth, _ = _maps(rng, _log_uniform(MN, 0.9 * MX))
    d, n = ks_log_depth(depth, MN, MX)
    assert d > ks_bound(n)


def test_wrong_slope_law_fails():
    rng = np.random.default_rng(3)
    for lo, hi in ((0.0, 1.0), (-2.0, 2.0)):
        depth, plane = _maps(rng, _log_uniform(MN, MX), lo, hi)
        d0, d1, _ = ks_slopes(depth, plane)
        assert d0 > ks_bound(N) and d1 > ks_bound(N)
    depth, plane = _maps(rng, _log_uniform(MN, MX), corr=True)
    assert ks_slopes(depth, plane)[2] > 0.99


def test_view_law():
    rng = np.random.default_rng(4)
    depth = np.ones(N, np.float32)
    p, _ = chi2_views(rng.integers(1, 9, N), depth, 9)
    assert p > 1e-4
    p, _ = chi2_views(rng.integers(1, 8, N), depth, 9)
    assert p < 1e-12
