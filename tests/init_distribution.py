"""Checks of RandomInitialization's distributions (depthmap.cc:237-256,
UniformRand :127-132), shared by the CPU and GPU tests.

The reference draws, per interior pixel, depth = exp(U(log min, log max))
and normal = (U(-1, 1), U(-1, 1), -1), and in sample mode a neighbour view
uniform in [1, N-1] (uni_, depthmap.cc:141, 247).  Given the initial maps
(no iterations, no post-process), these functions return the
Kolmogorov-Smirnov distance of each against its law; the plane's slopes are
recovered as (-v0 / v2, -v1 / v2) of the plane vector v.
"""
from __future__ import annotations

import numpy as np
from scipy import stats


def ks_log_depth(depth, min_depth, max_depth):
    """KS distance of log(depth) over live pixels from U(log min, log max)."""
    d = depth[depth > 0].astype(np.float64)
    u = (np.log(d) - np.log(min_depth)) / (np.log(max_depth) - np.log(min_depth))
    return float(stats.kstest(u, "uniform").statistic), int(d.size)


def ks_slopes(depth, plane):
    """KS distances of the two slopes over live pixels from U(-1, 1)."""
    live = depth > 0
    v = plane[live].astype(np.float64)
    s0 = -v[:, 0] / v[:, 2]
    s1 = -v[:, 1] / v[:, 2]
    d0 = stats.kstest((s0 + 1) / 2, "uniform").statistic
    d1 = stats.kstest((s1 + 1) / 2, "uniform").statistic
    return float(d0), float(d1), float(np.corrcoef(s0, s1)[0, 1])


def chi2_views(nghbr, depth, n_views):
    """p-value of a chi-square test that the live pixels' views are uniform
    over 1..n_views-1."""
    counts = np.bincount(nghbr[depth > 0], minlength=n_views)[1:]
    return float(stats.chisquare(counts).pvalue), counts


def ks_bound(n, alpha=1e-4):
    """Asymptotic KS critical distance for n samples at level alpha."""
    return float(np.sqrt(-0.5 * np.log(alpha / 2) / n))
