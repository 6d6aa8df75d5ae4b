"""Host logic of the cluster estimator surface (no device): the mapping of
the reference's hierarchy_levels (estimator.cc:140, 168-172: N levels at
scales 1 / 2^(N - 1) ... 1) onto the engine's pyramid, dropping levels at
which the smallest view would be narrower than one patch."""
from opensfm_amd.cluster import DepthmapClusterEstimator


def _levels(hierarchy_levels, min_side, patch):
    ce = DepthmapClusterEstimator.__new__(DepthmapClusterEstimator)
    ce._params = dict(hierarchy_levels=hierarchy_levels)
    return ce.pyramid_levels(min_side, patch)


def test_levels_follow_hierarchy_levels():
    assert _levels(1, 1000, 7) == 1
    assert _levels(4, 1000, 7) == 4
    assert _levels(4, 480, 11) == 4        # 60 px at 1/8
    assert _levels(12, 100000, 3) == 8     # at most 8 levels
    assert _levels(0, 1000, 7) == 1


def test_levels_dropped_for_small_views():
    # 120 >> 3 = 15 >= 7; 60 >> 3 = 7 >= 7; 56 >> 3 = 7; 55 >> 3 = 6 < 7
    assert _levels(4, 120, 7) == 4
    assert _levels(4, 56, 7) == 4
    assert _levels(4, 55, 7) == 3
    assert _levels(4, 20, 7) == 2
    assert _levels(4, 13, 7) == 1
    assert _levels(3, 40, 11) == 2         # 40 >> 2 = 10 < 11, 40 >> 1 = 20
