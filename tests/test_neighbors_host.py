"""Host-side logic of neighbour selection (no device): the covisibility
graph's super-points as a lazy sequence over the arrays the device returns."""
import numpy as np
import pytest

from opensfm_amd.neighbors import SuperPoint, SuperPointList


def _list():
    coord = np.arange(9.0).reshape(3, 3)
    moff = np.array([0, 2, 3, 5])
    members = np.array([4, 1, 0, 2, 3], np.int32)
    voff = np.array([0, 1, 3, 4])
    vis = np.array([1, 0, 2, 1], np.int32)
    return SuperPointList(coord, moff, members, voff, vis, ["a", "b", "c"], [10, 11, 12, 13, 14])


def test_super_point_list_items():
    L = _list()
    assert len(L) == 3
    assert np.array_equal(L[0].coord, [0.0, 1.0, 2.0])
    assert L[0].vis == ["b"] and L[0].tracks == [14, 11]
    assert L[1].vis == ["a", "c"] and L[1].tracks == [10]
    assert L[-1].tracks == [12, 13]
    assert [sp.vis for sp in L[1:]] == [["a", "c"], ["b"]]
    with pytest.raises(IndexError):
        L[3]
    # items are copies: changing one leaves the sequence as it was
    L[0].coord[0] = 99.0
    assert L[0].coord[0] == 0.0


def test_super_point_list_equality():
    L = _list()
    assert L == [L[0], L[1], L[2]]
    assert not (L == [L[0], L[1]])
    other = [SuperPoint(np.array([0.0, 1.0, 2.0]), ["b"], [14, 11]), L[1], L[2]]
    assert L == other
    other[0] = SuperPoint(np.array([0.0, 1.0, 2.5]), ["b"], [14, 11])
    assert not (L == other)
    empty = SuperPointList(np.zeros((0, 3)), np.array([0]), np.zeros(1, np.int32),
                           np.array([0]), np.zeros(1, np.int32), [], [])
    assert len(empty) == 0 and empty == []
