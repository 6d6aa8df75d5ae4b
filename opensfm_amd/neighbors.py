"""Neighbour selection before the depthmaps, on the GPU.

Mirrors ``pysfm.build_covisibility_graph`` / ``pysfm.select_neighbors``
(opensfm/src/lib/sfm/python/pybind.cc:80-117; C++ in
opensfm/src/lib/sfm/dense_helpers.h:40-67 and src/dense_helpers.cc) with the
reconstruction given as plain data instead of OpenSfM's ``TracksManager`` /
``Map``: landmark coordinates, each landmark's observing shots and track id,
and the camera centres.  The returned objects have the reference's
attributes (``super_points`` with ``coord`` / ``vis`` / ``tracks``,
``edges``, ``weights``, ``shot_order``; ``best_neighbors`` /
``all_neighbors``).

Orders: super-points are numbered by their smallest landmark (input order)
and edges listed in (lo, hi) order; the reference's come from hash-map
iteration and are unspecified.  Neighbour ties go to the earlier shot of
``processable``; the angle sums are accumulated in fixed point (2^-32), so
results do not depend on the order of the device's atomics.
"""
from __future__ import annotations

import collections.abc
import ctypes as C
import dataclasses
from typing import Dict, List, Sequence

import numpy as np

from . import _capi


@dataclasses.dataclass
class SuperPoint:
    coord: np.ndarray          # (3,) float64 centroid
    vis: list                  # sorted observing shot ids (by processable order)
    tracks: list               # track ids of the merged landmarks


class SuperPointList(collections.abc.Sequence):
    """The graph's super-points as a read-only sequence of SuperPoint, built
    on access from the arrays the device returned (one Python object per
    super-point only when it is asked for; select_neighbors reads the
    arrays directly)."""

    def __init__(self, coord, moff, members, voff, vis, processable, tids):
        self._coord, self._moff, self._members = coord, moff, members
        self._voff, self._vis = voff, vis
        self._processable, self._tids = processable, tids

    def __len__(self):
        return len(self._moff) - 1

    def __getitem__(self, s):
        if isinstance(s, slice):
            return [self[k] for k in range(*s.indices(len(self)))]
        n = len(self)
        if s < 0:
            s += n
        if not 0 <= s < n:
            raise IndexError("super-point index out of range")
        return SuperPoint(coord=self._coord[s].copy(),
                          vis=[self._processable[v] for v in self._vis[self._voff[s]:self._voff[s + 1]]],
                          tracks=[self._tids[m] for m in self._members[self._moff[s]:self._moff[s + 1]]])

    def __eq__(self, other):
        if isinstance(other, (list, tuple, SuperPointList)):
            return len(self) == len(other) and all(
                np.array_equal(a.coord, b.coord) and list(a.vis) == list(b.vis)
                and list(a.tracks) == list(b.tracks) for a, b in zip(self, other))
        return NotImplemented


@dataclasses.dataclass
class CovisibilityGraph:
    super_points: Sequence[SuperPoint]
    edges: List[tuple]         # (shot_idx_i, shot_idx_j), i < j
    weights: List[float]       # shared super-points per edge
    shot_order: list           # index -> shot id
    scale: float = 0.0         # median k-nearest distance of the landmarks
    point_sp: np.ndarray = None  # per landmark: its super-point, -1 if unused
    device_ms: float = 0.0     # device time of the construction


@dataclasses.dataclass
class NeighborResult:
    best_neighbors: Dict[object, list]
    all_neighbors: Dict[object, list]
    device_ms: float = 0.0


def _csr(lists, index):
    """(offsets, processable indices) of per-item id lists; a pair of arrays
    (offsets, indices) is taken as already in that form."""
    if isinstance(lists, tuple) and len(lists) == 2 and isinstance(lists[0], np.ndarray):
        return (np.ascontiguousarray(lists[0], np.int64),
                np.ascontiguousarray(lists[1], np.int32))
    off = np.zeros(len(lists) + 1, np.int64)
    off[1:] = np.cumsum([len(x) for x in lists])
    flat = np.fromiter((index.get(s, -1) for x in lists for s in x), np.int32, int(off[-1]))
    return off, flat


class _Handle:
    def __init__(self, device: int = 0):
        self.lib = _capi.load()
        h = C.c_void_p()
        _capi.check(self.lib.pm_covis_create(int(device), C.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.pm_covis_destroy(h)
            self.h = None


def build_covisibility_graph(coords, observers: Sequence[Sequence], processable: Sequence,
                             fuse_knn: int = 15, fuse_radius_factor: float = 0.5,
                             track_ids: Sequence = None, device: int = 0) -> CovisibilityGraph:
    """BuildCovisibilityGraph (dense_helpers.cc:30-326).

    coords: (n, 3) landmark positions; observers: per landmark the ids of the
    shots observing it (ids outside ``processable`` are ignored), or a pair
    of arrays (offsets (n + 1,), processable indices) in CSR form; track_ids:
    per landmark its track id (default: its index)."""
    processable = list(processable)
    index = {sid: i for i, sid in enumerate(processable)}
    coords = np.ascontiguousarray(np.asarray(coords, np.float64).reshape(-1, 3))
    n = len(coords)
    off, shots = _csr(observers, index)
    if len(off) != n + 1:
        raise ValueError("one observer list per landmark is needed")
    tids = range(n) if track_ids is None else list(track_ids)
    h = _Handle(device)
    nsp, nv, ne = C.c_int64(0), C.c_int64(0), C.c_int64(0)
    _capi.check(h.lib.pm_covis_build(
        h.h, n, _capi.ptr(coords, C.c_double), _capi.ptr(off, C.c_int64),
        _capi.ptr(shots, C.c_int32), len(processable), int(fuse_knn), float(fuse_radius_factor),
        C.byref(nsp), C.byref(nv), C.byref(ne)))
    nsp, nv, ne = nsp.value, nv.value, ne.value
    point_sp = np.zeros(max(n, 1), np.int32)
    sp_coord = np.zeros((max(nsp, 1), 3))
    moff = np.zeros(nsp + 1, np.int64)
    members = np.zeros(max(n, 1), np.int32)
    voff = np.zeros(nsp + 1, np.int64)
    vis = np.zeros(max(nv, 1), np.int32)
    edges = np.zeros((max(ne, 1), 2), np.int32)
    weights = np.zeros(max(ne, 1), np.int32)
    scale = C.c_float(0)
    _capi.check(h.lib.pm_covis_download(
        h.h, _capi.ptr(point_sp, C.c_int32), _capi.ptr(sp_coord, C.c_double),
        _capi.ptr(moff, C.c_int64), _capi.ptr(members, C.c_int32), _capi.ptr(voff, C.c_int64),
        _capi.ptr(vis, C.c_int32), _capi.ptr(edges, C.c_int32), _capi.ptr(weights, C.c_int32),
        C.byref(scale)))
    ms = C.c_float(0)
    _capi.check(h.lib.pm_covis_last_ms(h.h, C.byref(ms)))
    sps = SuperPointList(sp_coord[:nsp], moff, members, voff, vis, processable, tids)
    return CovisibilityGraph(super_points=sps, edges=[tuple(e) for e in edges[:ne].tolist()],
                             weights=[float(w) for w in weights[:ne]],
                             shot_order=processable, scale=scale.value,
                             point_sp=point_sp[:n].copy(), device_ms=ms.value)


def select_neighbors(origins, super_points: Sequence[SuperPoint], processable: Sequence,
                     num_neighbors: int, min_point_best: int = 20, min_point_all: int = 40,
                     theta_min_deg: float = 3.0, theta_max_deg: float = 60.0,
                     device: int = 0) -> NeighborResult:
    """SelectNeighbors (dense_helpers.cc:331-512).  origins: per processable
    shot its camera centre (array (N, 3) in processable order, or a dict
    shot id -> centre)."""
    processable = list(processable)
    index = {sid: i for i, sid in enumerate(processable)}
    if isinstance(origins, dict):
        origins = [origins[sid] for sid in processable]
    org = np.ascontiguousarray(np.asarray(origins, np.float64).reshape(-1, 3))
    if len(org) != len(processable):
        raise ValueError("one camera centre per processable shot is needed")
    nsp = len(super_points)
    if isinstance(super_points, SuperPointList) and list(super_points._processable) == processable:
        # a graph's own super-points: its arrays, visibility already as
        # indices into the same shot order
        coords = np.ascontiguousarray(super_points._coord, np.float64).reshape(-1, 3)
        voff = np.ascontiguousarray(super_points._voff, np.int64)
        vis = np.ascontiguousarray(super_points._vis[:max(int(voff[-1]), 1)], np.int32)
        ntr = np.ascontiguousarray(np.diff(super_points._moff), np.int32)
    else:
        coords = np.ascontiguousarray(
            np.asarray([sp.coord for sp in super_points], np.float64).reshape(-1, 3))
        voff, vis = _csr([sp.vis for sp in super_points], index)
        ntr = np.ascontiguousarray(np.asarray([len(sp.tracks) for sp in super_points],
                                              np.int32).reshape(-1))
    h = _Handle(device)
    nb, na = C.c_int64(0), C.c_int64(0)
    _capi.check(h.lib.pm_covis_select_neighbors(
        h.h, len(processable), _capi.ptr(org, C.c_double), nsp, _capi.ptr(coords, C.c_double),
        _capi.ptr(voff, C.c_int64), _capi.ptr(vis, C.c_int32), _capi.ptr(ntr, C.c_int32),
        int(num_neighbors), int(min_point_best), int(min_point_all), float(theta_min_deg),
        float(theta_max_deg), C.byref(nb), C.byref(na)))
    N = len(processable)
    boff = np.zeros(N + 1, np.int64)
    aoff = np.zeros(N + 1, np.int64)
    best = np.zeros(max(nb.value, 1), np.int32)
    allv = np.zeros(max(na.value, 1), np.int32)
    _capi.check(h.lib.pm_covis_download_neighbors(
        h.h, _capi.ptr(boff, C.c_int64), _capi.ptr(best, C.c_int32), _capi.ptr(aoff, C.c_int64),
        _capi.ptr(allv, C.c_int32)))
    ms = C.c_float(0)
    _capi.check(h.lib.pm_covis_last_ms(h.h, C.byref(ms)))
    return NeighborResult(
        device_ms=ms.value,
        best_neighbors={processable[i]: [processable[v] for v in best[boff[i]:boff[i + 1]]]
                        for i in range(N)},
        all_neighbors={processable[i]: [processable[v] for v in allv[aoff[i]:aoff[i + 1]]]
                       for i in range(N)})
