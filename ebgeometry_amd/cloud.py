"""Point-cloud nearest-neighbour queries on the GPU, behind the reference's interface.

    PointCloudBVH(positions, metadata=None, leaf_size)    Source/EBGeometry_PointCloudBVH.hpp:56-161
    PointCloudHashGrid(positions, metadata=None, ...)     Source/EBGeometry_PointCloudHashGrid.hpp:54-135
        closest_point(q)          -> Hit(index, distance_squared)      (closestPoint)
        closest_points(q, k)      -> list of Hit, nearest first         (closestPoints)
        nearest_neighbor(i)       -> Hit of the nearest other point     (nearestNeighbor)
        nearest_neighbors(i, k)   -> list of Hit                        (nearestNeighbors)
        all_nearest_neighbors(k)  -> (index, distance_squared) arrays   (allNearestNeighbors)
        knn(points, k)            -> (index, distance_squared) arrays of many query points at once

Both classes answer the same queries with the same results (the k nearest points by squared
distance, nearest first; exact ties in any order), so on the GPU they share one structure: a
K-wide tree over the points walked by the packet kernel (libebgpu.so, ebg_cloud_*). The hash
grid's cell parameters are accepted and recorded but do not change the answer.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import BuildOpts, QueryOpts, check, lib

NONE = 0xFFFFFFFF


@dataclass
class Hit:
    """One query result: the cloud index of a matched point and its squared distance (the
    reference's Hit)."""
    index: int
    distance_squared: float


class PointCloudBVH:
    def __init__(self, positions, metadata=None, leaf_size: int = 8, branching: int = 4):
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        h = C.c_void_p()
        opts = BuildOpts(0, int(branching), int(leaf_size), 0)
        check(lib().ebg_cloud_create(p.ctypes.data, p.shape[0], C.byref(opts), C.byref(h)))
        self._h = h
        self._positions = p
        self._metadata = list(metadata) if metadata is not None else None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().ebg_cloud_destroy(h)
            except Exception:  # pragma: no cover - interpreter teardown
                pass
            self._h = None

    def num_points(self) -> int:
        return self._positions.shape[0]

    def position(self, i: int) -> np.ndarray:
        return self._positions[i]

    def metadata(self, i: int):
        return self._metadata[i] if self._metadata is not None else i

    @staticmethod
    def _opts(presort=True, stream=None) -> QueryOpts:
        o = QueryOpts()
        lib().ebg_query_opts_default(C.byref(o))
        o.presort, o.stream = int(presort), stream
        return o

    def knn(self, points, k: int = 1, **kw):
        """(index, distance_squared) of the k nearest cloud points of every query point: arrays
        (n, k), nearest first. numpy in -> numpy out; a CUDA float32 tensor -> CUDA tensors
        (enqueued on torch's current stream)."""
        from .sdf import _is_tensor, _torch
        if _is_tensor(points):
            torch = _torch()
            q = points.contiguous().view(-1, 3)
            idx = torch.empty((q.shape[0], k), dtype=torch.int32, device=q.device)
            d2 = torch.empty((q.shape[0], k), dtype=torch.float32, device=q.device)
            kw.setdefault("stream", torch.cuda.current_stream(q.device).cuda_stream)
            check(lib().ebg_cloud_knn(self._h, q.data_ptr(), q.shape[0], int(k), idx.data_ptr(), d2.data_ptr(),
                                      C.byref(self._opts(**kw))))
            return idx, d2
        q = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        idx = np.empty((q.shape[0], k), dtype=np.uint32)
        d2 = np.empty((q.shape[0], k), dtype=np.float32)
        check(lib().ebg_cloud_knn_host(self._h, q.ctypes.data, q.shape[0], int(k), idx.ctypes.data, d2.ctypes.data,
                                       C.byref(self._opts(**kw))))
        return idx, d2

    def all_nearest_neighbors(self, k: int = 1):
        """(index, distance_squared) arrays (num_points, k): for every point its k nearest other
        points, nearest first (the k-nearest-neighbour graph)."""
        n = self.num_points()
        idx = np.empty((n, k), dtype=np.uint32)
        d2 = np.empty((n, k), dtype=np.float32)
        check(lib().ebg_cloud_knn_self_host(self._h, int(k), idx.ctypes.data, d2.ctypes.data, None))
        return idx, d2

    @staticmethod
    def _hits(idx_row, d2_row):
        return [Hit(int(i), float(d)) for i, d in zip(idx_row, d2_row) if i != NONE]

    def closest_point(self, q) -> Hit:
        i, d = self.knn(np.asarray(q, np.float32).reshape(1, 3), 1)
        return Hit(int(i[0, 0]), float(d[0, 0]))

    def closest_points(self, q, k: int):
        i, d = self.knn(np.asarray(q, np.float32).reshape(1, 3), k)
        return self._hits(i[0], d[0])

    def nearest_neighbor(self, i: int) -> Hit:
        hits = self.nearest_neighbors(i, 1)
        return hits[0] if hits else Hit(NONE, float("inf"))

    def nearest_neighbors(self, i: int, k: int):
        if k + 1 > 16:  # the whole graph's row (the single query below would need k + 1 entries)
            idx, d2 = self.all_nearest_neighbors(k)
            return self._hits(idx[i], d2[i])
        # the query is the cloud point itself, its own index dropped from k + 1 results
        idx, d2 = self.knn(self._positions[i:i + 1], k + 1)
        return [h for h in self._hits(idx[0], d2[0]) if h.index != i][:k]

    def check(self):
        check(lib().ebg_cloud_check(self._h))


class PointCloudHashGrid(PointCloudBVH):
    """The reference's uniform-grid point cloud (PointCloudHashGrid.hpp:54-135): the same queries
    and results as PointCloudBVH, answered by the same GPU structure."""

    def __init__(self, positions, metadata=None, target_per_cell: float = 1.0):
        super().__init__(positions, metadata)
        self.target_per_cell = float(target_per_cell)
