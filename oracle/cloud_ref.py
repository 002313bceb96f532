"""Brute-force k nearest neighbours in FP64 (numpy) -- test infrastructure only.

Restates what PointCloudBVH::closestPoints / allNearestNeighbors return
(Source/EBGeometry_PointCloudBVHImplem.hpp:235-507): the k cloud points of smallest squared
distance, nearest first, excluding the query's own index for self queries; exact ties in index
order (the reference's order among exact ties depends on its traversal, so comparisons treat
them as interchangeable). Pinned in tests/test_cloud_oracle.py against the compiled reference's
fixture and against closed-form lattices.
"""
from __future__ import annotations

import numpy as np


def knn(points, queries, k, exclude=None, chunk=2048):
    """(index (n, k) int64, d2 (n, k) float64); exclude: per-query index to skip, or None."""
    P = np.asarray(points, np.float64).reshape(-1, 3)
    Q = np.asarray(queries, np.float64).reshape(-1, 3)
    n = Q.shape[0]
    kk = min(k, P.shape[0] - (1 if exclude is not None else 0))
    idx = np.full((n, k), -1, np.int64)
    d2 = np.full((n, k), np.inf)
    for s in range(0, n, chunk):
        q = Q[s:s + chunk]
        D = ((q[:, None, :] - P[None, :, :]) ** 2).sum(-1)
        if exclude is not None:
            D[np.arange(q.shape[0]), np.asarray(exclude)[s:s + chunk]] = np.inf
        o = np.argsort(D, axis=1, kind="stable")[:, :kk]
        idx[s:s + chunk, :kk] = o
        d2[s:s + chunk, :kk] = np.take_along_axis(D, o, 1)
    return idx, d2
