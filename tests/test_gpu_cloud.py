"""GPU parity of the point-cloud nearest-neighbour queries (ebg_cloud_*) against the compiled
reference's PointCloudBVH (fixture tests/golden/cloud_knn.npz) and the FP64 brute force
(oracle/cloud_ref.py).

Bar: at every rank the GPU's squared distance equals the reference's within FP32 rounding
(relative 1e-6; 0 for coincident points), and the returned index is the reference's except
where the two candidates' FP64 distances are that close (near-ties and exact ties); the FP64
distance of every returned index equals the returned d^2 within the same tolerance; self queries
never return the query itself.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests"))

import cloud_cases as cc  # noqa: E402
from ebgeometry_amd import PointCloudBVH, PointCloudHashGrid  # noqa: E402
from oracle import cloud_ref  # noqa: E402

pytestmark = pytest.mark.gpu
REL = 1e-6


def check_knn(P, Q, idx, d2, ref_d2, label, exclude=None):
    idx = idx.astype(np.int64)
    assert (idx >= 0).all() and (idx < P.shape[0]).all(), label
    if exclude is not None:
        assert not (idx == np.asarray(exclude)[:, None]).any(), f"{label}: a query returned itself"
    # the GPU's own distances are the FP64 distances of the indices it returned, up to rounding
    d64 = ((P[idx].astype(np.float64) - Q.astype(np.float64)[:, None, :]) ** 2).sum(-1)
    assert np.allclose(d2, d64, rtol=REL, atol=1e-12), label
    # rank by rank they are the reference's k smallest (ties and near-ties may swap indices)
    assert np.allclose(d64, ref_d2, rtol=REL, atol=1e-12), (label, np.abs(d64 - ref_d2).max())
    assert (np.diff(d2, axis=1) >= 0).all(), f"{label}: not nearest first"


@pytest.fixture(scope="module")
def g():
    return load_golden("cloud_knn.npz")


@pytest.mark.parametrize("k", [1, 4, 16])
@pytest.mark.parametrize("branching,leaf", [(4, 8), (8, 16), (4, 1)])
def test_closest_points_match_reference(g, cuda, k, branching, leaf):
    P, Q = cc.cloud(), cc.queries()
    c = PointCloudBVH(P, leaf_size=leaf, branching=branching)
    idx, d2 = c.knn(Q, k)
    check_knn(P, Q, idx, d2, g[f"knn{k}.d2"], f"knn k={k} K{branching} L{leaf}")
    c.check()


@pytest.mark.parametrize("k", [1, 8])
def test_all_nearest_neighbors_match_reference(g, cuda, k):
    P = cc.cloud()
    c = PointCloudBVH(P)
    idx, d2 = c.all_nearest_neighbors(k)
    check_knn(P, P, idx, d2, g[f"self{k}.d2"], f"self k={k}", exclude=np.arange(P.shape[0]))


def test_single_point_api_and_hash_grid(cuda):
    P, Q = cc.cloud(), cc.queries(50, seed=9)
    c, h = PointCloudBVH(P), PointCloudHashGrid(P)
    ref_i, ref_d = cloud_ref.knn(P, Q, 4)
    for j in range(Q.shape[0]):
        hit = c.closest_point(Q[j])
        assert hit.distance_squared == pytest.approx(ref_d[j, 0], rel=REL, abs=1e-12)
        hits = h.closest_points(Q[j], 4)
        assert np.allclose([x.distance_squared for x in hits], ref_d[j], rtol=REL, atol=1e-12)
    for i in (0, 17, 3000, P.shape[0] - 1):  # the last points duplicate earlier ones: d^2 = 0
        nn = c.nearest_neighbors(i, 5)
        ri, rd = cloud_ref.knn(P, P[i:i + 1], 5, exclude=[i])
        assert all(x.index != i for x in nn)
        assert np.allclose([x.distance_squared for x in nn], rd[0], rtol=REL, atol=1e-12)
        assert c.nearest_neighbor(i).distance_squared == pytest.approx(rd[0, 0], rel=REL, abs=1e-12)


def test_large_cloud_against_brute_force(cuda):
    """200k points, 20k queries (device tensors): against the FP64 brute force on a sample."""
    import torch
    rng = np.random.default_rng(12)
    P = rng.normal(size=(200_000, 3)).astype(np.float32)
    Q = rng.normal(size=(20_000, 3)).astype(np.float32) * 1.2
    c = PointCloudBVH(P, leaf_size=16)
    idx, d2 = c.knn(torch.from_numpy(Q).to(cuda), 8)
    torch.cuda.synchronize()
    idx, d2 = idx.cpu().numpy(), d2.cpu().numpy()
    sel = np.arange(0, Q.shape[0], 10)
    ri, rd = cloud_ref.knn(P, Q[sel], 8)
    check_knn(P, Q[sel], idx[sel], d2[sel], rd, "200k cloud")


def test_fewer_points_than_k(cuda):
    P = cc.cloud()[:5]
    c = PointCloudBVH(P)
    idx, d2 = c.knn(cc.queries(10), 8)
    assert (idx[:, 5:] == 0xFFFFFFFF).all() and np.isinf(d2[:, 5:]).all()
    i2, d22 = c.all_nearest_neighbors(6)
    assert (i2[:, 4:] == 0xFFFFFFFF).all()
