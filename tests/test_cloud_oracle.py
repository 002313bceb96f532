"""CPU checks of the nearest-neighbour oracle (oracle/cloud_ref.py, brute force in FP64) against
the compiled reference's PointCloudBVH (fixture tests/golden/cloud_knn.npz) and a closed form."""
import hashlib
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests"))

import cloud_cases as cc  # noqa: E402
from oracle import cloud_ref  # noqa: E402


def same_up_to_ties(idx_a, d2_a, idx_b, d2_b, rel=0.0):
    """Equal sorted distances at every rank; indices equal except among equal distances."""
    assert np.allclose(d2_a, d2_b, rtol=rel, atol=0.0)
    for r in np.nonzero((idx_a != idx_b).any(1))[0]:
        for d in np.unique(d2_b[r]):
            ga, gb = idx_a[r][np.isclose(d2_a[r], d, rtol=rel, atol=0)], idx_b[r][np.isclose(d2_b[r], d, rtol=rel, atol=0)]
            # the last distance group may be cut by k differently: compare what fits
            m = min(len(ga), len(gb))
            if d < d2_b[r].max():
                assert sorted(ga) == sorted(gb), (r, d)
            else:
                assert m == len(ga) == len(gb)


@pytest.fixture(scope="module")
def g():
    return load_golden("cloud_knn.npz")


@pytest.mark.parametrize("k", [1, 4, 16])
def test_brute_force_matches_reference_closest_points(g, k):
    P, Q = cc.cloud(), cc.queries()
    assert hashlib.sha256(P.tobytes()).digest() == g["cloud_sha256"].tobytes()
    assert hashlib.sha256(Q.tobytes()).digest() == g["queries_sha256"].tobytes()
    i, d = cloud_ref.knn(P, Q, k)
    same_up_to_ties(i, d, g[f"knn{k}.index"], g[f"knn{k}.d2"])


@pytest.mark.parametrize("k", [1, 8])
def test_brute_force_matches_reference_all_nearest_neighbors(g, k):
    P = cc.cloud()
    i, d = cloud_ref.knn(P, P, k, exclude=np.arange(P.shape[0]))
    same_up_to_ties(i, d, g[f"self{k}.index"], g[f"self{k}.d2"])
    assert not (g[f"self{k}.index"] == np.arange(P.shape[0])[:, None]).any()


def test_lattice_closed_form():
    """On the integer lattice the nearest other points of an interior point are its 6 face
    neighbours at d^2 = 1, then 12 edge neighbours at d^2 = 2."""
    g3 = np.stack(np.meshgrid(*[np.arange(5.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    c = int(np.nonzero((g3 == 2).all(1))[0][0])
    i, d = cloud_ref.knn(g3, g3[c:c + 1], 19, exclude=[c])
    assert np.array_equal(d[0, :6], np.ones(6)) and np.array_equal(d[0, 6:18], np.full(12, 2.0)) and d[0, 18] == 3.0
