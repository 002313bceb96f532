"""GPU Morton BVH build (the reference's Build::Morton, Source/EBGeometry_BVHImplem.hpp:207-312).

The triangle order must be the stable sort of the Morton codes of the triangle box centroids,
with the codes bit-exact against the oracle (pinned to the reference's SFC codes); the tree it
builds must answer exactly like the SAH tree (the tree shape changes only the visit order)."""
import numpy as np
import pytest

from ebgeometry_amd import Build, TriMeshSDF
from ebgeometry_amd import workloads as wl
from oracle import port

pytestmark = pytest.mark.gpu


def centroid_order(v, f):
    x = np.asarray(v, np.float32).astype(np.float64)[np.asarray(f, np.int64)]  # (nf, 3, 3)
    cen = 0.5 * (x.min(1) + x.max(1))
    codes = port.morton_encode(port.bins_f64(cen))
    return port.sort_codes(codes), codes


@pytest.mark.parametrize("level", [1, 3, 5])
@pytest.mark.parametrize("K", [4, 8])
def test_leaf_order_is_stable_morton_sort(cuda, level, K):
    v, f = wl.displaced_icosphere(level, 17)
    m = TriMeshSDF(v, f, build=Build.Morton, branching=K, leaf_size=1)
    order, _ = centroid_order(v, f)
    assert np.array_equal(m.leaf_order(), order)
    info = m.info()
    depth = int(np.floor(np.log(f.shape[0]) / np.log(K)))
    assert info["num_leaves"] == K ** depth


@pytest.mark.parametrize("name", ["dodecahedron.npz", "tetrahedron.npz", "ico3_displaced.npz"])
@pytest.mark.parametrize("K,leaf", [(4, 1), (8, 1), (4, 4), (8, 8)])
def test_morton_tree_parity_on_fixtures(golden, cuda, name, K, leaf):
    g = golden(name)
    m = TriMeshSDF(g["verts"], g["faces"], build=Build.Morton, branching=K, leaf_size=leaf)
    d = m.signed_distance(g["queries"]).astype(np.float64)
    diag = float(np.linalg.norm(g["verts"].max(0).astype(np.float64) - g["verts"].min(0)))
    assert np.abs(d - g["d"]).max() <= 1e-5 * diag
    nz = g["d"] != 0
    assert np.array_equal(np.sign(d[nz]), np.sign(g["d"][nz]))


def test_morton_tree_equals_sah_tree_large(cuda):
    import torch
    v, f = wl.displaced_icosphere(6, 3)
    q = torch.from_numpy(wl.mixed_queries(200_000, v, f, 4)).to(cuda)
    a = TriMeshSDF(v, f, build=Build.SAH).query(q)
    b = TriMeshSDF(v, f, build=Build.Morton, leaf_size=1).query(q)
    c = TriMeshSDF(v, f, build=Build.Morton, branching=8, leaf_size=4).query(q)
    torch.cuda.synchronize()
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    assert (a - b).abs().max().item() <= 1e-6 * diag
    assert (a - c).abs().max().item() <= 1e-6 * diag
    assert torch.equal(torch.sign(a), torch.sign(b)) and torch.equal(torch.sign(a), torch.sign(c))
