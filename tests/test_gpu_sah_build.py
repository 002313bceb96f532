"""The SAH build on the GPU (Build.SAH_GPU: pseudo-normals, binned-SAH tree and packed records
computed on the device) against the host SAH build and the reference.

Its splits are the host builder's (same bins, same cost, same arithmetic), so the K-wide tree has
the same node count, leaf count and stack need, and a walk visits the same interior nodes per
query (the order of triangles inside a leaf may differ: its partitions are stable, the host's are
not). Its results meet the mesh bar against the reference's golden vectors and the oracle.
"""
import numpy as np
import pytest

from ebgeometry_amd import Build, MeshSDF, TriMeshSDF
from ebgeometry_amd import workloads as wl
from ebgeometry_amd._lib import EbgError
from oracle import port
from test_gpu_mesh import FIXTURES, check_parity, diag_of

pytestmark = pytest.mark.gpu


def _work(m, q, cuda):
    import torch
    w = torch.zeros(8, dtype=torch.int64, device=cuda)
    m.query(torch.from_numpy(q).to(cuda), work=w.data_ptr())
    torch.cuda.synchronize()
    return w.cpu().numpy()


@pytest.mark.parametrize("level,seed,K,leaf", [(3, 1, 4, 4), (5, 2, 4, 16), (6, 3, 8, 8), (7, 1, 4, 16)])
def test_same_tree_as_host_sah(cuda, level, seed, K, leaf):
    v, f = wl.displaced_icosphere(level, seed)
    q = wl.mixed_queries(20_000, v, f, seed + 10)
    host = TriMeshSDF(v, f, build=Build.SAH, branching=K, leaf_size=leaf)
    dev = TriMeshSDF(v, f, build=Build.SAH_GPU, branching=K, leaf_size=leaf)
    hi, di = host.info(), dev.info()
    for key in ("num_nodes", "num_leaves", "stack_need", "num_triangles"):
        assert hi[key] == di[key], (key, hi[key], di[key])
    assert np.array_equal(np.sort(dev.leaf_order()), np.arange(f.shape[0]))
    wh, wd = _work(host, q, cuda), _work(dev, q, cuda)
    assert wh[0] == wd[0], ("interior nodes visited", wh[0], wd[0])
    dh, dd = host.query(q), dev.query(q)
    diag = diag_of(v)
    assert np.abs(dh.astype(np.float64) - dd).max() <= 1e-6 * diag
    big = np.abs(dh) > 1e-6 * diag
    assert np.array_equal(np.sign(dh[big]), np.sign(dd[big]))
    print(f"L{level} K{K} leaf {leaf}: {di['num_nodes']} nodes; build {hi['build_seconds'] * 1e3:.1f} ms on the "
          f"host, {di['build_seconds'] * 1e3:.1f} ms on the device", flush=True)


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_parity_device_build(golden, cuda, name):
    g = golden(name)
    m = TriMeshSDF(g["verts"], g["faces"], build=Build.SAH_GPU, branching=4, leaf_size=4)
    d, tri = m.closest_triangle(g["queries"])
    check_parity(g["verts"], g["faces"], g["queries"], g["d"], g["tri"], d, tri, f"{name} device build")


def test_large_fixture_and_oracle_device_build(golden, cuda):
    g = golden("ico7_displaced_sample.npz")
    v, f = wl.displaced_icosphere(int(g["level"]), int(g["mesh_seed"]))
    m = TriMeshSDF(v, f, build=Build.SAH_GPU, leaf_size=16)
    d, tri = m.closest_triangle(g["queries"])
    check_parity(v, f, g["queries"], g["d"], g["tri"], d, tri, "L7 displaced, device build")
    v, f = wl.displaced_icosphere(6, 77)
    q = wl.mixed_queries(40_000, v, f, 78)
    d_ref, t_ref = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    d, t = TriMeshSDF(v, f, build=Build.SAH_GPU, branching=8, leaf_size=8).closest_triangle(q)
    check_parity(v, f, q, d_ref, t_ref, d, t, "L6 displaced, device build K8")


def test_small_meshes_and_errors(cuda):
    # one triangle (a leaf root) and a tetrahedron (one small subtree)
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    f = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32)
    q = np.random.default_rng(0).normal(size=(500, 3)).astype(np.float32)
    a = TriMeshSDF(v, f, build=Build.SAH_GPU, leaf_size=1).query(q)
    b = TriMeshSDF(v, f).query(q)
    assert np.abs(a - b).max() <= 1e-7
    TriMeshSDF(v, f[:1], build=Build.SAH_GPU).query(q)
    bad = f.copy()
    bad[2] = [0, 3, 3]
    with pytest.raises(EbgError, match="degenerate triangle 2"):
        TriMeshSDF(v, bad, build=Build.SAH_GPU)


def test_polygon_mesh_device_build(golden, cuda):
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import poly_cases as pc
    g = golden("polymesh.npz")
    for name in ("star_prism", "grid_cube"):
        V, F = pc.mesh(name)
        Q = g[f"{name}.queries"]
        d = MeshSDF(V, F, build=Build.SAH_GPU).query(Q)
        ref = g[f"{name}.d"]
        assert np.abs(d - ref).max() <= 1e-5 * diag_of(V)
        assert np.array_equal(np.sign(d[ref != 0]), np.sign(ref[ref != 0]))
