"""Lattice queries (ebg_query_grid, ebg_scene_query_grid): the batched form of the per-node
value() calls an embedded-boundary generator makes through the reference
(Integrations/AMReX/MeshSDF/main.cpp:23-69, Integrations/Chombo/MeshSDF/main.cpp:21-50).

Bar: the lattice result equals ebg_query_sdf / ebg_scene_query on the same points made on the
host (grid_points: the FP64 coordinate rounded to FP32), in the FAB layout (x fastest): bit for
bit on lattices below 2^14 points, and within 1e-6 diag with equal signs above (there expensive
packets are shared among subtree wavefronts, and which of them are shared -- at most 1/64 of the
packets -- depends on the order in which wavefronts claim list slots, so the FP32 result can move
by an ulp); against the oracle's FP64 SDF, |d - d_ref| <= 1e-5 diag and equal signs.
"""
import numpy as np
import pytest

from ebgeometry_amd import TriMeshSDF, grid_points
from ebgeometry_amd import workloads as wl
from ebgeometry_amd._lib import EbgError
from oracle import port

pytestmark = pytest.mark.gpu


def same(a, b, diag, label):
    diff = np.abs(a.astype(np.float64) - b)
    print(f"{label}: {np.count_nonzero(diff)} of {b.size} differ, max {diff.max():.3e}", flush=True)
    if b.size < (1 << 14):
        assert np.array_equal(a, b), label
    assert diff.max() <= 1e-6 * diag, label
    assert np.array_equal(np.sign(a), np.sign(b)), label


def box_of(v, pad=0.1):
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    ext = hi - lo
    return lo - pad * ext, hi + pad * ext


@pytest.mark.parametrize("dims", [(37, 29, 23), (1, 1, 1), (64, 1, 3), (130, 120, 110)])
def test_mesh_grid_equals_point_queries(cuda, dims):
    import torch
    v, f = wl.displaced_icosphere(5, 81)
    m = TriMeshSDF(v, f)
    lo, hi = box_of(v)
    spacing = (hi - lo) / np.maximum(np.array(dims) - 1, 1)
    print(f"lattice {dims} over the 20,480-face mesh", flush=True)
    g = m.query_grid(lo, spacing, dims)
    assert g.shape == (dims[2], dims[1], dims[0])
    pts = grid_points(lo, spacing, dims)
    d = m.query(pts)
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    same(g.ravel(), d, diag, "lattice vs points")
    # device output on torch's stream
    gd = m.query_grid(lo, spacing, dims, device=cuda)
    torch.cuda.synchronize()
    same(gd.cpu().numpy().ravel(), g.ravel(), diag, "device vs host lattice")
    # FAB layout: [k, j, i] is the point origin + (i, j, k) * spacing
    k, j, i = dims[2] - 1, dims[1] // 2, dims[0] // 3
    p = (lo + np.array([i, j, k]) * spacing).astype(np.float32)
    assert g[k, j, i] == m.query(p.reshape(1, 3))[0]
    # against the oracle on a strided sample
    sel = np.arange(0, pts.shape[0], max(1, pts.shape[0] // 3000))
    ref, _ = port.Mesh(v.astype(np.float64), f).sdf(pts[sel].astype(np.float64))
    assert np.abs(g.ravel()[sel] - ref).max() <= 1e-5 * diag
    nz = ref != 0
    assert np.array_equal(np.sign(g.ravel()[sel][nz]), np.sign(ref[nz]))


def test_scene_grid_equals_point_queries(cuda):
    from ebgeometry_amd import AABB, BoxSDF, BVHUnionIF, SphereSDF
    sc = wl.csg_scene(9, n_spheres=200, n_boxes=50, n_meshes=0, extent=4.0)
    prims, bvs = [], []
    for c, r in zip(sc.sphere_center.astype(np.float64), sc.sphere_radius.astype(np.float64)):
        prims.append(SphereSDF(c, r))
        bvs.append(AABB(c - r, c + r))
    for lo, hi in zip(sc.box_lo.astype(np.float64), sc.box_hi.astype(np.float64)):
        prims.append(BoxSDF(lo, hi))
        bvs.append(AABB(lo, hi))
    u = BVHUnionIF(prims, bvs)
    dims = (41, 33, 27)
    origin, spacing = np.zeros(3), np.array([0.1, 0.125, 0.15])
    g = u.query_grid(origin, spacing, dims)
    pts = grid_points(origin, spacing, dims)
    assert np.array_equal(g.ravel(), u.value(pts))  # the scene pipeline shares no packets


def test_grid_arguments(cuda):
    v, f = wl.icosphere(2)
    m = TriMeshSDF(v, f)
    assert m.query_grid([0, 0, 0], [1, 1, 1], (0, 5, 5)).shape == (5, 5, 0)
    with pytest.raises(EbgError):
        m.query_grid([0, 0, 0], [float("nan"), 1, 1], (2, 2, 2))
    with pytest.raises(EbgError):
        m.query_grid([0, 0, 0], [1, 1, 1], (1 << 20, 1 << 12, 2))
