"""GPU parity of MeshSDF over polygon faces (ebg_mesh_create_polygons) against the compiled
reference's MeshSDF<double, uint32_t, K> (fixture tests/golden/polymesh.npz) and the FP64
restatement of its face algorithm (oracle/poly_ref.py).

Bar: |d - d_ref| <= 1e-5 diag with equal signs wherever d_ref != 0; the closest face equals the
reference's getClosestFaces(p, true) first entry except where the two faces' FP64 distances tie
(within 1e-12 diag, the rounding of the restatement), and the count of such ties is reported.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests"))

import poly_cases as pc  # noqa: E402
from ebgeometry_amd import Build, MeshSDF  # noqa: E402
from oracle import poly_ref  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    return load_golden("polymesh.npz")


def _diag(V):
    V = np.asarray(V, np.float64)
    return float(np.linalg.norm(V.max(0) - V.min(0)))


def check(V, F, Q, d, face, ref_d, ref_face, label):
    diag = _diag(V)
    err = float(np.abs(d.astype(np.float64) - ref_d).max())
    assert err <= 1e-5 * diag, (label, err / diag)
    nz = ref_d != 0
    assert np.array_equal(np.sign(d[nz]), np.sign(ref_d[nz])), label
    if face is not None:
        fd = poly_ref.face_distances(V, F, Q)
        m = np.arange(Q.shape[0])
        diff = face.astype(np.int64) != ref_face
        tied = np.abs(fd[m, face] - fd[m, ref_face]) <= 1e-12 * diag
        assert (tied | ~diff).all(), f"{label}: {int((diff & ~tied).sum())} faces differ beyond ties"
        print(f"{label}: max |d - d_ref| / diag = {err / diag:.2e}; {int(diff.sum())} of {Q.shape[0]} faces "
              f"differ from the reference's, all exact ties")


@pytest.mark.parametrize("name", pc.NAMES)
@pytest.mark.parametrize("branching,leaf", [(4, 4), (8, 16)])
def test_polygon_mesh_matches_reference(g, cuda, name, branching, leaf):
    V, F = pc.mesh(name)
    Q = g[f"{name}.queries"]
    m = MeshSDF(V, F, branching=branching, leaf_size=leaf)
    d, face = m.closest_face(Q)
    check(V, F, Q, d, face, g[f"{name}.d"], g[f"{name}.face"].astype(np.int64), f"{name} K{branching} L{leaf}")
    # the distance-only query (no face decision) meets the same bar
    check(V, F, Q, m.query(Q), None, g[f"{name}.d"], None, f"{name} K{branching} L{leaf} distance only")
    assert m.triangle_faces().max() == len(F) - 1


@pytest.mark.parametrize("name", ["dodecahedron", "star_prism", "grid_cube"])
def test_morton_build_and_device_tensors(g, cuda, name):
    import torch
    V, F = pc.mesh(name)
    Q = g[f"{name}.queries"]
    m = MeshSDF(V, F, build=Build.Morton)
    d, face = m.closest_face(torch.from_numpy(Q).to(cuda))
    torch.cuda.synchronize()
    check(V, F, Q, d.cpu().numpy(), face.cpu().numpy().view(np.uint32).astype(np.int64), g[f"{name}.d"],
          g[f"{name}.face"].astype(np.int64), f"{name} Morton")


def test_lattice_and_large_batch_against_restatement(cuda):
    """A lattice over the star prism and 200k random points on the grid cube (presorted, split
    packets) against the FP64 restatement."""
    V, F = pc.mesh("star_prism")
    m = MeshSDF(V, F)
    lo = V.min(0).astype(np.float64) - 0.2
    sp = np.array([0.05, 0.05, 0.04])
    dims = (int((V[:, 0].max() - lo[0] + 0.2) / sp[0]) + 1, int((V[:, 1].max() - lo[1] + 0.2) / sp[1]) + 1, 30)
    grid = m.query_grid(lo, sp, dims)
    from ebgeometry_amd import grid_points
    P = grid_points(lo, sp, dims)
    ref, _ = poly_ref.sdf(V, F, P)
    check(V, F, P, grid.reshape(-1), None, ref, None, "star lattice")

    V, F = pc.mesh("grid_cube")
    m = MeshSDF(V, F, leaf_size=8)
    rng = np.random.default_rng(4)
    Q = (rng.random((200_000, 3)) * 1.6 - 0.3).astype(np.float32)
    d = m.query(Q)
    sel = np.arange(0, Q.shape[0], 40)
    ref, _ = poly_ref.sdf(V, F, Q[sel])
    check(V, F, Q[sel], d[sel], None, ref, None, "grid cube 200k")
