"""CPU checks of the polygon-mesh path (MeshSDF over polygon faces).

1. The FP64 restatement (oracle/poly_ref.py), which follows the reference's DCEL face algorithm
   with no triangulation, equals the compiled reference's MeshSDF (fixture
   tests/golden/polymesh.npz) on every mesh of tests/poly_cases.py.
2. The library's host clipping of polygons into triangles (ebg_polygons_to_triangles_host) covers
   every face exactly (n-2 triangles, same winding, triangle areas summing to the face area), and
   gives each triangle the polygon mesh's normals (the restatement's, to 1e-14).
3. The triangles with those normals, through the oracle's triangle SDF (the reference's lane
   formula, oracle/ebg_oracle.c), give the reference's polygon signed distances: the identity the
   GPU path rests on.
4. Faces the clipping cannot take are refused with a message.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests"))

import poly_cases as pc  # noqa: E402
from ebgeometry_amd import polygons_to_triangles  # noqa: E402
from ebgeometry_amd._lib import EbgError  # noqa: E402
from oracle import poly_ref, port  # noqa: E402


@pytest.fixture(scope="module")
def g():
    return load_golden("polymesh.npz")


def _diag(V):
    V = np.asarray(V, np.float64)
    return float(np.linalg.norm(V.max(0) - V.min(0)))


@pytest.mark.parametrize("name", pc.NAMES)
def test_fixture_matches_cases(g, name):
    V, F = pc.mesh(name)
    from oracle.ref import flatten_faces
    off, idx = flatten_faces(F)
    assert np.array_equal(g[f"{name}.verts"], V)
    assert np.array_equal(g[f"{name}.offsets"], off) and np.array_equal(g[f"{name}.indices"], idx)
    assert np.array_equal(g[f"{name}.queries"], pc.queries(V, 3000, seed=0))


@pytest.mark.parametrize("name", pc.NAMES)
def test_restatement_matches_reference(g, name):
    V, F = pc.mesh(name)
    Q = g[f"{name}.queries"]
    d, face = poly_ref.sdf(V, F, Q)
    diag = _diag(V)
    assert np.abs(d - g[f"{name}.d"]).max() <= 1e-12 * diag
    nz = g[f"{name}.d"] != 0
    assert np.array_equal(np.sign(d[nz]), np.sign(g[f"{name}.d"][nz]))
    # the reference's closest face is at the minimum distance (ties in any order)
    fd = poly_ref.face_distances(V, F, Q)
    m = np.arange(Q.shape[0])
    assert np.allclose(fd[m, g[f"{name}.face"]], fd.min(1), rtol=0, atol=1e-12 * diag)
    assert np.allclose(g[f"{name}.dist"], fd.min(1), rtol=0, atol=1e-12 * diag)


def _area_vec(P):
    """Vector area of a closed polygon (n, 3)."""
    return 0.5 * np.cross(P, np.roll(P, -1, axis=0)).sum(0)


@pytest.mark.parametrize("name", pc.NAMES)
def test_clipping_covers_faces_with_polygon_normals(name):
    V, F = pc.mesh(name)
    tri, tf, fn, vn, en = polygons_to_triangles(V, F)
    Vd = V.astype(np.float64)
    assert tri.shape[0] == sum(len(f) - 2 for f in F)
    pfn, _, pen, pvn = poly_ref.normals(Vd, F)
    for f, face in enumerate(F):
        mine = np.nonzero(tf == f)[0]
        assert len(mine) == len(face) - 2
        assert set(tri[mine].reshape(-1)) <= set(face)
        A = _area_vec(Vd[face])
        tri_area = sum(_area_vec(Vd[tri[t]]) for t in mine)
        assert np.allclose(tri_area, A, rtol=0, atol=1e-12 * max(1.0, np.linalg.norm(A)))
        for t in mine:
            # same winding as the face; each triangle non-degenerate
            assert _area_vec(Vd[tri[t]]) @ A > 0
            assert np.allclose(fn[t], pfn[f], rtol=0, atol=1e-14)
            for k in range(3):
                assert np.allclose(vn[t, k], pvn[tri[t, k]], rtol=0, atol=1e-14)
                a, b = tri[t, k], tri[t, (k + 1) % 3]
                i = face.index(int(a))
                if face[(i + 1) % len(face)] == b:  # a face edge: its half-edge normal
                    assert np.allclose(en[t, k], pen[f][i], rtol=0, atol=1e-14)
                else:  # a diagonal inside the face: the face normal
                    assert np.allclose(en[t, k], pfn[f], rtol=0, atol=1e-14)


@pytest.mark.parametrize("name", [n for n in pc.NAMES if n != "grid_cube"])
def test_clipped_triangles_give_reference_distance(g, name):
    V, F = pc.mesh(name)
    tri, tf, fn, vn, en = polygons_to_triangles(V, F)
    Vd = V.astype(np.float64)
    Q = g[f"{name}.queries"][::6].astype(np.float64)
    ref = g[f"{name}.d"][::6]
    D = np.array([[port.triangle_sdf(np.ascontiguousarray(Vd[t].reshape(-1)), np.ascontiguousarray(fn[i]),
                                     np.ascontiguousarray(vn[i].reshape(-1)), np.ascontiguousarray(en[i].reshape(-1)),
                                     np.ascontiguousarray(q)) for i, t in enumerate(tri)] for q in Q])
    d = D[np.arange(Q.shape[0]), np.argmin(np.abs(D), axis=1)]
    # float vertices of a pentagon are coplanar only to rounding: plane offsets through a corner
    # and through the centroid differ by ~1e-9 diag
    assert np.abs(d - ref).max() <= 1e-8 * _diag(V)
    nz = ref != 0
    assert np.array_equal(np.sign(d[nz]), np.sign(ref[nz]))


def test_grid_cube_triangles_through_oracle_mesh(g):
    """The 3,072-triangle clipping of the 1,536-quad cube through the oracle's BVH mesh SDF (its
    own normals of the triangles equal the polygon mesh's for planar faces, up to rounding)."""
    V, F = pc.mesh("grid_cube")
    tri, *_ = polygons_to_triangles(V, F)
    d, _ = port.Mesh(V.astype(np.float64), tri).sdf(g["grid_cube.queries"].astype(np.float64))
    assert np.abs(d - g["grid_cube.d"]).max() <= 1e-12 * _diag(V)


@pytest.mark.parametrize("faces,msg", [
    ([[0, 1, 2, 3]], "not planar"),           # a twisted quad
    ([[0, 1, 4, 2]], "not a simple polygon"),  # a bow tie
    ([[0, 1]], "fewer than 3"),
    ([[0, 1, 2, 1]], "repeats a vertex"),
    ([[0, 1, 9]], "out of range"),
])
def test_bad_faces_are_refused(faces, msg):
    V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.3], [0, 1, 0]], np.float32)
    with pytest.raises(EbgError, match=msg):
        polygons_to_triangles(V, faces)
