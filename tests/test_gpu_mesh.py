"""GPU parity of the mesh signed-distance path against the reference and the oracle.

Bar (from the reference's own checks, Tests/TestBVH.cpp:112-241, restated for FP32):
  |d_gpu - d_ref| <= 1e-5 * diag, sign equal wherever d_ref != 0, and the returned face is the
  reference's closest face (getClosestTriangle, Source/EBGeometry_MeshDistanceFunctionsImplem.hpp
  :511-549) up to exact ties: where it differs, its own FP64 distance (the oracle's lane kernel)
  equals the minimum within EXACT_TIE * diag -- faces meeting at the closest vertex or edge, whose
  FP64 distances are the same quantity up to the last bits. Every query must meet it; the number
  of such exact ties is printed.
"""
import numpy as np
import pytest

from ebgeometry_amd import TriMeshSDF
from ebgeometry_amd import workloads as wl
from oracle import port

pytestmark = pytest.mark.gpu

D_TOL = 1e-5       # |d_gpu - d_ref| / diag
EXACT_TIE = 1e-12  # FP64 distances of faces sharing the closest feature / diag


def diag_of(v):
    v = np.asarray(v, np.float64)
    return float(np.linalg.norm(v.max(0) - v.min(0)))


def face_dist(v, f, fn, vn, en, tri, p):
    x = np.asarray(v, np.float64)[np.asarray(f, np.int64)[tri]].ravel()
    return abs(port.triangle_sdf(x, fn[tri], vn[tri], en[tri], p))


def face_ties(v, f, q, d_ref, tri_ref, tri, label=""):
    """Count the queries whose face differs from the reference's; each must be an exact tie."""
    diag = diag_of(v)
    om = port.Mesh(np.asarray(v, np.float64), f)
    fn, vn, en = om.normals()
    qq = np.asarray(q, np.float64)
    exact = 0
    for i in np.nonzero(tri != tri_ref)[0]:
        a = face_dist(v, f, fn, vn, en, tri[i], qq[i])
        b = abs(d_ref[i])
        assert a - b <= EXACT_TIE * diag + 4e-16 * b, \
            f"{label}: query {i} face {tri[i]} at {a!r} vs the reference's {tri_ref[i]} at {b!r} (gap {(a - b) / diag:.3e} diag)"
        exact += 1
    print(f"{label}: {exact} of {len(tri)} faces differ from the reference's, all exact ties", flush=True)
    return exact


def check_parity(v, f, q, d_ref, tri_ref, d, tri, label="", rel=0.0):
    """rel: extra tolerance relative to |d_ref| -- only for points far outside the workload box,
    where one float32 ulp of the output distance alone exceeds D_TOL * diag."""
    diag = diag_of(v)
    err = np.abs(d.astype(np.float64) - d_ref)
    lim = D_TOL * diag + rel * np.abs(d_ref)
    assert np.all(err <= lim), f"{label}: max |d-d_ref| = {err.max():.3e} (diag {diag:.3f})"
    nz = d_ref != 0
    sgn_bad = np.nonzero(nz & (np.sign(d) != np.sign(d_ref)))[0]
    assert sgn_bad.size == 0, f"{label}: {sgn_bad.size} sign mismatches, e.g. q={q[sgn_bad[0]]} d={d[sgn_bad[0]]} ref={d_ref[sgn_bad[0]]}"
    if tri_ref is not None:
        face_ties(v, f, q, d_ref, tri_ref, tri, label)


FIXTURES = ["dodecahedron.npz", "tetrahedron.npz", "ico3_displaced.npz"]


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("branching,leaf", [(4, 4), (8, 4), (4, 1), (8, 8)])
def test_fixture_parity(golden, cuda, name, branching, leaf):
    g = golden(name)
    m = TriMeshSDF(g["verts"], g["faces"], branching=branching, leaf_size=leaf)
    d, tri = m.closest_triangle(g["queries"])
    check_parity(g["verts"], g["faces"], g["queries"], g["d"], g["tri"], d, tri, f"{name} K{branching} L{leaf}")


@pytest.mark.parametrize("name", ["ico5_unit.npz", "ico7_displaced_sample.npz"])
def test_large_mesh_fixture_parity(golden, cuda, name):
    g = golden(name)
    if "mesh_seed" in g:
        v, f = wl.displaced_icosphere(int(g["level"]), int(g["mesh_seed"]))
    else:
        v, f = wl.icosphere(int(g["level"]))
    m = TriMeshSDF(v, f)
    d, tri = m.closest_triangle(g["queries"])
    check_parity(v, f, g["queries"], g["d"], g["tri"], d, tri, name)


def test_oracle_parity_displaced_l6(cuda):
    """40k mixed queries on an 81,920-face displaced sphere against the oracle."""
    v, f = wl.displaced_icosphere(6, 77)
    q = wl.mixed_queries(40_000, v, f, 78)
    om = port.Mesh(v.astype(np.float64), f)
    d_ref, t_ref = om.sdf(q.astype(np.float64))
    m = TriMeshSDF(v, f)
    d, t = m.closest_triangle(q)
    check_parity(v, f, q, d_ref, t_ref, d, t, "L6 displaced")


def test_fp64_all_mode_matches_oracle_tightly(cuda):
    v, f = wl.displaced_icosphere(4, 5)
    q = wl.mixed_queries(20_000, v, f, 6)
    d_ref, t_ref = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    m = TriMeshSDF(v, f)
    d, t = m.closest_triangle(q, fp64_all=True)
    diag = diag_of(v)
    # only float rounding of the output remains
    assert np.abs(d - d_ref).max() <= 1e-7 * diag
    assert np.all(np.sign(d[d_ref != 0]) == np.sign(d_ref[d_ref != 0]))
    check_parity(v, f, q, d_ref, t_ref, d, t, "fp64_all")


def test_near_surface_recheck_changes_only_small_distances(cuda):
    v, f = wl.displaced_icosphere(5, 9)
    q = wl.near_surface(30_000, v, f, 1e-4, 10)
    m = TriMeshSDF(v, f)
    d_fast = m.query(q, fp64_recheck=False)
    d_chk = m.query(q)
    diag = diag_of(v)
    changed = d_fast != d_chk
    assert np.all(np.abs(d_chk[changed]) < 1e-4 * diag * 1.0001)
    d_ref, _ = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    assert np.all(np.sign(d_chk[d_ref != 0]) == np.sign(d_ref[d_ref != 0]))


def test_presort_is_invisible(cuda):
    v, f = wl.displaced_icosphere(5, 11)
    q = wl.mixed_queries(50_000, v, f, 12)
    m = TriMeshSDF(v, f)
    a = m.query(q, presort=True)
    b = m.query(q, presort=False)
    diag = diag_of(v)
    assert np.abs(a - b).max() <= 1e-6 * diag
    assert np.array_equal(np.sign(a), np.sign(b))


def test_device_tensor_path_equals_host_path(cuda):
    import torch
    v, f = wl.displaced_icosphere(4, 13)
    q = wl.mixed_queries(10_001, v, f, 14)
    m = TriMeshSDF(v, f)
    d_host, t_host = m.closest_triangle(q)
    qt = torch.from_numpy(q).to(cuda)
    d_dev, t_dev = m.closest_triangle(qt)
    torch.cuda.synchronize()
    assert np.array_equal(d_dev.cpu().numpy(), d_host)
    assert np.array_equal(t_dev.cpu().numpy().astype(np.uint32), t_host)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 4097])
def test_ragged_sizes(cuda, n):
    v, f = wl.displaced_icosphere(3, 15)
    q = wl.mixed_queries(n, v, f, 16) if n > 1 else wl.uniform_box(1, [-1] * 3, [1] * 3, 16)
    d_ref, t_ref = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    m = TriMeshSDF(v, f)
    d, t = m.closest_triangle(q)
    check_parity(v, f, q, d_ref, t_ref, d, t, f"n={n}")


def test_empty_query_is_noop(cuda):
    v, f = wl.icosphere(2)
    m = TriMeshSDF(v, f)
    d = m.query(np.zeros((0, 3), np.float32))
    assert d.shape == (0,)


def test_points_on_vertices_and_far_away(cuda):
    v, f = wl.displaced_icosphere(4, 17)
    far = wl.uniform_box(500, [-1e3] * 3, [1e3] * 3, 18)
    q = np.concatenate([v[:500], far]).astype(np.float32)
    d_ref, t_ref = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    m = TriMeshSDF(v, f)
    d, t = m.closest_triangle(q)
    assert np.all(np.abs(d[:500]) <= 1e-6)
    # |d| ~ 1e3: float32 resolution is 6e-5, so the bar is 1e-5*diag + 4 ulp(|d|)
    check_parity(v, f, q, d_ref, t_ref, d, t, "vertices+far", rel=4 * 2.0 ** -24)


def test_signed_distance_single_point_api(cuda):
    v, f = wl.icosphere(3)
    m = TriMeshSDF(v, f)
    assert m.value([0.0, 0.0, 0.0]) < -0.9
    assert m.value([2.0, 0.0, 0.0]) == pytest.approx(1.0, abs=0.01)
    assert isinstance(m.signed_distance([0.0, 0.0, 3.0]), float)


def test_bad_mesh_is_rejected(cuda):
    from ebgeometry_amd._lib import EbgError
    with pytest.raises(EbgError):
        TriMeshSDF(np.zeros((3, 3)), np.array([[0, 1, 2]]))
    with pytest.raises(EbgError):
        TriMeshSDF(np.eye(3), np.array([[0, 1, 3]]))
    with pytest.raises(EbgError):
        TriMeshSDF(np.eye(3), np.array([[0, 1, 2]]), branching=5)


def test_pair_leaves_equal_wavefront_leaves(cuda):
    """The mesh queries evaluate leaves pair-compacted (every (query, triangle) pair the screen
    passes is listed and evaluated one per lane, on the triangle's in-plane frame); the union scene
    walks the same mesh with every triangle evaluated by the whole wavefront, on its edge vectors.
    Both minimise an FP32 distance function over a conservative superset of the triangles that can
    win; the two functions agree to rounding, so a result may differ by the FP32 evaluation error
    of a distance, 2^-21 (|d| + 2 hmax) (hmax the longest edge), never by a missed triangle, and
    the signs are equal."""
    from ebgeometry_amd import AABB, BVHUnionIF
    v, f = wl.displaced_icosphere(5, 21)
    q = wl.mixed_queries(50_000, v, f, 22)
    m = TriMeshSDF(v, f)
    d_pairs = m.query(q, fp64_recheck=False)
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    u = BVHUnionIF([m], [AABB(lo, hi)])
    d_wave = u.value(q, fp64_recheck=False)
    x = v.astype(np.float64)[f.astype(np.int64)]
    hmax = max(np.linalg.norm(x[:, (i + 1) % 3] - x[:, i], axis=1).max() for i in range(3))
    gap = np.abs(d_pairs.astype(np.float64) - d_wave)
    tol = 2.0 ** -20 * (np.abs(d_wave).astype(np.float64) + 2 * hmax)
    bad = np.nonzero(gap > tol)[0]
    assert bad.size == 0, (f"{bad.size} of {q.shape[0]} differ beyond the FP32 distance error, e.g. "
                           f"{d_pairs[bad[0]]!r} vs {d_wave[bad[0]]!r} at {q[bad[0]]}")
    assert np.array_equal(np.sign(d_pairs), np.sign(d_wave))


def sliver_sphere(level, eps):
    """Icosphere whose every face (a, b, c) is split at p = m + eps (c - m), m the midpoint of ab,
    into the slivers (a, b, p) and (c, a, p) (aspect about 1/eps) and the ordinary (b, c, p):
    closed, consistently wound, no zero-area face."""
    v, f = wl.icosphere(level)
    v = np.asarray(v, np.float64)
    out_v, out_f = list(v), []
    for a, b, c in np.asarray(f, np.int64):
        mid = 0.5 * (v[a] + v[b])
        out_v.append(mid + eps * (v[c] - mid))
        p = len(out_v) - 1
        out_f += [(a, b, p), (b, c, p), (c, a, p)]
    return np.asarray(out_v, np.float32), np.asarray(out_f, np.uint32)


def test_sliver_triangles(cuda):
    """Pairs are evaluated on each triangle's in-plane frame, built in FP64 from its longest edge;
    on a mesh of slivers (aspect ~1000) beside ordinary triangles the FP32 pass alone keeps the
    1e-5 * diag bar and the signs outside the FP64 recheck band, and with the recheck every sign."""
    v, f = sliver_sphere(2, 1e-3)
    q = wl.mixed_queries(10_000, v, f, 31)
    d_ref, _ = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    m = TriMeshSDF(v, f)
    diag = diag_of(v)
    for kw in ({}, {"fp64_recheck": False}):
        d = np.asarray(m.query(q, **kw))
        err = np.abs(d.astype(np.float64) - d_ref)
        assert err.max() <= D_TOL * diag, (kw, err.max() / diag)
        checked = (np.abs(d_ref) > 1e-4 * diag) if kw else (d_ref != 0)
        assert np.array_equal(np.sign(d[checked]), np.sign(d_ref[checked])), kw


def test_packet_composition_does_not_change_results(cuda):
    """Presorted and unsorted queries form different packets (different lanes share pairs and
    bounds, and the presorted call splits its costliest packets into subtree walks whose winners
    a separate kernel resolves); each query's result is the same FP32 minimum either way, up to
    the FP32 evaluation error of a distance where the resolving kernels differ (seen: 6 of 50,000
    queries, a few ulp)."""
    v, f = wl.displaced_icosphere(5, 23)
    q = wl.mixed_queries(50_000, v, f, 24)
    m = TriMeshSDF(v, f)
    a = m.query(q, presort=True, fp64_recheck=False)
    b = m.query(q, presort=False, fp64_recheck=False)
    x = v.astype(np.float64)[f.astype(np.int64)]
    hmax = max(np.linalg.norm(x[:, (i + 1) % 3] - x[:, i], axis=1).max() for i in range(3))
    gap = np.abs(a.astype(np.float64) - b)
    bad = np.nonzero(gap > 2.0 ** -20 * (np.abs(b).astype(np.float64) + 2 * hmax))[0]
    assert bad.size == 0, f"{bad.size} differ beyond the FP32 distance error, e.g. {a[bad[0]]!r} vs {b[bad[0]]!r}"
    assert np.count_nonzero(gap) <= q.shape[0] // 1000
    assert np.array_equal(np.sign(a), np.sign(b))
