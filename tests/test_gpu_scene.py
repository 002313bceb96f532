"""GPU parity of CSG union scenes (BVHUnionIF over spheres, boxes and mesh instances).

Bar: |v_gpu - v_ref| <= 1e-5 * scene diagonal and equal sign wherever v_ref != 0, against
  * the reference's own BVHUnionIF::value on a seeded scene (tests/golden/csg_small.npz),
  * the oracle's union (brute-force minimum over every primitive, FP64) on larger scenes and on
    near-surface points, where the FP64 pass decides the sign,
  * closed forms: sphere and box values (Tests/TestAnalyticSDF.cpp:21-113 style), and a mesh
    under Translate / Rotate / Scale against the same mesh with its vertices transformed.
"""
import ctypes as C

import numpy as np
import pytest

from ebgeometry_amd import (AABB, BoxSDF, BVHUnionIF, Rotate, Scale, SphereSDF, Translate, TriMeshSDF, _lib)
from ebgeometry_amd import workloads as wl
from ebgeometry_amd.scene import instance_bounds, instance_matrix
from oracle import port

pytestmark = pytest.mark.gpu

D_TOL = 1e-5


def gpu_scene(sc, branching=4, leaf=4, prim_leaf=4, meshes=None):
    if meshes is None:
        meshes = [TriMeshSDF(v, f, branching=branching, leaf_size=leaf) for v, f in zip(sc.mesh_verts, sc.mesh_faces)]
    prims, bvs = [], []
    for c, r in zip(sc.sphere_center.astype(np.float64), sc.sphere_radius.astype(np.float64)):
        prims.append(SphereSDF(c, r))
        bvs.append(AABB(c - r, c + r))
    for lo, hi in zip(sc.box_lo.astype(np.float64), sc.box_hi.astype(np.float64)):
        prims.append(BoxSDF(lo, hi))
        bvs.append(AABB(lo, hi))
    mlo, mhi = instance_bounds(sc)
    for i, m in enumerate(meshes):
        rx, ry, rz = (float(a) for a in sc.mesh_rot_deg[i])
        f = Translate(Rotate(Rotate(Rotate(Scale(m, float(sc.mesh_scale[i])), rx, 0), ry, 1), rz, 2),
                      sc.mesh_translate[i].astype(np.float64))
        prims.append(f)
        bvs.append(AABB(mlo[i], mhi[i]))
    return BVHUnionIF(prims, bvs, leaf_size=prim_leaf), meshes


def oracle_scene(sc):
    om = [port.Mesh(v.astype(np.float64), f) for v, f in zip(sc.mesh_verts, sc.mesh_faces)]
    return port.Scene(sc.sphere_center, sc.sphere_radius, sc.box_lo, sc.box_hi, om, sc.mesh_translate,
                      sc.mesh_rot_deg, sc.mesh_scale)


def check(v, ref, diag, label):
    v = np.asarray(v, np.float64)
    err = np.abs(v - ref)
    assert np.all(err <= D_TOL * diag), f"{label}: max err {err.max():.3e} at {err.argmax()} (diag {diag:.3f})"
    nz = ref != 0
    bad = np.nonzero(nz & (np.sign(v) != np.sign(ref)))[0]
    assert bad.size == 0, f"{label}: {bad.size} sign mismatches, e.g. {v[bad[0]]} vs {ref[bad[0]]}"


def near_sphere_points(sc, n, seed, sigma=1e-4):
    s = wl.Stream(seed, "near-sphere")
    k = (s.uniform(n) * len(sc.sphere_radius)).astype(np.int64) % len(sc.sphere_radius)
    u = s.unit_vectors(n)
    rr = sc.sphere_radius[k].astype(np.float64) + sigma * s.normal(n)
    return (sc.sphere_center[k].astype(np.float64) + rr[:, None] * u).astype(np.float32)


def test_csg_small_matches_reference(golden, cuda):
    g = golden("csg_small.npz")
    sc = wl.csg_scene(int(g["scene_seed"]), n_spheres=64, n_boxes=16, n_meshes=2, mesh_level=3,
                      extent=float(g["extent"]))
    u, _ = gpu_scene(sc)
    _, _, diag = u.bounds()
    check(u.value(g["queries"]), g["value"], diag, "csg_small")


@pytest.mark.parametrize("branching,leaf,prim_leaf", [(8, 4, 4), (4, 1, 1), (4, 8, 16)])
def test_csg_small_build_variants(golden, cuda, branching, leaf, prim_leaf):
    g = golden("csg_small.npz")
    sc = wl.csg_scene(int(g["scene_seed"]), n_spheres=64, n_boxes=16, n_meshes=2, mesh_level=3,
                      extent=float(g["extent"]))
    u, _ = gpu_scene(sc, branching, leaf, prim_leaf)
    _, _, diag = u.bounds()
    check(u.value(g["queries"]), g["value"], diag, f"K{branching} L{leaf} P{prim_leaf}")


def test_large_scene_matches_oracle(cuda):
    sc = wl.csg_scene(5, n_spheres=4096, n_boxes=1024, n_meshes=8, mesh_level=4, extent=10.0)
    u, _ = gpu_scene(sc)
    _, _, diag = u.bounds()
    o = oracle_scene(sc)
    q = np.concatenate([wl.uniform_box(30000, [0.0] * 3, [10.0] * 3, 6), near_sphere_points(sc, 10000, 7)])
    check(u.value(q), o.value(q.astype(np.float64)), diag, "spheres + boxes + meshes scene")


def test_near_mesh_instances(cuda):
    sc = wl.csg_scene(9, n_spheres=32, n_boxes=8, n_meshes=3, mesh_level=4, extent=6.0)
    u, _ = gpu_scene(sc)
    _, _, diag = u.bounds()
    pts = []
    for i, (v, f) in enumerate(zip(sc.mesh_verts, sc.mesh_faces)):
        local = wl.near_surface(4000, v, f, 0.01, 11 + i)
        M = instance_matrix(sc.mesh_rot_deg[i])
        pts.append(sc.mesh_translate[i] + float(sc.mesh_scale[i]) * (local.astype(np.float64) @ M))
    q = np.concatenate(pts).astype(np.float32)
    check(u.value(q), oracle_scene(sc).value(q.astype(np.float64)), diag, "near mesh instances")


def test_transformed_mesh_equals_mesh_of_transformed_vertices(cuda):
    v, f = wl.displaced_icosphere(4, 3)
    rot, t, s = (20.0, 75.0, 200.0), np.array([0.5, -1.0, 2.0]), 0.7
    m = TriMeshSDF(v, f)
    u = BVHUnionIF([Translate(Rotate(Rotate(Rotate(Scale(m, s), rot[0], 0), rot[1], 1), rot[2], 2), t)])
    w = (t + s * (v.astype(np.float64) @ instance_matrix(rot))).astype(np.float32)
    mw = TriMeshSDF(w, f)
    q = np.concatenate([wl.uniform_box(20000, w.min(0) - 0.3, w.max(0) + 0.3, 4), wl.near_surface(5000, w, f, 0.02, 5)])
    ref, _ = port.Mesh(w.astype(np.float64), f).sdf(q.astype(np.float64))
    _, _, diag = u.bounds()
    a = u.value(q)
    check(a, ref, diag, "instance vs baked vertices")
    check(mw.signed_distance(q), ref, diag, "baked mesh")


def test_analytic_known_answers(cuda):
    box = BoxSDF([-1.0, -1.0, -1.0], [1.0, 1.0, 1.0])
    sph = SphereSDF([3.0, 0.0, 0.0], 0.5)
    pts = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [2, 2, 2], [0.5, 0, 0], [3, 0, 0], [3, 1, 0], [4, 0, 0]],
                   np.float32)
    vb = box.value(pts)
    np.testing.assert_allclose(vb, [-1, 1, np.sqrt(2), np.sqrt(3), -0.5, 2, np.sqrt(4 + 0), 3], rtol=1e-6)
    vs = sph.value(pts)
    ref_s = np.linalg.norm(pts.astype(np.float64) - [3, 0, 0], axis=1) - 0.5
    np.testing.assert_allclose(vs, ref_s, rtol=1e-6, atol=1e-6)
    u = BVHUnionIF([box, sph])
    np.testing.assert_allclose(u.value(pts), np.minimum(vb, vs), rtol=0, atol=0)
    assert abs(u.value([0.0, 0.0, 0.0]) + 1.0) < 1e-7


def test_options_and_device_path(cuda):
    import torch
    sc = wl.csg_scene(21, n_spheres=256, n_boxes=64, n_meshes=2, mesh_level=3, extent=5.0)
    u, _ = gpu_scene(sc)
    _, _, diag = u.bounds()
    q = np.concatenate([wl.uniform_box(5000, [-0.5] * 3, [5.5] * 3, 3), near_sphere_points(sc, 3000, 4)])
    ref = oracle_scene(sc).value(q.astype(np.float64))
    a = u.value(q)
    check(a, ref, diag, "default")
    check(u.value(q, presort=False), ref, diag, "no presort")
    check(u.value(q, fp64_all=True), ref, diag, "fp64 all")
    d = u.value(torch.from_numpy(q).to(cuda))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), a)
    # ragged and tiny batches
    for n in (1, 63, 65, 257):
        check(u.value(q[:n]), ref[:n], diag, f"n={n}")
    assert u.value(np.zeros((0, 3), np.float32)).shape == (0,)


def test_work_counters(cuda):
    import torch
    sc = wl.csg_scene(22, n_spheres=512, n_boxes=128, n_meshes=2, mesh_level=3, extent=6.0)
    u, _ = gpu_scene(sc)
    q = torch.from_numpy(wl.uniform_box(8192, [0.0] * 3, [6.0] * 3, 5)).to(cuda)
    w = torch.zeros(8, dtype=torch.int64, device=cuda)
    out = u.value(q, work=w.data_ptr())
    ref = u.value(q)
    torch.cuda.synchronize()
    c = w.cpu().numpy()
    assert torch.equal(out, ref)
    assert c[0] > 0 and c[4] > 0   # interior nodes (per query, per wavefront)
    assert c[6] > 0 and c[7] > 0   # primitives (per query, per wavefront)


def test_scene_keeps_destroyed_mesh_alive(cuda):
    v, f = wl.displaced_icosphere(3, 8)
    m = TriMeshSDF(v, f)
    u = BVHUnionIF([Translate(m, [1.0, 0.0, 0.0]), SphereSDF([-2.0, 0.0, 0.0], 0.5)])
    q = wl.uniform_box(2000, [-3.0] * 3, [3.0] * 3, 9)
    before = u.value(q)
    _lib.lib().ebg_mesh_destroy(m.handle)  # the caller's handle goes away; the scene retains the mesh
    m._h = None
    np.testing.assert_array_equal(u.value(q), before)


def test_mesh_instances_in_separate_leaves(cuda):
    """One instance per scene leaf: after the first instance lowers some lanes' bound, the next
    instance is walked by a wavefront in which only some lanes need it (regression: the other
    lanes must not take a value from the triangles the wavefront evaluated for them)."""
    sc = wl.csg_scene(41, n_spheres=0, n_boxes=0, n_meshes=4, mesh_level=3, extent=4.0)
    meshes = [TriMeshSDF(v, f) for v, f in zip(sc.mesh_verts, sc.mesh_faces)]
    prims = [Translate(Rotate(Rotate(Rotate(Scale(m, float(sc.mesh_scale[i])), float(sc.mesh_rot_deg[i][0]), 0),
                                    float(sc.mesh_rot_deg[i][1]), 1), float(sc.mesh_rot_deg[i][2]), 2),
                       sc.mesh_translate[i].astype(np.float64)) for i, m in enumerate(meshes)]
    q = wl.uniform_box(8192, [-0.5] * 3, [4.5] * 3, 43)
    ref = oracle_scene(sc).value(q.astype(np.float64))
    for leaf in (1, 2, 4):
        u = BVHUnionIF(prims, leaf_size=leaf)
        _, _, diag = u.bounds()
        check(u.value(q), ref, diag, f"meshes only, leaf {leaf}")


def test_learned_scene_bound_is_exact_and_reports_work(cuda):
    """The per-scene learned bound (C5's): train the union's network on points labelled by the
    scene's own walk, then seed every query's walk with it; the values must equal the unseeded
    walk's bit for bit (queries none of whose primitives falls inside the seed are redone), and
    the work and time of both are reported."""
    import time

    import torch
    sc = wl.csg_scene(5, n_spheres=4096, n_boxes=1024, n_meshes=8, mesh_level=4, extent=10.0)
    u, _ = gpu_scene(sc)
    t0 = time.perf_counter()
    loss = u.train_bound(num_samples=400_000, epochs=4, batch=2048, seed=7)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    assert np.isfinite(loss)
    q = torch.from_numpy(np.concatenate([wl.uniform_box(150_000, [0.0] * 3, [10.0] * 3, 8),
                                         near_sphere_points(sc, 50_000, 9)])).to(cuda)
    plain = u.value(q)
    learned = u.value(q, learned_bound=True)
    torch.cuda.synchronize()
    assert torch.equal(plain, learned)
    # a radius far below every answer: every query is redone, still exact
    tiny = torch.full((q.shape[0],), 1e-9, dtype=torch.float32, device=cuda)
    assert torch.equal(u.value(q, seed_radius=tiny.data_ptr()), plain)

    def work(**kw):
        w = torch.zeros(8, dtype=torch.int64, device=cuda)
        u.value(q, work=w.data_ptr(), **kw)
        torch.cuda.synchronize()
        return w.cpu().numpy()

    def ms(**kw):
        for _ in range(2):
            u.value(q, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            u.value(q, **kw)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / 5 * 1e3

    wp, wl_ = work(), work(learned_bound=True)
    n = q.shape[0]
    pred = u.predict_bound(q)
    rel = ((pred - plain.abs()).abs() / (plain.abs() + 0.01)).median().item()
    print(f"scene bound trained on 400000 points in {train_s:.2f}s, loss {loss:.4f}, median relative error {rel:.3f}; "
          f"per query plain {wp[6] / n:.2f} primitives {wp[0] / n:.2f} nodes, learned {wl_[6] / n:.2f} primitives "
          f"{wl_[0] / n:.2f} nodes ({int(wl_[3])} redone); {ms():.3f} ms plain, {ms(learned_bound=True):.3f} ms learned "
          f"for {n} queries", flush=True)
    assert wl_[3] < n // 10
