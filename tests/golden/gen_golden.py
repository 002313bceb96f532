"""Write the golden fixtures under tests/golden/ from the compiled reference.

Run in a container where the reference sources are available (EBGEOMETRY_HOME, default
/root/reference): ``make -C oracle ref && python tests/golden/gen_golden.py``.

The reference is run through oracle/_ref/libebgref.so (its portable build,
TriMeshSDF<double, uint32_t, K=4, W=4>(mesh, SAH, 4) via Soup::soupToDCEL). Every
fixture is data only: inputs (or the seed that regenerates them, with a checksum of the
regenerated arrays) and the reference's outputs. Meshes from the reference's own test data
(Tests/data/dodecahedron.obj, Tests/data/tetrahedron.stl) are stored as vertex/face
arrays.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ebgeometry_amd import workloads as wl  # noqa: E402
from oracle import ref  # noqa: E402

EBG_HOME = os.environ.get("EBGEOMETRY_HOME", "/root/reference")


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def read_obj(path):
    v, f = [], []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                v.append([float(x) for x in t[1:4]])
            elif t[0] == "f":
                f.append([int(x.split("/")[0]) - 1 for x in t[1:]])
    return np.array(v), np.array(f, dtype=np.uint32)


def read_ascii_stl(path):
    pts = []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if t and t[0] == "vertex":
                pts.append([float(x) for x in t[1:4]])
    pts = np.array(pts)
    uniq, inv = np.unique(pts, axis=0, return_inverse=True)
    return uniq, inv.reshape(-1, 3).astype(np.uint32)


def mesh_fixture(name, verts, faces, queries, extra=None):
    v32 = np.asarray(verts, dtype=np.float32)
    q32 = np.asarray(queries, dtype=np.float32)
    m = ref.Mesh(v32.astype(np.float64), faces, max_leaf_groups=4)
    d, tri = m.closest(q32.astype(np.float64))
    d_sdf = m.sdf(q32.astype(np.float64))
    fn, vn, en = m.normals()
    out = dict(verts=v32, faces=np.asarray(faces, dtype=np.uint32), queries=q32, d=d, tri=tri,
               d_sdf=d_sdf, face_n=fn, vert_n=vn, edge_n=en)
    if extra:
        out.update(extra)
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, v32.shape, faces.shape, q32.shape)


def mesh_fixture_seeded(name, verts, faces, queries, meta):
    """Large meshes: store the seed/recipe and a checksum instead of the arrays."""
    v32 = np.asarray(verts, dtype=np.float32)
    q32 = np.asarray(queries, dtype=np.float32)
    m = ref.Mesh(v32.astype(np.float64), faces, max_leaf_groups=4)
    d, tri = m.closest(q32.astype(np.float64))
    out = dict(queries=q32, d=d, tri=tri, mesh_sha=np.array(sha(v32, faces)), **meta)
    np.savez_compressed(os.path.join(HERE, name), **out)
    print(name, v32.shape, faces.shape, q32.shape)


def main():
    info = ref.info("portable")
    assert info["K"] == 4 and info["W"] == 4, info

    # reference test data meshes
    v, f = read_obj(os.path.join(EBG_HOME, "Tests/data/dodecahedron.obj"))
    lo, hi = v.min(0), v.max(0)
    q = np.concatenate([wl.uniform_box(2000, lo - 1.0, hi + 1.0, 11),
                        wl.near_surface(1000, v, f, 0.05, 12)])
    mesh_fixture("dodecahedron.npz", v, f, q)

    v, f = read_ascii_stl(os.path.join(EBG_HOME, "Tests/data/tetrahedron.stl"))
    q = np.concatenate([wl.uniform_box(1500, [-1, -1, -1], [2, 2, 2], 13),
                        wl.near_surface(500, v, f, 0.02, 14)])
    mesh_fixture("tetrahedron.npz", v, f, q)

    # small displaced icosphere
    v, f = wl.displaced_icosphere(3, 21)
    q = wl.mixed_queries(4096, v, f, 22)
    mesh_fixture("ico3_displaced.npz", v, f, q)

    # the unit icosphere of the CPU FP64 workload (L5, 20,480 triangles): first 8192 of its
    # 1M seed-0 uniform queries in [-1.5, 1.5]^3
    v, f = wl.icosphere(5)
    q = wl.uniform_box(1_000_000, [-1.5] * 3, [1.5] * 3, 0)[:8192]
    mesh_fixture_seeded("ico5_unit.npz", v, f, q, dict(level=np.array(5), seed=np.array(0)))

    # the 327,680-triangle displaced sphere: a 4096-query sample of the mixed workload
    v, f = wl.displaced_icosphere(7, 1)
    qa = wl.mixed_queries(16_000_000, v, f, 2)
    idx = np.sort(np.concatenate([np.arange(0, 8_000_000, 8_000_000 // 2048),
                                  np.arange(8_000_000, 16_000_000, 8_000_000 // 2048)]))
    mesh_fixture_seeded("ico7_displaced_sample.npz", v, f, qa[idx],
                        dict(level=np.array(7), mesh_seed=np.array(1), query_seed=np.array(2),
                             query_index=idx.astype(np.int64)))

    # SFC: bins, Morton codes, order on query clouds and triangle centroids
    v, f = wl.displaced_icosphere(4, 31)
    vv = v.astype(np.float64)
    cent64 = 0.5 * (np.minimum(np.minimum(vv[f[:, 0]], vv[f[:, 1]]), vv[f[:, 2]])
                    + np.maximum(np.maximum(vv[f[:, 0]], vv[f[:, 1]]), vv[f[:, 2]]))
    pts32 = wl.mixed_queries(20000, v, f, 32)
    # duplicates and a planar slab exercise equal codes and zero-extent axes
    dup = np.repeat(pts32[:50], 3, axis=0)
    slab = pts32[:300].copy()
    slab[:, 2] = 0.25
    b32 = ref.bins_f32(pts32)
    b64 = ref.bins_f64(cent64)
    bd = ref.bins_f32(dup)
    bs = ref.bins_f32(slab)
    np.savez_compressed(os.path.join(HERE, "morton.npz"),
                        pts32=pts32, bins32=b32, codes32=ref.morton_encode(b32),
                        order32=ref.morton_order_f32(pts32),
                        cent64=cent64, bins64=b64, codes64=ref.morton_encode(b64),
                        dup32=dup, bins_dup=bd, codes_dup=ref.morton_encode(bd), order_dup=ref.morton_order_f32(dup),
                        slab32=slab, bins_slab=bs, codes_slab=ref.morton_encode(bs))
    print("morton.npz")

    # CSG: small union of spheres, boxes and two transformed meshes
    sc = wl.csg_scene(41, n_spheres=64, n_boxes=16, n_meshes=2, mesh_level=3, extent=4.0)
    meshes = [ref.Mesh(mv.astype(np.float64), mf) for mv, mf in zip(sc.mesh_verts, sc.mesh_faces)]
    from ebgeometry_amd.scene import instance_bounds  # noqa: E402
    blo, bhi = instance_bounds(sc)
    R = ref.Scene(sc.sphere_center, sc.sphere_radius, sc.box_lo, sc.box_hi, meshes,
                  sc.mesh_translate, sc.mesh_rot_deg, sc.mesh_scale, blo, bhi)
    q = wl.uniform_box(4096, sc.lo - 0.5, sc.hi + 0.5, 42)
    val = R.value(q.astype(np.float64))
    np.savez_compressed(os.path.join(HERE, "csg_small.npz"), queries=q, value=val,
                        scene_seed=np.array(41), extent=np.array(4.0))
    print("csg_small.npz")


if __name__ == "__main__":
    main()
