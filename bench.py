"""Benchmark: signed-distance queries/s on one MI355X, learned-bound vs plain BVH traversal.

Workload (default, "displaced-sphere-16M"): icosphere subdivided 7x (327,680 faces), vertices
displaced radially by 0.15 sum_k a_k sin(f_k (n.u_k) + phi_k) (seed 1); 16,000,000 queries,
half uniform in the 1.2x bbox, half at area-uniform surface points offset along the normal by
N(0, 0.02) (seed 2). Queries are Morton-presorted on the GPU inside the timed pipeline; near-
surface results (|d| < 1e-4 diag) are redone in FP64. The learned bound is trained on the GPU on
4,000,000 (point, |d|) samples from the mesh's own BVH (same mix, different seed) before timing.

One step = one full query batch (all 16M points, inputs resident in HBM). `value` is the
learned-bound throughput; the plain traversal's throughput is reported beside it.

The same JSON line also carries, measured in the same run:
  correctness  the GPU results (plain and learned) against the reference's CPU SDF on the
               sample the CPU baseline times (|d - d_ref| / diag, signs), and the closest faces
               against getClosestTriangle on a sub-sample (differences split into exact ties and
               others); the run exits with status 3 if the 1e-5 diag bar or a sign fails;
  c1           the 20,480-face unit icosphere with 1M queries uniform in [-1.5,1.5]^3 (seed 0):
               the reference on all host cores over all 1M queries, the GPU beside it;
  batched      the 256-mesh workload (seed 3) through one captured plan;
  learned_c1   the learned bound on the C1 mesh (4M training pairs, 16M queries);
  scene        the CSG union (4,096 spheres, 1,024 boxes, 8 meshes; 32M queries).
`roofline` is priced on the bytes the packet traversal fetches (each record once per
wavefront of 64 queries), against the HBM peak; the rocprofv3 summaries under profiles/ give
the measured DRAM traffic and the counters of the resource that binds (instruction issue).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--skip-cpu] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_FLOPS = 2.5e15  # MI355X dense BF16 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--workload", default="displaced-sphere-16M", choices=["displaced-sphere-16M"])
    ap.add_argument("--queries", type=int, default=16_000_000)
    ap.add_argument("--train-samples", type=int, default=4_000_000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--branching", type=int, default=4)
    ap.add_argument("--leaf", type=int, default=16)
    ap.add_argument("--no-scene", action="store_true", help="skip the CSG scene measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-mesh batch measurement")
    ap.add_argument("--no-learned-c1", action="store_true", help="skip the learned bound on the C1 mesh")
    ap.add_argument("--seeded-only", metavar="BOUND_FILE", default=None,
                    help="profiling aid: the learned-bound pipeline alone, with the network read from BOUND_FILE "
                         "(trained and written there first when the file does not exist), so that a profiled "
                         "run launches no training kernel")
    ap.add_argument("--plain-only", action="store_true",
                    help="profiling aid: time only the plain pipeline (no training, no learned bound)")
    ap.add_argument("--batched", action="store_true",
                    help="instead: the many-mesh workload (256 displaced meshes L3-L7 under affine maps, seed 3, "
                         "10k-1M queries each) through one captured batch plan")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="run only the FP32 FMA kernel that sets the VALU issue ceiling (for rocprofv3 --pmc)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 2 GiB device copy (to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def timed(fn, steps, warmup, kernel_ms=None):
    """Seconds per step over exactly `steps` steps after `warmup` untimed ones. kernel_ms(), if
    given, is read after every timed step: the HIP-event time of that step's main traversal
    launch (it waits for that launch only; the step's later kernels keep running)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ks = []
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        if kernel_ms is not None:
            ks.append(kernel_ms())
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, (float(np.mean(ks)) if ks else None)


def profiled_traffic(name, summary="sdf32_displaced_sphere_16M_hbm_traffic.txt"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by the same counter's reading on a copy kernel of known
    size). None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch.{name}:"
    for line in open(path):
        if line.startswith(key):
            return int(float(line.split(":", 1)[1]))
    return None


def valu_calibration(iters=2048, launches=4):
    """A kernel of independent FP32 FMA chains (libebgcalib.so) that keeps every SIMD's VALU busy:
    its VALU wavefront-instructions per SIMD per cycle are the issue ceiling the traversal
    kernels' rocprofv3 VALU counts are set against."""
    import ctypes as C
    L = C.CDLL(os.path.join(ROOT, "ebgeometry_amd", "libebgcalib.so"))
    L.ebgcal_valu_fma.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_double)]
    ms = (C.c_float * launches)()
    w = C.c_double()
    if L.ebgcal_valu_fma(iters, launches, ms, C.byref(w)) != 0:
        raise RuntimeError("VALU calibration kernel failed")
    best = min(ms)
    print(json.dumps({"valu_fma_wave_instructions_per_launch": w.value, "launch_ms": [round(x, 4) for x in ms],
                      "per_simd_per_cycle_at_2400MHz": round(w.value / (best * 1e-3) / 1024 / 2.4e9, 4)}),
          flush=True)


def copy_calibration(gib=2):
    """A plain device copy of known size, run under rocprofv3 --pmc to scale FETCH_SIZE and
    WRITE_SIZE: each launch reads and writes exactly `gib` GiB."""
    import torch
    n = gib << 30
    a = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n, "launches": 3}), flush=True)


class CpuReference:
    """The reference's own CPU SDF (compiled from its sources into oracle/_ref), or where that
    library is absent the C restatement of its algorithm, built over (v, f) in FP64 from the same
    float32 vertices. Queries split statically over all host cores."""

    def __init__(self, v, f):
        self.nthreads = len(os.sched_getaffinity(0))
        try:
            from oracle import ref
            variants = ref.variants()
        except Exception:
            variants = []
        t0 = time.perf_counter()
        if variants:
            var = variants[0]
            info = ref.info(var)
            self.mesh = ref.Mesh(v.astype(np.float64), f, max_leaf_groups=4, variant=var)
            self.kind = "reference"
            self.label = f"TriMeshSDF<double,uint32_t,K={info['K']},W={info['W']}> ({var} build, SAH, 4 leaf groups)"
        else:
            from oracle import port
            self.mesh = port.Mesh(v.astype(np.float64), f, K=4, W=4, max_leaf_groups=4)
            self.kind = "port"
            self.label = "C restatement of the reference algorithm (K=4, W=4)"
        self.build_s = time.perf_counter() - t0

    def sdf(self, p):
        r = self.mesh.sdf(np.asarray(p, np.float64), nthreads=self.nthreads)
        return r if self.kind == "reference" else r[0]

    def closest(self, p):
        """(d, face) -- getClosestTriangle (reference) or the restatement's closest face."""
        p = np.asarray(p, np.float64)
        if self.kind == "reference":
            return self.mesh.closest(p, nthreads=self.nthreads)
        return self.mesh.sdf(p, nthreads=self.nthreads)


def cpu_baseline(cref, q, budget_s=12.0):
    """Time the CPU reference on a bounded strided sample of the workload (about budget_s of
    work on all host cores). Returns (baseline record, sample indices, reference distances)."""
    # calibrate on a strided slice, then run a sample sized for ~budget_s of CPU work
    stride = lambda m: np.arange(0, q.shape[0], max(1, q.shape[0] // m))[:m]  # noqa: E731
    idx = stride(20_000)
    t0 = time.perf_counter()
    cref.sdf(q[idx])
    rate = idx.size / (time.perf_counter() - t0)
    m = int(min(q.shape[0], max(20_000, rate * budget_s)))
    idx = stride(m)
    t0 = time.perf_counter()
    d_ref = cref.sdf(q[idx])
    dt = time.perf_counter() - t0
    rec = {"value": idx.size / dt, "unit": "queries/s", "cores": cref.nthreads, "kind": cref.kind,
           "sample": f"{idx.size} queries (every {max(1, q.shape[0] // idx.size)}th of the workload, same "
                     f"float32 points promoted to double), {cref.label}, build {cref.build_s:.2f}s not timed"}
    return rec, idx, d_ref


def face_check(v, f, q, d_ref, tri_ref, tri, diag):
    """Closest faces against the reference's: where they differ, the GPU face's own FP64
    distance (the oracle's lane kernel) is compared with the minimum. Exact ties: within 1e-12
    diag (faces meeting at the closest vertex or edge); the rest is reported."""
    from oracle import port
    bad = np.nonzero(tri != tri_ref)[0]
    out = {"queries": int(q.shape[0]), "differing": int(bad.size), "exact_ties": 0, "other": 0,
           "max_other_gap_over_diag": 0.0}
    if bad.size == 0:
        return out
    om = port.Mesh(v.astype(np.float64), f)
    fn, vn, en = om.normals()
    vv, ff = v.astype(np.float64), f.astype(np.int64)
    for i in bad:
        t = int(tri[i])
        a = abs(port.triangle_sdf(vv[ff[t]].ravel(), fn[t], vn[t], en[t], q[i].astype(np.float64)))
        gap = a - abs(d_ref[i])
        if gap <= 1e-12 * diag + 4e-16 * abs(d_ref[i]):
            out["exact_ties"] += 1
        else:
            out["other"] += 1
            out["max_other_gap_over_diag"] = max(out["max_other_gap_over_diag"], gap / diag)
    return out


def compare(d, d_ref, diag):
    """|d - d_ref| / diag and sign agreement (where d_ref != 0) of a GPU result on a sample."""
    err = np.abs(d.astype(np.float64) - d_ref)
    nz = d_ref != 0
    return {"max_abs_err_over_diag": float(err.max() / diag) if err.size else 0.0,
            "sign_mismatches": int(np.count_nonzero(np.sign(d[nz]) != np.sign(d_ref[nz]))),
            "within_1e-5_diag": bool(np.all(err <= 1e-5 * diag))}


def packet_roofline(n, w, K, kernel_ms, kernel, traffic):
    """Roofline of a packet-traversal launch, priced on the bytes it must fetch: per wavefront,
    each interior node record (32 K bytes: K child boxes, references and counts), each triangle
    record it copies for evaluation (128 B), each culling-record load of the screen (4
    triangles, 128 B) and each leaf-slab load (4 B), fetched once for the wavefront's 64
    queries, plus per query its point (12 B) and its result (4 B). The counts come from the
    counting variant of the same launch (work[4], work[5], pair counters [2], [3]). The per-query figure beside it
    (every query's own node and triangle reads, 16 + 28 K per node + 36 per triangle) is what 64
    separate traversals would read; a packet shares those reads, so that figure is not an HBM
    rate and is reported for reference only. With w[8], w[9] (the pair counters) the evaluation
    work is reported too: (query, triangle) pairs per query, 64-lane evaluation batches per
    wavefront, and the share of batch lanes used."""
    byts = 16 * n + w[4] * 32 * K + w[5] * 128 + (w[10] * 128 + w[11] * 4 if len(w) >= 12 else 0)
    gbs = byts / (kernel_ms * 1e-3) / 1e9
    lane = 16 * n + w[0] * 28 * K + w[1] * 36
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "kernel_ms": round(kernel_ms, 3), "algorithmic_bytes_per_launch": int(byts),
            "per_wavefront": {"interior_nodes": round(w[4] / (n / 64), 1), "triangles": round(w[5] / (n / 64), 1)},
            "per_query": {"interior_nodes": round(w[0] / n, 2), "triangles": round(w[1] / n, 2),
                          "fp64_rechecked": w[2], "fallbacks": w[3],
                          "fallback_interior_nodes": round(w[6] / n, 3), "fallback_triangles": round(w[7] / n, 3)},
            "per_query_bytes_if_unshared_GBs": round(lane / (kernel_ms * 1e-3) / 1e9, 1),
            "per_wavefront_culling_loads": (None if len(w) < 12 else round(w[10] / (n / 64), 1)),
            "pair_evaluation": (None if len(w) < 10 or not w[9] else
                                {"pairs_per_query": round(w[8] / n, 2),
                                 "batches_per_wavefront": round(w[9] / (n / 64), 1),
                                 "lanes_used_per_batch": round(w[8] / (64 * w[9]), 3)})}


def c1_bench(a, dev, cref_cls):
    """The 20,480-face unit icosphere (level 5) with 1,000,000 queries uniform in [-1.5,1.5]^3
    (seed 0): the reference's CPU SDF over all of them on every host core, the GPU pipeline
    beside it, and the GPU results against the reference's."""
    import torch

    from ebgeometry_amd import TriMeshSDF
    from ebgeometry_amd import workloads as wl

    v, f = wl.icosphere(5)
    q = wl.uniform_box(1_000_000, [-1.5] * 3, [1.5] * 3, 0)
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    mesh = TriMeshSDF(v, f, branching=a.branching, leaf_size=a.leaf)
    qt = torch.from_numpy(q).to(dev)
    out = torch.empty(q.shape[0], dtype=torch.float32, device=dev)
    dt, _ = timed(lambda: mesh.query(qt, out=out), a.steps, a.warmup)
    res = {"workload": "unit-icosphere-L5-1M", "faces": int(f.shape[0]), "queries": int(q.shape[0]),
           "gpu": {"value": round(q.shape[0] / dt, 1), "unit": "queries/s", "ms_per_step": round(dt * 1e3, 3)}}
    if a.skip_cpu:
        return res
    cref = cref_cls(v, f)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        d_ref = cref.sdf(q)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    res["cpu"] = {"value": round(q.shape[0] / med, 1), "unit": "queries/s", "cores": cref.nthreads,
                  "kind": cref.kind, "sample": f"all {q.shape[0]} queries, median of 3 runs, {cref.label}",
                  "build_seconds": round(cref.build_s, 3), "query_seconds": round(med, 3)}
    res["correctness"] = compare(out.cpu().numpy(), d_ref, diag)
    return res


def learned_c1_bench(a, dev):
    """The learned bound as C4 is written: the C1 mesh (unit icosphere, 20,480 faces), the network
    trained on 4,000,000 BVH-labelled pairs drawn by the 16M-query law around this mesh (half
    uniform in its 1.2x bbox, half near the surface, N(0, 0.02)), evaluated on 16,000,000 queries
    drawn by the same law (another seed), with and without the bound."""
    import torch

    from ebgeometry_amd import TriMeshSDF
    from ebgeometry_amd import workloads as wl

    v, f = wl.icosphere(5)
    q = wl.mixed_queries(a.queries, v, f, 2)
    n = q.shape[0]
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    mesh = TriMeshSDF(v, f, branching=a.branching, leaf_size=a.leaf)
    t0 = time.perf_counter()
    loss = mesh.train_bound(num_samples=a.train_samples, epochs=a.epochs, seed=3)
    train_s = time.perf_counter() - t0
    qt = torch.from_numpy(q).to(dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    out_l = torch.empty(n, dtype=torch.float32, device=dev)
    t_plain, _ = timed(lambda: mesh.query(qt, out=out), a.steps, a.warmup)
    t_learn, _ = timed(lambda: mesh.query(qt, out=out_l, learned_bound=True), a.steps, a.warmup)
    wk = {}
    for name, kw in (("plain", {}), ("learned", {"learned_bound": True})):
        w = torch.zeros(8, dtype=torch.int64, device=dev)
        mesh.query(qt, work=w.data_ptr(), **kw)
        torch.cuda.synchronize()
        wk[name] = [int(x) for x in w.cpu().numpy()]
    torch.cuda.synchronize()
    return {"workload": "unit-icosphere-L5-learned-16M", "faces": int(f.shape[0]), "queries": n,
            "train_samples": a.train_samples, "epochs": a.epochs, "train_seconds": round(train_s, 2),
            "final_loss": round(loss, 5),
            "learned": {"value": round(n / t_learn, 1), "ms_per_step": round(t_learn * 1e3, 3)},
            "plain": {"value": round(n / t_plain, 1), "ms_per_step": round(t_plain * 1e3, 3)},
            "speedup_learned_vs_plain": round(t_plain / t_learn, 3),
            "max_abs_diff_vs_plain_over_diag": (out - out_l).abs().max().item() / diag,
            "sign_agreement_vs_plain": bool(torch.equal(torch.sign(out), torch.sign(out_l))),
            "per_query": {k: {"interior_nodes": round(w[0] / n, 2), "triangles": round(w[1] / n, 2),
                              "fallbacks": w[3]} for k, w in wk.items()},
            "per_wavefront": {k: {"interior_nodes": round(w[4] / (n / 64), 1), "triangles": round(w[5] / (n / 64), 1)}
                              for k, w in wk.items()}}


def scene_bench(a, dev):
    """CSG union of 4,096 spheres, 1,024 boxes and 8 displaced L5 meshes under similarity maps
    (seed 5), 32M queries uniform in [0,10]^3 (seed 77): the plain scene traversal."""
    import torch

    from ebgeometry_amd import AABB, BoxSDF, BVHUnionIF, Rotate, Scale, SphereSDF, Translate, TriMeshSDF
    from ebgeometry_amd import workloads as wl
    from ebgeometry_amd.scene import instance_bounds

    t0 = time.perf_counter()
    sc = wl.csg_scene(5, n_spheres=4096, n_boxes=1024, n_meshes=8, mesh_level=5, extent=10.0)
    # the mesh instances' trees as the 16M-query mesh's (K=4, leaves of 16): 6.5 ms of scene
    # traversal against 8.5 ms with leaves of 4
    meshes = [TriMeshSDF(v, f, branching=4, leaf_size=16) for v, f in zip(sc.mesh_verts, sc.mesh_faces)]
    prims, bvs = [], []
    for c, r in zip(sc.sphere_center.astype(np.float64), sc.sphere_radius.astype(np.float64)):
        prims.append(SphereSDF(c, r))
        bvs.append(AABB(c - r, c + r))
    for lo, hi in zip(sc.box_lo.astype(np.float64), sc.box_hi.astype(np.float64)):
        prims.append(BoxSDF(lo, hi))
        bvs.append(AABB(lo, hi))
    mlo, mhi = instance_bounds(sc)
    for i, m in enumerate(meshes):
        rx, ry, rz = (float(x) for x in sc.mesh_rot_deg[i])
        prims.append(Translate(Rotate(Rotate(Rotate(Scale(m, float(sc.mesh_scale[i])), rx, 0), ry, 1), rz, 2),
                               sc.mesh_translate[i].astype(np.float64)))
        bvs.append(AABB(mlo[i], mhi[i]))
    u = BVHUnionIF(prims, bvs)
    n = 32_000_000
    q = torch.from_numpy(wl.uniform_box(n, [0.0] * 3, [10.0] * 3, 77)).to(dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    log(f"scene built in {time.perf_counter() - t0:.1f}s")
    dt, kms = timed(lambda: u.value(q, out=out), a.steps, a.warmup, u.last_traversal_ms)
    w = torch.zeros(8, dtype=torch.int64, device=dev)
    u.value(q, out=out, work=w.data_ptr())
    torch.cuda.synchronize()
    w = [int(x) for x in w.cpu().numpy()]
    # bytes the packets fetch: per wavefront each node record (128 B, scene and K=4 mesh nodes),
    # primitive record (128 B) and triangle record (128 B) once; per query point and result
    byts = 16 * n + w[4] * 128 + w[7] * 128 + w[5] * 128
    gbs = byts / (kms * 1e-3) / 1e9
    log(f"scene {dt * 1e3:.2f} ms/step (traversal {kms:.2f} ms)")
    # the per-scene learned bound: trained on 8M points labelled by the scene's own walk, then
    # every query's walk seeded with it (exact: queries with nothing inside the seed are redone)
    t0 = time.perf_counter()
    loss = u.train_bound(num_samples=8_000_000, epochs=6, seed=3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    out_l = torch.empty_like(out)
    dtl, kml = timed(lambda: u.value(q, out=out_l, learned_bound=True), a.steps, a.warmup, u.last_traversal_ms)
    torch.cuda.synchronize()
    exact = bool(torch.equal(out, out_l))
    wl_ = torch.zeros(8, dtype=torch.int64, device=dev)
    u.value(q, out=out_l, work=wl_.data_ptr(), learned_bound=True)
    torch.cuda.synchronize()
    wl_ = [int(x) for x in wl_.cpu().numpy()]
    log(f"scene learned {dtl * 1e3:.2f} ms/step (traversal {kml:.2f} ms), bound trained in {train_s:.1f}s")
    learned = {"value": round(n / dtl, 1), "ms_per_step": round(dtl * 1e3, 3), "traversal_ms": round(kml, 3),
               "speedup_learned_vs_plain": round(dt / dtl, 3), "train_samples": 8_000_000, "epochs": 6,
               "train_seconds": round(train_s, 2), "final_loss": round(loss, 5), "equal_to_plain": exact,
               "per_query": {"interior_nodes": round(wl_[0] / n, 2), "primitives": round(wl_[6] / n, 2),
                             "triangles": round(wl_[1] / n, 2), "redone_unseeded": wl_[3]}}
    return {"workload": "csg-union-4096-spheres-1024-boxes-8-meshes-32M", "queries": n, "mesh_branching": 4,
            "mesh_leaf_size": 16, "learned_bound": learned,
            "value": round(n / dt, 1), "unit": "queries/s", "ms_per_step": round(dt * 1e3, 3),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "scene32_kernel<4,1,false>",
                         "kernel_ms": round(kms, 3), "algorithmic_bytes_per_launch": int(byts),
                         "per_wavefront": {"interior_nodes": round(w[4] / (n / 64), 1),
                                           "primitives": round(w[7] / (n / 64), 1),
                                           "triangles": round(w[5] / (n / 64), 1)},
                         "per_query": {"interior_nodes": round(w[0] / n, 2), "primitives": round(w[6] / n, 2),
                                       "triangles": round(w[1] / n, 2), "fp64_rechecked": w[2]}}}


def batched_bench(a):
    """256 displaced icospheres (levels 3-7 uniform, affine maps with det > 0, seed 3), query counts
    log-uniform in [1e4, 1e6] per mesh (uniform/near-surface mix): one autotuned batch, one plan
    captured as a hipGraph; a step = one replay of the whole pipeline."""
    import torch

    from ebgeometry_amd import MeshBatch
    from ebgeometry_amd import workloads as wl

    t0 = time.perf_counter()
    w = wl.batched_meshes(256, 3)
    nq = sum(q.shape[0] for q in w.queries)
    nt = sum(f.shape[0] for f in w.faces)
    log(f"batched workload generated in {time.perf_counter() - t0:.1f}s: {nt} faces, {nq} queries")
    t0 = time.perf_counter()
    b = MeshBatch(list(zip(w.verts, w.faces)), autotune=True)
    info = b.info()
    log(f"batch built in {time.perf_counter() - t0:.1f}s: K={info['branching']} leaf={info['leaf_size']}")
    items = [(i, torch.from_numpy(q).to("cuda:0")) for i, q in enumerate(w.queries)]
    plan = b.plan(items, capture=True)
    dt, replay_ms = timed(plan.launch, a.steps, a.warmup, plan.last_traversal_ms)
    # the traversal kernel alone: the same items and workgroup through a plan without a graph
    # (events inside a graph are not timed), timed the same way
    eager = b.plan(items, capture=False, workgroup=plan.workgroup, persistent=plan.persistent)
    _, kms = timed(eager.launch, a.steps, a.warmup, eager.last_traversal_ms)
    del eager
    wk = torch.zeros(12, dtype=torch.int64, device="cuda:0")  # work[0..7], then the pair counters
    b.query(items, work=wk.data_ptr(), pair_work=wk[8:].data_ptr())
    torch.cuda.synchronize()
    c = [int(x) for x in wk.cpu().numpy()]
    K = info["branching"]
    # correctness of the captured plan's results: every 16th mesh, 200 strided queries each,
    # against the oracle
    plan.launch()
    torch.cuda.synchronize()
    worst, signs, checked = 0.0, 0, 0
    from oracle import port
    for i in range(0, len(w.verts), 16):
        vv, ff, qq = w.verts[i], w.faces[i], w.queries[i]
        sel = np.arange(0, qq.shape[0], max(1, qq.shape[0] // 200))[:200]
        d_ref, _ = port.Mesh(vv.astype(np.float64), ff).sdf(qq[sel].astype(np.float64))
        dg = plan.out[i].cpu().numpy()[sel]
        dia = float(np.linalg.norm(vv.max(0).astype(np.float64) - vv.min(0)))
        worst = max(worst, float(np.abs(dg - d_ref).max() / dia))
        nz = d_ref != 0
        signs += int(np.count_nonzero(np.sign(dg[nz]) != np.sign(d_ref[nz])))
        checked += sel.size
    res = {"metric": "SDF queries/s on 1 MI355X, many-mesh batch (one captured plan)", "value": round(nq / dt, 1),
           "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batched-256-meshes", "meshes": 256, "faces": nt, "queries": nq, "seed": 3,
                      "branching": K, "leaf_size": info["leaf_size"], "workgroup": plan.workgroup,
                      "launch": "persistent" if plan.persistent else "one wavefront per packet",
                      "tune_ms": info["tune_ms"]},
           "graph_replay_ms": round(replay_ms, 3),
           "roofline": packet_roofline(nq, c, K, kms, f"batch32_kernel<{K},{info['stack_capacity'] // 64},false,"
                                       f"{str(plan.persistent).lower()}>",
                                       profiled_traffic("batch32", "batch32_many_meshes_256_counters.txt")),
           "correctness": {"sample": f"{checked} queries (200 strided of every 16th mesh) against the oracle",
                           "max_abs_err_over_diag": worst, "sign_mismatches": signs,
                           "within_1e-5_diag": bool(worst <= 1e-5)}}
    return res


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        copy_calibration()
        return
    if a.valu_calibration:
        valu_calibration()
        return
    if a.batched:
        print(json.dumps(batched_bench(a)), flush=True)
        return

    import torch

    from ebgeometry_amd import TriMeshSDF
    from ebgeometry_amd import workloads as wl

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    v, f = wl.displaced_icosphere(7, 1)
    q = wl.mixed_queries(a.queries, v, f, 2)
    log(f"workload generated in {time.perf_counter() - t0:.1f}s: {f.shape[0]} faces, {q.shape[0]} queries")
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    n = q.shape[0]

    mesh = TriMeshSDF(v, f, branching=a.branching, leaf_size=a.leaf)
    info = mesh.info()
    log(f"mesh built in {info['build_seconds']:.2f}s: {info['num_nodes']} nodes, stack {info['stack_need']}")

    qt = torch.from_numpy(q).to(dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    out_l = torch.empty(n, dtype=torch.float32, device=dev)

    plain = lambda: mesh.query(qt, out=out)  # noqa: E731
    learned = lambda: mesh.query(qt, out=out_l, learned_bound=True)  # noqa: E731
    K = info["branching"]
    kname = f"sdf32_kernel<{K},{info['stack_capacity'] // 64},false,{{mode}}>"

    def work(**kw):
        w = torch.zeros(12, dtype=torch.int64, device=dev)  # work[0..7], then the pair counters
        mesh.query(qt, out=out_l, work=w.data_ptr(), pair_work=w[8:].data_ptr(), **kw)
        torch.cuda.synchronize()
        return [int(x) for x in w.cpu().numpy()]

    if a.plain_only:
        # profiling aid: the plain pipeline alone (no training)
        t_plain, trav_plain = timed(plain, a.steps, a.warmup, mesh.last_traversal_ms)
        rp = packet_roofline(n, work(), K, trav_plain, kname.format(mode=0), None)
        print(json.dumps({"profile_run": "plain-only", "ms_per_step": round(t_plain * 1e3, 3), "roofline": rp}),
              flush=True)
        return
    if a.seeded_only:
        if os.path.exists(a.seeded_only):
            mesh.load_bound(a.seeded_only)
        else:
            mesh.train_bound(num_samples=a.train_samples, epochs=a.epochs, seed=3)
            mesh.save_bound(a.seeded_only)
            print(json.dumps({"profile_run": "seeded-only", "bound_written": a.seeded_only}), flush=True)
            return
        t_learn, trav_learn = timed(learned, a.steps, a.warmup, mesh.last_traversal_ms)
        rl = packet_roofline(n, work(learned_bound=True), K, trav_learn, kname.format(mode=1), None)
        print(json.dumps({"profile_run": "seeded-only", "ms_per_step": round(t_learn * 1e3, 3), "roofline": rl}),
              flush=True)
        return

    t0 = time.perf_counter()
    loss = mesh.train_bound(num_samples=a.train_samples, epochs=a.epochs, seed=3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    log(f"bound trained in {train_s:.1f}s, final loss {loss:.4f}")

    # MFMA utilisation of the bound network: inference over the 16M queries alone (HIP events on
    # torch's stream, which the library uses), 2 * (51*128 + 3*128*128 + 128) FLOP per query
    mlp_flop_q = 2 * (51 * 128 + 3 * 128 * 128 + 128)
    for _ in range(2):
        mesh.predict_bound(qt)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        mesh.predict_bound(qt)
    e1.record()
    torch.cuda.synchronize()
    infer_ms = e0.elapsed_time(e1) / 5
    mfma_util = mlp_flop_q * n / (infer_ms * 1e-3) / BF16_PEAK_FLOPS
    log(f"bound inference {infer_ms:.2f} ms for {n} queries ({mfma_util:.3f} of the BF16 dense peak)")

    t_plain, trav_plain = timed(plain, a.steps, a.warmup, mesh.last_traversal_ms)
    t_learn, trav_learn = timed(learned, a.steps, a.warmup, mesh.last_traversal_ms)
    log(f"plain {t_plain * 1e3:.2f} ms/step (traversal {trav_plain:.2f} ms), learned {t_learn * 1e3:.2f} ms/step "
        f"(traversal {trav_learn:.2f} ms)")

    # exactness of the learned bound on this run (untimed)
    torch.cuda.synchronize()
    dmax = (out - out_l).abs().max().item()
    sign_ok = bool(torch.equal(torch.sign(out), torch.sign(out_l)))
    d_plain = out.cpu().numpy()
    d_learn = out_l.cpu().numpy()

    # work counters (untimed counting variants) for the roofline's algorithmic bytes
    wp, wl_ = work(), work(learned_bound=True)
    rl = packet_roofline(n, wl_, K, trav_learn, kname.format(mode=1), profiled_traffic("seeded"))
    rp = packet_roofline(n, wp, K, trav_plain, kname.format(mode=0), profiled_traffic("plain"))

    # closest faces of the whole workload (untimed), for the face check below
    d_t, t_gpu = mesh.closest_triangle(qt)
    torch.cuda.synchronize()
    t_gpu = t_gpu.cpu().numpy().astype(np.uint32)

    from_cpu = {}
    cref = CpuReference(v, f)
    if a.skip_cpu:
        idx = np.arange(0, n, max(1, n // 50_000))[:50_000]
        d_ref = cref.sdf(q[idx])
        cpu = None
    else:
        t0 = time.perf_counter()
        cpu, idx, d_ref = cpu_baseline(cref, q)
        log(f"cpu baseline {cpu['value']:.3g} q/s ({time.perf_counter() - t0:.1f}s)")
    from_cpu["sample"] = f"{idx.size} strided queries of the workload against {cref.label}"
    from_cpu["plain"] = compare(d_plain[idx], d_ref, diag)
    from_cpu["learned"] = compare(d_learn[idx], d_ref, diag)
    # faces against getClosestTriangle on a 100,000-query sub-sample
    fidx = idx[:: max(1, idx.size // 100_000)][:100_000]
    dref_f, tref_f = cref.closest(q[fidx])
    from_cpu["faces"] = face_check(v, f, q[fidx], dref_f, tref_f, t_gpu[fidx], diag)
    ok = from_cpu["plain"]["within_1e-5_diag"] and from_cpu["learned"]["within_1e-5_diag"] and \
        from_cpu["plain"]["sign_mismatches"] == 0 and from_cpu["learned"]["sign_mismatches"] == 0
    from_cpu["pass"] = bool(ok)
    log(f"correctness: {from_cpu}")
    del cref

    c1 = c1_bench(a, dev, CpuReference)
    log(f"c1: {c1}")
    batched = None if a.no_batched else batched_bench(a)
    learned_c1 = None if a.no_learned_c1 else learned_c1_bench(a, dev)
    scene = None if a.no_scene else scene_bench(a, dev)

    value = n / t_learn
    res = {
        "metric": "SDF queries/s on 1 MI355X, learned-bound vs plain BVH at <=1e-5*diag error",
        "value": round(value, 1), "unit": "queries/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t_learn * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": a.workload, "faces": int(f.shape[0]), "queries": n, "mesh_seed": 1, "query_seed": 2,
                   "branching": K, "leaf_size": info["leaf_size"], "parallelism": "single-gpu",
                   "fp64_recheck_below": "1e-4*diag"},
        "plain": {"value": round(n / t_plain, 1), "ms_per_step": round(t_plain * 1e3, 3)},
        "speedup_learned_vs_plain": round(t_plain / t_learn, 3),
        "learned_bound": {"inference_ms": round(infer_ms, 3),
                          "inference_mfma_utilisation": round(mfma_util, 4),
                          "train_mfma_utilisation_incl_labelling": round(
                              3 * mlp_flop_q * a.train_samples * a.epochs / train_s / BF16_PEAK_FLOPS, 4),
                          "train_samples": a.train_samples, "epochs": a.epochs, "train_seconds": round(train_s, 2),
                          "final_loss": round(loss, 5), "max_abs_diff_vs_plain_over_diag": dmax / diag,
                          "sign_agreement_vs_plain": sign_ok},
        "correctness": from_cpu,
        "roofline": rl,
        "roofline_plain": rp,
        "cpu_baseline": cpu,
        "c1": c1,
        "batched": batched,
        "learned_c1": learned_c1,
        "scene": scene,
    }
    print(json.dumps(res), flush=True)
    bad = not ok or (c1.get("correctness") and not (c1["correctness"]["within_1e-5_diag"] and
                                                    c1["correctness"]["sign_mismatches"] == 0)) or \
        (batched and not batched["correctness"]["within_1e-5_diag"])
    if bad:
        log("correctness bar failed")
        sys.exit(3)


if __name__ == "__main__":
    main()
