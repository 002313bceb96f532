"""GPU parity of the many-mesh path: every item of a batch must equal the single-mesh query of
the same mesh (same kernels, so the distances agree bit for bit up to exact ties), and the oracle
within 1e-5 * diag with equal signs; captured graphs must give the eager result on every replay."""
import numpy as np
import pytest

from ebgeometry_amd import MeshBatch, TriMeshSDF
from ebgeometry_amd import workloads as wl
from oracle import port

pytestmark = pytest.mark.gpu

D_TOL = 1e-5


def diag_of(v):
    v = np.asarray(v, np.float64)
    return float(np.linalg.norm(v.max(0) - v.min(0)))


@pytest.fixture(scope="module")
def small_batch():
    w = wl.batched_meshes(10, 5, level_range=(1, 4), query_range=(1, 5000))
    return w, MeshBatch(list(zip(w.verts, w.faces)), branching=4, leaf_size=8)


def items_for(w, cuda, extra=()):
    import torch
    items = [(i, torch.from_numpy(q).to(cuda)) for i, q in enumerate(w.queries)]
    for mesh, q in extra:
        items.append((mesh, torch.from_numpy(q).to(cuda)))
    return items


def test_batch_equals_single_mesh_queries(small_batch, cuda):
    import torch
    w, b = small_batch
    # a repeated mesh, an empty item and a ragged one on top of one item per mesh
    extra = [(3, w.queries[7][:37]), (0, np.zeros((0, 3), np.float32)), (9, w.queries[1][:65])]
    items = items_for(w, cuda, extra)
    outs, tris = b.query(items, with_triangle=True)
    torch.cuda.synchronize()
    for k, (mesh, q) in enumerate(items):
        if q.shape[0] == 0:
            assert outs[k].shape[0] == 0
            continue
        single = TriMeshSDF(w.verts[mesh], w.faces[mesh], branching=4, leaf_size=8)
        d1, t1 = single.closest_triangle(q)
        torch.cuda.synchronize()
        d0 = outs[k].cpu().numpy()
        d1 = d1.cpu().numpy()
        diag = diag_of(w.verts[mesh])
        assert np.abs(d0 - d1).max() <= 1e-6 * diag, (k, mesh)
        assert np.array_equal(np.sign(d0), np.sign(d1)), (k, mesh)


@pytest.mark.parametrize("mesh", [0, 4, 9])
def test_batch_matches_oracle(small_batch, cuda, mesh):
    import torch
    w, b = small_batch
    items = items_for(w, cuda)
    outs = b.query(items)
    torch.cuda.synchronize()
    d = outs[mesh].cpu().numpy().astype(np.float64)
    ref, _ = port.Mesh(w.verts[mesh].astype(np.float64), w.faces[mesh]).sdf(w.queries[mesh].astype(np.float64))
    diag = diag_of(w.verts[mesh])
    assert np.abs(d - ref).max() <= D_TOL * diag
    nz = ref != 0
    assert np.array_equal(np.sign(d[nz]), np.sign(ref[nz]))


def test_plan_graph_replays_match_eager(small_batch, cuda):
    import torch
    w, b = small_batch
    items = items_for(w, cuda)
    eager = [o.clone() for o in b.query(items)]
    plan = b.plan(items, capture=True)
    assert plan.workgroup in (64, 128, 256)
    for rep in range(3):
        for o in plan.out:
            o.fill_(12345.0)
        plan.launch()
        torch.cuda.synchronize()
        for k in range(len(items)):
            assert torch.equal(plan.out[k], eager[k]), (rep, k)
        # unrelated eager work between replays
        _ = b.query(items[:2])
        torch.cuda.synchronize()
    assert plan.last_traversal_ms() > 0


def test_plan_sees_new_point_values(small_batch, cuda):
    """The plan fixes the arrays, not their contents."""
    import torch
    w, b = small_batch
    items = items_for(w, cuda)
    plan = b.plan(items, workgroup=128, capture=True)
    plan.launch()
    for _, q in items:
        q.mul_(1.01)
    plan.launch()
    torch.cuda.synchronize()
    fresh = b.query(items)
    torch.cuda.synchronize()
    for k in range(len(items)):
        assert torch.equal(plan.out[k], fresh[k])


def test_autotuned_batch(cuda):
    import torch
    w = wl.batched_meshes(4, 11, level_range=(2, 4), query_range=(1000, 3000))
    b = MeshBatch(list(zip(w.verts, w.faces)), autotune=True)
    info = b.info()
    assert info["tuned"] == 1
    assert (info["branching"], info["leaf_size"]) in [(4, 8), (4, 16), (8, 8), (8, 16)]
    assert min(t for t in info["tune_ms"] if t >= 0) > 0
    items = items_for(w, cuda)
    outs = b.query(items)
    torch.cuda.synchronize()
    for k in range(len(items)):
        ref, _ = port.Mesh(w.verts[k].astype(np.float64), w.faces[k]).sdf(w.queries[k].astype(np.float64))
        d = outs[k].cpu().numpy()
        assert np.abs(d - ref).max() <= D_TOL * diag_of(w.verts[k])
        assert np.array_equal(np.sign(d[ref != 0]), np.sign(ref[ref != 0]))


def test_work_counters_and_fp64_all(small_batch, cuda):
    import torch
    w, b = small_batch
    items = items_for(w, cuda)
    wk = torch.zeros(8, dtype=torch.int64, device=cuda)
    outs = b.query(items, work=wk.data_ptr())
    exact = b.query(items, fp64_all=True)
    torch.cuda.synchronize()
    c = wk.cpu().numpy()
    assert c[0] > 0 and c[1] > 0 and c[4] > 0 and c[5] > 0
    for k in range(len(items)):
        diag = diag_of(w.verts[k])
        assert (outs[k] - exact[k]).abs().max().item() <= D_TOL * diag


def test_batch_rejects_bad_items(small_batch, cuda):
    import torch
    from ebgeometry_amd import _lib
    w, b = small_batch
    q = torch.zeros((10, 3), dtype=torch.float32, device=cuda)
    with pytest.raises(_lib.EbgError):
        b.query([(99, q)])


def test_plan_with_fixed_workgroup_times_and_leaves_no_error(cuda):
    """A captured plan with a caller-chosen workgroup (no timing launches before the capture)
    still reports its traversal time, and a second plan can be made right after."""
    import torch
    w = wl.batched_meshes(3, 21, level_range=(2, 3), query_range=(1000, 2000))
    b = MeshBatch(list(zip(w.verts, w.faces)))
    items = [(i, torch.from_numpy(q).to(cuda)) for i, q in enumerate(w.queries)]
    for wg in (64, 256):
        plan = b.plan(items, capture=True, workgroup=wg)
        plan.launch()
        torch.cuda.synchronize()
        assert plan.last_traversal_ms() >= 0.0
        assert plan.workgroup == wg
    torch.empty(16, device=cuda).fill_(1.0)  # no stale HIP error reported here
    torch.cuda.synchronize()


def test_early_packets_in_a_batch(cuda):
    """Meshes of >= 16,384 faces get a cost grid; packets of queries deep inside them start
    first. The results must equal the single-mesh queries and the oracle."""
    import torch
    rng = np.random.default_rng(11)
    meshes, queries = [], []
    for k in range(3):
        v, f = wl.displaced_icosphere(5, 40 + k)  # 20,480 faces
        meshes.append((v, f))
        d = rng.normal(size=(6000, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        deep = (d * (0.3 * rng.random(6000) ** (1 / 3))[:, None]).astype(np.float32)
        queries.append(np.concatenate([deep, wl.mixed_queries(6000, v, f, 50 + k)]))
    b = MeshBatch(meshes, branching=4, leaf_size=16)
    items = [(k, torch.from_numpy(q).to(cuda)) for k, q in enumerate(queries)]
    plan = b.plan(items, capture=True)
    outs = plan.launch()
    torch.cuda.synchronize()
    assert plan.last_early_packets() > 0
    for k, (v, f) in enumerate(meshes):
        diag = diag_of(v)
        single = TriMeshSDF(v, f, branching=4, leaf_size=16)
        d1 = single.query(items[k][1]).cpu().numpy()
        d0 = outs[k].cpu().numpy()
        assert np.abs(d0 - d1).max() <= 1e-6 * diag, k
        assert np.array_equal(np.sign(d0), np.sign(d1)), k
        sub = np.arange(0, queries[k].shape[0], 20)
        ref, _ = port.Mesh(v.astype(np.float64), f).sdf(queries[k][sub].astype(np.float64))
        assert np.abs(d0[sub] - ref).max() <= D_TOL * diag, k
        assert np.array_equal(np.sign(d0[sub]), np.sign(ref)), k


def test_more_than_256_items_use_wide_keys(cuda):
    """Plans with more than 256 items sort 64-bit keys (item bits above 8): results must still
    equal the single-mesh queries."""
    import torch
    w = wl.batched_meshes(3, 23, level_range=(2, 3), query_range=(100, 300))
    b = MeshBatch(list(zip(w.verts, w.faces)), branching=4, leaf_size=8)
    rng = np.random.default_rng(4)
    items = []
    for k in range(300):
        mesh = int(rng.integers(0, 3))
        q = w.queries[mesh][: int(rng.integers(1, w.queries[mesh].shape[0] + 1))]
        items.append((mesh, torch.from_numpy(np.ascontiguousarray(q)).to(cuda)))
    outs = b.query(items)
    torch.cuda.synchronize()
    singles = [TriMeshSDF(w.verts[k], w.faces[k], branching=4, leaf_size=8) for k in range(3)]
    for (mesh, q), d0 in zip(items, outs):
        d1 = singles[mesh].query(q).cpu().numpy()
        diag = diag_of(w.verts[mesh])
        assert np.abs(d0.cpu().numpy() - d1).max() <= 1e-6 * diag
        assert np.array_equal(np.sign(d0.cpu().numpy()), np.sign(d1))


@pytest.mark.parametrize("wg", [64, 256])
def test_persistent_plan_equals_grid_plan(cuda, wg):
    """The persistent launch (a grid that fits the device at once, wavefronts taking packets from
    a queue: the early list, then the table) gives the one-wavefront-per-packet launch's results
    bit for bit, on replays of a captured plan, with early packets in play."""
    import torch
    rng = np.random.default_rng(12)
    meshes, queries = [], []
    for k in range(3):
        v, f = wl.displaced_icosphere(5 if k else 3, 60 + k)
        meshes.append((v, f))
        d = rng.normal(size=(4000, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        deep = (d * (0.3 * rng.random(4000) ** (1 / 3))[:, None]).astype(np.float32)
        queries.append(np.concatenate([deep, wl.mixed_queries(7001, v, f, 70 + k)]))
    b = MeshBatch(meshes, branching=4, leaf_size=16)
    items = [(k, torch.from_numpy(q).to(cuda)) for k, q in enumerate(queries)]
    grid = b.plan(items, capture=True, workgroup=wg)
    pers = b.plan(items, capture=True, workgroup=wg, persistent=True)
    assert not grid.persistent and pers.persistent and pers.workgroup == wg
    ref = [o.clone() for o in grid.launch()]
    torch.cuda.synchronize()
    for rep in range(3):
        outs = pers.launch()
        torch.cuda.synchronize()
        for k in range(len(meshes)):
            assert torch.equal(outs[k], ref[k]), (rep, k)
        torch.empty(1024, device=cuda).fill_(float(rep))  # eager work between replays
    assert pers.last_early_packets() > 0
