"""GPU parity of split packets: expensive packets (queries deep inside a closed surface, where
many triangles lie at almost the same distance) are shared among wavefronts that walk disjoint
subtrees; their results must be those of the plain traversal.

Bar: against the same queries in chunks below the split size (2^14, so nothing is split), equal
sign everywhere and |d - d_chunked| <= 1e-6 * diag; against the oracle's FP64 SDF on a strided
subset, |d - d_ref| <= 1e-5 * diag and equal sign. Closest-face queries split no packets (each
lane decides its face from its own whole walk); their faces equal the oracle's up to exact ties.
"""
import numpy as np
import pytest

from ebgeometry_amd import TriMeshSDF
from ebgeometry_amd import workloads as wl
from oracle import port

pytestmark = pytest.mark.gpu


def deep_queries(n, seed):
    """Half inside a ball of radius 0.3 about the centre of the unit-ish displaced sphere (the
    expensive packets), half the usual uniform / near-surface mix."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n // 2, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = 0.3 * rng.random(n // 2) ** (1.0 / 3.0)
    return (d * r[:, None]).astype(np.float32)


@pytest.mark.parametrize("branching,leaf", [(4, 16), (8, 8), (4, 4)])
def test_split_packets_match_plain(cuda, branching, leaf):
    import torch
    v, f = wl.displaced_icosphere(5, 17)
    m = TriMeshSDF(v, f, branching=branching, leaf_size=leaf)
    n = 120_000
    q = np.concatenate([deep_queries(n, 3), wl.mixed_queries(n - n // 2, v, f, 4)])
    qt = torch.from_numpy(q).to(cuda)
    d = m.query(qt)
    split = m.last_split_packets()
    assert split > 0, "the deep queries should have produced split packets"

    chunks = [m.query(qt[i:i + 16_000]) for i in range(0, n, 16_000)]
    assert m.last_split_packets() == 0
    d_ref = torch.cat(chunks)
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    assert torch.equal(torch.sign(d), torch.sign(d_ref))
    assert (d - d_ref).abs().max().item() <= 1e-6 * diag
    om = port.Mesh(v.astype(np.float64), f)

    # the same queries with closest faces: no packet is split, and every face is the oracle's up
    # to an exact tie (faces meeting at the closest vertex or edge)
    d2, tri = m.query(qt, with_triangle=True)
    assert m.last_split_packets() == 0
    assert torch.equal(torch.sign(d2), torch.sign(d))
    assert (d2 - d).abs().max().item() <= 1e-6 * diag
    sub = np.arange(0, n, 7)
    dmin, tmin = om.sdf(q[sub].astype(np.float64))
    t = tri.cpu().numpy()[sub]
    diff = np.nonzero(t != tmin)[0]
    if diff.size:
        fn, vn, en = om.normals()
        for i in diff:
            x = v.astype(np.float64)[f.astype(np.int64)[t[i]]].ravel()
            a = abs(port.triangle_sdf(x, fn[t[i]], vn[t[i]], en[t[i]], q[sub[i]].astype(np.float64)))
            assert a - abs(dmin[i]) <= 1e-12 * diag + 4e-16 * abs(dmin[i]), (sub[i], a, dmin[i])
    print(f"K{branching} L{leaf}: {split} split packets; {diff.size} of {sub.size} faces differ from the "
          f"oracle's, all exact ties", flush=True)

    sub = np.arange(0, n, 60)
    ref, _ = om.sdf(q[sub].astype(np.float64))
    dd = d.cpu().numpy()[sub]
    assert np.abs(dd - ref).max() <= 1e-5 * diag
    assert np.array_equal(np.sign(dd), np.sign(ref))


def test_split_packets_with_learned_bound(cuda):
    """A seeded launch leaves split packets unseeded: results equal the plain launch's."""
    import torch
    v, f = wl.displaced_icosphere(4, 19)
    m = TriMeshSDF(v, f)
    m.train_bound(num_samples=200_000, epochs=2, seed=5)
    n = 80_000
    q = np.concatenate([deep_queries(n, 7), wl.mixed_queries(n - n // 2, v, f, 8)])
    qt = torch.from_numpy(q).to(cuda)
    d0 = m.query(qt)
    d1 = m.query(qt, learned_bound=True)
    assert m.last_split_packets() > 0
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    assert torch.equal(torch.sign(d0), torch.sign(d1))
    assert (d0 - d1).abs().max().item() <= 1e-6 * diag


@pytest.mark.parametrize("name", ["tetrahedron.npz", "dodecahedron.npz"])
def test_split_on_tiny_meshes(cuda, golden, name):
    """Tiny meshes: the subtree frontier has fewer than 16 roots, some of them leaves. A query
    set of 2^15 points (the golden queries tiled, their centroid-heavy part included) must give
    the reference's golden distances whether or not packets are split."""
    import torch
    g = golden(name)
    v, f = g["verts"], g["faces"]
    m = TriMeshSDF(v, f)
    reps = (1 << 15) // g["queries"].shape[0] + 1
    q = np.tile(g["queries"].astype(np.float32), (reps, 1))
    d = m.query(torch.from_numpy(q).to(cuda)).cpu().numpy()
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    ref = np.tile(g["d_sdf"], reps)
    assert np.abs(d - ref).max() <= 1e-5 * diag
    nz = ref != 0
    assert np.array_equal(np.sign(d[nz]), np.sign(ref[nz]))


def test_unaligned_points_give_the_same_results(cuda):
    """A point array that starts 12 bytes into its allocation (a view from the second point) is
    read point by point by the presort key kernel; results must equal those of a contiguous,
    aligned copy of the same points, with and without the FP64 recheck."""
    import torch
    v, f = wl.displaced_icosphere(5, 21)
    m = TriMeshSDF(v, f, branching=4, leaf_size=16)
    n = 70_001  # above the split size, not a multiple of 4, 64 or 4096
    q = np.concatenate([deep_queries(n, 8), wl.mixed_queries(n - n // 2 + 1, v, f, 9)])
    base = torch.from_numpy(q).to(cuda)
    view = base[1:]
    assert view.data_ptr() % 16 != 0
    copy = view.clone()
    assert copy.data_ptr() % 16 == 0
    for kw in ({}, {"fp64_recheck": False}):
        d0 = m.query(view, **kw)
        d1 = m.query(copy, **kw)
        torch.cuda.synchronize()
        assert torch.equal(d0, d1), kw
