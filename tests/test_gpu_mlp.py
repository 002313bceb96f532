"""GPU learned bound: MFMA forward/backward against the CPU oracle, training, exactness."""
import ctypes as C

import numpy as np
import pytest

from ebgeometry_amd import TriMeshSDF
from ebgeometry_amd import workloads as wl
from ebgeometry_amd._lib import check, lib
from oracle import mlp_ref, port

pytestmark = pytest.mark.gpu


def norm_box(mesh, box_scale=1.2):
    info = mesh.info()
    lo, hi = np.array(info["bbox_lo"]), np.array(info["bbox_hi"])
    ctr, h = 0.5 * (lo + hi), 0.5 * box_scale * (hi - lo)
    return (ctr - h).astype(np.float32), (1.0 / h).astype(np.float32), np.float32(info["diag"])


def small_theta(seed):
    s = wl.Stream(seed, "theta-gpu")
    parts = []
    for shape in mlp_ref.SHAPES:
        fan_in = shape[0] if len(shape) == 2 else 128
        a = np.sqrt(6.0 / fan_in)
        parts.append(((s.uniform(int(np.prod(shape))) * 2 - 1) * a).reshape(shape))
    parts[1] += 0.01
    parts[8] *= 0.1  # keep y = log-distance estimate O(1) so exp() stays well conditioned
    return mlp_ref.pack(parts).astype(np.float32)


@pytest.fixture(scope="module")
def mesh():
    v, f = wl.displaced_icosphere(4, 3)
    return TriMeshSDF(v, f), v, f


def test_forward_matches_oracle(cuda, mesh):
    import torch
    m, v, f = mesh
    theta = small_theta(1)
    m.set_bound_params(theta)
    assert np.array_equal(m.get_bound_params(), theta)
    q = wl.mixed_queries(3000, v, f, 5)
    pred = m.predict_bound(torch.from_numpy(q).to(cuda)).cpu().numpy()
    lo, sc, diag = norm_box(m)
    ref = mlp_ref.predict(theta.astype(np.float64), q, lo, sc, float(diag), bf16_mode=True)
    scale = np.abs(ref).max()
    err = np.abs(pred - ref) / scale
    # Same rounding points as the oracle: most outputs agree to FP32 accumulation order; a few
    # move by one BF16 step where an activation sits on a rounding boundary (sin() and the FP32
    # sums differ in the last bits), which a layout or permutation mistake could never mimic.
    # Measured on MI355X: ~90% of queries agree to 1e-7, ~9% move by 1e-4..2e-2 of the scale.
    assert np.median(err) <= 1e-5, np.median(err)
    assert np.quantile(err, 0.8) <= 1e-5, np.quantile(err, 0.8)
    assert np.quantile(err, 0.98) <= 5e-3, np.quantile(err, 0.98)
    assert err.max() <= 3e-2, err.max()
    # and against exact FP32 math the BF16 network stays within BF16 precision everywhere
    ref32 = mlp_ref.predict(theta.astype(np.float64), q, lo, sc, float(diag), bf16_mode=False)
    assert np.abs(pred - ref32).max() <= 5e-2 * scale


@pytest.mark.parametrize("samples", [1000, 4100])
def test_gradient_matches_oracle(cuda, mesh, samples):
    """1000 samples: each weight gradient is one K loop per output tile; 4100 samples (K = 4128
    rows): the weight gradients split K over the grid and add with atomics."""
    import torch
    m, v, f = mesh
    theta = small_theta(2)
    m.set_bound_params(theta)
    q = wl.mixed_queries(samples, v, f, 6)
    d_ref, _ = port.Mesh(v.astype(np.float64), f).sdf(q.astype(np.float64))
    qt = torch.from_numpy(q).to(cuda)
    tt = torch.from_numpy(np.abs(d_ref).astype(np.float32)).to(cuda)
    n = int(lib().ebg_mlp_param_count())
    g = torch.zeros(n, dtype=torch.float32, device=cuda)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    check(lib().ebg_mlp_loss_grad(m.handle, qt.data_ptr(), tt.data_ptr(), q.shape[0], g.data_ptr(),
                                  loss.data_ptr(), None))
    torch.cuda.synchronize()
    lo, sc, diag = norm_box(m)
    L, gref = mlp_ref.loss_and_grad(theta.astype(np.float64), q, np.abs(d_ref).astype(np.float32), lo, sc,
                                    float(diag), bf16_mode=True)
    assert loss.item() / q.shape[0] == pytest.approx(L, rel=2e-3)
    gg = g.cpu().numpy().astype(np.float64)
    o = 0
    for shape in mlp_ref.SHAPES:
        k = int(np.prod(shape))
        a, b = gg[o:o + k], gref[o:o + k]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert err < 2e-2, (shape, o, err)
        o += k


def test_training_learns_and_bound_is_exact(cuda):
    """Train the bound on 400,000 BVH-labelled pairs (batches of 2,048, the last one 640 rows),
    then query 200,000 points with and without it: the answer is unchanged, the network's
    predictions follow |d|, few queries fall back, and the work of the seeded launch (its own
    per-query counts, the fallbacks' re-runs counted apart) is never above the plain traversal's.
    The counts and times are printed: what the trained network gains over the plain method."""
    import time

    import torch
    v, f = wl.displaced_icosphere(5, 21)
    m = TriMeshSDF(v, f)
    assert lib().ebg_mlp_check_schedule(400_000, 2048) == 0
    print("training 400000 pairs, batch 2048", flush=True)
    loss = m.train_bound(num_samples=400_000, epochs=4, batch=2048, seed=7)
    assert np.isfinite(loss) and loss < 1.0
    q = wl.mixed_queries(200_000, v, f, 22)
    qt = torch.from_numpy(q).to(cuda)
    pred = m.predict_bound(qt).cpu().numpy()
    d0, t0 = m.closest_triangle(qt)
    w = torch.zeros(8, dtype=torch.int64, device=cuda)
    d1, t1 = m.closest_triangle(qt, learned_bound=True, work=w.data_ptr())
    wp = torch.zeros(8, dtype=torch.int64, device=cuda)
    m.closest_triangle(qt, work=wp.data_ptr())
    torch.cuda.synchronize()
    d0, d1 = d0.cpu().numpy(), d1.cpu().numpy()
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    # the bound never changes the answer: only exact ties may resolve to another face
    assert np.abs(d0 - d1).max() <= 1e-6 * diag
    assert np.array_equal(np.sign(d0), np.sign(d1))
    # predictions are close to |d| for most queries
    rel = np.abs(pred - np.abs(d0)) / (np.abs(d0) + 0.01 * diag)
    assert np.median(rel) < 0.5, np.median(rel)
    w, wp = w.cpu().numpy(), wp.cpu().numpy()
    fallbacks = int(w[3])
    assert fallbacks < 0.10 * q.shape[0], fallbacks
    n = q.shape[0]
    # the seeded launch's own work stays at the plain traversal's or below (each lane's radius
    # starts below +inf; the visit order of a packet may differ, hence the 1% allowance)
    assert w[0] <= 1.01 * wp[0] and w[1] <= 1.01 * wp[1], (w, wp)
    times = {}
    for name, kw in (("plain", {}), ("learned", {"learned_bound": True})):
        m.query(qt, **kw)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(5):
            m.query(qt, **kw)
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t) / 5 * 1e3
    print(f"learned bound on {n} queries: loss {loss:.4f}, median |pred-|d||/(|d|+0.01 diag) {np.median(rel):.3f}; "
          f"per query plain {wp[0] / n:.2f} nodes {wp[1] / n:.2f} triangles, seeded {w[0] / n:.2f} nodes "
          f"{w[1] / n:.2f} triangles + fallback re-runs {w[6] / n:.3f} nodes {w[7] / n:.3f} triangles "
          f"({fallbacks} fallbacks); per wavefront plain {wp[4] / (n / 64):.1f} nodes {wp[5] / (n / 64):.1f} "
          f"triangles, learned {w[4] / (n / 64):.1f} nodes {w[5] / (n / 64):.1f} triangles; "
          f"pipeline {times['plain']:.3f} ms plain, {times['learned']:.3f} ms learned", flush=True)
    assert lib().ebg_mesh_check(m.handle) == 0


@pytest.mark.parametrize("n", [16_000, 60_000])
def test_poisoned_bound_falls_back_everywhere(cuda, n):
    """A network predicting 0 gives radius 1e-3*diag: every query beyond it must fall back and
    still return the exact answer (also above 2^14 queries, where uncounted launches split
    expensive packets: a counting launch splits none)."""
    import torch
    v, f = wl.displaced_icosphere(4, 31)
    m = TriMeshSDF(v, f)
    theta = np.zeros(int(lib().ebg_mlp_param_count()), np.float32)
    theta[-1] = -20.0  # y = -20: prediction exp(-20) - 0.005 < 0, clamped to a zero radius
    m.set_bound_params(theta)
    q = wl.uniform_box(n, [-1.5] * 3, [1.5] * 3, 32)
    qt = torch.from_numpy(q).to(cuda)
    d0 = m.query(qt)
    w = torch.zeros(8, dtype=torch.int64, device=cuda)
    d1 = m.query(qt, learned_bound=True, work=w.data_ptr())
    torch.cuda.synchronize()
    diag = float(np.linalg.norm(v.max(0).astype(np.float64) - v.min(0)))
    far = (torch.abs(d0) > 2e-3 * diag).sum().item()
    assert w[3].item() >= far
    assert torch.equal(torch.sign(d0), torch.sign(d1))
    assert (d0 - d1).abs().max().item() <= 1e-6 * diag


def test_caller_seeded_radius_is_exact(cuda):
    import torch
    v, f = wl.displaced_icosphere(4, 41)
    m = TriMeshSDF(v, f)
    q = wl.mixed_queries(50_000, v, f, 42)
    qt = torch.from_numpy(q).to(cuda)
    d0 = m.query(qt)
    for fac in [0.5, 1.0001, 3.0]:
        seed = (d0.abs() * fac + 1e-6).contiguous()
        d1 = m.query(qt, seed_radius=seed.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(torch.sign(d0), torch.sign(d1))
        assert (d0 - d1).abs().max().item() <= 1e-6 * 4


def test_bound_save_and_load(cuda, mesh, tmp_path):
    """A saved network loads into a mesh built from the same arrays (same parameters, same
    predictions, and its learned-bound queries equal the plain ones); a mesh of other arrays
    refuses the file, and so does a truncated file."""
    import torch
    from ebgeometry_amd._lib import EbgError
    m, v, f = mesh
    theta = small_theta(7)
    m.set_bound_params(theta)
    path = tmp_path / "bound.bin"
    m.save_bound(path)
    m2 = TriMeshSDF(v, f)
    m2.load_bound(path)
    assert np.array_equal(m2.get_bound_params(), theta)
    q = torch.from_numpy(wl.mixed_queries(20_000, v, f, 8)).to(cuda)
    assert torch.equal(m.predict_bound(q), m2.predict_bound(q))
    assert torch.equal(m2.query(q, learned_bound=True), m2.query(q))
    v3, f3 = wl.displaced_icosphere(3, 4)
    with pytest.raises(EbgError, match="another mesh"):
        TriMeshSDF(v3, f3).load_bound(path)
    short = tmp_path / "short.bin"
    short.write_bytes(path.read_bytes()[:100])
    with pytest.raises(EbgError):
        TriMeshSDF(v, f).load_bound(short)
