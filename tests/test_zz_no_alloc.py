"""The allocation-free query entry point (ebg_query_sdf_noalloc) against ebg_query_sdf, through
the C ABI: identical distances and faces on two calls in a row (the second smaller, so the
shared scratch array holds the first call's contents), plain and with the learned bound, and
the size limit."""
import ctypes as C

import numpy as np
import pytest

from ebgeometry_amd import TriMeshSDF
from ebgeometry_amd import workloads as wl
from ebgeometry_amd._lib import lib

pytestmark = pytest.mark.gpu

MAX_Q = 1 << 20


def run(entry, m, qt, **kw):
    import torch
    n = qt.shape[0]
    d = torch.full((n,), float("nan"), dtype=torch.float32, device=qt.device)
    t = torch.full((n,), -1, dtype=torch.int32, device=qt.device)
    o = m.query_opts(stream=torch.cuda.current_stream(qt.device).cuda_stream, **kw)
    rc = entry(m.handle, qt.data_ptr(), n, d.data_ptr(), t.data_ptr(), C.byref(o))
    torch.cuda.synchronize()
    assert rc == 0, lib().ebg_last_error()
    return d.cpu().numpy(), t.cpu().numpy()


def test_noalloc_matches_query_sdf(cuda):
    import torch
    v, f = wl.displaced_icosphere(5, 61)
    m = TriMeshSDF(v, f)
    L = lib()
    # a large presorted call first builds the mesh's cost grid, so the allocation-free calls
    # below also take the split-packet path
    big = torch.from_numpy(wl.mixed_queries(300_000, v, f, 62)).to(cuda)
    small = torch.from_numpy(wl.mixed_queries(5_003, v, f, 63)).to(cuda)
    for label, qt in (("300k", big), ("5k", small)):
        print(f"noalloc vs query_sdf, {label} queries", flush=True)
        d0, t0 = run(L.ebg_query_sdf, m, qt)
        d1, t1 = run(L.ebg_query_sdf_noalloc, m, qt)
        assert np.isfinite(d1).all()
        assert np.array_equal(d0, d1), (label, np.abs(d0 - d1).max())
        assert np.array_equal(t0, t1), label
    # learned bound: the network runs on the caller's stream inside the allocation-free call
    rng = np.random.default_rng(5)
    theta = (rng.standard_normal(int(L.ebg_mlp_param_count())) * 0.05).astype(np.float32)
    m.set_bound_params(theta)
    for label, qt in (("300k learned", big), ("5k learned", small)):
        print(f"noalloc vs query_sdf, {label}", flush=True)
        d0, t0 = run(L.ebg_query_sdf, m, qt, learned_bound=True)
        d1, t1 = run(L.ebg_query_sdf_noalloc, m, qt, learned_bound=True)
        assert np.array_equal(d0, d1), (label, np.abs(d0 - d1).max())
        assert np.array_equal(t0, t1), label
    assert L.ebg_mesh_check(m.handle) == 0


def test_noalloc_rejects_batches_beyond_its_scratch(cuda):
    import torch
    v, f = wl.icosphere(3)
    m = TriMeshSDF(v, f)
    n = MAX_Q + 1
    qt = torch.zeros((n, 3), dtype=torch.float32, device=cuda)
    d = torch.empty(n, dtype=torch.float32, device=cuda)
    o = m.query_opts()
    assert lib().ebg_query_sdf_noalloc(m.handle, qt.data_ptr(), n, d.data_ptr(), None, C.byref(o)) == 1
    assert b"at most" in lib().ebg_last_error()
