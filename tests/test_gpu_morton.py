"""GPU Morton codes and radix sort: bit-exact against the reference's SFC code and order."""
import numpy as np
import pytest

from ebgeometry_amd import morton_codes, morton_sort
from ebgeometry_amd import workloads as wl
from oracle import port

pytestmark = pytest.mark.gpu


def to_u64(t):
    return t.cpu().numpy().view(np.uint64)


def test_codes_match_reference_fixture(golden, cuda):
    import torch
    g = golden("morton.npz")
    for pts, codes in [("pts32", "codes32"), ("dup32", "codes_dup"), ("slab32", "codes_slab")]:
        c = morton_codes(torch.from_numpy(g[pts]).to(cuda))
        assert np.array_equal(to_u64(c), g[codes]), pts
    c = morton_codes(torch.from_numpy(g["cent64"]).to(cuda))
    assert np.array_equal(to_u64(c), g["codes64"])


def test_order_matches_reference_fixture(golden, cuda):
    import torch
    g = golden("morton.npz")
    for pts, order in [("pts32", "order32"), ("dup32", "order_dup")]:
        c = morton_codes(torch.from_numpy(g[pts]).to(cuda))
        perm, sc = morton_sort(c)
        perm = perm.cpu().numpy().astype(np.int64)
        codes = to_u64(c)
        ref = g[order].astype(np.int64)
        assert np.array_equal(codes[perm], codes[ref])
        assert np.array_equal(to_u64(sc), codes[ref])
        assert np.array_equal(perm, port.sort_codes(codes).astype(np.int64))  # stable (code, index)


@pytest.mark.parametrize("n", [1, 2, 100, 4095, 4096, 4097, 100_003])
def test_radix_sort_stable_random(cuda, n):
    import torch
    s = wl.Stream(n, "keys")
    keys = s.u64(n) & np.uint64((1 << 63) - 1)
    keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # plenty of duplicates
    kt = torch.from_numpy(keys.view(np.int64)).to(cuda)
    perm, sk = morton_sort(kt)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm.cpu().numpy().astype(np.int64), ref)
    assert np.array_equal(to_u64(sk), keys[ref])


def test_codes_and_sort_full_size(cuda):
    """16M query points of the large workload: codes equal the oracle's, order is a stable sort."""
    import torch
    v, f = wl.displaced_icosphere(5, 1)
    q = wl.mixed_queries(16_000_000, v, f, 2)
    c = morton_codes(torch.from_numpy(q).to(cuda))
    codes = to_u64(c)
    sub = np.arange(0, q.shape[0], 997)
    # bins depend on the global bbox: compare with the oracle on the full cloud's box via a
    # 2-point cloud carrying that box
    lo, hi = q.min(0), q.max(0)
    probe = np.concatenate([lo[None], hi[None], q[sub]]).astype(np.float32)
    ref = port.morton_encode(port.bins_f32(probe))[2:]
    assert np.array_equal(codes[sub], ref)
    perm, sk = morton_sort(c)
    p = perm.cpu().numpy().astype(np.int64)
    assert np.array_equal(np.bincount(p, minlength=q.shape[0]), np.ones(q.shape[0], np.int64))
    sc = to_u64(sk)
    assert np.all(sc[1:] >= sc[:-1])
    assert np.array_equal(codes[p], sc)
    eq = sc[1:] == sc[:-1]
    assert np.all(p[1:][eq] > p[:-1][eq])  # stable
