"""GPU parity of CSG programs: every analytic SDF, transform and CSG operator of the reference
(tests/csg_cases.py) evaluated by the GPU interpreter against the reference's own value()
(golden fixture tests/golden/csg_algebra.npz, written by the compiled reference) and against the
FP64 restatement (oracle/csg_ref.py).

Bar: |v_gpu - v_ref| <= 1e-5 * L (L the diagonal of the case's query box) and equal signs wherever
|v_ref| > 1e-6 L (FP32 evaluation; values below 1e-4 L are redone in FP64); with every query in
FP64 (fp64_all), |v_gpu - v_ref| <= 1e-7 L (only the float rounding of the output remains).
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests"))

import csg_cases as cc  # noqa: E402
from oracle import csg_ref  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def golden_csg():
    return load_golden("csg_algebra.npz")


@pytest.mark.parametrize("name", cc.NAMES)
def test_program_matches_reference(golden_csg, cuda, name):
    f, meshes, lo, hi = cc.case(name, gpu=True)
    q = cc.queries(name)
    ref = golden_csg[f"{name}.value"]
    L = cc.scale(name)
    print(f"{name}: {q.shape[0]} queries", flush=True)
    v = f.value(q)
    err = np.abs(v.astype(np.float64) - ref)
    assert err.max() <= 1e-5 * L, (name, err.max() / L)
    big = np.abs(ref) > 1e-6 * L
    assert np.array_equal(np.sign(v[big]), np.sign(ref[big])), name
    v64 = f.value(q, fp64_all=True)
    assert np.abs(v64.astype(np.float64) - ref).max() <= 1e-7 * L, name


@pytest.mark.parametrize("name", ["smooth_union", "bvh_union_mixed", "mesh_bvh_union", "mollify"])
def test_program_matches_oracle_on_fresh_points(cuda, name):
    """More points than the fixture, a different seed: against the FP64 restatement."""
    f, meshes, lo, hi = cc.case(name, gpu=True)
    q = cc.queries(name, 20_000, seed=5)
    o = csg_ref.evaluate(f, q, meshes)
    v = f.value(q)
    L = cc.scale(name)
    assert np.abs(v - o).max() <= 1e-5 * L
    big = np.abs(o) > 1e-6 * L
    assert np.array_equal(np.sign(v[big]), np.sign(o[big]))


def test_device_tensor_and_lattice(cuda):
    import torch
    from ebgeometry_amd import grid_points
    f, meshes, lo, hi = cc.case("smooth_difference", gpu=True)
    q = cc.queries("smooth_difference", 50_000, seed=3)
    v_host = f.value(q)
    v_dev = f.value(torch.from_numpy(q).to(cuda))
    torch.cuda.synchronize()
    assert np.array_equal(v_dev.cpu().numpy(), v_host)
    dims = (33, 29, 31)
    spacing = (hi - lo) / (np.array(dims) - 1)
    g = f.query_grid(lo, spacing, dims)
    assert np.array_equal(g.ravel(), f.value(grid_points(lo, spacing, dims)))


def test_bvh_union_of_general_primitives_falls_back_to_a_program(cuda):
    """BVHUnionIF over primitives that are not spheres, boxes or mesh instances is evaluated as a
    program with a BVH union; a union of only spheres and boxes keeps the scene path. Both equal
    the oracle."""
    import ebgeometry_amd as E
    f, meshes, lo, hi = cc.case("bvh_union_mixed", gpu=True)
    q = cc.queries("bvh_union_mixed", 5000, seed=8)
    v = f.value(q)
    assert f._program is not None and f._h is None
    o = csg_ref.evaluate(f, q, meshes)
    assert np.abs(v - o).max() <= 1e-5 * cc.scale("bvh_union_mixed")
    u = E.BVHUnion([E.SphereSDF([0, 0, 0], 1.0), E.BoxSDF([1, 1, 1], [2, 2, 2])])
    w = u.value(q)
    assert u._h is not None and u._program is None
    assert np.abs(w - csg_ref.evaluate(u, q, {})).max() <= 1e-5 * cc.scale("bvh_union_mixed")
