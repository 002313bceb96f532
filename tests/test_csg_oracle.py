"""CPU checks of the CSG-program path's host side: the FP64 restatement (oracle/csg_ref.py) run on
the compiled programs against the reference's own value() (golden fixture
tests/golden/csg_algebra.npz, and the compiled reference live where it was built), and the C ABI's
validation of malformed programs (it rejects them before any device work)."""
import ctypes as C
import hashlib
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests"))

import csg_cases as cc  # noqa: E402
from oracle import csg_ref  # noqa: E402


@pytest.fixture(scope="module")
def golden_csg():
    return load_golden("csg_algebra.npz")


@pytest.mark.parametrize("name", cc.NAMES)
def test_oracle_matches_reference_fixture(golden_csg, name):
    f, meshes, lo, hi = cc.case(name)
    q = cc.queries(name)
    assert hashlib.sha256(q.tobytes()).digest() == golden_csg[f"{name}.queries_sha256"].tobytes()
    ref = golden_csg[f"{name}.value"]
    got = csg_ref.evaluate(f, q, meshes)
    assert np.abs(got - ref).max() <= 1e-12 * cc.scale(name), np.abs(got - ref).max()


def _ref_available():
    try:
        from oracle import ref
        return ref.available()
    except Exception:
        return False


@pytest.mark.skipif(not _ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("name", ["smooth_union", "blur", "bvh_union_mixed", "bvh_smooth_union", "mesh_bvh_union",
                                  "repetition", "cone"])
def test_oracle_matches_live_reference_on_fresh_points(name):
    from oracle import ref
    f, meshes, lo, hi = cc.case(name)
    q = cc.queries(name, 3000, seed=99)
    rm, idx = [], {}
    for k, (v, fa) in meshes.items():
        idx[k] = len(rm)
        rm.append(ref.Mesh(np.asarray(v, np.float64), fa))
    r = ref.expr_value(f, q.astype(np.float64), rm, idx)
    assert np.abs(csg_ref.evaluate(f, q, meshes) - r).max() <= 1e-12 * cc.scale(name)


def test_smooth_min_closed_forms():
    """SmoothMin equals min outside the blend band and min - s/4 where the values meet
    (CSG.hpp:256-263); ExpMin of equal values is v - s ln 2 (CSG.hpp:236-245)."""
    from ebgeometry_amd.csg import ExpMin, SmoothMin
    assert SmoothMin(0.3, 1.0, 0.2) == 0.3
    assert SmoothMin(0.5, 0.5, 0.2) == pytest.approx(0.5 - 0.05)
    assert ExpMin(0.5, 0.5, 0.1) == pytest.approx(0.5 - 0.1 * np.log(2.0))


def _create(ops, main=None, prims=()):
    from ebgeometry_amd import _lib
    arr = (_lib.CsgOp * len(ops))()
    for i, (op, n, a, b, p) in enumerate(ops):
        arr[i].op, arr[i].n, arr[i].a, arr[i].b = op, n, a, b
        arr[i].p[:len(p)] = list(p)
    parr = (_lib.CsgPrim * max(1, len(prims)))()
    for j, (s, c, lo, hi) in enumerate(prims):
        parr[j].start, parr[j].count = s, c
        parr[j].bv_lo[:], parr[j].bv_hi[:] = lo, hi
    h = C.c_void_p()
    rc = _lib.lib().ebg_csg_create(arr, len(ops), len(ops) if main is None else main, parr, len(prims), None, 0,
                                   1.0, C.byref(h))
    return rc, _lib.lib().ebg_last_error()


def test_malformed_programs_are_rejected():
    from ebgeometry_amd import program as pg
    sph = (pg.SPHERE, 0, 0, 0, (0, 0, 0, 1))
    assert _create([(pg.MIN_ACC, 0, 0, 0, ())])[0] == 1                       # stack underflow
    assert _create([sph, sph])[0] == 1                                         # two values left
    assert _create([(99, 0, 0, 0, ())])[0] == 1                                # unknown opcode
    assert _create([(pg.TRANSLATE, 0, 0, 0, (1, 0, 0)), sph])[0] == 1         # point never restored
    assert _create([(pg.ROTATE, 5, 0, 0, (1, 0)), sph, (pg.POP_POINT, 0, 0, 0, ())])[0] == 1  # bad axis
    assert _create([(pg.SPHERE, 0, 0, 0, (0, 0, float("nan"), 1))])[0] == 1  # non-finite
    rc, msg = _create([sph] * 17 + [(pg.MIN_ACC, 0, 0, 0, ())] * 16)          # 17 values at once
    assert rc == 6 and b"stack values" in msg
    # a BVH union whose primitive holds another BVH union
    ops = [(pg.BVH_UNION, 0, 0, 1, ()), (pg.BVH_UNION, 0, 0, 1, ())]
    rc, msg = _create(ops, main=1, prims=[(1, 1, (0, 0, 0), (1, 1, 1))])
    assert rc == 1 and b"inside a BVH union" in msg
    # primitive range beyond the list
    assert _create([(pg.BVH_UNION, 0, 0, 2, ()), sph], main=1, prims=[(1, 1, (0, 0, 0), (1, 1, 1))])[0] == 1
