"""Host-side CSG logic: how Translate / Rotate / Scale chains reduce to the (M, t, s) form the
scene kernel evaluates, and which chains are rejected. No GPU needed."""
import ctypes as C

import numpy as np
import pytest

from ebgeometry_amd import csg
from ebgeometry_amd.scene import instance_matrix


def apply_chain_pointwise(ops, p):
    """The reference's evaluation order, applied to a point: outermost op first
    (TranslateIF: p - t, RotateIF: R p, ScaleIF: p / s, then times s on the way out)."""
    scale = 1.0
    for kind, arg in ops:
        if kind == "T":
            p = p - arg
        elif kind == "R":
            p = csg.rotation_point_map(*arg) @ p
        else:
            p = p / arg
            scale *= arg
    return p, scale


def build_chain(base, ops):
    f = base
    for kind, arg in reversed(ops):  # innermost first
        if kind == "T":
            f = csg.Translate(f, arg)
        elif kind == "R":
            f = csg.Rotate(f, *arg)
        else:
            f = csg.Scale(f, arg)
    return f


@pytest.mark.parametrize("seed", range(5))
def test_canonical_form_matches_pointwise_chain(seed):
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(6):
        k = rng.integers(3)
        if k == 0:
            ops.append(("T", rng.normal(size=3)))
        elif k == 1:
            ops.append(("R", (float(rng.uniform(0, 360)), int(rng.integers(3)))))
        else:
            ops.append(("S", float(rng.uniform(0.2, 3.0))))
    base = csg.SphereSDF([0.0, 0.0, 0.0], 1.0)
    f = build_chain(base, ops)
    b, M, t, s = csg._canonical(f)
    assert b is base
    np.testing.assert_allclose(M @ M.T, np.eye(3), atol=1e-12)
    for _ in range(10):
        p = rng.normal(size=3) * 3
        q_ref, s_ref = apply_chain_pointwise(ops, p)
        np.testing.assert_allclose(M @ (p - t) / s, q_ref, rtol=1e-12, atol=1e-12)
        assert abs(s - s_ref) < 1e-12 * s_ref


def test_instance_chain_is_rx_ry_rz():
    rot = (30.0, 110.0, 250.0)
    m = csg.SphereSDF()
    f = csg.Translate(csg.Rotate(csg.Rotate(csg.Rotate(csg.Scale(m, 0.5), rot[0], 0), rot[1], 1), rot[2], 2),
                      [1.0, 2.0, 3.0])
    _, M, t, s = csg._canonical(f)
    np.testing.assert_allclose(M, instance_matrix(rot), atol=1e-15)
    np.testing.assert_allclose(t, [1, 2, 3])
    assert s == 0.5


def test_sphere_under_similarity_is_a_sphere():
    f = csg.Translate(csg.Rotate(csg.Scale(csg.SphereSDF([1.0, 0.0, 0.0], 0.5), 2.0), 90.0, 2), [0.0, 0.0, 1.0])
    p = csg._to_prim(f, {}, [])
    assert p.type == csg.PRIM_SPHERE
    # Scale: centre (2,0,0) r 1; Rotate z by 90: f(Rp) -> centre R^T (2,0,0) = (0,2,0); Translate z
    np.testing.assert_allclose(list(p.p)[:4], [0.0, 2.0, 1.0, 1.0], atol=1e-12)


def test_box_rotations():
    b = csg.BoxSDF([0.0, 0.0, 0.0], [1.0, 2.0, 3.0])
    p = csg._to_prim(csg.Rotate(b, 90.0, 2), {}, [])
    assert p.type == csg.PRIM_BOX
    lo, hi = np.array(p.p[:3]), np.array(p.p[3:6])
    # f(Rp) with R the z-rotation point map: world box = R^T [lo, hi]
    np.testing.assert_allclose(lo, [-2.0, 0.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(hi, [0.0, 1.0, 3.0], atol=1e-12)
    with pytest.raises(NotImplementedError):
        csg._to_prim(csg.Rotate(b, 30.0, 0), {}, [])


def test_unsupported_primitive():
    class Other(csg.ImplicitFunction):
        pass
    with pytest.raises(TypeError):
        csg._to_prim(Other(), {}, [])
    with pytest.raises(ValueError):
        csg.Scale(csg.SphereSDF(), 0.0)


def test_scene_create_validates_without_gpu():
    """Argument validation happens before any device call."""
    from ebgeometry_amd import _lib
    L = _lib.lib()
    p = (_lib.Prim * 1)()
    p[0].type = 2
    p[0].mesh = 3
    p[0].scale = 1.0
    p[0].bv_lo[0], p[0].bv_hi[0] = 1.0, 0.0
    h = C.c_void_p()
    assert L.ebg_scene_create(p, 1, None, 0, None, C.byref(h)) == 1
    p[0].type = 0
    p[0].p[3] = -1.0
    assert L.ebg_scene_create(p, 1, None, 0, None, C.byref(h)) == 1
    p[0].type = 7
    assert L.ebg_scene_create(p, 1, None, 0, None, C.byref(h)) == 1
    assert L.ebg_scene_create(p, 0, None, 0, None, C.byref(h)) == 1
