"""Batch argument validation happens on the host, before any device call (no GPU needed)."""
import ctypes as C

import numpy as np


def test_batch_create_validates_without_gpu():
    from ebgeometry_amd import _lib
    L = _lib.lib()
    v = np.zeros((4, 3), np.float32)
    f = np.array([[0, 1, 9]], np.uint32)  # index out of range
    xyz = (C.c_void_p * 1)(v.ctypes.data)
    tri = (C.c_void_p * 1)(f.ctypes.data)
    nv = (C.c_uint64 * 1)(4)
    nt = (C.c_uint64 * 1)(1)
    h = C.c_void_p()
    assert L.ebg_batch_create(xyz, nv, tri, nt, 1, None, C.byref(h)) == 1
    assert L.ebg_batch_create(xyz, nv, tri, nt, 0, None, C.byref(h)) == 1
    o = _lib.BatchOpts(0, 5, 8, 0)  # branching 5
    f2 = np.array([[0, 1, 2]], np.uint32)
    tri2 = (C.c_void_p * 1)(f2.ctypes.data)
    assert L.ebg_batch_create(xyz, nv, tri2, nt, 1, C.byref(o), C.byref(h)) == 1
    assert L.ebg_batch_plan_create(None, None, 0, None, 0, 0, C.byref(C.c_void_p())) == 1
