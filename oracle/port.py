"""ctypes front-end of the C oracle (liboracle.so) -- test infrastructure only.

See ebg_oracle.h for what each function restates (reference file:line).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_i64 = C.POINTER(C.c_int64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_mesh_create.restype = C.c_void_p
        L.orc_mesh_create.argtypes = [_d, C.c_int64, _u32, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.orc_mesh_destroy.argtypes = [C.c_void_p]
        L.orc_mesh_num_nodes.restype = C.c_int64
        L.orc_mesh_num_nodes.argtypes = [C.c_void_p]
        L.orc_mesh_normals.argtypes = [C.c_void_p, _d, _d, _d]
        L.orc_mesh_sdf.argtypes = [C.c_void_p, _d, C.c_int64, _d, _u32, C.c_int]
        L.orc_mesh_sdf_brute.argtypes = [C.c_void_p, _d, C.c_int64, _d, _u32, C.c_int]
        L.orc_mesh_work.argtypes = [C.c_void_p, _d, C.c_int64, _i64, _i64, _i64]
        L.orc_triangle_sdf.restype = C.c_double
        L.orc_triangle_sdf.argtypes = [_d, _d, _d, _d, _d]
        L.orc_triangle_normal.argtypes = [_d, _d]
        L.orc_bins_f32.argtypes = [_f, C.c_int64, _u32]
        L.orc_bins_f64.argtypes = [_d, C.c_int64, _u32]
        L.orc_morton_encode.argtypes = [_u32, C.c_int64, _u64]
        L.orc_morton_decode.argtypes = [_u64, C.c_int64, _u32]
        L.orc_sort_codes.argtypes = [_u64, C.c_int64, _u32]
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.c_int64, _d, _d, C.c_int64, _d, _d, C.c_int64,
                                       C.POINTER(C.c_void_p), _d, _d, _d]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_scene_value.argtypes = [C.c_void_p, _d, C.c_int64, _d, C.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


class Mesh:
    """Oracle mesh SDF (double precision, reference SAH tree, K-ary, W-wide leaf groups)."""

    def __init__(self, verts, faces, K: int = 4, W: int = 4, max_leaf_groups: int = 4):
        v = _f64(verts, (-1, 3))
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        self.nv, self.nt = v.shape[0], f.shape[0]
        self._h = lib().orc_mesh_create(_p(v, _d), self.nv, _p(f, _u32), self.nt, K, W, max_leaf_groups)
        if not self._h:
            raise ValueError("oracle rejected the mesh")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mesh_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def num_nodes(self) -> int:
        return int(lib().orc_mesh_num_nodes(self._h))

    def normals(self):
        fn = np.zeros((self.nt, 3))
        vn = np.zeros((self.nt, 3, 3))
        en = np.zeros((self.nt, 3, 3))
        lib().orc_mesh_normals(self._h, _p(fn, _d), _p(vn, _d), _p(en, _d))
        return fn, vn, en

    def sdf(self, pts, nthreads: int | None = None, brute: bool = False):
        p = _f64(pts, (-1, 3))
        n = p.shape[0]
        d = np.zeros(n)
        t = np.zeros(n, dtype=np.uint32)
        nthreads = nthreads or os.cpu_count() or 1
        fn = lib().orc_mesh_sdf_brute if brute else lib().orc_mesh_sdf
        fn(self._h, _p(p, _d), n, _p(d, _d), _p(t, _u32), nthreads)
        return d, t

    def work(self, pts):
        p = _f64(pts, (-1, 3))
        n = p.shape[0]
        a, b, c = (np.zeros(n, dtype=np.int64) for _ in range(3))
        lib().orc_mesh_work(self._h, _p(p, _d), n, _p(a, _i64), _p(b, _i64), _p(c, _i64))
        return a, b, c


def triangle_sdf(v, fn, vn, en, p) -> float:
    v, fn, vn, en, p = (_f64(x).ravel() for x in (v, fn, vn, en, p))
    return float(lib().orc_triangle_sdf(_p(v, _d), _p(fn, _d), _p(vn, _d), _p(en, _d), _p(p, _d)))


def triangle_normal(v):
    v = _f64(v).ravel()
    n = np.zeros(3)
    lib().orc_triangle_normal(_p(v, _d), _p(n, _d))
    return n


def bins_f32(pts):
    p = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
    b = np.zeros(p.shape, dtype=np.uint32)
    lib().orc_bins_f32(_p(p, _f), p.shape[0], _p(b, _u32))
    return b


def bins_f64(pts):
    p = _f64(pts, (-1, 3))
    b = np.zeros(p.shape, dtype=np.uint32)
    lib().orc_bins_f64(_p(p, _d), p.shape[0], _p(b, _u32))
    return b


def morton_encode(bins):
    b = np.ascontiguousarray(bins, dtype=np.uint32).reshape(-1, 3)
    c = np.zeros(b.shape[0], dtype=np.uint64)
    lib().orc_morton_encode(_p(b, _u32), b.shape[0], _p(c, _u64))
    return c


def morton_decode(codes):
    c = np.ascontiguousarray(codes, dtype=np.uint64)
    b = np.zeros((c.shape[0], 3), dtype=np.uint32)
    lib().orc_morton_decode(_p(c, _u64), c.shape[0], _p(b, _u32))
    return b


def sort_codes(codes):
    c = np.ascontiguousarray(codes, dtype=np.uint64)
    p = np.zeros(c.shape[0], dtype=np.uint32)
    lib().orc_sort_codes(_p(c, _u64), c.shape[0], _p(p, _u32))
    return p


class Scene:
    """Oracle CSG union: spheres, boxes and Translate(Rz(Ry(Rx(Scale(mesh))))) instances."""

    def __init__(self, sphere_c, sphere_r, box_lo, box_hi, meshes, mt, mrot, ms):
        self._keep = [_f64(sphere_c, (-1, 3)), _f64(sphere_r).ravel(), _f64(box_lo, (-1, 3)),
                      _f64(box_hi, (-1, 3)), _f64(mt, (-1, 3)), _f64(mrot, (-1, 3)), _f64(ms).ravel()]
        sc, sr, blo, bhi, t, r, s = self._keep
        self._meshes = list(meshes)
        arr = (C.c_void_p * max(1, len(meshes)))(*[m.handle for m in meshes])
        self._h = lib().orc_scene_create(sc.shape[0], _p(sc, _d), _p(sr, _d), blo.shape[0], _p(blo, _d),
                                         _p(bhi, _d), len(meshes), arr, _p(t, _d), _p(r, _d), _p(s, _d))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_scene_destroy(self._h)
            self._h = None

    def value(self, pts, nthreads: int | None = None):
        p = _f64(pts, (-1, 3))
        v = np.zeros(p.shape[0])
        lib().orc_scene_value(self._h, _p(p, _d), p.shape[0], _p(v, _d), nthreads or os.cpu_count() or 1)
        return v
