"""Compilation of implicit-function expressions into CSG programs (ebg_csg_*, include/ebgpu.h).

An expression tree of the reference's objects (analytic SDFs, meshes, transforms, CSG operators;
Source/EBGeometry_AnalyticDistanceFunctions.hpp, Transform.hpp, CSG.hpp) is emitted in post order
as a stack program: leaves push their value at the current point, point transforms replace the
current point until a matching pop, value transforms change the top value, and n-ary operators
become folds over their operands. BVH unions carry their primitives as subprograms placed after
the main program.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import CsgOp, CsgPrim, QueryOpts, check, lib

# opcodes (enum ebg_csg_opcode)
(SPHERE, BOX, PLANE, TORUS, CYLINDER, INF_CYLINDER, CAPSULE, INF_CONE, CONE, ROUNDED_BOX, ROUNDED_CYLINDER, MESH,
 TRANSLATE, ROTATE, SCALE_POINT, ELONGATE, REFLECT, REPEAT, POP_POINT,
 SCALE_VALUE, NEGATE, OFFSET, ANNULAR,
 MIN_ACC, MAX_ACC, ADD_ACC, TWO_MIN_ACC, TWO_MAX_ACC, SMOOTH_MIN_END, SMOOTH_MAX_END, DIFFERENCE, PUSH_CONST,
 BVH_UNION, BVH_SMOOTH_UNION) = range(34)

# the reference clamps smoothing lengths to the smallest normal double (CSGImplem.hpp:383, 542, 663)
DBL_MIN = 2.2250738585072014e-308


class Program:
    """Instructions under construction: the main list, BVH-union primitives and meshes."""

    def __init__(self):
        self.main: list = []
        self.subs: list = []     # one instruction list per BVH-union primitive
        self.prims: list = []    # (sub index, bv_lo, bv_hi)
        self.meshes: list = []
        self._mesh_index: dict = {}
        self._target = self.main

    def emit(self, op, n=0, a=0, b=0, p=()):
        self._target.append((int(op), int(n), int(a), int(b), tuple(float(x) for x in p)))

    def mesh(self, m) -> int:
        if id(m) not in self._mesh_index:
            self._mesh_index[id(m)] = len(self.meshes)
            self.meshes.append(m)
        return self._mesh_index[id(m)]

    def bvh(self, prims, bvs, smooth=None, kind=0):
        """A BVH union (smooth when `smooth` is a length) over `prims` with bounding volumes `bvs`."""
        if self._target is not self.main:
            raise NotImplementedError("a BVH union inside a BVH union's primitive")
        first = len(self.prims)
        for f, (lo, hi) in zip(prims, bvs):
            sub: list = []
            self._target = sub
            emit(f, self)
            self._target = self.main
            self.subs.append(sub)
            self.prims.append((len(self.subs) - 1, np.asarray(lo, np.float64), np.asarray(hi, np.float64)))
        if smooth is None:
            self.emit(BVH_UNION, 0, first, len(prims))
        else:
            self.emit(BVH_SMOOTH_UNION, kind, first, len(prims), [max(float(smooth), DBL_MIN)])

    def build(self, length_scale: float = 0.0):
        """Upload: returns the ebg_csg handle."""
        ops = list(self.main)
        starts = []
        for sub in self.subs:
            starts.append(len(ops))
            ops.extend(sub)
        arr = (CsgOp * len(ops))()
        for i, (op, n, a, b, p) in enumerate(ops):
            arr[i].op, arr[i].n, arr[i].a, arr[i].b = op, n, a, b
            arr[i].p[:len(p)] = list(p)
        parr = (CsgPrim * max(1, len(self.prims)))()
        for j, (k, lo, hi) in enumerate(self.prims):
            parr[j].start, parr[j].count = starts[k], len(self.subs[k])
            parr[j].bv_lo[:] = [float(x) for x in lo]
            parr[j].bv_hi[:] = [float(x) for x in hi]
        mh = (C.c_void_p * max(1, len(self.meshes)))(*[m.handle for m in self.meshes])
        h = C.c_void_p()
        check(lib().ebg_csg_create(arr, len(ops), len(self.main), parr, len(self.prims), mh, len(self.meshes),
                                   float(length_scale), C.byref(h)))
        return h


def emit(f, prog: Program):
    """Emit f into prog: objects provide _emit(prog); the scene-union primitives are covered here."""
    e = getattr(f, "_emit", None)
    if e is None:
        raise TypeError(f"{type(f).__name__} cannot be compiled into a GPU program")
    e(prog)


def rotation_cos_sin(angle_deg: float):
    """cos and sin of the angle as RotateIF's constructor forms them (TransformImplem.hpp:174-177)."""
    th = float(angle_deg) * math.pi / 180.0
    return math.cos(th), math.sin(th)


class CompiledProgram:
    """An uploaded program: value() at points (numpy -> numpy, CUDA tensor -> CUDA tensor) and on
    lattices."""

    def __init__(self, f, length_scale: float = 0.0):
        prog = Program()
        emit(f, prog)
        self._meshes = prog.meshes  # keep the meshes alive with the program
        self._h = prog.build(length_scale)
        self.num_ops = len(prog.main) + sum(len(s) for s in prog.subs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().ebg_csg_destroy(h)
            except Exception:  # pragma: no cover - interpreter teardown
                pass
            self._h = None

    @staticmethod
    def query_opts(presort=True, fp64_recheck=True, recheck_eps=1e-4, fp64_all=False, stream=None) -> QueryOpts:
        o = QueryOpts()
        lib().ebg_query_opts_default(C.byref(o))
        o.presort, o.fp64_recheck, o.recheck_eps, o.fp64_all, o.stream = (int(presort), int(fp64_recheck),
                                                                            float(recheck_eps), int(fp64_all), stream)
        return o

    def value(self, points, out=None, **kw):
        from .sdf import _is_tensor, _torch
        if _is_tensor(points):
            torch = _torch()
            if points.dtype != torch.float32 or not points.is_cuda:
                raise TypeError("device queries take a CUDA float32 tensor of shape (n, 3)")
            p = points.contiguous().view(-1, 3)
            v = out if out is not None else torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
            if "stream" not in kw:
                kw["stream"] = torch.cuda.current_stream(p.device).cuda_stream
            o = self.query_opts(**kw)
            check(lib().ebg_csg_query(self._h, p.data_ptr(), p.shape[0], v.data_ptr(), C.byref(o)))
            return v
        single = np.asarray(points).shape == (3,)
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        v = np.empty(p.shape[0], dtype=np.float32)
        o = self.query_opts(**kw)
        check(lib().ebg_csg_query_host(self._h, p.ctypes.data, p.shape[0], v.ctypes.data, C.byref(o)))
        return float(v[0]) if single else v

    def query_grid(self, origin, spacing, dims, device=None, out=None, **kw):
        from .sdf import _grid_args, _torch
        o3, s3, n3 = _grid_args(origin, spacing, dims)
        shape = (int(n3[2]), int(n3[1]), int(n3[0]))
        if device is not None:
            torch = _torch()
            v = out if out is not None else torch.empty(shape, dtype=torch.float32, device=device)
            if "stream" not in kw:
                kw["stream"] = torch.cuda.current_stream(v.device).cuda_stream
            o = self.query_opts(**kw)
            check(lib().ebg_csg_query_grid(self._h, o3, s3, n3, v.data_ptr(), C.byref(o)))
            return v
        v = np.empty(shape, dtype=np.float32)
        o = self.query_opts(**kw)
        check(lib().ebg_csg_query_grid_host(self._h, o3, s3, n3, v.ctypes.data, C.byref(o)))
        return v

    def check(self):
        check(lib().ebg_csg_check(self._h))
