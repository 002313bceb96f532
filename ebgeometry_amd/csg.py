"""CSG union scenes on the GPU: analytic primitives, transforms and BVHUnionIF.

Mirrors the reference's interface:

    SphereSDF(center, radius), BoxSDF(lo, hi)      Source/EBGeometry_AnalyticDistanceFunctions.hpp:137-423
    Translate(f, t), Rotate(f, angle, axis), Scale(f, s)
                                                   Source/EBGeometry_Transform.hpp:44-70,
                                                   TransformImplem.hpp:140-270
    BVHUnionIF(prims, bvs) / BVHUnion(prims, bvs)  Source/EBGeometry_CSG.hpp:99-101, 401-440;
                                                   value(): CSGImplem.hpp:466-501

A union is evaluated as one GPU scene (ebg_scene_*): every primitive is reduced to a sphere, an
axis-aligned box or a mesh instance ``s * mesh(M (p - t) / s)``. A chain of Translate / Rotate
/ Scale always has that form (each step composes into (M, t, s) below); over a sphere it is a
sphere again, over a box it is a box when M permutes axes (possibly with sign flips).
Single primitives evaluated on their own go through a one-primitive scene.

Everything else of the reference's algebra -- the other analytic SDFs, the other transforms, the
other CSG operators, and BVH unions of arbitrary expressions -- compiles into a CSG program
(ebg_csg_*, see program.py) evaluated by the GPU interpreter:

    PlaneSDF, TorusSDF, CylinderSDF, InfiniteCylinderSDF, CapsuleSDF, InfiniteConeSDF, ConeSDF,
    RoundedBoxSDF, RoundedCylinderSDF          AnalyticDistanceFunctions.hpp:41-1750
    Complement, Offset, Annular, Blur, Mollify, Elongate, Reflect
                                                   Transform.hpp:78-934, TransformImplem.hpp
    Union, SmoothUnion, Intersection, SmoothIntersection, Difference, SmoothDifference,
    FiniteRepetition, BVHSmoothUnion            CSG.hpp:40-866, CSGImplem.hpp
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import program as pg
from ._lib import BuildOpts, Prim, QueryOpts, check, lib
from .scene import rotation_point_map
from .sdf import ImplicitFunction, SignedDistanceFunction, TriMeshSDF, _is_tensor, _torch

PRIM_SPHERE, PRIM_BOX, PRIM_MESH = 0, 1, 2


class AABB:
    """Axis-aligned bounding volume (the reference's BoundingVolumes::AABBT)."""

    def __init__(self, lo, hi):
        self.lo = np.asarray(lo, np.float64).reshape(3)
        self.hi = np.asarray(hi, np.float64).reshape(3)


class _GpuScalarField(ImplicitFunction):
    """value() through a (lazily built) one-primitive GPU scene, or a CSG program when the
    function is not a scene primitive."""

    _evaluator = None

    def value(self, points, **kw):
        if self._evaluator is None:
            try:
                _to_prim(self, {}, [])  # a sphere, box or mesh instance: a one-primitive scene
                self._evaluator = BVHUnionIF([self])
            except (TypeError, NotImplementedError):
                self._evaluator = pg.CompiledProgram(self)
        return self._evaluator.value(points, **kw)

    def query_grid(self, origin, spacing, dims, **kw):
        if self._evaluator is None:
            self.value(np.zeros((1, 3), np.float32))
        return self._evaluator.query_grid(origin, spacing, dims, **kw)


class SphereSDF(SignedDistanceFunction, _GpuScalarField):
    """|p - c| - r (AnalyticDistanceFunctions.hpp:240-247)."""

    def __init__(self, center=(0.0, 0.0, 0.0), radius: float = 1.0):
        self.center = np.asarray(center, np.float64).reshape(3)
        self.radius = float(radius)

    def signed_distance(self, points):
        return _GpuScalarField.value(self, points)

    def bounding_volume(self) -> AABB:
        return AABB(self.center - self.radius, self.center + self.radius)

    def _emit(self, prog):
        prog.emit(pg.SPHERE, p=[*self.center, self.radius])


class BoxSDF(SignedDistanceFunction, _GpuScalarField):
    """Axis-aligned box [lo, hi] (AnalyticDistanceFunctions.hpp:377-399)."""

    def __init__(self, lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5)):
        self.lo = np.asarray(lo, np.float64).reshape(3)
        self.hi = np.asarray(hi, np.float64).reshape(3)

    def signed_distance(self, points):
        return _GpuScalarField.value(self, points)

    def bounding_volume(self) -> AABB:
        return AABB(self.lo, self.hi)

    def _emit(self, prog):
        prog.emit(pg.BOX, p=[*self.lo, *self.hi])


class _Transformed(_GpuScalarField):
    """f under a similarity map: value(p) = s * f(M (p - t) / s)."""

    def __init__(self, f: ImplicitFunction):
        self.f = f

    def canonical(self):
        """(base, M, t, s) with value(p) = s * base(M (p - t) / s)."""
        base, M, t, s = _canonical(self.f)
        return self._apply(base, M, t, s)


class TranslateIF(_Transformed):
    """value(p) = f(p - shift) (TransformImplem.hpp:140-160)."""

    def __init__(self, f, shift):
        super().__init__(f)
        self.shift = np.asarray(shift, np.float64).reshape(3)

    def _apply(self, base, M, t, s):
        return base, M, t + self.shift, s

    def _emit(self, prog):
        prog.emit(pg.TRANSLATE, p=self.shift)
        pg.emit(self.f, prog)
        prog.emit(pg.POP_POINT)


class RotateIF(_Transformed):
    """value(p) = f(R p), R the axis rotation by angle degrees (TransformImplem.hpp:162-226)."""

    def __init__(self, f, angle_deg: float, axis: int):
        super().__init__(f)
        self.R = rotation_point_map(angle_deg, axis)
        self.angle, self.axis = float(angle_deg), int(axis)

    def _apply(self, base, M, t, s):
        return base, M @ self.R, self.R.T @ t, s

    def _emit(self, prog):
        prog.emit(pg.ROTATE, n=self.axis, p=pg.rotation_cos_sin(self.angle))
        pg.emit(self.f, prog)
        prog.emit(pg.POP_POINT)


class ScaleIF(_Transformed):
    """value(p) = s * f(p / s) (TransformImplem.hpp:245-270)."""

    def __init__(self, f, scale: float):
        super().__init__(f)
        if not scale > 0:
            raise ValueError("scale must be > 0")
        self.scale = float(scale)

    def _apply(self, base, M, t, s):
        return base, M, self.scale * t, self.scale * s

    def _emit(self, prog):
        prog.emit(pg.SCALE_POINT, p=[self.scale])
        pg.emit(self.f, prog)
        prog.emit(pg.POP_POINT)
        prog.emit(pg.SCALE_VALUE, p=[self.scale])


def Translate(f, shift) -> TranslateIF:
    return TranslateIF(f, shift)


def Rotate(f, angle_deg: float, axis: int) -> RotateIF:
    return RotateIF(f, angle_deg, axis)


def Scale(f, scale: float) -> ScaleIF:
    return ScaleIF(f, scale)


def _canonical(f):
    if isinstance(f, _Transformed):
        return f.canonical()
    return f, np.eye(3), np.zeros(3), 1.0


def _signed_permutation(M, tol=1e-12) -> bool:
    A = np.abs(M)
    return bool(np.all((np.abs(A - 1.0) < tol) | (A < tol)) and np.all(np.abs(A.sum(0) - 1) < tol))


def _to_prim(f, mesh_index: dict, meshes: list) -> Prim:
    base, M, t, s = _canonical(f)
    p = Prim()
    p.bv_lo[0], p.bv_hi[0] = 1.0, 0.0  # computed by the library unless the caller gives one
    if isinstance(base, SphereSDF):
        c = t + s * (M.T @ base.center)
        p.type = PRIM_SPHERE
        p.p[0:4] = [*c, s * base.radius]
    elif isinstance(base, BoxSDF):
        if not _signed_permutation(M):
            raise NotImplementedError("a box under a rotation that does not permute axes is not axis-aligned")
        a = t + s * (M.T @ base.lo)
        b = t + s * (M.T @ base.hi)
        p.type = PRIM_BOX
        p.p[0:6] = [*np.minimum(a, b), *np.maximum(a, b)]
    elif isinstance(base, TriMeshSDF):
        if id(base) not in mesh_index:
            mesh_index[id(base)] = len(meshes)
            meshes.append(base)
        p.type = PRIM_MESH
        p.mesh = mesh_index[id(base)]
        p.translate[:] = list(t)
        p.matrix[:] = list(M.reshape(9))
        p.scale = float(s)
    else:
        raise TypeError(f"unsupported primitive {type(base).__name__} in a GPU union")
    return p


class BVHUnionIF(ImplicitFunction):
    """Union (minimum) of primitives accelerated by a BVH over their bounding volumes.

    prims: SphereSDF / BoxSDF / TriMeshSDF, optionally wrapped in Translate / Rotate / Scale.
    bvs: one bounding volume per primitive (AABB or a (lo, hi) pair), as the reference's
    constructor takes them; None computes them (exact for spheres and boxes, the transformed
    vertex bounds for meshes). value(p) = min over primitives, as BVHUnionIF::value returns.

    The device side is built at the first query: a scene when every primitive is a sphere, a box
    or a mesh instance, else a CSG program with a BVH union. Inside another expression the union
    is compiled into that expression's program instead.
    """

    def __init__(self, prims, bvs=None, leaf_size: int = 4):
        prims = list(prims)
        if bvs is not None and len(bvs) != len(prims):
            raise ValueError("one bounding volume per primitive")
        self._prims = prims
        self._bvs = None if bvs is None else [(b.lo, b.hi) if isinstance(b, AABB) else b for b in bvs]
        self._leaf = int(leaf_size)
        self._h = None
        self._program = None
        self.num_primitives = len(prims)

    def _ensure(self):
        if self._h is not None or self._program is not None:
            return
        try:
            self._build_scene(self._prims, self._bvs, self._leaf)
        except (TypeError, NotImplementedError):
            # some primitive is not a sphere, box or mesh instance: a CSG program with a BVH union
            self._program = pg.CompiledProgram(self)

    def _bounds(self):
        if self._bvs is not None:
            return self._bvs
        out = []
        for f in self._prims:
            bv = getattr(f, "bounding_volume", None)
            if bv is None:
                raise ValueError(f"{type(f).__name__}: give BVHUnionIF a bounding volume per primitive")
            b = bv()
            out.append((b.lo, b.hi))
        return out

    def _emit(self, prog):
        prog.bvh(self._prims, self._bounds())

    def _build_scene(self, prims, bvs, leaf_size):
        mesh_index: dict = {}
        self._meshes: list = []
        arr = (Prim * max(1, len(prims)))()
        for i, f in enumerate(prims):
            q = _to_prim(f, mesh_index, self._meshes)
            if bvs is not None:
                lo, hi = bvs[i]
                q.bv_lo[:] = [float(x) for x in np.asarray(lo, np.float64).reshape(3)]
                q.bv_hi[:] = [float(x) for x in np.asarray(hi, np.float64).reshape(3)]
            arr[i] = q
        mh = (C.c_void_p * max(1, len(self._meshes)))(*[m.handle for m in self._meshes])
        opts = BuildOpts(0, 4, int(leaf_size), 0)
        h = C.c_void_p()
        check(lib().ebg_scene_create(arr, len(prims), mh, len(self._meshes), C.byref(opts), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().ebg_scene_destroy(h)
            except Exception:  # pragma: no cover - interpreter teardown
                pass
            self._h = None

    @property
    def handle(self):
        self._ensure()
        return self._h

    def bounds(self):
        self._ensure()
        lo = (C.c_double * 3)()
        hi = (C.c_double * 3)()
        d = C.c_double()
        check(lib().ebg_scene_bounds(self._h, lo, hi, C.byref(d)))
        return np.array(lo[:]), np.array(hi[:]), d.value

    @staticmethod
    def query_opts(presort=True, fp64_recheck=True, recheck_eps=1e-4, fp64_all=False, stream=None,
                   work=None, learned_bound=False, bound_scale=1.25, bound_offset=1e-3,
                   seed_radius=None) -> QueryOpts:
        o = QueryOpts()
        lib().ebg_query_opts_default(C.byref(o))
        o.presort = int(presort)
        o.fp64_recheck = int(fp64_recheck)
        o.recheck_eps = float(recheck_eps)
        o.fp64_all = int(fp64_all)
        o.stream = stream
        o.work = work
        o.learned_bound = int(learned_bound)
        o.bound_scale = float(bound_scale)
        o.bound_offset = float(bound_offset)
        o.seed_radius = seed_radius
        return o

    def train_bound(self, num_samples: int = 8_000_000, epochs: int = 6, batch: int = 4096, lr: float = 2e-3,
                    seed: int = 12345, near_sigma: float = 0.02, box_scale: float = 1.2, under_weight: float = 16.0,
                    stream=None) -> float:
        """Train the union's learned cull bound on the GPU (the scene's own walk labels the
        samples); value(..., learned_bound=True) then seeds each query's walk. Returns the last
        epoch's mean loss."""
        self._ensure()
        if self._program is not None:
            raise NotImplementedError("the learned bound is for unions evaluated as a scene")
        from . import _lib
        o = _lib.MlpOpts()
        lib().ebg_mlp_opts_default(C.byref(o))
        o.num_samples, o.epochs, o.batch, o.lr, o.seed = num_samples, epochs, batch, lr, seed
        o.near_sigma, o.box_scale, o.stream, o.under_weight = near_sigma, box_scale, stream, under_weight
        loss = C.c_float(0.0)
        check(lib().ebg_scene_mlp_train(self._h, C.byref(o), C.byref(loss)))
        return float(loss.value)

    def predict_bound(self, pts_tensor):
        """The trained network's predicted |value| at CUDA float32 points (n, 3)."""
        torch = _torch()
        p = pts_tensor.contiguous().view(-1, 3)
        out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
        check(lib().ebg_scene_mlp_predict(self._h, p.data_ptr(), p.shape[0], out.data_ptr(),
                                          torch.cuda.current_stream(p.device).cuda_stream))
        return out

    def value(self, points, out=None, **kw):
        """Union value at a point (float) or at every row of an (n, 3) array / CUDA tensor."""
        self._ensure()
        if self._program is not None:
            kw.pop("work", None)
            return self._program.value(points, out=out, **kw)
        if _is_tensor(points):
            torch = _torch()
            if points.dtype != torch.float32 or not points.is_cuda:
                raise TypeError("device queries take a CUDA float32 tensor of shape (n, 3)")
            p = points.contiguous().view(-1, 3)
            v = out if out is not None else torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
            if "stream" not in kw:
                kw["stream"] = torch.cuda.current_stream(p.device).cuda_stream
            o = self.query_opts(**kw)
            check(lib().ebg_scene_query(self._h, p.data_ptr(), p.shape[0], v.data_ptr(), C.byref(o)))
            return v
        single = np.asarray(points).shape == (3,)
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        v = np.empty(p.shape[0], dtype=np.float32)
        o = self.query_opts(**kw)
        check(lib().ebg_scene_query_host(self._h, p.ctypes.data, p.shape[0], v.ctypes.data, C.byref(o)))
        return float(v[0]) if single else v

    def query_grid(self, origin, spacing, dims, device=None, out=None, **kw):
        """Union value at every lattice point origin + (i, j, k) * spacing, shape
        (dims[2], dims[1], dims[0]) (x fastest). CUDA tensor with device, else numpy."""
        self._ensure()
        if self._program is not None:
            return self._program.query_grid(origin, spacing, dims, device=device, out=out, **kw)
        from .sdf import _grid_args
        o3, s3, n3 = _grid_args(origin, spacing, dims)
        shape = (int(n3[2]), int(n3[1]), int(n3[0]))
        if device is not None:
            torch = _torch()
            v = out if out is not None else torch.empty(shape, dtype=torch.float32, device=device)
            if "stream" not in kw:
                kw["stream"] = torch.cuda.current_stream(v.device).cuda_stream
            o = self.query_opts(**kw)
            check(lib().ebg_scene_query_grid(self._h, o3, s3, n3, v.data_ptr(), C.byref(o)))
            return v
        v = np.empty(shape, dtype=np.float32)
        o = self.query_opts(**kw)
        check(lib().ebg_scene_query_grid_host(self._h, o3, s3, n3, v.ctypes.data, C.byref(o)))
        return v

    def last_traversal_ms(self) -> float:
        self._ensure()
        return float(lib().ebg_scene_last_traversal_ms(self._h)) if self._h else -1.0


def BVHUnion(prims, bvs=None, leaf_size: int = 4) -> BVHUnionIF:
    return BVHUnionIF(prims, bvs, leaf_size)


# ---------------------------------------------------------------------------------------------
# The rest of the algebra: evaluated as CSG programs
# ---------------------------------------------------------------------------------------------

class _Programmed(ImplicitFunction):
    """value() / query_grid() through a (lazily compiled) CSG program of this expression."""

    _compiled = None
    length_scale = 0.0

    def compiled(self) -> "pg.CompiledProgram":
        if self._compiled is None:
            self._compiled = pg.CompiledProgram(self, self.length_scale)
        return self._compiled

    def value(self, points, **kw):
        return self.compiled().value(points, **kw)

    def query_grid(self, origin, spacing, dims, **kw):
        return self.compiled().query_grid(origin, spacing, dims, **kw)


class _AnalyticSDF(SignedDistanceFunction, _Programmed):
    def value(self, points, **kw):
        return _Programmed.value(self, points, **kw)

    def signed_distance(self, points, **kw):
        return _Programmed.value(self, points, **kw)


def _v3(x):
    return np.asarray(x, np.float64).reshape(3)


class PlaneSDF(_AnalyticSDF):
    """(p - point) . n, n the unit normal (AnalyticDistanceFunctions.hpp:41-128)."""

    def __init__(self, point=(0.0, 0.0, 0.0), normal=(0.0, 1.0, 0.0)):
        self.point = _v3(point)
        n = _v3(normal)
        self.normal = n / math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])

    def _emit(self, prog):
        prog.emit(pg.PLANE, p=[*self.point, *self.normal])


class TorusSDF(_AnalyticSDF):
    """Torus about the z axis (AnalyticDistanceFunctions.hpp:424-579)."""

    def __init__(self, center=(0.0, 0.0, 0.0), major_radius: float = 1.0, minor_radius: float = 0.5):
        self.center, self.major, self.minor = _v3(center), float(major_radius), float(minor_radius)

    def bounding_volume(self) -> AABB:
        r = self.major + self.minor
        return AABB(self.center - [r, r, self.minor], self.center + [r, r, self.minor])

    def _emit(self, prog):
        prog.emit(pg.TORUS, p=[*self.center, self.major, self.minor])


class CylinderSDF(_AnalyticSDF):
    """Capped cylinder between two centres (AnalyticDistanceFunctions.hpp:590-766)."""

    def __init__(self, center1=(0.0, -0.5, 0.0), center2=(0.0, 0.5, 0.0), radius: float = 1.0):
        c1, c2 = _v3(center1), _v3(center2)
        self.center1, self.center2, self.radius = c1, c2, float(radius)
        self.center = (c2 + c1) * 0.5
        d = c2 - c1
        self.length = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        self.axis = d / self.length

    def bounding_volume(self) -> AABB:
        r = self.radius
        return AABB(np.minimum(self.center1, self.center2) - r, np.maximum(self.center1, self.center2) + r)

    def _emit(self, prog):
        prog.emit(pg.CYLINDER, p=[*self.center, *self.axis, self.length, self.radius])


class InfiniteCylinderSDF(_AnalyticSDF):
    """Cylinder along a coordinate axis (AnalyticDistanceFunctions.hpp:777-873)."""

    def __init__(self, center=(0.0, 0.0, 0.0), radius: float = 1.0, axis: int = 1):
        if axis not in (0, 1, 2):
            raise ValueError("axis must be 0, 1 or 2")
        self.center, self.radius, self.axis = _v3(center), float(radius), int(axis)

    def _emit(self, prog):
        prog.emit(pg.INF_CYLINDER, n=self.axis, p=[*self.center, self.radius])


class CapsuleSDF(_AnalyticSDF):
    """Capsule with tips tip1, tip2 (AnalyticDistanceFunctions.hpp:892-1001)."""

    def __init__(self, tip1=(0.0, -0.5, 0.0), tip2=(0.0, 0.5, 0.0), radius: float = 0.5):
        t1, t2, r = _v3(tip1), _v3(tip2), float(radius)
        self.tip1, self.tip2 = t1, t2
        d = t2 - t1
        axis = d / math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        self.center1, self.center2, self.radius = t1 + r * axis, t2 - r * axis, r
        self._lo, self._hi = np.minimum(t1, t2) - r, np.maximum(t1, t2) + r

    def bounding_volume(self) -> AABB:
        return AABB(self._lo, self._hi)

    def _emit(self, prog):
        prog.emit(pg.CAPSULE, p=[*self.center1, *self.center2, self.radius])


class InfiniteConeSDF(_AnalyticSDF):
    """Cone opening along -z from its tip, full opening angle in degrees
    (AnalyticDistanceFunctions.hpp:1018-1111)."""

    def __init__(self, tip=(0.0, 0.0, 0.0), angle_deg: float = 45.0):
        self.tip, self.angle = _v3(tip), float(angle_deg)
        a = 0.5 * float(angle_deg) * math.pi / 180.0
        self.c = (math.sin(a), math.cos(a))

    def _emit(self, prog):
        prog.emit(pg.INF_CONE, p=[*self.tip, *self.c])


class ConeSDF(_AnalyticSDF):
    """Capped cone of the given height (AnalyticDistanceFunctions.hpp:1125-1239)."""

    def __init__(self, tip=(0.0, 0.0, 0.0), height: float = 1.0, angle_deg: float = 45.0):
        self.tip, self.height, self.angle = _v3(tip), float(height), float(angle_deg)
        a = 0.5 * float(angle_deg) * math.pi / 180.0
        self.c = (math.sin(a), math.cos(a))

    def _emit(self, prog):
        prog.emit(pg.CONE, p=[*self.tip, *self.c, self.height])


class RoundedBoxSDF(_AnalyticSDF):
    """Box of the given dimensions about the origin, rounded by `curvature`
    (AnalyticDistanceFunctions.hpp:1251-1339)."""

    def __init__(self, dimensions=(1.0, 1.0, 1.0), curvature: float = 0.1):
        self.dimensions = _v3(dimensions)
        self.half, self.curvature = 0.5 * self.dimensions, float(curvature)

    def bounding_volume(self) -> AABB:
        r = self.half + self.curvature
        return AABB(-r, r)

    def _emit(self, prog):
        prog.emit(pg.ROUNDED_BOX, p=[*self.half, self.curvature])


class RoundedCylinderSDF(_AnalyticSDF):
    """Cylinder about the y axis with rounded rims (AnalyticDistanceFunctions.hpp:1652-1750)."""

    def __init__(self, radius: float = 1.0, curvature: float = 0.1, height: float = 1.0):
        self.args = (float(radius), float(curvature), float(height))
        self.major = float(radius) - float(curvature)
        self.minor = float(curvature)
        self.half_height = 0.5 * float(height) - float(curvature)

    def _emit(self, prog):
        prog.emit(pg.ROUNDED_CYLINDER, p=[self.major, self.minor, self.half_height])


class PerlinSDF(_AnalyticSDF):
    """Not on the GPU: the reference's noise field (AnalyticDistanceFunctions.hpp:1358-1640) depends
    on std::shuffle of a seeded std::mt19937, which this library does not restate."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("PerlinSDF is not supported on the GPU")


def _sdf_emit_mesh(self, prog):
    prog.emit(pg.MESH, n=prog.mesh(self))


TriMeshSDF._emit = _sdf_emit_mesh


class _Unary(_Programmed):
    def __init__(self, f):
        self.f = f


class ComplementIF(_Unary):
    """-f (TransformImplem.hpp:129-137)."""

    def _emit(self, prog):
        pg.emit(self.f, prog)
        prog.emit(pg.NEGATE)


class OffsetIF(_Unary):
    """f - offset (TransformImplem.hpp:239-247)."""

    def __init__(self, f, offset: float):
        super().__init__(f)
        self.offset = float(offset)

    def _emit(self, prog):
        pg.emit(self.f, prog)
        prog.emit(pg.OFFSET, p=[self.offset])


class AnnularIF(_Unary):
    """|f| - delta (TransformImplem.hpp:284-292)."""

    def __init__(self, f, delta: float):
        super().__init__(f)
        self.delta = float(delta)

    def _emit(self, prog):
        pg.emit(self.f, prog)
        prog.emit(pg.ANNULAR, p=[self.delta])


class ElongateIF(_Unary):
    """f(p - clamp(p, -e, e)) (TransformImplem.hpp:427-435)."""

    def __init__(self, f, elongation):
        super().__init__(f)
        self.e = _v3(elongation)

    def _emit(self, prog):
        prog.emit(pg.ELONGATE, p=self.e)
        pg.emit(self.f, prog)
        prog.emit(pg.POP_POINT)


class ReflectIF(_Unary):
    """f(p with coordinate `plane` negated) (TransformImplem.hpp:438-462)."""

    def __init__(self, f, plane: int):
        super().__init__(f)
        self.plane = int(plane)
        self.r = np.ones(3)
        if 0 <= int(plane) <= 2:
            self.r[int(plane)] = -1.0

    def _emit(self, prog):
        prog.emit(pg.REFLECT, p=self.r)
        pg.emit(self.f, prog)
        prog.emit(pg.POP_POINT)


class BlurIF(_Unary):
    """Weighted mean of f over the 27 points p + {-d, 0, d}^3 (TransformImplem.hpp:312-345)."""

    def __init__(self, f, blur_distance: float, alpha: float = 0.5):
        super().__init__(f)
        self.d, self.alpha = float(blur_distance), float(alpha)

    def _emit(self, prog):
        a = self.alpha
        w = [a ** 3, a ** 2 * ((1.0 - a) / 2.0), a * ((1.0 - a) / 2.0) ** 2, ((1.0 - a) / 2.0) ** 3]
        prog.emit(pg.PUSH_CONST, p=[0.0])
        for i in (-1, 0, 1):
            for j in (-1, 0, 1):
                for k in (-1, 0, 1):
                    o = self.d * np.array([i, j, k], np.float64)
                    prog.emit(pg.TRANSLATE, p=-o)  # f(p + o)
                    pg.emit(self.f, prog)
                    prog.emit(pg.POP_POINT)
                    prog.emit(pg.SCALE_VALUE, p=[w[abs(i) + abs(j) + abs(k)]])
                    prog.emit(pg.ADD_ACC)


class MollifyIF(_Unary):
    """Sum of f(p - x_s) w_s over an n^3 lattice of offsets x_s in [-max, max]^3 with weights the
    mollifier's values there, normalised (TransformImplem.hpp:347-407)."""

    def __init__(self, f, mollifier, max_value: float, num_points: int = 2):
        super().__init__(f)
        self.mollifier, self.max_value, self.num_points = mollifier, float(max_value), int(num_points)
        m = abs(float(max_value))
        if m > 0.0 and num_points > 1:
            dx = 2 * m / (num_points - 1)
            pos = np.array([[-m + i * dx, -m + j * dx, -m + k * dx] for i in range(num_points)
                            for j in range(num_points) for k in range(num_points)], np.float64)
            if isinstance(mollifier, SphereSDF):  # the Mollify() factory's mollifier: exact in FP64
                w = np.linalg.norm(pos - mollifier.center, axis=1) - mollifier.radius
            else:
                w = np.asarray(mollifier.value(pos.astype(np.float32)), np.float64)
            self.samples = list(zip(pos, w / w.sum()))
        else:
            self.samples = [(np.zeros(3), 1.0)]

    def _emit(self, prog):
        prog.emit(pg.PUSH_CONST, p=[0.0])
        for pos, w in self.samples:
            prog.emit(pg.TRANSLATE, p=pos)
            pg.emit(self.f, prog)
            prog.emit(pg.POP_POINT)
            prog.emit(pg.SCALE_VALUE, p=[w])
            prog.emit(pg.ADD_ACC)


class FiniteRepetitionIF(_Unary):
    """f repeated with the given period, -repeat_lo .. repeat_hi copies per axis
    (CSGImplem.hpp:797-826)."""

    def __init__(self, f, period, repeat_lo, repeat_hi):
        super().__init__(f)
        self.period, self.lo, self.hi = _v3(period), _v3(repeat_lo), _v3(repeat_hi)

    def _emit(self, prog):
        prog.emit(pg.REPEAT, p=[*self.period, *self.lo, *self.hi])
        pg.emit(self.f, prog)
        prog.emit(pg.POP_POINT)


class _Nary(_Programmed):
    def __init__(self, fs):
        self.fs = list(fs)


class UnionIF(_Nary):
    """min over the functions (CSGImplem.hpp:355-366)."""

    def _emit(self, prog):
        if not self.fs:
            prog.emit(pg.PUSH_CONST, p=[math.inf])
            return
        pg.emit(self.fs[0], prog)
        for f in self.fs[1:]:
            pg.emit(f, prog)
            prog.emit(pg.MIN_ACC)


class IntersectionIF(_Nary):
    """max over the functions (CSGImplem.hpp:624-635)."""

    def _emit(self, prog):
        if not self.fs:
            prog.emit(pg.PUSH_CONST, p=[-math.inf])
            return
        pg.emit(self.fs[0], prog)
        for f in self.fs[1:]:
            pg.emit(f, prog)
            prog.emit(pg.MAX_ACC)


def ExpMin(a, b, s):
    """The reference's exponential smooth minimum (CSG.hpp:236-245): pass as smooth_min."""
    return -math.log(math.exp(-a / s) + math.exp(-b / s)) * s


def SmoothMin(a, b, s):
    """The reference's quadratic smooth minimum (CSG.hpp:256-263), the default."""
    h = max(s - abs(a - b), 0.0) / s
    return min(a, b) - 0.25 * h * h * s


class SmoothUnionIF(_Nary):
    """Smooth minimum of the two smallest values (CSGImplem.hpp:390-414)."""

    def __init__(self, fs, smooth_len: float, smooth_min=SmoothMin):
        super().__init__(fs)
        self.s = max(float(smooth_len), pg.DBL_MIN)
        if smooth_min not in (SmoothMin, ExpMin):
            raise NotImplementedError("the GPU smooth union takes SmoothMin or ExpMin")
        self.kind = 1 if smooth_min is ExpMin else 0

    def _emit(self, prog):
        if len(self.fs) == 0:
            prog.emit(pg.PUSH_CONST, p=[math.inf])
        elif len(self.fs) == 1:
            pg.emit(self.fs[0], prog)
        else:
            prog.emit(pg.PUSH_CONST, p=[math.inf])
            prog.emit(pg.PUSH_CONST, p=[math.inf])
            for f in self.fs:
                pg.emit(f, prog)
                prog.emit(pg.TWO_MIN_ACC)
            prog.emit(pg.SMOOTH_MIN_END, n=self.kind, p=[self.s])


class SmoothIntersectionIF(_Nary):
    """Smooth maximum of the two largest values (CSGImplem.hpp:678-702)."""

    def __init__(self, fs, smooth_len: float):
        super().__init__(fs)
        self.s = max(float(smooth_len), pg.DBL_MIN)

    def _emit(self, prog):
        if len(self.fs) == 0:
            prog.emit(pg.PUSH_CONST, p=[math.inf])
        elif len(self.fs) == 1:
            pg.emit(self.fs[0], prog)
        else:
            prog.emit(pg.PUSH_CONST, p=[-math.inf])
            prog.emit(pg.PUSH_CONST, p=[-math.inf])
            for f in self.fs:
                pg.emit(f, prog)
                prog.emit(pg.TWO_MAX_ACC)
            prog.emit(pg.SMOOTH_MAX_END, p=[self.s])


class DifferenceIF(_Programmed):
    """max(a, -b), b a union when given a list (CSGImplem.hpp:712-741)."""

    def __init__(self, a, b):
        self.a, self.b_args = a, b
        self.b = UnionIF(b) if isinstance(b, (list, tuple)) else b

    def _emit(self, prog):
        pg.emit(self.a, prog)
        pg.emit(self.b, prog)
        prog.emit(pg.DIFFERENCE)


class SmoothDifferenceIF(_Programmed):
    """Smooth intersection of a with the complements of b (CSGImplem.hpp:743-792)."""

    def __init__(self, a, b, smooth_len: float):
        bs = list(b) if isinstance(b, (list, tuple)) else [b]
        self.a, self.b_args, self.s = a, b, float(smooth_len)
        self.inner = SmoothIntersectionIF([a] + [ComplementIF(x) for x in bs], smooth_len)

    def _emit(self, prog):
        self.inner._emit(prog)


class BVHSmoothUnionIF(_Programmed):
    """Smooth minimum of the two smallest primitive values, found with a BVH over the primitives'
    bounding volumes (CSGImplem.hpp:510-608)."""

    def __init__(self, prims, bvs, smooth_len: float, smooth_min=SmoothMin):
        self.prims = list(prims)
        self.bvs = [(b.lo, b.hi) if isinstance(b, AABB) else b for b in bvs]
        if len(self.bvs) != len(self.prims):
            raise ValueError("one bounding volume per primitive")
        self.s = max(float(smooth_len), pg.DBL_MIN)
        if smooth_min not in (SmoothMin, ExpMin):
            raise NotImplementedError("the GPU smooth union takes SmoothMin or ExpMin")
        self.kind = 1 if smooth_min is ExpMin else 0

    def _emit(self, prog):
        prog.bvh(self.prims, self.bvs, smooth=self.s, kind=self.kind)


# the reference's factory functions (CSG.hpp:40-225, Transform.hpp:40-130)
def Union(*fs):
    return UnionIF(fs[0] if len(fs) == 1 and isinstance(fs[0], (list, tuple)) else fs)


def SmoothUnion(fs, smooth_len, smooth_min=SmoothMin):
    return SmoothUnionIF(fs, smooth_len, smooth_min)


def Intersection(*fs):
    return IntersectionIF(fs[0] if len(fs) == 1 and isinstance(fs[0], (list, tuple)) else fs)


def SmoothIntersection(fs, smooth_len):
    return SmoothIntersectionIF(fs, smooth_len)


def Difference(a, b):
    return DifferenceIF(a, b)


def SmoothDifference(a, b, smooth_len):
    return SmoothDifferenceIF(a, b, smooth_len)


def FiniteRepetition(f, period, repeat_lo, repeat_hi):
    return FiniteRepetitionIF(f, period, repeat_lo, repeat_hi)


def BVHSmoothUnion(prims, bvs, smooth_len, smooth_min=SmoothMin):
    return BVHSmoothUnionIF(prims, bvs, smooth_len, smooth_min)


def Complement(f):
    return ComplementIF(f)


def Offset(f, offset):
    return OffsetIF(f, offset)


def Annular(f, delta):
    return AnnularIF(f, delta)


def Blur(f, blur_distance, alpha: float = 0.5):
    return BlurIF(f, blur_distance, alpha)


def Mollify(f, dist, mollifier_samples: int = 2):
    """Mollify(f, dist, samples): a sphere mollifier of radius |dist| (TransformImplem.hpp:90-99)."""
    return MollifyIF(f, SphereSDF((0.0, 0.0, 0.0), abs(float(dist))), abs(float(dist)), mollifier_samples)


def Elongate(f, elongation):
    return ElongateIF(f, elongation)


def Reflect(f, plane):
    return ReflectIF(f, plane)
