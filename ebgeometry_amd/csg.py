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
"""
from __future__ import annotations

import ctypes as C

import numpy as np

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
    """value() through a (lazily built) one-primitive GPU scene."""

    _scene = None

    def value(self, points):
        if self._scene is None:
            self._scene = BVHUnionIF([self])
        return self._scene.value(points)


class SphereSDF(SignedDistanceFunction, _GpuScalarField):
    """|p - c| - r (AnalyticDistanceFunctions.hpp:240-247)."""

    def __init__(self, center=(0.0, 0.0, 0.0), radius: float = 1.0):
        self.center = np.asarray(center, np.float64).reshape(3)
        self.radius = float(radius)

    def signed_distance(self, points):
        return _GpuScalarField.value(self, points)

    def bounding_volume(self) -> AABB:
        return AABB(self.center - self.radius, self.center + self.radius)


class BoxSDF(SignedDistanceFunction, _GpuScalarField):
    """Axis-aligned box [lo, hi] (AnalyticDistanceFunctions.hpp:377-399)."""

    def __init__(self, lo=(-0.5, -0.5, -0.5), hi=(0.5, 0.5, 0.5)):
        self.lo = np.asarray(lo, np.float64).reshape(3)
        self.hi = np.asarray(hi, np.float64).reshape(3)

    def signed_distance(self, points):
        return _GpuScalarField.value(self, points)

    def bounding_volume(self) -> AABB:
        return AABB(self.lo, self.hi)


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


class RotateIF(_Transformed):
    """value(p) = f(R p), R the axis rotation by angle degrees (TransformImplem.hpp:162-226)."""

    def __init__(self, f, angle_deg: float, axis: int):
        super().__init__(f)
        self.R = rotation_point_map(angle_deg, axis)

    def _apply(self, base, M, t, s):
        return base, M @ self.R, self.R.T @ t, s


class ScaleIF(_Transformed):
    """value(p) = s * f(p / s) (TransformImplem.hpp:245-270)."""

    def __init__(self, f, scale: float):
        super().__init__(f)
        if not scale > 0:
            raise ValueError("scale must be > 0")
        self.scale = float(scale)

    def _apply(self, base, M, t, s):
        return base, M, self.scale * t, self.scale * s


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
    """

    def __init__(self, prims, bvs=None, leaf_size: int = 4):
        prims = list(prims)
        if bvs is not None and len(bvs) != len(prims):
            raise ValueError("one bounding volume per primitive")
        mesh_index: dict = {}
        self._meshes: list = []
        arr = (Prim * max(1, len(prims)))()
        for i, f in enumerate(prims):
            q = _to_prim(f, mesh_index, self._meshes)
            if bvs is not None:
                bv = bvs[i]
                lo, hi = (bv.lo, bv.hi) if isinstance(bv, AABB) else bv
                q.bv_lo[:] = [float(x) for x in np.asarray(lo, np.float64).reshape(3)]
                q.bv_hi[:] = [float(x) for x in np.asarray(hi, np.float64).reshape(3)]
            arr[i] = q
        mh = (C.c_void_p * max(1, len(self._meshes)))(*[m.handle for m in self._meshes])
        opts = BuildOpts(0, 4, int(leaf_size), 0)
        h = C.c_void_p()
        check(lib().ebg_scene_create(arr, len(prims), mh, len(self._meshes), C.byref(opts), C.byref(h)))
        self._h = h
        self.num_primitives = len(prims)

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
        return self._h

    def bounds(self):
        lo = (C.c_double * 3)()
        hi = (C.c_double * 3)()
        d = C.c_double()
        check(lib().ebg_scene_bounds(self._h, lo, hi, C.byref(d)))
        return np.array(lo[:]), np.array(hi[:]), d.value

    @staticmethod
    def query_opts(presort=True, fp64_recheck=True, recheck_eps=1e-4, fp64_all=False, stream=None,
                   work=None) -> QueryOpts:
        o = QueryOpts()
        lib().ebg_query_opts_default(C.byref(o))
        o.presort = int(presort)
        o.fp64_recheck = int(fp64_recheck)
        o.recheck_eps = float(recheck_eps)
        o.fp64_all = int(fp64_all)
        o.stream = stream
        o.work = work
        return o

    def value(self, points, out=None, **kw):
        """Union value at a point (float) or at every row of an (n, 3) array / CUDA tensor."""
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
        return float(lib().ebg_scene_last_traversal_ms(self._h))


def BVHUnion(prims, bvs=None, leaf_size: int = 4) -> BVHUnionIF:
    return BVHUnionIF(prims, bvs, leaf_size)
