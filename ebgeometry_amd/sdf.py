"""Python mirror of the reference's signed-distance interface over the C ABI.

Reference interface (Source/EBGeometry_ImplicitFunction.hpp:51-52,
Source/EBGeometry_SignedDistanceFunction.hpp:54-62, Source/EBGeometry_MeshDistanceFunctions.hpp:309-489):

    ImplicitFunction<T>::value(const Vec3T<T>& p) -> T           (negative inside)
    SignedDistanceFunction<T>::signedDistance(p) -> T
    TriMeshSDF(mesh, BVH::Build, maxLeafGroups)
    TriMeshSDF::getClosestTriangle(p) -> {signedDistance, metaData}

Here the same names take arrays of points. ``value`` / ``signed_distance`` accept a single
point (returns a float) or an (n, 3) array / tensor (returns an array / tensor). Host
(numpy) inputs go through the synchronous host entry point; CUDA tensors stay on the device.
"""
from __future__ import annotations

import ctypes as C
import os
from enum import IntEnum

import numpy as np

from . import _lib
from ._lib import BuildOpts, MeshInfo, QueryOpts, check, lib


class Build(IntEnum):
    """BVH build strategy (the reference's BVH::Build, Source/EBGeometry_BVH.hpp:47-55)."""
    SAH = 0
    Morton = 1
    SAH_GPU = 2  # the SAH build on the device (same tree), normals and records too


def _torch():
    import torch  # plumbing only: device memory and streams
    return torch


def _grid_args(origin, spacing, dims):
    """ctypes arrays (double[3], double[3], uint64[3]) of a lattice."""
    o = (C.c_double * 3)(*[float(x) for x in origin])
    s = (C.c_double * 3)(*[float(x) for x in spacing])
    n = (C.c_uint64 * 3)(*[int(x) for x in dims])
    return o, s, n


def grid_points(origin, spacing, dims) -> np.ndarray:
    """The lattice points of query_grid as an (n, 3) float32 array, x fastest: each coordinate
    origin + index * spacing in FP64 (product and sum each rounded), then rounded to FP32."""
    nx, ny, nz = (int(x) for x in dims)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float64)
    return (np.asarray(origin, np.float64) + idx * np.asarray(spacing, np.float64)).astype(np.float32)


def _is_tensor(x) -> bool:
    try:
        torch = _torch()
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


class ImplicitFunction:
    """Base class: value(p) < 0 inside."""

    def value(self, points):
        raise NotImplementedError

    def __call__(self, points):
        return self.value(points)


class SignedDistanceFunction(ImplicitFunction):
    def value(self, points):
        return self.signed_distance(points)

    def signed_distance(self, points):
        raise NotImplementedError


class TriMeshSDF(SignedDistanceFunction):
    """GPU triangle-mesh SDF (drop-in for TriMeshSDF<T, Meta, K, W>).

    vertices: (nv, 3) float array; faces: (nf, 3) integer array, outward oriented, closed.
    ``branching`` and ``leaf_size`` replace the reference's compile-time K and
    maxLeafGroups*W (the GPU layout is not tied to a host SIMD width).
    """

    def __init__(self, vertices, faces, build: Build = Build.SAH, branching: int = 4, leaf_size: int = 4):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        opts = BuildOpts(int(build), int(branching), int(leaf_size), 0)
        h = C.c_void_p()
        check(lib().ebg_mesh_create(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0], C.byref(opts), C.byref(h)))
        self._h = h
        self.num_vertices = v.shape[0]
        self.num_triangles = f.shape[0]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owned", True):
            try:
                lib().ebg_mesh_destroy(h)
            except Exception:  # pragma: no cover - interpreter teardown
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        i = MeshInfo()
        check(lib().ebg_mesh_get_info(self._h, C.byref(i)))
        return {f: (list(getattr(i, f)) if f.startswith("bbox") else getattr(i, f)) for f, _ in MeshInfo._fields_}

    def leaf_order(self) -> np.ndarray:
        """Original face index of each triangle in BVH leaf order."""
        out = np.empty(self.num_triangles, dtype=np.uint32)
        check(lib().ebg_mesh_leaf_order(self._h, out.ctypes.data))
        return out

    @staticmethod
    def query_opts(presort=True, fp64_recheck=True, recheck_eps=1e-4, learned_bound=False, fp64_all=False,
                   bound_scale=1.25, bound_offset=1e-3, stream=None, work=None, seed_radius=None,
                   pair_work=None) -> QueryOpts:
        o = QueryOpts()
        lib().ebg_query_opts_default(C.byref(o))
        o.presort = int(presort)
        o.fp64_recheck = int(fp64_recheck)
        o.recheck_eps = float(recheck_eps)
        o.learned_bound = int(learned_bound)
        o.fp64_all = int(fp64_all)
        o.bound_scale = float(bound_scale)
        o.bound_offset = float(bound_offset)
        o.stream = stream
        o.work = work
        o.seed_radius = seed_radius
        o.pair_work = pair_work
        return o

    # -- queries --------------------------------------------------------------------------

    def query(self, points, with_triangle: bool = False, **kw):
        """Signed distance (and closest face index) of every point.

        numpy (n,3) -> numpy results (synchronous host path); CUDA tensor (n,3) float32 ->
        CUDA tensors, enqueued on torch's current stream.
        """
        if _is_tensor(points):
            return self._query_device(points, with_triangle, **kw)
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        n = p.shape[0]
        d = np.empty(n, dtype=np.float32)
        t = np.empty(n, dtype=np.uint32) if with_triangle else None
        o = self.query_opts(**kw)
        check(lib().ebg_query_sdf_host(self._h, p.ctypes.data, n, d.ctypes.data,
                                       t.ctypes.data if t is not None else None, C.byref(o)))
        return (d, t) if with_triangle else d

    def _query_device(self, pts, with_triangle, out=None, tri_out=None, **kw):
        torch = _torch()
        if pts.dtype != torch.float32 or not pts.is_cuda:
            raise TypeError("device queries take a CUDA float32 tensor of shape (n, 3)")
        p = pts.contiguous().view(-1, 3)
        n = p.shape[0]
        d = out if out is not None else torch.empty(n, dtype=torch.float32, device=p.device)
        t = None
        if with_triangle:
            t = tri_out if tri_out is not None else torch.empty(n, dtype=torch.int32, device=p.device)
        if "stream" not in kw:
            kw["stream"] = torch.cuda.current_stream(p.device).cuda_stream
        o = self.query_opts(**kw)
        check(lib().ebg_query_sdf(self._h, p.data_ptr(), n, d.data_ptr(), t.data_ptr() if t is not None else None,
                                  C.byref(o)))
        return (d, t) if with_triangle else d

    def query_grid(self, origin, spacing, dims, device=None, out=None, **kw):
        """Signed distance at every lattice point origin + (i, j, k) * spacing, as an array of
        shape (dims[2], dims[1], dims[0]) (x fastest: a FAB's layout). With device (a torch
        device) the result is a CUDA tensor made on torch's current stream; else numpy."""
        o3, s3, n3 = _grid_args(origin, spacing, dims)
        shape = (int(n3[2]), int(n3[1]), int(n3[0]))
        if device is not None:
            torch = _torch()
            d = out if out is not None else torch.empty(shape, dtype=torch.float32, device=device)
            if "stream" not in kw:
                kw["stream"] = torch.cuda.current_stream(d.device).cuda_stream
            o = self.query_opts(**kw)
            check(lib().ebg_query_grid(self._h, o3, s3, n3, d.data_ptr(), C.byref(o)))
            return d
        d = np.empty(shape, dtype=np.float32)
        o = self.query_opts(**kw)
        check(lib().ebg_query_grid_host(self._h, o3, s3, n3, d.ctypes.data, C.byref(o)))
        return d

    def signed_distance(self, points, **kw):
        single = not _is_tensor(points) and np.asarray(points).shape == (3,)
        r = self.query(np.asarray(points, dtype=np.float32).reshape(1, 3) if single else points, **kw)
        return float(r[0]) if single else r

    def closest_triangle(self, points, **kw):
        """(signed distance, face index) -- TriMeshSDF::getClosestTriangle with face metadata."""
        return self.query(points, with_triangle=True, **kw)

    def last_traversal_ms(self) -> float:
        return float(lib().ebg_mesh_last_traversal_ms(self._h))

    def last_split_packets(self) -> int:
        """Packets of the last query that were shared among subtree wavefronts (waits)."""
        return int(lib().ebg_mesh_last_split_packets(self._h))

    # -- learned bound --------------------------------------------------------------------

    def train_bound(self, num_samples: int = 4_000_000, epochs: int = 8, batch: int = 4096, lr: float = 2e-3,
                    seed: int = 12345, near_sigma: float = 0.02, box_scale: float = 1.2, under_weight: float = 16.0,
                    stream=None) -> float:
        """Train the mesh's learned cull bound on the GPU; returns the final mean loss."""
        o = _lib.MlpOpts()
        lib().ebg_mlp_opts_default(C.byref(o))
        o.num_samples, o.epochs, o.batch, o.lr, o.seed = num_samples, epochs, batch, lr, seed
        o.near_sigma, o.box_scale, o.stream, o.under_weight = near_sigma, box_scale, stream, under_weight
        loss = C.c_float(0.0)
        check(lib().ebg_mlp_train(self._h, C.byref(o), C.byref(loss)))
        return float(loss.value)

    def predict_bound(self, pts_tensor):
        torch = _torch()
        p = pts_tensor.contiguous().view(-1, 3)
        out = torch.empty(p.shape[0], dtype=torch.float32, device=p.device)
        check(lib().ebg_mlp_predict(self._h, p.data_ptr(), p.shape[0], out.data_ptr(),
                                    torch.cuda.current_stream(p.device).cuda_stream))
        return out

    def save_bound(self, path) -> None:
        """Write the trained network (parameters, input box, loss settings) to a file, tagged with
        a hash of the vertex and triangle arrays the mesh was built from (ebg_mlp_save)."""
        check(lib().ebg_mlp_save(self._h, os.fsencode(path)))

    def load_bound(self, path) -> None:
        """Read a network written by save_bound for a mesh of the same arrays and arm the learned
        bound (ebg_mlp_load; a file of another mesh is refused)."""
        check(lib().ebg_mlp_load(self._h, os.fsencode(path)))

    def get_bound_params(self) -> np.ndarray:
        n = int(lib().ebg_mlp_param_count())
        a = np.empty(n, dtype=np.float32)
        check(lib().ebg_mlp_get_params(self._h, a.ctypes.data))
        return a

    def set_bound_params(self, params) -> None:
        a = np.ascontiguousarray(params, dtype=np.float32)
        if a.size != int(lib().ebg_mlp_param_count()):
            raise ValueError("wrong parameter count")
        check(lib().ebg_mlp_set_params(self._h, a.ctypes.data))


def flatten_faces(faces):
    """Polygon faces (a sequence of vertex-index sequences, or an (nf, k) array) as the C ABI's
    (offsets uint32 (nf + 1), vertex indices uint32)."""
    if isinstance(faces, np.ndarray) and faces.ndim == 2:
        nf, k = faces.shape
        return (np.arange(nf + 1, dtype=np.uint32) * k, np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1))
    sizes = np.array([len(f) for f in faces], dtype=np.int64)
    off = np.zeros(len(sizes) + 1, dtype=np.uint32)
    off[1:] = np.cumsum(sizes)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(f, dtype=np.uint32).reshape(-1) for f in faces]),
                               dtype=np.uint32)
    return off, idx


class MeshSDF(TriMeshSDF):
    """GPU signed distance to a mesh of planar polygon faces (drop-in for the reference's
    MeshSDF<T, Meta, K>, Source/EBGeometry_MeshDistanceFunctions.hpp:149-281).

    vertices: (nv, 3); faces: a sequence of vertex-index lists (each a planar, simple polygon,
    outward oriented), or an (nf, k) array. The faces are clipped into triangles that carry the
    polygon mesh's own pseudo-normals, so the signed distance is the reference's
    FaceT::signedDistance minimum; every query of TriMeshSDF is available, and the closest-face
    outputs (``closest_face``, ``query(..., with_triangle=True)``) hold polygon face indices.
    """

    def __init__(self, vertices, faces, build: Build = Build.SAH, branching: int = 4, leaf_size: int = 4):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        off, idx = flatten_faces(faces)
        opts = BuildOpts(int(build), int(branching), int(leaf_size), 0)
        h = C.c_void_p()
        check(lib().ebg_mesh_create_polygons(v.ctypes.data, v.shape[0], off.ctypes.data, idx.ctypes.data,
                                             len(off) - 1, C.byref(opts), C.byref(h)))
        self._h = h
        self.num_vertices = v.shape[0]
        self.num_faces = len(off) - 1
        self.num_triangles = int(off[-1]) - 2 * self.num_faces

    def closest_face(self, points, **kw):
        """(signed distance, index of the closest face): getClosestFaces(p, true)'s first entry
        (MeshDistanceFunctionsImplem.hpp:297-354), whose other entries are the faces its walk
        happened to improve on and are not a property of the mesh."""
        return self.query(points, with_triangle=True, **kw)

    def triangle_faces(self) -> np.ndarray:
        """The face each clipped triangle belongs to."""
        out = np.empty(self.num_triangles, dtype=np.uint32)
        check(lib().ebg_mesh_triangle_faces(self._h, out.ctypes.data))
        return out


def polygons_to_triangles(vertices, faces):
    """The host clipping of polygon faces into triangles: (tri (nt, 3), tri_face (nt,), face
    normals (nt, 3), vertex normals (nt, 3, 3), edge normals (nt, 3, 3)), FP64 normals."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    off, idx = flatten_faces(faces)
    nf = len(off) - 1
    nt = int(off[-1]) - 2 * nf
    tri = np.zeros((nt, 3), dtype=np.uint32)
    tf = np.zeros(nt, dtype=np.uint32)
    fn = np.zeros((nt, 3))
    vn = np.zeros((nt, 3, 3))
    en = np.zeros((nt, 3, 3))
    check(lib().ebg_polygons_to_triangles_host(v.ctypes.data, v.shape[0], off.ctypes.data, idx.ctypes.data, nf,
                                               tri.ctypes.data, tf.ctypes.data, fn.ctypes.data, vn.ctypes.data,
                                               en.ctypes.data))
    return tri, tf, fn, vn, en


def pseudo_normals(vertices, faces):
    """Per-face (face normal, vertex normals, edge normals), FP64, computed on the host."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    fn = np.zeros((f.shape[0], 3))
    vn = np.zeros((f.shape[0], 3, 3))
    en = np.zeros((f.shape[0], 3, 3))
    check(lib().ebg_pseudo_normals_host(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0], fn.ctypes.data,
                                        vn.ctypes.data, en.ctypes.data))
    return fn, vn, en


def morton_codes(points):
    """63-bit Morton codes (SFC::computeBins + Morton::encode) of a CUDA float32/float64 (n,3) tensor."""
    torch = _torch()
    p = points.contiguous().view(-1, 3)
    n = p.shape[0]
    codes = torch.empty(n, dtype=torch.int64, device=p.device)
    s = torch.cuda.current_stream(p.device).cuda_stream
    if p.dtype == torch.float32:
        check(lib().ebg_morton_codes_f32(p.data_ptr(), n, codes.data_ptr(), s))
    elif p.dtype == torch.float64:
        check(lib().ebg_morton_codes_f64(p.data_ptr(), n, codes.data_ptr(), s))
    else:
        raise TypeError("float32 or float64 points")
    return codes


def morton_sort(codes):
    """Stable sort of codes (int64 CUDA tensor holding uint64 values): (perm int32, sorted codes)."""
    torch = _torch()
    n = codes.shape[0]
    perm = torch.empty(n, dtype=torch.int32, device=codes.device)
    out = torch.empty(n, dtype=torch.int64, device=codes.device)
    check(lib().ebg_morton_sort(codes.data_ptr(), n, perm.data_ptr(), out.data_ptr(),
                                torch.cuda.current_stream(codes.device).cuda_stream))
    return perm, out
