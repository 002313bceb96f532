"""Many meshes, one traversal launch (the batched path of the C ABI).

The reference evaluates one TriMeshSDF per call (Source/EBGeometry_MeshDistanceFunctions.hpp:
376-387); a caller with many meshes loops over them. ``MeshBatch`` builds all of them on the GPU
(optionally with the branching factor and leaf size chosen by timing), and ``query`` / ``plan``
answer any number of (mesh, points) items with one traversal launch; a plan can replay the
whole pipeline as one captured hipGraph.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BatchInfo, BatchItem, BatchOpts, check, lib
from .sdf import TriMeshSDF, _torch


class MeshBatch:
    """meshes: sequence of (vertices (nv,3), faces (nf,3)) pairs."""

    def __init__(self, meshes, autotune: bool = False, branching: int = 4, leaf_size: int = 8, threads: int = 0):
        self._v = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3) for v, _ in meshes]
        self._f = [np.ascontiguousarray(f, dtype=np.uint32).reshape(-1, 3) for _, f in meshes]
        n = len(self._v)
        xyz = (C.c_void_p * n)(*[v.ctypes.data for v in self._v])
        tri = (C.c_void_p * n)(*[f.ctypes.data for f in self._f])
        nv = (C.c_uint64 * n)(*[v.shape[0] for v in self._v])
        nt = (C.c_uint64 * n)(*[f.shape[0] for f in self._f])
        o = BatchOpts(int(autotune), int(branching), int(leaf_size), int(threads))
        h = C.c_void_p()
        check(lib().ebg_batch_create(xyz, nv, tri, nt, n, C.byref(o), C.byref(h)))
        self._h = h
        self.num_meshes = n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().ebg_batch_destroy(h)
            except Exception:  # pragma: no cover - interpreter teardown
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        i = BatchInfo()
        check(lib().ebg_batch_get_info(self._h, C.byref(i)))
        return {f: (list(getattr(i, f)) if f == "tune_ms" else getattr(i, f)) for f, _ in BatchInfo._fields_}

    def _items(self, items, outs, tris):
        torch = _torch()
        arr = (BatchItem * max(1, len(items)))()
        for k, (mesh, pts) in enumerate(items):
            if not (isinstance(pts, torch.Tensor) and pts.is_cuda and pts.dtype == torch.float32):
                raise TypeError("batch items take CUDA float32 tensors of shape (n, 3)")
            if not pts.is_contiguous():
                raise ValueError("point tensors must be contiguous")
            arr[k].mesh = int(mesh)
            arr[k].n = pts.shape[0]
            arr[k].d_pts = pts.data_ptr()
            arr[k].d_dist = outs[k].data_ptr()
            arr[k].d_tri = tris[k].data_ptr() if tris is not None else None
        return arr

    def _outputs(self, items, with_triangle):
        torch = _torch()
        outs = [torch.empty(p.shape[0], dtype=torch.float32, device=p.device) for _, p in items]
        tris = [torch.empty(p.shape[0], dtype=torch.int32, device=p.device) for _, p in items] if with_triangle else None
        return outs, tris

    def query(self, items, with_triangle: bool = False, **kw):
        """items: list of (mesh index, CUDA float32 (n,3) tensor). Returns a list of distance
        tensors (and of face-index tensors), enqueued on torch's current stream."""
        from .sdf import TriMeshSDF as _T
        torch = _torch()
        outs, tris = self._outputs(items, with_triangle)
        arr = self._items(items, outs, tris)
        if "stream" not in kw:
            kw["stream"] = torch.cuda.current_stream().cuda_stream
        o = _T.query_opts(**kw)
        check(lib().ebg_batch_query(self._h, arr, len(items), C.byref(o)))
        return (outs, tris) if with_triangle else outs

    def plan(self, items, with_triangle: bool = False, workgroup: int = 0, capture: bool = True,
             persistent: bool = False, **kw) -> "BatchPlan":
        """workgroup 0 times the launch shapes and keeps the fastest; else 64/128/256 lanes, with
        persistent=True for the persistent grid whose wavefronts take packets from a queue."""
        return BatchPlan(self, items, with_triangle, workgroup | (PERSISTENT if persistent else 0), capture, **kw)

    def mesh(self, i: int) -> TriMeshSDF:
        """Mesh i of the batch as a TriMeshSDF view (the batch keeps it alive)."""
        h = C.c_void_p()
        check(lib().ebg_batch_mesh(self._h, int(i), C.byref(h)))
        m = TriMeshSDF.__new__(TriMeshSDF)
        m._h = h
        m._owned = False  # the batch destroys it
        m._batch = self
        m.num_vertices = self._v[i].shape[0]
        m.num_triangles = self._f[i].shape[0]
        return m


PERSISTENT = 0x10000  # EBG_PLAN_PERSISTENT


class BatchPlan:
    """A fixed set of items with its outputs, schedule and (optionally) captured hipGraph."""

    def __init__(self, batch: MeshBatch, items, with_triangle=False, workgroup=0, capture=True, **kw):
        from .sdf import TriMeshSDF as _T
        torch = _torch()
        self.batch = batch
        self.items = list(items)
        self.out, self.tri = batch._outputs(self.items, with_triangle)
        arr = batch._items(self.items, self.out, self.tri)
        if "stream" not in kw:
            kw["stream"] = torch.cuda.current_stream().cuda_stream
        o = _T.query_opts(**kw)
        h = C.c_void_p()
        check(lib().ebg_batch_plan_create(batch.handle, arr, len(self.items), C.byref(o), int(workgroup),
                                          int(capture), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().ebg_batch_plan_destroy(h)
            except Exception:  # pragma: no cover
                pass
            self._h = None

    @property
    def workgroup(self) -> int:
        """Lanes per workgroup of the traversal launch."""
        return int(lib().ebg_batch_plan_workgroup(self._h)) & 0xFFFF

    @property
    def persistent(self) -> bool:
        """True when the traversal runs as the persistent grid."""
        return bool(int(lib().ebg_batch_plan_workgroup(self._h)) & PERSISTENT)

    def launch(self, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().ebg_batch_plan_launch(self._h, s))
        return self.out

    def last_traversal_ms(self) -> float:
        return float(lib().ebg_batch_plan_last_traversal_ms(self._h))

    def last_early_packets(self) -> int:
        """Packets of the last launch that were started ahead of the table order (waits)."""
        return int(lib().ebg_batch_plan_last_early_packets(self._h))
