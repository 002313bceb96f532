"""ctypes binding of libebgpu.so (the C ABI declared in include/ebgpu.h).

The library is the product's only compute path: if it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libebgpu.so")

_vp = C.c_void_p
_u64 = C.c_uint64
_i32 = C.c_int32


class BuildOpts(C.Structure):
    _fields_ = [("builder", _i32), ("branching", _i32), ("leaf_size", _i32), ("reserved", _i32)]


class MeshInfo(C.Structure):
    _fields_ = [("bbox_lo", C.c_double * 3), ("bbox_hi", C.c_double * 3), ("diag", C.c_double),
                ("num_vertices", _u64), ("num_triangles", _u64), ("num_nodes", _u64), ("num_leaves", _u64),
                ("branching", _i32), ("leaf_size", _i32), ("stack_need", _i32), ("stack_capacity", _i32),
                ("device_bytes", _u64), ("build_seconds", C.c_double)]


class QueryOpts(C.Structure):
    _fields_ = [("presort", _i32), ("fp64_recheck", _i32), ("recheck_eps", C.c_double),
                ("learned_bound", _i32), ("fp64_all", _i32), ("bound_scale", C.c_float),
                ("bound_offset", C.c_float), ("stream", _vp), ("work", _vp), ("seed_radius", _vp),
                ("pair_work", _vp)]


class MlpOpts(C.Structure):
    _fields_ = [("num_samples", _u64), ("epochs", _i32), ("batch", _i32), ("lr", C.c_float), ("seed", _u64),
                ("near_sigma", C.c_float), ("box_scale", C.c_float), ("stream", _vp),
                ("under_weight", C.c_float), ("reserved", _i32)]


class Prim(C.Structure):
    _fields_ = [("type", _i32), ("mesh", _i32), ("p", C.c_double * 6), ("translate", C.c_double * 3),
                ("matrix", C.c_double * 9), ("scale", C.c_double), ("bv_lo", C.c_double * 3),
                ("bv_hi", C.c_double * 3)]


class CsgOp(C.Structure):
    _fields_ = [("op", _i32), ("n", _i32), ("a", _i32), ("b", _i32), ("p", C.c_double * 12)]


class CsgPrim(C.Structure):
    _fields_ = [("start", C.c_uint32), ("count", C.c_uint32), ("bv_lo", C.c_double * 3), ("bv_hi", C.c_double * 3)]


class BatchOpts(C.Structure):
    _fields_ = [("autotune", _i32), ("branching", _i32), ("leaf_size", _i32), ("threads", _i32)]


class BatchInfo(C.Structure):
    _fields_ = [("num_meshes", _u64), ("num_triangles", _u64), ("branching", _i32), ("leaf_size", _i32),
                ("stack_capacity", _i32), ("tuned", _i32), ("tune_ms", C.c_double * 4), ("build_seconds", C.c_double)]


class BatchItem(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("reserved", C.c_uint32), ("n", _u64), ("d_pts", _vp), ("d_dist", _vp),
                ("d_tri", _vp)]


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("ebg_last_error", C.c_char_p, []),
    ("ebg_mesh_create", C.c_int, [_vp, _u64, _vp, _u64, C.POINTER(BuildOpts), C.POINTER(_vp)]),
    ("ebg_mesh_create_polygons", C.c_int, [_vp, _u64, _vp, _vp, _u64, C.POINTER(BuildOpts), C.POINTER(_vp)]),
    ("ebg_mesh_triangle_faces", C.c_int, [_vp, _vp]),
    ("ebg_polygons_to_triangles_host", C.c_int, [_vp, _u64, _vp, _vp, _u64, _vp, _vp, _vp, _vp, _vp]),
    ("ebg_mesh_destroy", C.c_int, [_vp]),
    ("ebg_mesh_get_info", C.c_int, [_vp, C.POINTER(MeshInfo)]),
    ("ebg_mesh_leaf_order", C.c_int, [_vp, _vp]),
    ("ebg_query_opts_default", None, [C.POINTER(QueryOpts)]),
    ("ebg_query_sdf", C.c_int, [_vp, _vp, _u64, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_query_sdf_host", C.c_int, [_vp, _vp, _u64, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_query_sdf_noalloc", C.c_int, [_vp, _vp, _u64, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_query_sdf_f64", C.c_int, [_vp, _vp, _u64, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_query_grid", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_query_grid_host", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_mesh_last_traversal_ms", C.c_double, [_vp]),
    ("ebg_mesh_last_split_packets", C.c_int64, [_vp]),
    ("ebg_mesh_check", C.c_int, [_vp]),
    ("ebg_scene_check", C.c_int, [_vp]),
    ("ebg_batch_plan_last_early_packets", C.c_int64, [_vp]),
    ("ebg_morton_codes_f32", C.c_int, [_vp, _u64, _vp, _vp]),
    ("ebg_morton_codes_f64", C.c_int, [_vp, _u64, _vp, _vp]),
    ("ebg_morton_sort", C.c_int, [_vp, _u64, _vp, _vp, _vp]),
    ("ebg_pseudo_normals_host", C.c_int, [_vp, _u64, _vp, _u64, _vp, _vp, _vp]),
    ("ebg_scene_create", C.c_int, [C.POINTER(Prim), _u64, _vp, _u64, C.POINTER(BuildOpts), C.POINTER(_vp)]),
    ("ebg_scene_destroy", C.c_int, [_vp]),
    ("ebg_scene_bounds", C.c_int, [_vp, _vp, _vp, _vp]),
    ("ebg_scene_query", C.c_int, [_vp, _vp, _u64, _vp, C.POINTER(QueryOpts)]),
    ("ebg_scene_query_host", C.c_int, [_vp, _vp, _u64, _vp, C.POINTER(QueryOpts)]),
    ("ebg_scene_query_grid", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_scene_query_grid_host", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_scene_last_traversal_ms", C.c_double, [_vp]),
    ("ebg_scene_mlp_train", C.c_int, [_vp, C.POINTER(MlpOpts), C.POINTER(C.c_float)]),
    ("ebg_scene_mlp_predict", C.c_int, [_vp, _vp, _u64, _vp, _vp]),
    ("ebg_csg_create", C.c_int, [C.POINTER(CsgOp), _u64, _u64, C.POINTER(CsgPrim), _u64, _vp, _u64, C.c_double,
                                 C.POINTER(_vp)]),
    ("ebg_csg_destroy", C.c_int, [_vp]),
    ("ebg_csg_query", C.c_int, [_vp, _vp, _u64, _vp, C.POINTER(QueryOpts)]),
    ("ebg_csg_query_host", C.c_int, [_vp, _vp, _u64, _vp, C.POINTER(QueryOpts)]),
    ("ebg_csg_query_grid", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_csg_query_grid_host", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_csg_check", C.c_int, [_vp]),
    ("ebg_cloud_create", C.c_int, [_vp, _u64, C.POINTER(BuildOpts), C.POINTER(_vp)]),
    ("ebg_cloud_destroy", C.c_int, [_vp]),
    ("ebg_cloud_knn", C.c_int, [_vp, _vp, _u64, C.c_uint32, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_cloud_knn_host", C.c_int, [_vp, _vp, _u64, C.c_uint32, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_cloud_knn_self", C.c_int, [_vp, C.c_uint32, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_cloud_knn_self_host", C.c_int, [_vp, C.c_uint32, _vp, _vp, C.POINTER(QueryOpts)]),
    ("ebg_cloud_check", C.c_int, [_vp]),
    ("ebg_batch_create", C.c_int, [_vp, _vp, _vp, _vp, _u64, C.POINTER(BatchOpts), C.POINTER(_vp)]),
    ("ebg_batch_destroy", C.c_int, [_vp]),
    ("ebg_batch_get_info", C.c_int, [_vp, C.POINTER(BatchInfo)]),
    ("ebg_batch_mesh", C.c_int, [_vp, _u64, C.POINTER(_vp)]),
    ("ebg_batch_plan_create", C.c_int, [_vp, C.POINTER(BatchItem), _u64, C.POINTER(QueryOpts), _i32, _i32,
                                        C.POINTER(_vp)]),
    ("ebg_batch_plan_launch", C.c_int, [_vp, _vp]),
    ("ebg_batch_plan_destroy", C.c_int, [_vp]),
    ("ebg_batch_plan_workgroup", _i32, [_vp]),
    ("ebg_batch_plan_last_traversal_ms", C.c_double, [_vp]),
    ("ebg_batch_query", C.c_int, [_vp, C.POINTER(BatchItem), _u64, C.POINTER(QueryOpts)]),
    ("ebg_mlp_opts_default", None, [C.POINTER(MlpOpts)]),
    ("ebg_mlp_train", C.c_int, [_vp, C.POINTER(MlpOpts), C.POINTER(C.c_float)]),
    ("ebg_mlp_predict", C.c_int, [_vp, _vp, _u64, _vp, _vp]),
    ("ebg_mlp_param_count", _u64, []),
    ("ebg_mlp_check_schedule", C.c_int, [_u64, _i32]),
    ("ebg_mlp_save", C.c_int, [_vp, C.c_char_p]),
    ("ebg_mlp_load", C.c_int, [_vp, C.c_char_p]),
    ("ebg_mlp_get_params", C.c_int, [_vp, _vp]),
    ("ebg_mlp_set_params", C.c_int, [_vp, _vp]),
    ("ebg_mlp_loss_grad", C.c_int, [_vp, _vp, _vp, _u64, _vp, _vp, _vp]),
]

STATUS = {0: "OK", 1: "INVALID_ARG", 2: "OOM", 3: "HIP_ERROR", 4: "STACK_OVERFLOW", 5: "NOT_READY",
          6: "UNSUPPORTED"}


class EbgError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_LIB = None


def lib():
    """Load libebgpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ebgeometry_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _LIB = L
    return _LIB


def check(status: int):
    if status != 0:
        raise EbgError(status, (lib().ebg_last_error() or b"").decode())
