"""Host-side helpers for CSG union scenes (spheres, boxes, transformed mesh instances).

A mesh instance is the reference composition
``Translate(Rotate(Rotate(Rotate(Scale(mesh, s), ax, 0), ay, 1), az, 2), t)``
(Source/EBGeometry_Transform.hpp:44-70, TransformImplem.hpp:140-270), so for a query p

    value(p) = s * mesh(Rx(ax) Ry(ay) Rz(az) (p - t) / s)

where R_axis is the map RotateIF::value applies to the point (angles in degrees,
TransformImplem.hpp:185-225).
"""
from __future__ import annotations

import math

import numpy as np


def rotation_point_map(angle_deg: float, axis: int) -> np.ndarray:
    """The 3x3 matrix RotateIF::value applies to its argument (TransformImplem.hpp:185-225)."""
    th = angle_deg * math.pi / 180.0
    c, s = math.cos(th), math.sin(th)
    if axis == 0:
        return np.array([[1, 0, 0], [0, c, s], [0, -s, c]], dtype=np.float64)
    if axis == 1:
        return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]], dtype=np.float64)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]], dtype=np.float64)


def instance_matrix(rot_deg) -> np.ndarray:
    """M = Rx Ry Rz: local point = M (p - t) / s."""
    ax, ay, az = (float(a) for a in rot_deg)
    return rotation_point_map(ax, 0) @ rotation_point_map(ay, 1) @ rotation_point_map(az, 2)


def instance_bounds(scene, pad_rel: float = 1e-6):
    """World-space AABB of every mesh instance (the BVH bounding volume of that primitive).

    world = t + s * M^T x for local vertices x, padded outward by pad_rel of the box size
    so that rounding in the transform never makes the box too small.
    """
    lo, hi = [], []
    for v, t, r, s in zip(scene.mesh_verts, scene.mesh_translate, scene.mesh_rot_deg, scene.mesh_scale):
        M = instance_matrix(r)
        w = np.asarray(t, np.float64) + float(s) * (np.asarray(v, np.float64) @ M)
        a, b = w.min(0), w.max(0)
        pad = pad_rel * (np.linalg.norm(b - a) + 1.0)
        lo.append(a - pad)
        hi.append(b + pad)
    return np.array(lo).reshape(-1, 3), np.array(hi).reshape(-1, 3)
