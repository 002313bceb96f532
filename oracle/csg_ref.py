"""CPU restatement of CSG-program evaluation in FP64 (numpy) -- test infrastructure only.

Evaluates a compiled program (ebgeometry_amd.program.Program: the instruction list the GPU
interpreter runs) at many points at once, each instruction restating the reference function it
stands for:

  leaves       Source/EBGeometry_AnalyticDistanceFunctions.hpp (signedDistance of PlaneSDF :107-116,
               SphereSDF :240-247, BoxSDF :377-399, TorusSDF :551-562, CylinderSDF :694-734,
               InfiniteCylinderSDF :844-856, CapsuleSDF :970-984, InfiniteConeSDF :1085-1099,
               ConeSDF :1195-1222, RoundedBoxSDF :1320-1327, RoundedCylinderSDF :1721-1733);
               meshes by the C restatement (oracle.port.Mesh)
  transforms   Source/EBGeometry_TransformImplem.hpp (Translate :154-162, Rotate :185-226, Scale
               :262-270, Complement :129-137, Offset :239-247, Annular :284-292, Elongate
               :427-435, Reflect :454-462); FiniteRepetitionIF, CSGImplem.hpp:815-826
  operators    Source/EBGeometry_CSGImplem.hpp (UnionIF :355-366, SmoothUnionIF :390-414,
               IntersectionIF :624-635, SmoothIntersectionIF :678-702, DifferenceIF :735-741);
               SmoothMin / SmoothMax / ExpMin, CSG.hpp:236-283
  BVH unions   BVHUnionIF / BVHSmoothUnionIF (CSGImplem.hpp:466-501, 550-584) over all primitives
               without a tree: the same minimum (two smallest) whenever every primitive's value is
               at least the distance to its bounding volume, which the pruning assumes.

Pinned in tests/test_csg_oracle.py against the compiled reference's value() of the same
expressions (oracle/ref.py expr_value) and against the golden fixtures.
"""
from __future__ import annotations

import numpy as np

from ebgeometry_amd import program as pg


def _leaf(op, n, p, x, y, z):
    if op == pg.SPHERE:
        return np.sqrt((x - p[0]) ** 2 + (y - p[1]) ** 2 + (z - p[2]) ** 2) - p[3]
    if op == pg.BOX:
        dx, dy, dz = np.maximum(p[0] - x, x - p[3]), np.maximum(p[1] - y, y - p[4]), np.maximum(p[2] - z, z - p[5])
        return np.minimum(0.0, np.maximum(dx, np.maximum(dy, dz))) + np.sqrt(
            np.maximum(0, dx) ** 2 + np.maximum(0, dy) ** 2 + np.maximum(0, dz) ** 2)
    if op == pg.PLANE:
        return (x - p[0]) * p[3] + (y - p[1]) * p[4] + (z - p[2]) * p[5]
    if op == pg.TORUS:
        ux, uy, uz = x - p[0], y - p[1], z - p[2]
        rho = np.sqrt(ux * ux + uy * uy) - p[3]
        return np.sqrt(rho * rho + uz * uz) - p[4]
    if op == pg.CYLINDER:
        if not (p[6] > 0 and p[7] > 0):
            return np.full_like(x, np.inf)
        ux, uy, uz = x - p[0], y - p[1], z - p[2]
        para = ux * p[3] + uy * p[4] + uz * p[5]
        w = np.sqrt((ux - para * p[3]) ** 2 + (uy - para * p[4]) ** 2 + (uz - para * p[5]) ** 2) - p[7]
        h = np.abs(para) - 0.5 * p[6]
        return np.where((w <= 0) & (h <= 0), np.where(np.abs(w) < np.abs(h), w, h),
                        np.where((w <= 0) & (h > 0), h, np.where((w > 0) & (h < 0), w, np.sqrt(w * w + h * h))))
    if op == pg.INF_CYLINDER:
        u = [x - p[0], y - p[1], z - p[2]]
        u[n] = np.zeros_like(x)
        return np.sqrt(u[0] ** 2 + u[1] ** 2 + u[2] ** 2) - p[3]
    if op == pg.CAPSULE:
        v1 = np.stack([x - p[0], y - p[1], z - p[2]], -1)
        v2 = np.array([p[3] - p[0], p[4] - p[1], p[5] - p[2]])
        h = np.clip((v1 @ v2) / (v2 @ v2), 0.0, 1.0)
        return np.linalg.norm(v1 - h[:, None] * v2, axis=1) - p[6]
    if op == pg.INF_CONE:
        dx, dy, dz = x - p[0], y - p[1], z - p[2]
        qx, qy = np.sqrt(dx * dx + dy * dy), -dz
        t = np.maximum(qx * p[3] + qy * p[4], 0.0)
        d1 = np.sqrt((qx - p[3] * t) ** 2 + (qy - p[4] * t) ** 2)
        return d1 * np.where(qx * p[4] - qy * p[3] < 0, -1.0, 1.0)
    if op == pg.CONE:
        dx, dy, dz = x - p[0], y - p[1], z - p[2]
        wx, wy = np.sqrt(dx * dx + dy * dy), dz
        qx, qy = p[5] * (p[3] / p[4]), -p[5]
        ta = np.clip((wx * qx + wy * qy) / (qx * qx + qy * qy), 0.0, 1.0)
        ax, ay = wx - ta * qx, wy - ta * qy
        bx, by = wx - qx * np.clip(wx / qx, 0.0, 1.0), wy - qy
        k = np.sign(qy)
        d = np.minimum(ax * ax + ay * ay, bx * bx + by * by)
        s = np.maximum(k * (wx * qy - wy * qx), k * (wy - qy))
        return np.sqrt(d) * np.sign(s)
    if op == pg.ROUNDED_BOX:
        return np.sqrt((x - np.clip(x, -p[0], p[0])) ** 2 + (y - np.clip(y, -p[1], p[1])) ** 2 +
                       (z - np.clip(z, -p[2], p[2])) ** 2) - p[3]
    if op == pg.ROUNDED_CYLINDER:
        xz = np.sqrt(z * z + x * x)
        d1x, d1y = xz - p[0], np.abs(y) - p[2]
        return np.minimum(np.maximum(d1x, d1y), 0.0) + np.sqrt(np.maximum(d1x, 0) ** 2 + np.maximum(d1y, 0) ** 2) - p[1]
    raise ValueError(op)


def _smooth_min(a, b, s, kind):
    if kind == 1:
        with np.errstate(over="ignore", divide="ignore"):
            return -np.log(np.exp(-a / s) + np.exp(-b / s)) * s
    h = np.maximum(s - np.abs(a - b), 0.0) / s
    return np.minimum(a, b) - 0.25 * h * h * s


def _smooth_max(a, b, s):
    h = np.maximum(s - np.abs(a - b), 0.0) / s
    return np.maximum(a, b) + 0.25 * h * h * s


def _two_min(a, b, v):
    lt = v < a
    nb = np.where(lt, a, np.where(v < b, v, b))
    return np.where(lt, v, a), nb


def _two_max(a, b, v):
    gt = v > a
    nb = np.where(gt, a, np.where(v > b, v, b))
    return np.where(gt, v, a), nb


def run(ops, x, y, z, prog, mesh_oracles, stack=None):
    """Run instructions `ops` at points (x, y, z) on the value stack `stack` (list of arrays)."""
    stack = [] if stack is None else stack
    pts = []
    with np.errstate(invalid="ignore"):
        for op, n, a, b, p in ops:
            p = list(p) + [0.0] * (12 - len(p))
            if op <= pg.ROUNDED_CYLINDER:
                stack.append(_leaf(op, n, p, x, y, z))
            elif op == pg.MESH:
                d, _ = mesh_oracles[n].sdf(np.stack([x, y, z], 1))
                stack.append(d)
            elif pg.TRANSLATE <= op <= pg.REPEAT:
                pts.append((x, y, z))
                if op == pg.TRANSLATE:
                    x, y, z = x - p[0], y - p[1], z - p[2]
                elif op == pg.ROTATE:
                    c, s = p[0], p[1]
                    if n == 0:
                        y, z = y * c + z * s, -y * s + z * c
                    elif n == 1:
                        x, z = x * c - z * s, x * s + z * c
                    else:
                        x, y = x * c + y * s, -x * s + y * c
                elif op == pg.SCALE_POINT:
                    x, y, z = x / p[0], y / p[0], z / p[0]
                elif op == pg.ELONGATE:
                    x, y, z = x - np.clip(x, -p[0], p[0]), y - np.clip(y, -p[1], p[1]), z - np.clip(z, -p[2], p[2])
                elif op == pg.REFLECT:
                    x, y, z = x * p[0], y * p[1], z * p[2]
                else:  # std::round: half away from zero
                    rnd = lambda t: np.sign(t) * np.floor(np.abs(t) + 0.5)  # noqa: E731
                    x = x - p[0] * rnd(np.clip(x / p[0], -p[3], p[6]))
                    y = y - p[1] * rnd(np.clip(y / p[1], -p[4], p[7]))
                    z = z - p[2] * rnd(np.clip(z / p[2], -p[5], p[8]))
            elif op == pg.POP_POINT:
                x, y, z = pts.pop()
            elif op == pg.SCALE_VALUE:
                stack[-1] = stack[-1] * p[0]
            elif op == pg.NEGATE:
                stack[-1] = -stack[-1]
            elif op == pg.OFFSET:
                stack[-1] = stack[-1] - p[0]
            elif op == pg.ANNULAR:
                stack[-1] = np.abs(stack[-1]) - p[0]
            elif op == pg.PUSH_CONST:
                stack.append(np.full_like(x, p[0]))
            elif op in (pg.MIN_ACC, pg.MAX_ACC, pg.ADD_ACC):
                v = stack.pop()
                stack[-1] = {pg.MIN_ACC: np.minimum, pg.MAX_ACC: np.maximum, pg.ADD_ACC: np.add}[op](stack[-1], v)
            elif op in (pg.TWO_MIN_ACC, pg.TWO_MAX_ACC):
                v = stack.pop()
                stack[-2], stack[-1] = (_two_min if op == pg.TWO_MIN_ACC else _two_max)(stack[-2], stack[-1], v)
            elif op == pg.SMOOTH_MIN_END:
                b_ = stack.pop()
                stack[-1] = _smooth_min(stack[-1], b_, p[0], n)
            elif op == pg.SMOOTH_MAX_END:
                b_ = stack.pop()
                stack[-1] = _smooth_max(stack[-1], b_, p[0])
            elif op == pg.DIFFERENCE:
                b_ = stack.pop()
                stack[-1] = np.maximum(stack[-1], -b_)
            elif op in (pg.BVH_UNION, pg.BVH_SMOOTH_UNION):
                va = np.full_like(x, np.inf)
                vb = np.full_like(x, np.inf)
                for j in range(a, a + b):
                    k = prog.prims[j][0]
                    v = run(prog.subs[k], x, y, z, prog, mesh_oracles)[-1]
                    if op == pg.BVH_UNION:
                        va = np.minimum(va, v)
                    else:
                        va, vb = _two_min(va, vb, v)
                stack.append(va if op == pg.BVH_UNION else _smooth_min(va, vb, p[0], n))
            else:
                raise ValueError(f"opcode {op}")
    return stack


def evaluate(f, pts, mesh_arrays=None):
    """FP64 value of the package expression f at pts (n, 3). mesh_arrays: id(TriMeshSDF) ->
    (vertices, faces) for the meshes inside f."""
    from oracle import port
    prog = pg.Program()
    pg.emit(f, prog)
    mo = [port.Mesh(np.asarray(mesh_arrays[id(m)][0], np.float64), mesh_arrays[id(m)][1]) for m in prog.meshes]
    p = np.asarray(pts, np.float64).reshape(-1, 3)
    out = run(prog.main, p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), prog, mo)
    assert len(out) == 1
    return out[0]
