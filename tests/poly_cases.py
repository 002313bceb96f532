"""Polygon meshes and query points for the MeshSDF (polygon face) tests.

Every mesh is closed, outward oriented (corners counter-clockwise seen from outside) and made of
planar simple polygons: convex polyhedra (cube, hexagonal prism, dodecahedron, truncated
octahedron: quads, pentagons and hexagons), extrusions of non-convex outlines (an L and a
five-pointed star: non-convex hexagon and decagon caps), and a cube whose faces are split into
a grid of quads (1,536 faces, for sizes where the BVH matters).
"""
from __future__ import annotations

import numpy as np

NAMES = ["cube", "hex_prism", "dodecahedron", "trunc_octahedron", "l_prism", "star_prism", "grid_cube"]


def _convex_faces(V):
    """Faces of the convex hull of V as CCW (outward) vertex cycles, coplanar hull facets merged."""
    from scipy.spatial import ConvexHull
    h = ConvexHull(V)
    planes = {}
    for eq, simplex in zip(h.equations, h.simplices):
        key = tuple(np.round(eq, 6))
        planes.setdefault(key, set()).update(int(i) for i in simplex)
    faces = []
    for key, ids in planes.items():
        n = np.array(key[:3])
        ids = sorted(ids)
        c = V[ids].mean(0)
        a = V[ids[0]] - c
        a /= np.linalg.norm(a)
        b = np.cross(n, a)
        ang = [np.arctan2((V[i] - c) @ b, (V[i] - c) @ a) for i in ids]
        faces.append([ids[j] for j in np.argsort(ang)])
    return faces


def _prism(outline, h=1.0):
    """Extrude a CCW 2D outline (k, 2) along z: bottom cap, top cap, k side quads."""
    k = outline.shape[0]
    V = np.concatenate([np.c_[outline, np.zeros(k)], np.c_[outline, np.full(k, h)]])
    bottom = list(range(k - 1, -1, -1))           # seen from below: reversed
    top = list(range(k, 2 * k))
    sides = [[i, (i + 1) % k, k + (i + 1) % k, k + i] for i in range(k)]
    return V, [bottom, top] + sides


def mesh(name):
    """(vertices float32 (nv, 3), faces list of lists)."""
    if name == "cube":
        V = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
        V, F = V, _convex_faces(V)
    elif name == "hex_prism":
        t = np.arange(6) * np.pi / 3
        V, F = _prism(np.c_[np.cos(t), np.sin(t)], 1.5)
    elif name == "dodecahedron":
        p = (1 + 5 ** 0.5) / 2
        pts = [[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
        for a in (-1 / p, 1 / p):
            for b in (-p, p):
                pts += [[0, a, b], [a, b, 0], [b, 0, a]]
        V = np.array(pts, float)
        F = _convex_faces(V)
    elif name == "trunc_octahedron":
        import itertools
        pts = set()
        for perm in itertools.permutations([0, 1, 2]):
            for s1 in (-1, 1):
                for s2 in (-1, 1):
                    v = [0, 0, 0]
                    v[perm[1]] = s1 * 1.0
                    v[perm[2]] = s2 * 2.0
                    pts.add(tuple(v))
        V = np.array(sorted(pts), float)
        F = _convex_faces(V)
    elif name == "l_prism":
        V, F = _prism(np.array([[0, 0], [2, 0], [2, 0.7], [0.7, 0.7], [0.7, 2], [0, 2]], float), 0.8)
    elif name == "star_prism":
        t = np.pi / 2 + np.arange(10) * np.pi / 5
        r = np.where(np.arange(10) % 2 == 0, 1.0, 0.42)
        V, F = _prism(np.c_[r * np.cos(t), r * np.sin(t)], 0.5)
    elif name == "grid_cube":
        n = 16
        ids = {}
        V = []
        F = []

        def vid(p):
            key = tuple(np.round(p, 9))
            if key not in ids:
                ids[key] = len(V)
                V.append(p)
            return ids[key]
        for axis in range(3):
            for side in (0.0, 1.0):
                u, w = [d for d in range(3) if d != axis]
                for i in range(n):
                    for j in range(n):
                        quad = []
                        for (a, b) in ((i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)):
                            p = np.zeros(3)
                            p[axis], p[u], p[w] = side, a / n, b / n
                            quad.append(vid(p))
                        # outward: (u, w, axis) right-handed for the +side, reversed for the -side
                        right = np.cross(np.eye(3)[u], np.eye(3)[w])[axis] > 0
                        if right != (side == 1.0):
                            quad = quad[::-1]
                        F.append(quad)
        V = np.array(V, float)
    else:
        raise KeyError(name)
    return np.asarray(V, np.float32), [list(map(int, f)) for f in F]


def queries(V, n=3000, seed=0):
    """Uniform points in the 1.3x box, points near faces, edges and vertices (FP32 values)."""
    rng = np.random.default_rng(seed)
    V = np.asarray(V, np.float64)
    lo, hi = V.min(0), V.max(0)
    c, e = (lo + hi) / 2, (hi - lo) / 2 * 1.3
    diag = np.linalg.norm(hi - lo)
    u = c + (rng.random((n // 2, 3)) * 2 - 1) * e
    near = V[rng.integers(0, V.shape[0], n // 4)] + rng.normal(size=(n // 4, 3)) * 0.03 * diag
    mid = (V[rng.integers(0, V.shape[0], n - n // 2 - n // 4)] + V[rng.integers(0, V.shape[0], n - n // 2 - n // 4)]) / 2
    mid += rng.normal(size=mid.shape) * 0.01 * diag
    return np.concatenate([u, near, mid]).astype(np.float32)
