"""Deterministic synthetic inputs for the signed-distance query path.

Everything here is seeded and platform independent: random numbers come from a
counter-based SplitMix64 stream (not from ``numpy.random``), so the same seed
gives bit-identical meshes and query clouds on every machine.  The reference
draws its test points with ``std::uniform_real_distribution``
(Source/EBGeometry_RandomImplem.hpp:29-44), whose output is implementation
defined; these generators replace it.

Meshes are indexed triangle lists (float32 vertices, uint32 faces) with
outward orientation, i.e. the face normal (x2-x0)x(x2-x1) used by the
reference (Source/EBGeometry_TriangleImplem.hpp:94-106) points out of the
solid, so signed distances are negative inside.

Workloads:
  * ``icosphere(level)``            unit icosphere, vertices on the sphere
  * ``displaced_icosphere(level, seed)``  radial sum-of-sines displacement
  * ``uniform_box(n, lo, hi, seed)``      uniform query points
  * ``near_surface(n, verts, faces, sigma, seed)``  area-uniform surface
    samples pushed along the face normal by N(0, sigma)
  * ``mixed_queries``               half uniform in the 1.2x bbox, half near
    the surface (the large single-mesh workload)
  * ``batched_meshes``              many displaced icospheres under affine maps
  * ``csg_scene``                   spheres + boxes + transformed meshes
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64 stream keyed by (seed, name)."""

    def __init__(self, seed: int, name: str):
        h = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        for ch in name.encode():
            with np.errstate(over="ignore"):
                h = _mix64(np.array([h ^ np.uint64(ch)], dtype=np.uint64))[0] + _GOLDEN
        self._key = np.uint64(h)
        self._ctr = 0

    def u64(self, n: int) -> np.ndarray:
        idx = np.arange(self._ctr + 1, self._ctr + 1 + n, dtype=np.uint64)
        self._ctr += n
        with np.errstate(over="ignore"):
            return _mix64(self._key + idx * _GOLDEN)

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1) with 53 random bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform_open(self, n: int) -> np.ndarray:
        """Doubles in (0, 1]."""
        return ((self.u64(n) >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform_open(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)

    def unit_vectors(self, n: int) -> np.ndarray:
        z = 2.0 * self.uniform(n) - 1.0
        phi = 2.0 * math.pi * self.uniform(n)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


# ----------------------------------------------------------------------------
# meshes
# ----------------------------------------------------------------------------

def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
                  (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
                  (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
                  (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                  (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v, f


def _orient_outward(v: np.ndarray, f: np.ndarray) -> np.ndarray:
    """Flip faces whose (x2-x0)x(x2-x1) normal points towards the origin."""
    x0, x1, x2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(x2 - x0, x2 - x1)
    c = (x0 + x1 + x2) / 3.0
    bad = np.einsum("ij,ij->i", n, c) < 0
    f = f.copy()
    f[bad] = f[bad][:, [0, 2, 1]]
    return f


def icosphere_unit_f64(level: int):
    """Icosphere subdivided ``level`` times; float64 vertices on the unit sphere."""
    v, f = _icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
        uk, inv = np.unique(key, return_inverse=True)
        i0, i1 = uk // nv, uk % nv
        mid = v[i0] + v[i1]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + nv
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
        v = np.concatenate([v, mid])
    f = _orient_outward(v, f)
    return v, f


def icosphere(level: int, radius: float = 1.0):
    """Unit icosphere (vertices normalised to the sphere), float32 vertices."""
    v, f = icosphere_unit_f64(level)
    return (v * radius).astype(np.float32), f.astype(np.uint32)


@dataclass
class Displacement:
    u: np.ndarray      # (8,3) unit vectors
    freq: np.ndarray   # (8,)  U[2,8]
    amp: np.ndarray    # (8,)  U[0,1]/8
    phase: np.ndarray  # (8,)  U[0,2pi]
    scale: float = 0.15

    @staticmethod
    def from_seed(seed: int, terms: int = 8) -> "Displacement":
        s = Stream(seed, "displacement")
        u = s.unit_vectors(terms)
        freq = 2.0 + 6.0 * s.uniform(terms)
        amp = s.uniform(terms) / 8.0
        phase = 2.0 * math.pi * s.uniform(terms)
        return Displacement(u, freq, amp, phase)

    def radius(self, n: np.ndarray) -> np.ndarray:
        arg = n @ self.u.T * self.freq[None, :] + self.phase[None, :]
        return 1.0 + self.scale * (np.sin(arg) * self.amp[None, :]).sum(axis=1)


def displaced_icosphere(level: int, seed: int):
    """Icosphere with each vertex moved radially by 0.15*sum_k a_k sin(f_k (n.u_k) + phi_k)."""
    v, f = icosphere_unit_f64(level)
    d = Displacement.from_seed(seed)
    v = v * d.radius(v)[:, None]
    return v.astype(np.float32), f.astype(np.uint32)


def bbox(verts: np.ndarray):
    v = np.asarray(verts, dtype=np.float64)
    return v.min(axis=0), v.max(axis=0)


def diag(verts: np.ndarray) -> float:
    lo, hi = bbox(verts)
    return float(np.linalg.norm(hi - lo))


# ----------------------------------------------------------------------------
# query clouds
# ----------------------------------------------------------------------------

def uniform_box(n: int, lo, hi, seed: int, name: str = "uniform") -> np.ndarray:
    s = Stream(seed, name)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    u = s.uniform(3 * n).reshape(n, 3)
    return (lo + (hi - lo) * u).astype(np.float32)


def near_surface(n: int, verts: np.ndarray, faces: np.ndarray, sigma: float, seed: int,
                 name: str = "surface") -> np.ndarray:
    """Area-uniform surface points offset along the face normal by N(0, sigma)."""
    s = Stream(seed, name)
    v = np.asarray(verts, dtype=np.float64)
    f = np.asarray(faces, dtype=np.int64)
    x0, x1, x2 = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    cr = np.cross(x2 - x0, x2 - x1)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    cdf = np.cumsum(area)
    t = np.searchsorted(cdf, s.uniform(n) * cdf[-1], side="right")
    t = np.minimum(t, f.shape[0] - 1)
    r1 = np.sqrt(s.uniform(n))
    r2 = s.uniform(n)
    p = ((1.0 - r1)[:, None] * x0[t] + (r1 * (1.0 - r2))[:, None] * x1[t]
         + (r1 * r2)[:, None] * x2[t])
    p += (sigma * s.normal(n))[:, None] * nrm[t]
    return p.astype(np.float32)


def mixed_queries(n: int, verts: np.ndarray, faces: np.ndarray, seed: int,
                  sigma: float = 0.02, box_scale: float = 1.2) -> np.ndarray:
    """Half uniform in the box_scale-scaled bbox, half near the surface (N(0, sigma))."""
    lo, hi = bbox(verts)
    c = 0.5 * (lo + hi)
    h = 0.5 * box_scale * (hi - lo)
    nu = n // 2
    a = uniform_box(nu, c - h, c + h, seed, "mixed-uniform")
    b = near_surface(n - nu, verts, faces, sigma, seed, "mixed-surface")
    return np.concatenate([a, b])


# ----------------------------------------------------------------------------
# multi-mesh workloads
# ----------------------------------------------------------------------------

def random_affine(seed: int, name: str = "affine", smin: float = 0.5, smax: float = 2.0,
                  shift: float = 4.0):
    """A det>0 affine map x -> A x + b with A = R1 diag(s) R2 (rotations from quaternions)."""
    s = Stream(seed, name)

    def rot():
        q = s.normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                         [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                         [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])

    sc = np.exp(math.log(smin) + (math.log(smax) - math.log(smin)) * s.uniform(3))
    A = rot() @ np.diag(sc) @ rot()
    b = shift * (2.0 * s.uniform(3) - 1.0)
    return A, b


@dataclass
class BatchedWorkload:
    verts: list = field(default_factory=list)    # float32 (nv,3) per mesh
    faces: list = field(default_factory=list)    # uint32 (nf,3) per mesh
    queries: list = field(default_factory=list)  # float32 (nq,3) per mesh
    levels: list = field(default_factory=list)


def batched_meshes(num_meshes: int, seed: int, level_range=(3, 7), query_range=(10_000, 1_000_000),
                   query_scale: float = 1.0) -> BatchedWorkload:
    """Displaced icospheres under random affine maps, with per-mesh query clouds.

    Levels are uniform over level_range (triangle counts log-uniform over 20*4^L),
    query counts log-uniform over query_range, each mesh with its own seed.
    """
    s = Stream(seed, "batched")
    lv = level_range[0] + (s.u64(num_meshes) % np.uint64(level_range[1] - level_range[0] + 1)).astype(int)
    lq0, lq1 = math.log(query_range[0]), math.log(query_range[1])
    nq = np.exp(lq0 + (lq1 - lq0) * s.uniform(num_meshes))
    nq = np.maximum(1, (nq * query_scale).astype(np.int64))
    out = BatchedWorkload()
    cache = {}
    for i in range(num_meshes):
        L = int(lv[i])
        mseed = seed * 1_000_003 + i
        if L not in cache:
            cache[L] = icosphere_unit_f64(L)
        v0, f = cache[L]
        d = Displacement.from_seed(mseed)
        v = v0 * d.radius(v0)[:, None]
        A, b = random_affine(mseed)
        v = (v @ A.T + b).astype(np.float32)
        f32 = f.astype(np.uint32)
        out.verts.append(v)
        out.faces.append(f32)
        out.queries.append(mixed_queries(int(nq[i]), v, f32, mseed))
        out.levels.append(L)
    return out


@dataclass
class CsgScene:
    sphere_center: np.ndarray   # (ns,3) float32
    sphere_radius: np.ndarray   # (ns,)  float32
    box_lo: np.ndarray          # (nb,3) float32
    box_hi: np.ndarray          # (nb,3) float32
    mesh_verts: list            # float32 (nv,3) per mesh (local frame)
    mesh_faces: list            # uint32 (nf,3) per mesh
    mesh_translate: np.ndarray  # (nm,3)
    mesh_rot_deg: np.ndarray    # (nm,3) rotation angles (deg) about x, then y, then z
    mesh_scale: np.ndarray      # (nm,)
    lo: np.ndarray
    hi: np.ndarray


def csg_scene(seed: int, n_spheres: int = 4096, n_boxes: int = 1024, n_meshes: int = 8,
              mesh_level: int = 5, extent: float = 10.0) -> CsgScene:
    """Union of analytic spheres, axis-aligned boxes and displaced meshes under similarity maps."""
    s = Stream(seed, "csg")
    sc = (extent * s.uniform(3 * n_spheres)).reshape(-1, 3)
    sr = 0.05 + 0.35 * s.uniform(n_spheres)
    bc = (extent * s.uniform(3 * n_boxes)).reshape(-1, 3)
    bh = (0.05 + 0.45 * s.uniform(3 * n_boxes)).reshape(-1, 3)
    mv, mf = [], []
    for i in range(n_meshes):
        v, f = displaced_icosphere(mesh_level, seed * 7919 + i)
        mv.append(v)
        mf.append(f)
    mt = (1.0 + (extent - 2.0) * s.uniform(3 * n_meshes)).reshape(-1, 3)
    ang = (360.0 * s.uniform(3 * n_meshes)).reshape(-1, 3)
    msc = 0.3 + 0.7 * s.uniform(n_meshes)
    return CsgScene(sc.astype(np.float32), sr.astype(np.float32),
                    (bc - bh).astype(np.float32), (bc + bh).astype(np.float32),
                    mv, mf, mt.astype(np.float32), ang.astype(np.float32), msc.astype(np.float32),
                    np.zeros(3), np.full(3, extent))
