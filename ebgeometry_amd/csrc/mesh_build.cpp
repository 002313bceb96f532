// Host-side mesh preparation: pseudo-normals (FP64) and the K-wide SAH BVH.
#include "mesh_build.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>

namespace ebgpu {

namespace {

struct D3
{
  double x, y, z;
};
inline D3 ld3f(const float* p) { return {double(p[0]), double(p[1]), double(p[2])}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(D3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline D3 cross(D3 a, D3 u) { return {a.y * u.z - a.z * u.y, a.z * u.x - a.x * u.z, a.x * u.y - a.y * u.x}; }
// Division of a vector by a scalar multiplies by the rounded reciprocal, as the reference's
// Vec3T::operator/(T) does (Source/EBGeometry_VecImplem.hpp:303-309).
inline D3 div_by(D3 a, double s)
{
  const double is = 1.0 / s;
  return {is * a.x, is * a.y, is * a.z};
}

} // namespace

bool compute_pseudo_normals(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt,
                            PseudoNormals& out, std::string& err)
{
  out.face.assign(3 * nt, 0.0);
  out.vert.assign(3 * nv, 0.0);
  out.edge.assign(9 * nt, 0.0);
  std::vector<D3> fn(nt);
  for (std::size_t f = 0; f < nt; f++) {
    const D3 x0 = ld3f(xyz + 3 * tri[3 * f]), x1 = ld3f(xyz + 3 * tri[3 * f + 1]), x2 = ld3f(xyz + 3 * tri[3 * f + 2]);
    const D3 n  = cross(x2 - x0, x2 - x1);
    const double l = length(n);
    if (!(l > std::numeric_limits<double>::epsilon())) {
      err = "degenerate triangle " + std::to_string(f);
      return false;
    }
    fn[f]            = div_by(n, l);
    out.face[3 * f]     = fn[f].x;
    out.face[3 * f + 1] = fn[f].y;
    out.face[3 * f + 2] = fn[f].z;
  }

  // Half-edge pairs: edge a->b of one face pairs with b->a of its neighbour.
  std::vector<std::pair<uint64_t, uint32_t>> keys(3 * nt);
  for (std::size_t f = 0; f < nt; f++)
    for (int i = 0; i < 3; i++) {
      const uint64_t a = tri[3 * f + i], b = tri[3 * f + (i + 1) % 3];
      keys[3 * f + i]  = {a * uint64_t(nv) + b, uint32_t(3 * f + i)};
    }
  std::sort(keys.begin(), keys.end());
  for (std::size_t f = 0; f < nt; f++)
    for (int i = 0; i < 3; i++) {
      const uint64_t a = tri[3 * f + i], b = tri[3 * f + (i + 1) % 3];
      const uint64_t q = b * uint64_t(nv) + a;
      auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(q, uint32_t(0)));
      D3 n = fn[f];
      if (it != keys.end() && it->first == q) n = n + fn[it->second / 3];
      const double l = length(n);
      const D3     e = (l > std::numeric_limits<double>::epsilon()) ? div_by(n, l) : D3{0, 0, 0};
      out.edge[9 * f + 3 * i]     = e.x;
      out.edge[9 * f + 3 * i + 1] = e.y;
      out.edge[9 * f + 3 * i + 2] = e.z;
    }

  // Angle-weighted vertex normals, accumulated in face order.
  std::vector<D3> vn(nv, D3{0, 0, 0});
  std::vector<uint8_t> touched(nv, 0);
  for (std::size_t f = 0; f < nt; f++)
    for (int i = 0; i < 3; i++) {
      const uint32_t o = tri[3 * f + i];
      uint32_t       nb[2];
      int            k = 0;
      for (int e = 0; e < 3; e++) {
        const uint32_t a = tri[3 * f + e], b = tri[3 * f + (e + 1) % 3];
        if (a == o && k < 2) nb[k++] = b;
        else if (b == o && k < 2) nb[k++] = a;
      }
      if (k != 2) {
        err = "face " + std::to_string(f) + " repeats a vertex";
        return false;
      }
      const D3 x0 = ld3f(xyz + 3 * o);
      D3       e1 = ld3f(xyz + 3 * nb[0]) - x0;
      D3       e2 = ld3f(xyz + 3 * nb[1]) - x0;
      e1          = div_by(e1, length(e1));
      e2          = div_by(e2, length(e2));
      const double c     = std::clamp(dot(e1, e2), -1.0, 1.0);
      const double alpha = std::acos(c);
      vn[o]              = vn[o] + alpha * fn[f];
      touched[o]         = 1;
    }
  for (std::size_t v = 0; v < nv; v++) {
    if (!touched[v]) continue; // unreferenced vertex: never a closest feature
    const D3 n = div_by(vn[v], length(vn[v]));
    out.vert[3 * v]     = n.x;
    out.vert[3 * v + 1] = n.y;
    out.vert[3 * v + 2] = n.z;
  }
  return true;
}

namespace {

// Signed area of a 2D polygon given by coordinates (u, v) of its vertices (twice the area).
double area2(const std::vector<double>& u, const std::vector<double>& v)
{
  double a = 0.0;
  for (std::size_t i = 0, n = u.size(); i < n; i++) {
    const std::size_t j = (i + 1) % n;
    a += u[i] * v[j] - u[j] * v[i];
  }
  return a;
}

// Twice the signed area of the 2D triangle (a, b, c).
inline double orient2(double au, double av, double bu, double bv, double cu, double cv)
{
  return (bu - au) * (cv - av) - (cu - au) * (bv - av);
}

// Ear clipping of one simple polygon given in 2D (vertex k at (u[k], v[k])): appends triangles as
// triples of positions 0..n-1 in the polygon's own vertex order (each triangle keeps the polygon's
// winding). An ear is a convex corner whose triangle holds no other vertex (points on its
// boundary count as inside, so collinear chains are not cut through). Returns false when no ear
// exists (a self-intersecting or degenerate polygon).
bool ear_clip(const std::vector<double>& u, const std::vector<double>& v, std::vector<uint32_t>& out)
{
  const std::size_t n    = u.size();
  // simple: no two edges that do not share a corner meet (touching counts)
  for (std::size_t i = 0; i < n; i++)
    for (std::size_t j = i + 1; j < n; j++) {
      const std::size_t i1 = (i + 1) % n, j1 = (j + 1) % n;
      if (j == i1 || i == j1) continue;
      const double d1 = orient2(u[i], v[i], u[i1], v[i1], u[j], v[j]);
      const double d2 = orient2(u[i], v[i], u[i1], v[i1], u[j1], v[j1]);
      const double d3 = orient2(u[j], v[j], u[j1], v[j1], u[i], v[i]);
      const double d4 = orient2(u[j], v[j], u[j1], v[j1], u[i1], v[i1]);
      if (((d1 > 0) != (d2 > 0) || d1 == 0 || d2 == 0) && ((d3 > 0) != (d4 > 0) || d3 == 0 || d4 == 0)) {
        // collinear pairs only meet when their extents overlap
        if (d1 == 0 && d2 == 0) {
          const bool ox = std::max(u[i], u[i1]) >= std::min(u[j], u[j1]) && std::max(u[j], u[j1]) >= std::min(u[i], u[i1]);
          const bool oy = std::max(v[i], v[i1]) >= std::min(v[j], v[j1]) && std::max(v[j], v[j1]) >= std::min(v[i], v[i1]);
          if (!(ox && oy)) continue;
        }
        return false;
      }
    }
  const double      sgn  = area2(u, v) > 0.0 ? 1.0 : -1.0;
  std::vector<uint32_t> ring(n);
  for (std::size_t i = 0; i < n; i++) ring[i] = uint32_t(i);
  while (ring.size() > 3) {
    const std::size_t m     = ring.size();
    bool              found = false;
    for (std::size_t i = 0; i < m && !found; i++) {
      const uint32_t a = ring[(i + m - 1) % m], b = ring[i], c = ring[(i + 1) % m];
      if (sgn * orient2(u[a], v[a], u[b], v[b], u[c], v[c]) <= 0.0) continue; // reflex or flat corner
      bool empty = true;
      for (std::size_t j = 0; j < m && empty; j++) {
        const uint32_t p = ring[j];
        if (p == a || p == b || p == c) continue;
        const double d0 = sgn * orient2(u[a], v[a], u[b], v[b], u[p], v[p]);
        const double d1 = sgn * orient2(u[b], v[b], u[c], v[c], u[p], v[p]);
        const double d2 = sgn * orient2(u[c], v[c], u[a], v[a], u[p], v[p]);
        if (d0 >= 0.0 && d1 >= 0.0 && d2 >= 0.0) empty = false;
      }
      if (!empty) continue;
      out.push_back(a);
      out.push_back(b);
      out.push_back(c);
      ring.erase(ring.begin() + std::ptrdiff_t(i));
      found = true;
    }
    if (!found) return false;
  }
  out.push_back(ring[0]);
  out.push_back(ring[1]);
  out.push_back(ring[2]);
  return true;
}

} // namespace

bool polygon_mesh(const float* xyz, std::size_t nv, const uint32_t* offsets, const uint32_t* idx, std::size_t nf,
                  std::vector<uint32_t>& tri, std::vector<uint32_t>& tri_face, PseudoNormals& pn, std::string& err)
{
  if (offsets[0] != 0) {
    err = "face offsets must start at 0";
    return false;
  }
  for (std::size_t f = 0; f < nf; f++) {
    if (offsets[f + 1] < offsets[f] + 3) {
      err = "face " + std::to_string(f) + " has fewer than 3 vertices";
      return false;
    }
    for (uint32_t k = offsets[f]; k < offsets[f + 1]; k++) {
      if (idx[k] >= nv) {
        err = "face " + std::to_string(f) + ": vertex index out of range";
        return false;
      }
      for (uint32_t j = offsets[f]; j < k; j++)
        if (idx[j] == idx[k]) {
          err = "face " + std::to_string(f) + " repeats a vertex";
          return false;
        }
    }
  }
  const std::size_t nh = offsets[nf]; // half-edges: one per face corner, k -> next corner
  auto corner = [&](std::size_t f, std::size_t i) { // vertex of corner i (mod size) of face f
    const std::size_t n = offsets[f + 1] - offsets[f];
    return idx[offsets[f] + i % n];
  };

  // Face normal from the first corner triple that spans a plane, normalised
  // (Source/EBGeometry_DCEL_FaceImplem.hpp:138-165).
  std::vector<D3> fn(nf);
  for (std::size_t f = 0; f < nf; f++) {
    const std::size_t n = offsets[f + 1] - offsets[f];
    D3                nrm{0, 0, 0};
    for (std::size_t i = 0; i < n; i++) {
      const D3 x0 = ld3f(xyz + 3 * corner(f, i)), x1 = ld3f(xyz + 3 * corner(f, i + 1)),
               x2 = ld3f(xyz + 3 * corner(f, i + 2));
      nrm = cross(x2 - x0, x2 - x1);
      if (length(nrm) > 0.0) break;
    }
    const double l = length(nrm);
    if (!(l > std::numeric_limits<double>::epsilon())) {
      err = "degenerate face " + std::to_string(f);
      return false;
    }
    fn[f] = div_by(nrm, l);
  }

  // The distance to a polygon is taken in its plane: refuse faces whose corners leave the plane
  // through their centroid by more than 1e-5 of the face's extent.
  for (std::size_t f = 0; f < nf; f++) {
    const std::size_t n = offsets[f + 1] - offsets[f];
    D3                c{0, 0, 0};
    for (std::size_t i = 0; i < n; i++) c = c + ld3f(xyz + 3 * corner(f, i));
    c             = (1.0 / double(n)) * c;
    double ext = 0.0, dev = 0.0;
    for (std::size_t i = 0; i < n; i++) {
      const D3 d = ld3f(xyz + 3 * corner(f, i)) - c;
      ext        = std::max(ext, length(d));
      dev        = std::max(dev, std::abs(dot(d, fn[f])));
    }
    if (dev > 1e-5 * ext) {
      err = "face " + std::to_string(f) + " is not planar";
      return false;
    }
  }

  // Half-edge pseudo-normals: normalise(n_face + n_pairface) (DCEL_EdgeImplem.hpp:130-165).
  std::vector<std::size_t> face_of(nh);
  for (std::size_t f = 0; f < nf; f++)
    for (uint32_t k = offsets[f]; k < offsets[f + 1]; k++) face_of[k] = f;
  std::vector<std::pair<uint64_t, uint32_t>> keys(nh);
  for (std::size_t f = 0; f < nf; f++) {
    const std::size_t n = offsets[f + 1] - offsets[f];
    for (std::size_t i = 0; i < n; i++)
      keys[offsets[f] + i] = {uint64_t(corner(f, i)) * uint64_t(nv) + corner(f, i + 1), uint32_t(offsets[f] + i)};
  }
  std::sort(keys.begin(), keys.end());
  std::vector<D3> hn(nh);
  for (std::size_t f = 0; f < nf; f++) {
    const std::size_t n = offsets[f + 1] - offsets[f];
    for (std::size_t i = 0; i < n; i++) {
      const uint64_t q  = uint64_t(corner(f, i + 1)) * uint64_t(nv) + corner(f, i);
      auto           it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(q, uint32_t(0)));
      D3             s  = fn[f];
      if (it != keys.end() && it->first == q) s = s + fn[face_of[it->second]];
      const double l = length(s);
      hn[offsets[f] + i] = (l > std::numeric_limits<double>::epsilon()) ? div_by(s, l) : D3{0, 0, 0};
    }
  }

  // Vertex pseudo-normals: sum over incident faces, in face order, of the corner angle times the
  // face normal, normalised (DCEL_VertexImplem.hpp:190-300).
  std::vector<D3>      vn(nv, D3{0, 0, 0});
  std::vector<uint8_t> touched(nv, 0);
  for (std::size_t f = 0; f < nf; f++) {
    const std::size_t n = offsets[f + 1] - offsets[f];
    for (std::size_t i = 0; i < n; i++) {
      const uint32_t o  = corner(f, i);
      const D3       x0 = ld3f(xyz + 3 * o);
      D3             e1 = ld3f(xyz + 3 * corner(f, i + 1)) - x0;
      D3             e2 = ld3f(xyz + 3 * corner(f, i + n - 1)) - x0;
      e1              = div_by(e1, length(e1));
      e2              = div_by(e2, length(e2));
      const double al = std::acos(std::clamp(dot(e1, e2), -1.0, 1.0));
      vn[o]           = vn[o] + al * fn[f];
      touched[o]      = 1;
    }
  }

  // Triangles: each face clipped into ears in the plane of its two non-dominant normal axes
  // (the projection Source/EBGeometry_Polygon2DImplem.hpp:72-103 uses for its inside test).
  tri.clear();
  tri_face.clear();
  std::vector<double>   u, w;
  std::vector<uint32_t> ears;
  std::vector<uint32_t> tri_corner; // per triangle corner: the face-corner index, or kInvalid
  for (std::size_t f = 0; f < nf; f++) {
    const std::size_t n = offsets[f + 1] - offsets[f];
    int               drop = 0;
    const double      a[3] = {std::abs(fn[f].x), std::abs(fn[f].y), std::abs(fn[f].z)};
    for (int d = 1; d < 3; d++)
      if (a[d] > a[drop]) drop = d;
    const int ax = drop == 0 ? 1 : 0, ay = drop == 2 ? 1 : 2;
    u.resize(n);
    w.resize(n);
    for (std::size_t i = 0; i < n; i++) {
      u[i] = double(xyz[3 * corner(f, i) + ax]);
      w[i] = double(xyz[3 * corner(f, i) + ay]);
    }
    ears.clear();
    if (!ear_clip(u, w, ears)) {
      err = "face " + std::to_string(f) + " is not a simple polygon";
      return false;
    }
    for (std::size_t t = 0; t < ears.size(); t += 3) {
      for (int k = 0; k < 3; k++) {
        const uint32_t i = ears[t + k], j = ears[t + (k + 1) % 3];
        tri.push_back(corner(f, i));
        // the triangle edge i -> j is the face's half-edge i -> i+1 when j follows i
        tri_corner.push_back((i + 1) % n == j ? uint32_t(offsets[f] + i) : kInvalid);
      }
      tri_face.push_back(uint32_t(f));
    }
  }

  const std::size_t nt = tri_face.size();
  pn.face.assign(3 * nt, 0.0);
  pn.vert.assign(3 * nv, 0.0);
  pn.edge.assign(9 * nt, 0.0);
  for (std::size_t t = 0; t < nt; t++) {
    const D3 n = fn[tri_face[t]];
    pn.face[3 * t] = n.x, pn.face[3 * t + 1] = n.y, pn.face[3 * t + 2] = n.z;
    for (int k = 0; k < 3; k++) {
      // an edge inside the face (a diagonal of the clipping) has the face's own normal
      const uint32_t h = tri_corner[3 * t + k];
      const D3       e = h == kInvalid ? n : hn[h];
      pn.edge[9 * t + 3 * k] = e.x, pn.edge[9 * t + 3 * k + 1] = e.y, pn.edge[9 * t + 3 * k + 2] = e.z;
    }
  }
  for (std::size_t v = 0; v < nv; v++) {
    if (!touched[v]) continue;
    const D3 n = div_by(vn[v], length(vn[v]));
    pn.vert[3 * v] = n.x, pn.vert[3 * v + 1] = n.y, pn.vert[3 * v + 2] = n.z;
  }
  return true;
}

void pack_triangles(const float* xyz, const uint32_t* tri, const std::vector<uint32_t>& order,
                    const PseudoNormals& pn, std::size_t nv, PackedMesh& out)
{
  const std::size_t nt = order.size();
  out.order = order;
  out.tris.resize(nt);
  out.tripos.assign(12 * nt, 0.0f);
  out.feat.assign(8 * nt, 0u);
  out.fn64.assign(4 * nt, 0.0);
  out.tsph.assign(8 * (nt + kSphPad), 0.0f);
  double hmax = 0.0;
  for (std::size_t i = 0; i < nt; i++) {
    const uint32_t f  = order[i];
    const D3       x0 = ld3f(xyz + 3 * tri[3 * f]), x1 = ld3f(xyz + 3 * tri[3 * f + 1]),
             x2 = ld3f(xyz + 3 * tri[3 * f + 2]);
    const D3 e[3] = {x1 - x0, x2 - x1, x0 - x2};
    const D3 n    = {pn.face[3 * f], pn.face[3 * f + 1], pn.face[3 * f + 2]};
    TriRec&  r    = out.tris[i];
    std::memset(&r, 0, sizeof(r));
    const float* p0 = xyz + 3 * tri[3 * f];
    for (int c = 0; c < 3; c++) r.v0[c] = p0[c];
    auto put = [](float* dst, D3 v) {
      dst[0] = float(v.x);
      dst[1] = float(v.y);
      dst[2] = float(v.z);
    };
    put(r.e0, e[0]);
    put(r.e1, e[1]);
    put(r.e2, e[2]);
    put(r.n, n);
    // edge normals in the face plane, pointing out of the triangle (the face normal of a polygon
    // may point against a clipped triangle's winding)
    const double sg = dot(cross(e[0], e[1]), n) >= 0.0 ? 1.0 : -1.0;
    put(r.c0, sg * cross(e[0], n));
    put(r.c1, sg * cross(e[1], n));
    put(r.c2, sg * cross(e[2], n));
    for (int c = 0; c < 3; c++) r.inv[c] = float(1.0 / dot(e[c], e[c]));
    r.face = f;
    // bounding sphere about the (float-rounded) centroid; the radius is measured in double from
    // that float centre and inflated so that FP32 culling tests stay conservative
    const D3 cen = {double(float((x0.x + x1.x + x2.x) / 3.0)), double(float((x0.y + x1.y + x2.y) / 3.0)),
                    double(float((x0.z + x1.z + x2.z) / 3.0))};
    const double rad = std::max(length(x0 - cen), std::max(length(x1 - cen), length(x2 - cen)));
    put(r.sc, cen);
    r.sr = float(rad * (1.0 + 1e-5) + 1e-7 * (length(cen) + rad));
    for (int c = 0; c < 3; c++) out.tsph[8 * i + c] = r.sc[c];
    out.tsph[8 * i + 3] = r.sr;
    for (int c = 0; c < 3; c++) out.tsph[8 * i + 4 + c] = r.n[c];
    out.tsph[8 * i + 7] = float(-(double(r.n[0]) * r.v0[0] + double(r.n[1]) * r.v0[1] + double(r.n[2]) * r.v0[2]));
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) out.tripos[12 * i + 3 * v + c] = xyz[3 * tri[3 * f + v] + c];
    for (int v = 0; v < 3; v++) {
      out.feat[8 * i + v]     = tri[3 * f + v];
      out.feat[8 * i + 3 + v] = uint32_t(3 * f + v);
    }
    out.feat[8 * i + 6] = f;
    for (int c = 0; c < 3; c++) out.fn64[4 * i + c] = pn.face[3 * f + c];
    hmax = std::max(hmax, std::max(length(e[0]), std::max(length(e[1]), length(e[2]))));
  }
  out.hmax = float(hmax * (1.0 + 1e-6));
  out.vn32.assign(4 * nv, 0.0f);
  out.vn64.assign(4 * nv, 0.0);
  for (std::size_t v = 0; v < nv; v++)
    for (int c = 0; c < 3; c++) {
      out.vn32[4 * v + c] = float(pn.vert[3 * v + c]);
      out.vn64[4 * v + c] = pn.vert[3 * v + c];
    }
  const std::size_t nh = pn.edge.size() / 3;
  out.en32.assign(4 * nh, 0.0f);
  out.en64.assign(4 * nh, 0.0);
  for (std::size_t h = 0; h < nh; h++)
    for (int c = 0; c < 3; c++) {
      out.en32[4 * h + c] = float(pn.edge[3 * h + c]);
      out.en64[4 * h + c] = pn.edge[3 * h + c];
    }
}

// ------------------------------------------------------------------------------------------
// Binned SAH binary build
// ------------------------------------------------------------------------------------------

namespace {

using BNode = BinNode;

inline float area(const float* lo, const float* hi)
{
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct SahBuilder
{
  const float*          blo;  // 3 per prim
  const float*          bhi;
  const float*          cen;
  std::vector<uint32_t> idx;
  std::vector<BNode>    nodes;
  std::atomic<int32_t>  next{0};
  int                   leaf_size = 4;

  int32_t alloc() { return next.fetch_add(1); }

  void build(int32_t ni, uint32_t b, uint32_t e, int depth)
  {
    BNode& nd = nodes[ni];
    float  lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float  clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = b; i < e; i++) {
      const uint32_t p = idx[i];
      for (int c = 0; c < 3; c++) {
        lo[c]  = std::min(lo[c], blo[3 * p + c]);
        hi[c]  = std::max(hi[c], bhi[3 * p + c]);
        clo[c] = std::min(clo[c], cen[3 * p + c]);
        chi[c] = std::max(chi[c], cen[3 * p + c]);
      }
    }
    for (int c = 0; c < 3; c++) {
      nd.lo[c] = lo[c];
      nd.hi[c] = hi[c];
    }
    const uint32_t n = e - b;
    if (n <= uint32_t(leaf_size)) {
      nd.left = nd.right = -1;
      nd.first           = b;
      nd.count           = n;
      return;
    }
    constexpr int BINS = 32;
    float         bestCost = INFINITY;
    int           bestAxis = -1, bestBin = 0;
    for (int axis = 0; axis < 3; axis++) {
      const float ext = chi[axis] - clo[axis];
      if (!(ext > 0.0f)) continue;
      const float scale = float(BINS) / ext;
      float       bl[BINS][3], bh[BINS][3];
      uint32_t    bc[BINS];
      for (int k = 0; k < BINS; k++) {
        bc[k] = 0;
        for (int c = 0; c < 3; c++) {
          bl[k][c] = INFINITY;
          bh[k][c] = -INFINITY;
        }
      }
      for (uint32_t i = b; i < e; i++) {
        const uint32_t p = idx[i];
        int            k = int((cen[3 * p + axis] - clo[axis]) * scale);
        k                = std::min(BINS - 1, std::max(0, k));
        bc[k]++;
        for (int c = 0; c < 3; c++) {
          bl[k][c] = std::min(bl[k][c], blo[3 * p + c]);
          bh[k][c] = std::max(bh[k][c], bhi[3 * p + c]);
        }
      }
      float    la[BINS];
      uint32_t lc[BINS];
      float    rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t rc = 0;
      for (int k = 0; k < BINS - 1; k++) {
        for (int c = 0; c < 3; c++) {
          rl[c] = std::min(rl[c], bl[k][c]);
          rh[c] = std::max(rh[c], bh[k][c]);
        }
        rc += bc[k];
        la[k] = rc ? area(rl, rh) : 0.0f;
        lc[k] = rc;
      }
      float    sl[3] = {INFINITY, INFINITY, INFINITY}, sh[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t sc = 0;
      for (int k = BINS - 1; k >= 1; k--) {
        for (int c = 0; c < 3; c++) {
          sl[c] = std::min(sl[c], bl[k][c]);
          sh[c] = std::max(sh[c], bh[k][c]);
        }
        sc += bc[k];
        if (lc[k - 1] && sc) {
          const float cost = la[k - 1] * float(lc[k - 1]) + area(sl, sh) * float(sc);
          if (cost < bestCost) {
            bestCost = cost;
            bestAxis = axis;
            bestBin  = k;
          }
        }
      }
    }
    uint32_t mid = b + n / 2;
    if (bestAxis >= 0) {
      const float lo_a  = clo[bestAxis];
      const float scale = float(BINS) / (chi[bestAxis] - lo_a);
      auto        it    = std::partition(idx.begin() + b, idx.begin() + e, [&](uint32_t p) {
        int k = int((cen[3 * p + bestAxis] - lo_a) * scale);
        k     = std::min(BINS - 1, std::max(0, k));
        return k < bestBin;
      });
      mid = uint32_t(it - idx.begin());
      if (mid == b || mid == e) mid = b + n / 2;
    }
    const int32_t l = alloc(), r = alloc();
    nd.left  = l;
    nd.right = r;
    nd.first = b;
    nd.count = n;
    if (n > 65536 && depth < 6) {
      std::thread th([this, l, b, mid, depth]() { build(l, b, mid, depth + 1); });
      build(r, mid, e, depth + 1);
      th.join();
    }
    else {
      build(l, b, mid, depth + 1);
      build(r, mid, e, depth + 1);
    }
  }
};

template <int K>
struct Collapser
{
  const std::vector<BNode>& bn;
  std::vector<Node<K>>      out;
  std::size_t               leaves = 0;

  explicit Collapser(const std::vector<BNode>& b) : bn(b) {}

  uint32_t emit(const std::vector<int32_t>& kids0)
  {
    const uint32_t me = uint32_t(out.size());
    out.emplace_back();
    std::vector<int32_t> kids = kids0;
    while (int(kids.size()) < K) {
      int   pick = -1;
      float best = -1.0f;
      for (int i = 0; i < int(kids.size()); i++) {
        const BNode& c = bn[kids[i]];
        if (c.left < 0) continue;
        const float a = area(c.lo, c.hi);
        if (a > best) {
          best = a;
          pick = i;
        }
      }
      if (pick < 0) break;
      const BNode& c = bn[kids[pick]];
      kids[pick]     = c.left;
      kids.insert(kids.begin() + pick + 1, c.right);
    }
    Node<K> nd;
    for (int k = 0; k < K; k++) {
      nd.lox[k] = nd.loy[k] = nd.loz[k] = 0.0f;
      nd.hix[k] = nd.hiy[k] = nd.hiz[k] = 0.0f;
      nd.ref[k] = 0;
      nd.cnt[k] = kInvalid;
    }
    for (int k = 0; k < int(kids.size()); k++) {
      const BNode& c = bn[kids[k]];
      nd.lox[k] = c.lo[0];
      nd.loy[k] = c.lo[1];
      nd.loz[k] = c.lo[2];
      nd.hix[k] = c.hi[0];
      nd.hiy[k] = c.hi[1];
      nd.hiz[k] = c.hi[2];
      if (c.left < 0) {
        nd.ref[k] = c.first;
        nd.cnt[k] = c.count;
        leaves++;
      }
      else {
        nd.cnt[k] = 0;
        nd.ref[k] = emit({c.left, c.right});
      }
    }
    out[me] = nd;
    return me;
  }
};

template <int K>
int stack_need_rec(const Node<K>* nodes, uint32_t i)
{
  int m = 0, deepest = 0;
  for (int k = 0; k < K; k++) {
    if (nodes[i].cnt[k] == kInvalid) continue;
    m++;
    if (nodes[i].cnt[k] == 0) deepest = std::max(deepest, stack_need_rec<K>(nodes, nodes[i].ref[k]));
  }
  return (m - 1) + deepest;
}

} // namespace

int compute_stack_need(const PackedMesh& m)
{
  if (m.num_nodes == 0) return 0;
  if (m.K == 4) return 1 + std::max(0, stack_need_rec<4>(reinterpret_cast<const Node<4>*>(m.nodes.data()), 0));
  return 1 + std::max(0, stack_need_rec<8>(reinterpret_cast<const Node<8>*>(m.nodes.data()), 0));
}

bool build_sah_mesh(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt, int K, int leaf_size,
                    const PseudoNormals& pn, PackedMesh& out, std::string& err)
{
  if (K != 4 && K != 8) {
    err = "branching must be 4 or 8";
    return false;
  }
  if (leaf_size < 1 || leaf_size > kMaxLeaf) {
    err = "leaf_size must be in 1..16";
    return false;
  }
  std::vector<float> blo(3 * nt), bhi(3 * nt);
  for (std::size_t f = 0; f < nt; f++)
    for (int c = 0; c < 3; c++) {
      const float a = xyz[3 * tri[3 * f] + c], b = xyz[3 * tri[3 * f + 1] + c], d = xyz[3 * tri[3 * f + 2] + c];
      blo[3 * f + c] = std::min(a, std::min(b, d));
      bhi[3 * f + c] = std::max(a, std::max(b, d));
    }
  WideBvh w;
  if (!build_wide_bvh(blo.data(), bhi.data(), nt, K, leaf_size, w, err)) return false;
  out.K          = K;
  out.leaf_size  = leaf_size;
  out.num_nodes  = w.num_nodes;
  out.num_leaves = w.num_leaves;
  out.nodes      = std::move(w.nodes);
  for (int c = 0; c < 3; c++) {
    out.lo[c] = w.lo[c];
    out.hi[c] = w.hi[c];
  }
  out.diag = w.diag;
  pack_triangles(xyz, tri, w.order, pn, nv, out);
  out.stack_need = w.stack_need;
  return true;
}

void collapse_binary(const std::vector<BinNode>& bn, int32_t root, int K, WideBvh& out)
{
  const BinNode& r = bn[root];
  for (int c = 0; c < 3; c++) {
    out.lo[c] = r.lo[c];
    out.hi[c] = r.hi[c];
  }
  const double dx = double(r.hi[0]) - r.lo[0], dy = double(r.hi[1]) - r.lo[1], dz = double(r.hi[2]) - r.lo[2];
  out.diag = std::sqrt(dx * dx + dy * dy + dz * dz);
  PackedMesh tmp; // only to reuse compute_stack_need
  tmp.K = K;
  auto collapse = [&](auto tag) {
    constexpr int KK = decltype(tag)::value;
    Collapser<KK> col(bn);
    if (r.left < 0) col.emit({root}); // single leaf: root with one leaf child
    else col.emit({r.left, r.right});
    out.num_nodes  = col.out.size();
    out.num_leaves = col.leaves;
    out.nodes.resize(col.out.size() * sizeof(Node<KK>));
    std::memcpy(out.nodes.data(), col.out.data(), out.nodes.size());
  };
  if (K == 4) collapse(std::integral_constant<int, 4>{});
  else collapse(std::integral_constant<int, 8>{});
  tmp.nodes      = out.nodes;
  tmp.num_nodes  = out.num_nodes;
  out.stack_need = compute_stack_need(tmp);
}

bool build_wide_bvh(const float* blo, const float* bhi, std::size_t n, int K, int leaf_size, WideBvh& out,
                    std::string& err)
{
  if (K != 4 && K != 8) {
    err = "branching must be 4 or 8";
    return false;
  }
  if (leaf_size < 1 || leaf_size > kMaxLeaf || n < 1) {
    err = "leaf_size must be in 1..16 and there must be at least one primitive";
    return false;
  }
  std::vector<float> cen(3 * n);
  for (std::size_t i = 0; i < 3 * n; i++) cen[i] = 0.5f * (blo[i] + bhi[i]);
  SahBuilder sb;
  sb.blo       = blo;
  sb.bhi       = bhi;
  sb.cen       = cen.data();
  sb.leaf_size = leaf_size;
  sb.idx.resize(n);
  for (std::size_t i = 0; i < n; i++) sb.idx[i] = uint32_t(i);
  sb.nodes.resize(std::max<std::size_t>(1, 2 * n));
  const int32_t root = sb.alloc();
  sb.build(root, 0, uint32_t(n), 0);
  sb.nodes.resize(std::size_t(sb.next.load()));
  collapse_binary(sb.nodes, root, K, out);
  out.order      = std::move(sb.idx);
  return true;
}

} // namespace ebgpu
