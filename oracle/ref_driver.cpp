// Thin extern "C" driver around the reference library (header-only C++17).
//
// This file is test infrastructure: it is compiled by oracle/Makefile against the
// reference headers found at $EBGEOMETRY_HOME (never copied into this repository)
// into oracle/_ref/libebgref*.so.  It is used to (1) write the golden fixtures under
// tests/golden/ and (2) time the reference CPU path as the CPU baseline.  Nothing in the
// product links it.
//
// What it wraps:
//   * mesh SDF      TriMeshSDF<double, uint32_t, K, W>(mesh, BVH::Build::SAH, 4) built through
//                   Soup::soupToDCEL (Source/EBGeometry_SoupImplem.hpp:148-226) and the mesh
//                   constructor (Source/EBGeometry_MeshDistanceFunctionsImplem.hpp:417-460);
//                   queries via signedDistance (:480-509) and getClosestTriangle (:511-549).
//                   Face metadata is set to the face index so the closest-triangle id is known.
//   * normals       DCEL face / edge / vertex pseudo-normals after MeshT::reconcile
//                   (Source/EBGeometry_DCEL_{Face,Edge,Vertex}Implem.hpp).
//   * SFC           SFC::computeBins<T>, SFC::Morton::encode, SFC::order<Morton>
//                   (Source/EBGeometry_SFCImplem.hpp:31-58, 228-300).
//   * CSG           BVHUnionIF<double, ImplicitFunction<double>, AABBT<double>, K> over spheres,
//                   boxes and Translate(Rotate(Rotate(Rotate(Scale(TriMeshSDF)))))) instances
//                   (Source/EBGeometry_CSGImplem.hpp:466-501, TransformImplem.hpp:140-270).
//   * point clouds  PointCloudBVH<double>: closestPoints (k nearest of external points) and
//                   allNearestNeighbors (Source/EBGeometry_PointCloudBVHImplem.hpp:235-507).
//   * polygon mesh  MeshSDF<double, uint32_t, K>(mesh, BVH::Build::SAH) over a DCEL mesh of
//                   general polygons built by Soup::soupToDCEL (Source/EBGeometry_
//                   MeshDistanceFunctionsImplem.hpp:250-354): signedDistance and the first
//                   entry of getClosestFaces(p, true) (face index in the metadata).
//   * expressions   any tree of the reference's analytic SDFs, transforms and CSG operators,
//                   given in post order as constructor calls (ebgref_expr_value below), built
//                   with the reference's own constructors and factories and evaluated by value().
#include <EBGeometry.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

using namespace EBGeometry;

#ifndef EBGREF_K
#define EBGREF_K (BVH::DefaultBranchingRatio<double>())
#endif
#ifndef EBGREF_W
#define EBGREF_W (TriangleSoA::DefaultWidth<double>())
#endif

namespace {
constexpr size_t K = EBGREF_K;
constexpr size_t W = EBGREF_W;

using Mesh   = DCEL::MeshT<double, uint32_t>;
using SDF    = TriMeshSDF<double, uint32_t, K, W>;
using Vec3   = Vec3T<double>;
using AABB   = BoundingVolumes::AABBT<double>;
using IF     = ImplicitFunction<double>;
using RefUnion = BVHUnionIF<double, IF, AABB, K>;

using PolySDF = MeshSDF<double, uint32_t, K>;

struct RefPolyMesh
{
  std::shared_ptr<Mesh>    mesh;
  std::shared_ptr<PolySDF> sdf;
};

struct RefMesh
{
  std::shared_ptr<Mesh> mesh;
  std::shared_ptr<SDF>  sdf;
};

template <class F>
void
parallel_for(int64_t n, int nthreads, F&& f)
{
  if (nthreads <= 1 || n < 1024) {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t            chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; t++) {
    const int64_t b = t * chunk;
    const int64_t e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([&f, b, e]() { f(b, e); });
  }
  for (auto& x : th) x.join();
}
} // namespace

extern "C" {

int
ebgref_info(int* k, int* w)
{
  *k = int(K);
  *w = int(W);
#if defined(__AVX512F__)
  return 512;
#elif defined(__AVX__)
  return 256;
#else
  return 128;
#endif
}

void*
ebgref_mesh_create(const double* xyz, int64_t nv, const uint32_t* tri, int64_t nt, int max_leaf_groups)
{
  std::vector<Vec3>                vertices(nv);
  std::vector<std::vector<size_t>> facets(nt);
  for (int64_t i = 0; i < nv; i++) vertices[i] = Vec3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  for (int64_t i = 0; i < nt; i++) facets[i] = {tri[3 * i], tri[3 * i + 1], tri[3 * i + 2]};

  auto* m = new RefMesh;
  m->mesh = std::make_shared<Mesh>();
  Soup::soupToDCEL(*m->mesh, vertices, facets, "ebgref");
  auto& faces = m->mesh->getFaces();
  for (size_t i = 0; i < faces.size(); i++) faces[i]->setMetaData(uint32_t(i));
  m->sdf = std::make_shared<SDF>(m->mesh, BVH::Build::SAH, size_t(max_leaf_groups));
  return m;
}

void*
ebgref_polymesh_create(const double* xyz, int64_t nv, const uint32_t* offsets, const uint32_t* idx, int64_t nf)
{
  std::vector<Vec3>                vertices(nv);
  std::vector<std::vector<size_t>> facets(nf);
  for (int64_t i = 0; i < nv; i++) vertices[i] = Vec3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  for (int64_t f = 0; f < nf; f++)
    for (uint32_t k = offsets[f]; k < offsets[f + 1]; k++) facets[f].push_back(idx[k]);
  auto* m = new RefPolyMesh;
  m->mesh = std::make_shared<Mesh>();
  Soup::soupToDCEL(*m->mesh, vertices, facets, "ebgref");
  auto& faces = m->mesh->getFaces();
  for (size_t i = 0; i < faces.size(); i++) faces[i]->setMetaData(uint32_t(i));
  m->sdf = std::make_shared<PolySDF>(m->mesh, BVH::Build::SAH);
  return m;
}

void
ebgref_polymesh_destroy(void* h)
{
  delete static_cast<RefPolyMesh*>(h);
}

// Signed distance per point; face[i] = metadata of the first face of getClosestFaces(p, true),
// dist[i] its unsigned distance.
void
ebgref_polymesh_query(void* h, const double* pts, int64_t n, double* d, uint32_t* face, double* dist, int nthreads)
{
  const PolySDF& sdf = *static_cast<RefPolyMesh*>(h)->sdf;
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const Vec3 p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
      d[i]         = sdf.signedDistance(p);
      const auto c = sdf.getClosestFaces(p, true);
      face[i]      = c.empty() ? 0xFFFFFFFFu : c.front().first->getMetaData();
      dist[i]      = c.empty() ? std::numeric_limits<double>::infinity() : c.front().second;
    }
  });
}

void
ebgref_mesh_destroy(void* h)
{
  delete static_cast<RefMesh*>(h);
}

void
ebgref_mesh_sdf(void* h, const double* pts, int64_t n, double* d, int nthreads)
{
  const SDF& sdf = *static_cast<RefMesh*>(h)->sdf;
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) d[i] = sdf.signedDistance(Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
  });
}

void
ebgref_mesh_closest(void* h, const double* pts, int64_t n, double* d, uint32_t* tri, int nthreads)
{
  const SDF& sdf = *static_cast<RefMesh*>(h)->sdf;
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) {
      const auto c = sdf.getClosestTriangle(Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
      d[i]         = c.signedDistance;
      tri[i]       = c.metaData;
    }
  });
}

// Per face: face normal (3), vertex pseudo-normals in facet order (9), half-edge pseudo-normals
// for the edges v0->v1, v1->v2, v2->v0 (9).
void
ebgref_mesh_normals(void* h, double* fn, double* vn, double* en)
{
  const auto& faces = static_cast<RefMesh*>(h)->mesh->getFaces();
  for (size_t i = 0; i < faces.size(); i++) {
    const auto n  = faces[i]->getNormal();
    const auto vs = faces[i]->gatherVertices();
    const auto es = faces[i]->gatherEdges();
    for (int c = 0; c < 3; c++) fn[3 * i + c] = n[c];
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) vn[9 * i + 3 * v + c] = vs[v]->getNormal()[c];
    for (int e = 0; e < 3; e++)
      for (int c = 0; c < 3; c++) en[9 * i + 3 * e + c] = es[e]->getNormal()[c];
  }
}

void
ebgref_bins_f32(const float* pts, int64_t n, uint32_t* bins)
{
  std::vector<Vec3T<float>> p(n);
  for (int64_t i = 0; i < n; i++) p[i] = Vec3T<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  const auto b = SFC::computeBins<float>(p);
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) bins[3 * i + c] = b[i][c];
}

void
ebgref_bins_f64(const double* pts, int64_t n, uint32_t* bins)
{
  std::vector<Vec3> p(n);
  for (int64_t i = 0; i < n; i++) p[i] = Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  const auto b = SFC::computeBins<double>(p);
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) bins[3 * i + c] = b[i][c];
}

void
ebgref_morton_encode(const uint32_t* bins, int64_t n, uint64_t* codes)
{
  for (int64_t i = 0; i < n; i++) codes[i] = SFC::Morton::encode(SFC::Index{bins[3 * i], bins[3 * i + 1], bins[3 * i + 2]});
}

void
ebgref_morton_order_f32(const float* pts, int64_t n, uint32_t* order)
{
  std::vector<Vec3T<float>> p(n);
  for (int64_t i = 0; i < n; i++) p[i] = Vec3T<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  const auto o = SFC::order<SFC::Morton, float>(p);
  std::copy(o.begin(), o.end(), order);
}

// CSG scene: spheres (center, radius), boxes (lo, hi) and mesh instances
// Translate(RotZ(RotY(RotX(Scale(mesh, s))))) -- i.e. value(p) = s * mesh(Rx Ry Rz (p - t) / s)
// with the reference's RotateIF conventions (angles in degrees).
struct RefScene
{
  std::vector<std::shared_ptr<RefMesh>> meshes;
  std::shared_ptr<RefUnion>              u;
};

void*
ebgref_scene_create(int64_t        ns,
                    const double*  sc,
                    const double*  sr,
                    int64_t        nb,
                    const double*  blo,
                    const double*  bhi,
                    int64_t        nm,
                    void* const*   meshes,
                    const double*  mt,
                    const double*  mrot,
                    const double*  ms,
                    const double*  mbvlo,
                    const double*  mbvhi)
{
  auto*                                              s = new RefScene;
  std::vector<std::pair<std::shared_ptr<const IF>, AABB>> pb;
  for (int64_t i = 0; i < ns; i++) {
    const Vec3 c(sc[3 * i], sc[3 * i + 1], sc[3 * i + 2]);
    const double r = sr[i];
    pb.emplace_back(std::make_shared<SphereSDF<double>>(c, r), AABB(c - r * Vec3::ones(), c + r * Vec3::ones()));
  }
  for (int64_t i = 0; i < nb; i++) {
    const Vec3 lo(blo[3 * i], blo[3 * i + 1], blo[3 * i + 2]);
    const Vec3 hi(bhi[3 * i], bhi[3 * i + 1], bhi[3 * i + 2]);
    pb.emplace_back(std::make_shared<BoxSDF<double>>(lo, hi), AABB(lo, hi));
  }
  for (int64_t i = 0; i < nm; i++) {
    auto* rm = static_cast<RefMesh*>(meshes[i]);
    std::shared_ptr<IF> f = rm->sdf;
    f                     = Scale<double>(f, ms[i]);
    f                     = Rotate<double>(f, mrot[3 * i + 0], 0);
    f                     = Rotate<double>(f, mrot[3 * i + 1], 1);
    f                     = Rotate<double>(f, mrot[3 * i + 2], 2);
    f                     = Translate<double>(f, Vec3(mt[3 * i], mt[3 * i + 1], mt[3 * i + 2]));
    const Vec3 lo(mbvlo[3 * i], mbvlo[3 * i + 1], mbvlo[3 * i + 2]);
    const Vec3 hi(mbvhi[3 * i], mbvhi[3 * i + 1], mbvhi[3 * i + 2]);
    pb.emplace_back(f, AABB(lo, hi));
  }
  s->u = std::make_shared<RefUnion>(pb);
  return s;
}

void
ebgref_scene_destroy(void* h)
{
  delete static_cast<RefScene*>(h);
}

// Expression nodes in post order: kind[i], n[i] (child count / axis / mesh index / samples) and
// 12 parameters p[12 i ..]; BVH unions take their children's bounding volumes from bv (6 per
// child: lo, hi), consumed in node order. Returns 0, or -(i + 1) for a malformed node i.
int
ebgref_expr_value(const int32_t* kind,
                  const int32_t* nn,
                  const double*  par,
                  int64_t        num_nodes,
                  const double*  bv,
                  void* const*   meshes,
                  const double*  pts,
                  int64_t        n,
                  double*        out,
                  int            nthreads)
{
  using Ptr = std::shared_ptr<IF>;
  std::vector<Ptr> st;
  int64_t          nb = 0;
  auto             v3 = [](const double* p) { return Vec3(p[0], p[1], p[2]); };
  for (int64_t i = 0; i < num_nodes; i++) {
    const double* p = par + 12 * i;
    const int     m = nn[i];
    auto          pop = [&st]() {
      Ptr f = st.back();
      st.pop_back();
      return f;
    };
    auto popn = [&st](int c) {
      std::vector<Ptr> v(st.end() - c, st.end());
      st.resize(st.size() - c);
      return v;
    };
    switch (kind[i]) {
    case 0: st.push_back(std::make_shared<SphereSDF<double>>(v3(p), p[3])); break;
    case 1: st.push_back(std::make_shared<BoxSDF<double>>(v3(p), v3(p + 3))); break;
    case 2: st.push_back(std::make_shared<PlaneSDF<double>>(v3(p), v3(p + 3))); break;
    case 3: st.push_back(std::make_shared<TorusSDF<double>>(v3(p), p[3], p[4])); break;
    case 4: st.push_back(std::make_shared<CylinderSDF<double>>(v3(p), v3(p + 3), p[6])); break;
    case 5: st.push_back(std::make_shared<InfiniteCylinderSDF<double>>(v3(p), p[3], size_t(m))); break;
    case 6: st.push_back(std::make_shared<CapsuleSDF<double>>(v3(p), v3(p + 3), p[6])); break;
    case 7: st.push_back(std::make_shared<InfiniteConeSDF<double>>(v3(p), p[3])); break;
    case 8: st.push_back(std::make_shared<ConeSDF<double>>(v3(p), p[3], p[4])); break;
    case 9: st.push_back(std::make_shared<RoundedBoxSDF<double>>(v3(p), p[3])); break;
    case 10: st.push_back(std::make_shared<RoundedCylinderSDF<double>>(p[0], p[1], p[2])); break;
    case 11: st.push_back(static_cast<RefMesh*>(meshes[m])->sdf); break;
    case 20: st.push_back(Translate<double>(pop(), v3(p))); break;
    case 21: st.push_back(Rotate<double>(pop(), p[0], size_t(m))); break;
    case 22: st.push_back(Scale<double>(pop(), p[0])); break;
    case 23: st.push_back(Complement<double>(pop())); break;
    case 24: st.push_back(Offset<double>(pop(), p[0])); break;
    case 25: st.push_back(Annular<double>(pop(), p[0])); break;
    case 26: st.push_back(std::make_shared<BlurIF<double>>(pop(), p[0], p[1])); break;
    case 27: st.push_back(Mollify<double>(pop(), p[0], size_t(m))); break;
    case 28: st.push_back(Elongate<double>(pop(), v3(p))); break;
    case 29: st.push_back(Reflect<double>(pop(), size_t(m))); break;
    case 30: st.push_back(FiniteRepetition<double>(pop(), v3(p), v3(p + 3), v3(p + 6))); break;
    case 40: st.push_back(Union<double>(popn(m))); break;
    case 41:
      if (p[1] != 0.0) st.push_back(std::make_shared<SmoothUnionIF<double>>(popn(m), p[0], ExpMin<double>));
      else st.push_back(SmoothUnion<double>(popn(m), p[0]));
      break;
    case 42: st.push_back(Intersection<double>(popn(m))); break;
    case 43: st.push_back(SmoothIntersection<double>(popn(m), p[0])); break;
    case 44: {
      std::vector<Ptr> v = popn(m);
      if (m == 2) st.push_back(Difference<double>(v[0], v[1]));
      else st.push_back(std::make_shared<DifferenceIF<double>>(v[0], std::vector<Ptr>(v.begin() + 1, v.end())));
      break;
    }
    case 45: {
      std::vector<Ptr> v = popn(m);
      if (m == 2) st.push_back(SmoothDifference<double>(v[0], v[1], p[0]));
      else
        st.push_back(
          std::make_shared<SmoothDifferenceIF<double>>(v[0], std::vector<Ptr>(v.begin() + 1, v.end()), p[0]));
      break;
    }
    case 46:
    case 47: {
      std::vector<Ptr>  v = popn(m);
      std::vector<AABB> b;
      for (int c = 0; c < m; c++, nb++) b.emplace_back(v3(bv + 6 * nb), v3(bv + 6 * nb + 3));
      if (kind[i] == 46) st.push_back(BVHUnion<double, IF, AABB, K>(v, b));
      else if (p[1] != 0.0)
        st.push_back(std::make_shared<BVHSmoothUnionIF<double, IF, AABB, K>>(v, b, p[0], ExpMin<double>));
      else st.push_back(BVHSmoothUnion<double, IF, AABB, K>(v, b, p[0]));
      break;
    }
    default: return int(-(i + 1));
    }
  }
  if (st.size() != 1) return int(-(num_nodes + 1));
  const IF& f = *st.back();
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) out[i] = f.value(Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
  });
  return 0;
}

void
ebgref_scene_value(void* h, const double* pts, int64_t n, double* v, int nthreads)
{
  const RefUnion& u = *static_cast<RefScene*>(h)->u;
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) v[i] = u.value(Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
  });
}

// k nearest cloud points of nq query points (self = 0) or of every cloud point excluding itself
// (self = 1, allNearestNeighbors; queries ignored): index (cloud order) and squared distance,
// missing entries index -1 / +inf.
int
ebgref_cloud_knn(const double* pts, int64_t n, const double* q, int64_t nq, int64_t k, int self, int64_t* idx,
                 double* d2, int nthreads)
{
  std::vector<Vec3> pos(n);
  for (int64_t i = 0; i < n; i++) pos[i] = Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  std::vector<std::size_t> meta(n);
  for (int64_t i = 0; i < n; i++) meta[i] = size_t(i);
  const PointCloudBVH<double> cloud(pos, meta);
  using Hit = PointCloudBVH<double>::Hit;
  if (self) {
    const std::vector<Hit> r = cloud.allNearestNeighbors(size_t(k));
    for (int64_t i = 0; i < n * k; i++) {
      const bool ok = r[i].index < size_t(n);
      idx[i]        = ok ? int64_t(r[i].index) : -1;
      d2[i]         = ok ? r[i].distanceSquared : std::numeric_limits<double>::infinity();
    }
    return 0;
  }
  parallel_for(nq, nthreads, [&](int64_t b, int64_t e) {
    std::vector<Hit> h(k);
    for (int64_t i = b; i < e; i++) {
      const size_t found = cloud.closestPoints(Vec3(q[3 * i], q[3 * i + 1], q[3 * i + 2]), size_t(k), h.data());
      for (int64_t j = 0; j < k; j++) {
        const bool ok         = size_t(j) < found;
        idx[i * k + j]        = ok ? int64_t(h[j].index) : -1;
        d2[i * k + j]         = ok ? h[j].distanceSquared : std::numeric_limits<double>::infinity();
      }
    }
  });
  return 0;
}

} // extern "C"
