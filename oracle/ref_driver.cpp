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
#include <EBGeometry.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
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

void
ebgref_scene_value(void* h, const double* pts, int64_t n, double* v, int nthreads)
{
  const RefUnion& u = *static_cast<RefScene*>(h)->u;
  parallel_for(n, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++) v[i] = u.value(Vec3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
  });
}

} // extern "C"
