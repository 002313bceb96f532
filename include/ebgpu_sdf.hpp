/*
 * ebgpu_sdf.hpp -- C++ adapters over the C ABI (ebgpu.h) for code written against the
 * reference's interface. Include the reference's EBGeometry.hpp before this header.
 *
 *   ebgpu::GpuTriMeshSDF<T>   a SignedDistanceFunction<T> (Source/EBGeometry_SignedDistanceFunction.hpp:
 *                             54-62) backed by an ebg_mesh: drops in wherever the reference's
 *                             TriMeshSDF<T, Meta, K, W> (Source/EBGeometry_MeshDistanceFunctions.hpp:
 *                             376-387) is held as a std::shared_ptr<ImplicitFunction<T>>.
 *   ebgpu::GpuMeshSDF<T>      the same over a mesh of planar polygon faces: the reference's
 *                             MeshSDF<T, Meta, K> (Source/EBGeometry_MeshDistanceFunctions.hpp:
 *                             149-281); closestTriangles gives face indices.
 *   ebgpu::GpuUnionIF<T>      an ImplicitFunction<T> backed by an ebg_scene (BVHUnionIF,
 *                             Source/EBGeometry_CSG.hpp:401-440).
 *
 * Besides the per-point calls of the reference interface (correct, but one device round trip
 * each), both give the batched calls a GPU needs: signedDistances / values over arrays of points,
 * and fillBox over a lattice in FAB layout (x fastest), which replaces the per-node value() calls
 * of an embedded-boundary generator (Integrations/AMReX/MeshSDF/main.cpp:23-69,
 * Integrations/Chombo/MeshSDF/main.cpp:21-50).
 *
 * Copies share the handle (std::shared_ptr), as the reference's TriMeshSDF copies share their
 * BVH (Source/EBGeometry_MeshDistanceFunctions.hpp:400). Construction failures throw
 * std::runtime_error with ebg_last_error(); the noexcept per-point calls return NaN on failure.
 * Points are rounded to float32 for the device (T = double callers: the result is that of the
 * rounded point).
 */
#ifndef EBGPU_SDF_HPP
#define EBGPU_SDF_HPP

#include "ebgpu.h"

#include <array>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ebgpu {

namespace detail {
inline void check(ebg_status s)
{
  if (s != EBG_OK) throw std::runtime_error(std::string("ebgpu: ") + ebg_last_error());
}

template <class T, class P>
std::vector<float> to_f32(const P* pts, std::size_t n)
{
  std::vector<float> q(3 * n);
  for (std::size_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) q[3 * i + c] = float(pts[i][c]);
  return q;
}
} // namespace detail

template <class T>
class GpuTriMeshSDF : public EBGeometry::SignedDistanceFunction<T>
{
public:
  using Vec3 = EBGeometry::Vec3T<T>;

  // vertices and (v0, v1, v2) vertex-index triples of a closed, outward oriented triangle mesh
  // (what Soup::soupToDCEL takes); branching K = 4 or 8, leaf_size 1..16 triangles
  GpuTriMeshSDF(const std::vector<Vec3>& vertices, const std::vector<std::array<uint32_t, 3>>& triangles,
                int branching = 4, int leaf_size = 8)
  {
    std::vector<float>    xyz = detail::to_f32<T>(vertices.data(), vertices.size());
    std::vector<uint32_t> tri(3 * triangles.size());
    for (std::size_t i = 0; i < triangles.size(); i++)
      for (int c = 0; c < 3; c++) tri[3 * i + c] = triangles[i][c];
    ebg_build_opts o{EBG_BUILD_SAH, branching, leaf_size, 0};
    ebg_mesh*      m = nullptr;
    detail::check(ebg_mesh_create(xyz.data(), vertices.size(), tri.data(), triangles.size(), &o, &m));
    m_mesh.reset(m, [](ebg_mesh* p) { (void)ebg_mesh_destroy(p); });
  }

  // one point (the reference signature); NaN if the device call fails
  T
  signedDistance(const Vec3& p) const noexcept override
  {
    const float q[3] = {float(p[0]), float(p[1]), float(p[2])};
    float       d    = std::numeric_limits<float>::quiet_NaN();
    if (ebg_query_sdf_host(m_mesh.get(), q, 1, &d, nullptr, nullptr) != EBG_OK)
      return std::numeric_limits<T>::quiet_NaN();
    return T(d);
  }

  // n points -> n signed distances (host arrays, one device round trip)
  void
  signedDistances(const Vec3* pts, std::size_t n, T* out) const
  {
    const std::vector<float> q = detail::to_f32<T>(pts, n);
    std::vector<float>       d(n);
    detail::check(ebg_query_sdf_host(m_mesh.get(), q.data(), n, d.data(), nullptr, nullptr));
    for (std::size_t i = 0; i < n; i++) out[i] = T(d[i]);
  }

  // TriMeshSDF::getClosestTriangle (Source/EBGeometry_MeshDistanceFunctionsImplem.hpp:511-549)
  // over n points: signed distance and the closest face's index in the input triangle list
  void
  closestTriangles(const Vec3* pts, std::size_t n, T* dist, uint32_t* face) const
  {
    const std::vector<float> q = detail::to_f32<T>(pts, n);
    std::vector<float>       d(n);
    detail::check(ebg_query_sdf_host(m_mesh.get(), q.data(), n, d.data(), face, nullptr));
    for (std::size_t i = 0; i < n; i++) dist[i] = T(d[i]);
  }

  // the lattice origin + (i, j, k) * spacing, i < dims[0], j < dims[1], k < dims[2], into
  // out[(k * dims[1] + j) * dims[0] + i]
  void
  fillBox(const T origin[3], const T spacing[3], const uint64_t dims[3], T* out) const
  {
    const double       o[3] = {double(origin[0]), double(origin[1]), double(origin[2])};
    const double       s[3] = {double(spacing[0]), double(spacing[1]), double(spacing[2])};
    const uint64_t     n    = dims[0] * dims[1] * dims[2];
    std::vector<float> d(n);
    detail::check(ebg_query_grid_host(m_mesh.get(), o, s, dims, d.data(), nullptr));
    for (uint64_t i = 0; i < n; i++) out[i] = T(d[i]);
  }

  ebg_mesh*
  handle() const noexcept
  {
    return m_mesh.get();
  }

protected:
  explicit GpuTriMeshSDF(ebg_mesh* m) { m_mesh.reset(m, [](ebg_mesh* p) { (void)ebg_mesh_destroy(p); }); }

private:
  std::shared_ptr<ebg_mesh> m_mesh;
};

template <class T>
class GpuMeshSDF : public GpuTriMeshSDF<T>
{
public:
  using Vec3 = EBGeometry::Vec3T<T>;

  // vertices and polygon faces (vertex-index lists, what Soup::soupToDCEL takes): planar, simple,
  // outward oriented, closed
  GpuMeshSDF(const std::vector<Vec3>& vertices, const std::vector<std::vector<std::size_t>>& faces,
             int branching = 4, int leaf_size = 8)
    : GpuTriMeshSDF<T>(create(vertices, faces, branching, leaf_size))
  {}

private:
  static ebg_mesh*
  create(const std::vector<Vec3>& vertices, const std::vector<std::vector<std::size_t>>& faces, int branching,
         int leaf_size)
  {
    std::vector<float>    xyz = detail::to_f32<T>(vertices.data(), vertices.size());
    std::vector<uint32_t> off{0u}, idx;
    for (const auto& f : faces) {
      for (std::size_t v : f) idx.push_back(uint32_t(v));
      off.push_back(uint32_t(idx.size()));
    }
    ebg_build_opts o{EBG_BUILD_SAH, branching, leaf_size, 0};
    ebg_mesh*      m = nullptr;
    detail::check(ebg_mesh_create_polygons(xyz.data(), vertices.size(), off.data(), idx.data(), faces.size(), &o, &m));
    return m;
  }
};

template <class T>
class GpuUnionIF : public EBGeometry::ImplicitFunction<T>
{
public:
  using Vec3 = EBGeometry::Vec3T<T>;

  // primitives as ebg_scene_create takes them (spheres, boxes, instances of `meshes` under
  // similarity maps); the meshes stay shared with their other holders
  GpuUnionIF(const std::vector<ebg_prim>& prims, const std::vector<GpuTriMeshSDF<T>>& meshes = {})
  {
    std::vector<ebg_mesh*> mh;
    for (const auto& m : meshes) mh.push_back(m.handle());
    ebg_scene* sc = nullptr;
    detail::check(ebg_scene_create(prims.data(), prims.size(), mh.data(), mh.size(), nullptr, &sc));
    m_scene.reset(sc, [](ebg_scene* p) { (void)ebg_scene_destroy(p); });
  }

  T
  value(const Vec3& p) const noexcept override
  {
    const float q[3] = {float(p[0]), float(p[1]), float(p[2])};
    float       v    = std::numeric_limits<float>::quiet_NaN();
    if (ebg_scene_query_host(m_scene.get(), q, 1, &v, nullptr) != EBG_OK) return std::numeric_limits<T>::quiet_NaN();
    return T(v);
  }

  void
  values(const Vec3* pts, std::size_t n, T* out) const
  {
    const std::vector<float> q = detail::to_f32<T>(pts, n);
    std::vector<float>       v(n);
    detail::check(ebg_scene_query_host(m_scene.get(), q.data(), n, v.data(), nullptr));
    for (std::size_t i = 0; i < n; i++) out[i] = T(v[i]);
  }

  void
  fillBox(const T origin[3], const T spacing[3], const uint64_t dims[3], T* out) const
  {
    const double       o[3] = {double(origin[0]), double(origin[1]), double(origin[2])};
    const double       s[3] = {double(spacing[0]), double(spacing[1]), double(spacing[2])};
    const uint64_t     n    = dims[0] * dims[1] * dims[2];
    std::vector<float> v(n);
    detail::check(ebg_scene_query_grid_host(m_scene.get(), o, s, dims, v.data(), nullptr));
    for (uint64_t i = 0; i < n; i++) out[i] = T(v[i]);
  }

private:
  std::shared_ptr<ebg_scene> m_scene;
};

} // namespace ebgpu

#endif
