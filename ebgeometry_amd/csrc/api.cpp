// C-ABI layer: handle lifecycle, device memory, the query pipeline.
//
// Query pipeline for one batch (all on the caller's stream, no host synchronisation):
//   1. Morton presort: query bbox -> 63-bit codes -> stable radix sort -> permutation
//   2. [learned bound] MLP forward -> per-query seeded cull radius (on a second stream,
//      concurrently with step 1)
//   3. FP32 packet traversal over the permutation (writes d, face id by query index)
//      with the exactness fallback inside the same wavefront: queries whose seeded radius held
//      no triangle walk the tree again unseeded
//   4. FP64 traversal for queries with |d| < recheck_eps * diag (sign is ill-conditioned in
//      FP32 there), seeded with |d32| + slack
#include "host_common.hpp"
#include "gpu_build.hpp"
#include "mesh_build.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {
thread_local std::string g_last_error;
} // namespace

using ebgpu::DevBuf;
using ebgpu::fail;
using ebgpu::upload;
using ebgpu::Workspace;

namespace ebgpu {
void set_last_error(const std::string& s) { g_last_error = s; }
} // namespace ebgpu

// Split packets. Queries near the medial axis (e.g. deep inside a closed surface) see many
// triangles at almost the same distance and cost one wavefront tens of times the median packet,
// so the last of them to finish sets the end of the launch. A 32^3 grid over the mesh box (grown
// by 10% per side) holds, per cell, the work of one packet of 64 queries spread over the cell
// (4x4x4 lattice), measured once by a counting traversal; from then on every presorted query of
// at least 2^14 points lists the packets whose most expensive cell reaches kHeavyFactor times the
// grid mean; the leading workgroups of the traversal launch share each listed packet among
// wavefronts that each walk one subtree of a frontier covering the tree, beside the bulk; the
// packets between kEarlyFactor and that threshold start next, before the bulk in curve order.
// Measured on the 327,680-face displaced sphere (K=4, leaf 16): kHeavyFactor 6 above 2^22
// queries (the launch is full, and splitting costs extra work), 3 below (spare capacity); more
// than 16 subtree roots only adds work.
constexpr uint64_t kHeavyMinQ   = 1u << 14;
constexpr uint32_t kSplitBlocks = 2048; // workgroups that stride over the split items
constexpr std::size_t kSplitRoots = 16;
// Packets at >= 1.5x the grid mean that are not split are listed too and start right after the
// split items, before the bulk (Hilbert-ordered 16M queries on the 327,680-face displaced sphere: 20.4 -> 17.4 ms).
constexpr double kEarlyFactor = 1.5;
static double heavy_factor(uint64_t n) { return n >= (1u << 22) ? 6.0 : 3.0; }


// Subtree roots for split packets: breadth-first from the root, each interior entry replaced by
// its children until there are at least `target` roots (leaves stay as they are).
template <int K>
static std::vector<ebgpu::SplitRoot> split_frontier(const ebgpu::PackedMesh& pm, std::size_t target)
{
  const auto*                   nodes = reinterpret_cast<const ebgpu::Node<K>*>(pm.nodes.data());
  std::vector<ebgpu::SplitRoot> cur{{0u, 0u, {pm.lo[0], pm.lo[1], pm.lo[2]}, {pm.hi[0], pm.hi[1], pm.hi[2]}}};
  while (cur.size() < target) {
    std::vector<ebgpu::SplitRoot> next;
    bool                          grew = false;
    for (std::size_t i = 0; i < cur.size(); i++) {
      const ebgpu::SplitRoot& r = cur[i];
      if (r.cnt != 0u || next.size() + (cur.size() - i) >= target) {
        next.push_back(r);
        continue;
      }
      const ebgpu::Node<K>& nd = nodes[r.ref];
      for (int k = 0; k < K; k++) {
        if (nd.cnt[k] == ebgpu::kInvalid) continue;
        next.push_back({nd.ref[k], nd.cnt[k], {nd.lox[k], nd.loy[k], nd.loz[k]}, {nd.hix[k], nd.hiy[k], nd.hiz[k]}});
      }
      grew = true;
    }
    cur.swap(next);
    if (!grew) break;
  }
  return cur;
}

extern "C" {

const char* ebg_last_error(void) { return g_last_error.c_str(); }

void ebg_query_opts_default(ebg_query_opts* o)
{
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->presort      = 1;
  o->fp64_recheck = 1;
  o->recheck_eps  = 1e-4;
  o->bound_scale  = 1.25f;
  o->bound_offset = 1e-3f;
}

static ebg_status finalize_mesh(ebg_mesh* m, const ebgpu::PackedMesh& pm);

// Upload a mesh packed on the host, then finish the handle.
static ebg_status upload_mesh(ebg_mesh* m, ebgpu::PackedMesh& pm)
{
  EBG_HIP(upload(m->nodes, pm.nodes));
  EBG_HIP(upload(m->tris, pm.tris));
  EBG_HIP(upload(m->tripos, pm.tripos));
  EBG_HIP(upload(m->tsph, pm.tsph));
  EBG_HIP(upload(m->feat, pm.feat));
  EBG_HIP(upload(m->vn32, pm.vn32));
  EBG_HIP(upload(m->en32, pm.en32));
  EBG_HIP(upload(m->vn64, pm.vn64));
  EBG_HIP(upload(m->en64, pm.en64));
  EBG_HIP(upload(m->fn64, pm.fn64));
  {
    const std::size_t  nt = pm.tris.size();
    std::vector<float> cdf(nt);
    double             acc = 0.0;
    std::vector<double> run(nt);
    for (std::size_t i = 0; i < nt; i++) {
      const float* x  = &pm.tripos[12 * i];
      const double ax = double(x[3]) - x[0], ay = double(x[4]) - x[1], az = double(x[5]) - x[2];
      const double bx = double(x[6]) - x[0], by = double(x[7]) - x[1], bz = double(x[8]) - x[2];
      const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
      acc += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
      run[i] = acc;
    }
    for (std::size_t i = 0; i < nt; i++) cdf[i] = float(run[i] / acc);
    if (nt) cdf[nt - 1] = 1.0f;
    EBG_HIP(upload(m->area_cdf, cdf));
  }
  return finalize_mesh(m, pm);
}

// Adopt a mesh packed on the device (its nodes come from the host collapse), then finish the handle.
static ebg_status adopt_gpu_mesh(ebg_mesh* m, ebgpu::GpuBuiltMesh& g)
{
  auto take = [](ebgpu::DevBuf& dst, ebgpu::DevBuf& src) {
    std::swap(dst.p, src.p);
    std::swap(dst.bytes, src.bytes);
  };
  EBG_HIP(upload(m->nodes, g.meta.nodes));
  take(m->tris, g.tris);
  take(m->tripos, g.tripos);
  take(m->tsph, g.tsph);
  take(m->feat, g.feat);
  take(m->vn32, g.vn32);
  take(m->en32, g.en32);
  take(m->vn64, g.vn64);
  take(m->en64, g.en64);
  take(m->fn64, g.fn64);
  take(m->area_cdf, g.area_cdf);
  return finalize_mesh(m, g.meta);
}

// Split frontier, counts, stack banks, bounds and the device view of an uploaded mesh.
static ebg_status finalize_mesh(ebg_mesh* m, const ebgpu::PackedMesh& pm)
{
  {
    const std::vector<ebgpu::SplitRoot> roots =
      pm.K == 4 ? split_frontier<4>(pm, kSplitRoots) : split_frontier<8>(pm, kSplitRoots);
    EBG_HIP(upload(m->split_roots, roots));
    m->num_split_roots = uint32_t(roots.size());
  }
  m->K          = pm.K;
  m->leaf       = pm.leaf_size;
  m->num_nodes  = pm.num_nodes;
  m->num_leaves = pm.num_leaves;
  m->stack_need = pm.stack_need;
  m->banks      = (pm.stack_need + 63) / 64;
  if (m->banks > 4)
    return fail(EBG_STACK_OVERFLOW, "tree needs a traversal stack of " + std::to_string(pm.stack_need) +
                                      " entries (max 256); use a smaller branching factor");
  if (m->banks == 3) m->banks = 4;
  for (int c = 0; c < 3; c++) {
    m->lo[c] = pm.lo[c];
    m->hi[c] = pm.hi[c];
  }
  m->diag = pm.diag;
  ebgpu::MeshDev& d = m->dev;
  d.nodes           = m->nodes.p;
  d.tris            = m->tris.as<ebgpu::TriRec>();
  d.tripos          = m->tripos.as<float>();
  d.tsph            = m->tsph.as<float>();
  d.feat            = m->feat.as<uint32_t>();
  d.vn32            = m->vn32.as<float>();
  d.en32            = m->en32.as<float>();
  d.vn64            = m->vn64.as<double>();
  d.en64            = m->en64.as<double>();
  d.fn64            = m->fn64.as<double>();
  for (int c = 0; c < 3; c++) {
    d.root_lo[c] = pm.lo[c];
    d.root_hi[c] = pm.hi[c];
  }
  d.diag          = float(pm.diag);
  d.num_tris      = uint32_t(m->nt);
  d.hmax          = pm.hmax;
  EBG_HIP(upload(m->status, std::vector<uint32_t>(16, 0u)));
  d.status = m->status.as<uint32_t>();
  // leaf slabs, from the device arrays (the same for every builder)
  EBG_HIP(m->lthick.reserve(4 * std::max<uint64_t>(m->nt, 1)));
  EBG_HIP(ebgpu::launch_leaf_slabs(m->nodes.p, pm.num_nodes, pm.K, d.tsph, d.tripos, m->lthick.as<float>(), nullptr));
  EBG_HIP(hipStreamSynchronize(nullptr));
  d.lthick        = m->lthick.as<float>();
  EBG_HIP(m->frames.reserve(sizeof(ebgpu::FrameRec) * std::max<uint64_t>(m->nt, 1)));
  EBG_HIP(ebgpu::launch_frames(d.tripos, m->nt, m->frames.as<ebgpu::FrameRec>(), nullptr));
  EBG_HIP(hipStreamSynchronize(nullptr));
  d.frames        = m->frames.as<ebgpu::FrameRec>();
  m->device_bytes = m->nodes.bytes + m->tris.bytes + m->tripos.bytes + m->tsph.bytes + m->feat.bytes + m->vn32.bytes +
                    m->en32.bytes + m->vn64.bytes + m->en64.bytes + m->fn64.bytes + m->lthick.bytes + m->frames.bytes;
  return EBG_OK;
}

// Build and upload a triangle mesh whose pseudo-normals are given (both creation entries).
static ebg_status create_mesh(const float* xyz, uint64_t nv, const uint32_t* tri, uint64_t nt,
                              const ebgpu::PseudoNormals* pn, const ebg_build_opts* opts,
                              std::chrono::steady_clock::time_point t0, ebg_mesh** out)
{
  ebg_build_opts o{};
  if (opts) o = *opts;
  const int K    = o.branching ? o.branching : 4;
  const int leaf = o.leaf_size ? o.leaf_size : 4;
  if (K != 4 && K != 8) return fail(EBG_INVALID_ARG, "branching must be 4 or 8");
  if (leaf < 1 || leaf > ebgpu::kMaxLeaf) return fail(EBG_INVALID_ARG, "leaf_size must be 1..16");
  ebgpu::PackedMesh pm;
  std::string       err;
  if (o.builder == EBG_BUILD_SAH_GPU) {
    ebgpu::GpuBuiltMesh g;
    const ebg_status    bs = ebgpu::build_mesh_gpu_sah(xyz, nv, tri, nt, K, leaf, pn, g, err);
    if (bs != EBG_OK) return fail(bs, err);
    auto* m = new ebg_mesh;
    m->nv   = nv;
    m->nt   = nt;
    m->hash = ebgpu::mesh_hash(xyz, nv, tri, nt);
    m->host_xyz.assign(xyz, xyz + 3 * nv);
    m->order = g.meta.order;
    const ebg_status st = adopt_gpu_mesh(m, g);
    if (st != EBG_OK) {
      delete m;
      return st;
    }
    m->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out             = m;
    return EBG_OK;
  }
  if (!pn) return fail(EBG_INVALID_ARG, "no pseudo-normals");
  if (o.builder == EBG_BUILD_MORTON) {
    if (!ebgpu::build_morton_mesh_gpu(xyz, nv, tri, nt, K, o.leaf_size ? o.leaf_size : 1, *pn, pm, err))
      return fail(EBG_HIP_ERROR, err);
  }
  else if (o.builder == EBG_BUILD_SAH) {
    if (!ebgpu::build_sah_mesh(xyz, nv, tri, nt, K, leaf, *pn, pm, err)) return fail(EBG_INVALID_ARG, err);
  }
  else {
    return fail(EBG_INVALID_ARG, "unknown builder");
  }

  auto* m = new ebg_mesh;
  m->nv   = nv;
  m->nt   = nt;
  m->hash = ebgpu::mesh_hash(xyz, nv, tri, nt);
  m->host_xyz.assign(xyz, xyz + 3 * nv);
  m->order = std::move(pm.order);
  const ebg_status st = upload_mesh(m, pm);
  if (st != EBG_OK) {
    delete m;
    return st;
  }
  m->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out             = m;
  return EBG_OK;
}

static ebg_status check_vertices(const float* xyz, uint64_t nv)
{
  for (uint64_t i = 0; i < 3 * nv; i++)
    if (!std::isfinite(xyz[i])) return fail(EBG_INVALID_ARG, "non-finite vertex coordinate");
  return EBG_OK;
}

ebg_status ebg_mesh_create(const float* xyz, uint64_t nv, const uint32_t* tri, uint64_t nt,
                           const ebg_build_opts* opts, ebg_mesh** out)
{
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (!xyz || !tri || nv < 3 || nt < 1) return fail(EBG_INVALID_ARG, "empty mesh");
  if (nt > 0x7FFFFFFFull / 3 || nv > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "mesh too large");
  if (check_vertices(xyz, nv) != EBG_OK) return EBG_INVALID_ARG;
  for (uint64_t i = 0; i < 3 * nt; i++)
    if (tri[i] >= nv) return fail(EBG_INVALID_ARG, "triangle index out of range");
  const auto t0 = std::chrono::steady_clock::now();
  if (opts && opts->builder == EBG_BUILD_SAH_GPU) return create_mesh(xyz, nv, tri, nt, nullptr, opts, t0, out);
  ebgpu::PseudoNormals pn;
  std::string          err;
  if (!ebgpu::compute_pseudo_normals(xyz, nv, tri, nt, pn, err)) return fail(EBG_INVALID_ARG, err);
  return create_mesh(xyz, nv, tri, nt, &pn, opts, t0, out);
}

ebg_status ebg_mesh_create_polygons(const float* xyz, uint64_t nv, const uint32_t* face_offsets,
                                    const uint32_t* face_vertices, uint64_t nf, const ebg_build_opts* opts,
                                    ebg_mesh** out)
{
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (!xyz || !face_offsets || !face_vertices || nv < 3 || nf < 1) return fail(EBG_INVALID_ARG, "empty mesh");
  if (nv > 0xFFFFFFFFull || nf > 0x7FFFFFFFull / 3 || face_offsets[nf] > 0x7FFFFFFFull / 3)
    return fail(EBG_INVALID_ARG, "mesh too large");
  if (check_vertices(xyz, nv) != EBG_OK) return EBG_INVALID_ARG;
  const auto            t0 = std::chrono::steady_clock::now();
  std::vector<uint32_t> tri, tri_face;
  ebgpu::PseudoNormals  pn;
  std::string           err;
  if (!ebgpu::polygon_mesh(xyz, nv, face_offsets, face_vertices, nf, tri, tri_face, pn, err))
    return fail(EBG_INVALID_ARG, err);
  const uint64_t   nt = tri_face.size();
  const ebg_status st = create_mesh(xyz, nv, tri.data(), nt, &pn, opts, t0, out);
  if (st != EBG_OK) return st;
  ebg_mesh* m = *out;
  m->nf       = nf;
  m->host_tri_face = tri_face;
  if (upload(m->tri_face, tri_face) != hipSuccess) {
    ebgpu::mesh_release(m);
    *out = nullptr;
    return fail(EBG_OOM, "face map upload");
  }
  return EBG_OK;
}

ebg_status ebg_mesh_triangle_faces(const ebg_mesh* m, uint32_t* tri_face)
{
  if (!m || !tri_face) return fail(EBG_INVALID_ARG, "null mesh or output");
  for (uint64_t t = 0; t < m->nt; t++) tri_face[t] = m->host_tri_face.empty() ? uint32_t(t) : m->host_tri_face[t];
  return EBG_OK;
}

ebg_status ebg_mesh_destroy(ebg_mesh* mesh)
{
  if (mesh) ebgpu::mesh_release(mesh);
  return EBG_OK;
}

ebg_status ebg_pseudo_normals_host(const float* xyz, uint64_t nv, const uint32_t* tri, uint64_t nt, double* fn,
                                   double* vn, double* en)
{
  if (!xyz || !tri || !fn || !vn || !en || nv < 3 || nt < 1) return fail(EBG_INVALID_ARG, "bad arguments");
  for (uint64_t i = 0; i < 3 * nt; i++)
    if (tri[i] >= nv) return fail(EBG_INVALID_ARG, "triangle index out of range");
  ebgpu::PseudoNormals pn;
  std::string          err;
  if (!ebgpu::compute_pseudo_normals(xyz, nv, tri, nt, pn, err)) return fail(EBG_INVALID_ARG, err);
  for (uint64_t f = 0; f < nt; f++) {
    for (int c = 0; c < 3; c++) fn[3 * f + c] = pn.face[3 * f + c];
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) {
        vn[9 * f + 3 * v + c] = pn.vert[3 * tri[3 * f + v] + c];
        en[9 * f + 3 * v + c] = pn.edge[9 * f + 3 * v + c];
      }
  }
  return EBG_OK;
}

ebg_status ebg_polygons_to_triangles_host(const float* xyz, uint64_t nv, const uint32_t* face_offsets,
                                         const uint32_t* face_vertices, uint64_t nf, uint32_t* tri, uint32_t* tri_face,
                                         double* fn, double* vn, double* en)
{
  if (!xyz || !face_offsets || !face_vertices || !tri || !tri_face || !fn || !vn || !en || nv < 3 || nf < 1)
    return fail(EBG_INVALID_ARG, "bad arguments");
  std::vector<uint32_t> t, tf;
  ebgpu::PseudoNormals  pn;
  std::string           err;
  if (!ebgpu::polygon_mesh(xyz, nv, face_offsets, face_vertices, nf, t, tf, pn, err)) return fail(EBG_INVALID_ARG, err);
  const uint64_t nt = tf.size();
  if (nt != uint64_t(face_offsets[nf]) - 2 * nf) return fail(EBG_INVALID_ARG, "unexpected triangle count");
  std::copy(t.begin(), t.end(), tri);
  std::copy(tf.begin(), tf.end(), tri_face);
  for (uint64_t f = 0; f < nt; f++) {
    for (int c = 0; c < 3; c++) fn[3 * f + c] = pn.face[3 * f + c];
    for (int v = 0; v < 3; v++)
      for (int c = 0; c < 3; c++) {
        vn[9 * f + 3 * v + c] = pn.vert[3 * t[3 * f + v] + c];
        en[9 * f + 3 * v + c] = pn.edge[9 * f + 3 * v + c];
      }
  }
  return EBG_OK;
}

ebg_status ebg_mesh_leaf_order(const ebg_mesh* m, uint32_t* faces)
{
  if (!m || !faces) return fail(EBG_INVALID_ARG, "null argument");
  std::copy(m->order.begin(), m->order.end(), faces);
  return EBG_OK;
}

ebg_status ebg_mesh_get_info(const ebg_mesh* m, ebg_mesh_info* info)
{
  if (!m || !info) return fail(EBG_INVALID_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  for (int c = 0; c < 3; c++) {
    info->bbox_lo[c] = m->lo[c];
    info->bbox_hi[c] = m->hi[c];
  }
  info->diag           = m->diag;
  info->num_vertices   = m->nv;
  info->num_triangles  = m->nt;
  info->num_nodes      = m->num_nodes;
  info->num_leaves     = m->num_leaves;
  info->branching      = m->K;
  info->leaf_size      = m->leaf;
  info->stack_need     = m->stack_need;
  info->stack_capacity = 64 * m->banks;
  info->device_bytes   = m->device_bytes;
  info->build_seconds  = m->build_seconds;
  return EBG_OK;
}


} // extern "C"

namespace ebgpu {
ebg_status ensure_cost_grid(ebg_mesh* m, hipStream_t s)
{
  if (m->grid_ready) return EBG_OK;
  const uint32_t g  = ebgpu::kCostGrid;
  const uint64_t ng = uint64_t(g) * g * g; // cells = packets of the measuring launch
  double         lo[3], step[3];
  for (int c = 0; c < 3; c++) {
    const double ext = std::max(m->hi[c] - m->lo[c], 1e-3 * m->diag);
    lo[c]            = m->lo[c] - 0.1 * ext;
    step[c]          = 1.2 * ext / g;
    m->grid_lo[c]    = float(lo[c]);
    m->grid_scale[c] = float(1.0 / step[c]);
  }
  std::vector<float> pts(3 * 64 * ng);
  for (uint32_t z = 0; z < g; z++)
    for (uint32_t y = 0; y < g; y++)
      for (uint32_t x = 0; x < g; x++)
        for (uint32_t l = 0; l < 64; l++) {
          float* p = &pts[3 * (64 * ((uint64_t(z) * g + y) * g + x) + l)];
          p[0]     = float(lo[0] + (x + ((l & 3) + 0.5) / 4) * step[0]);
          p[1]     = float(lo[1] + (y + (((l >> 2) & 3) + 0.5) / 4) * step[1]);
          p[2]     = float(lo[2] + (z + ((l >> 4) + 0.5) / 4) * step[2]);
        }
  ebgpu::DevBuf dp, dd;
  EBG_HIP(upload(dp, pts));
  EBG_HIP(dd.reserve(4 * 64 * ng));
  EBG_HIP(m->cost_grid.reserve(4 * ng));
  ebgpu::QueryArgs q{};
  q.pts         = dp.as<float>();
  q.n           = 64 * ng;
  q.out_d       = dd.as<float>();
  q.packet_cost = m->cost_grid.as<uint32_t>();
  EBG_HIP(ebgpu::launch_sdf32(m->dev, m->K, m->banks, q, true, s));
  std::vector<uint32_t> cost(ng);
  EBG_HIP(hipMemcpyAsync(cost.data(), m->cost_grid.p, 4 * ng, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  double sum = 0.0;
  for (uint32_t c : cost) sum += c;
  m->grid_mean      = sum / double(ng);
  m->grid_ready     = true;
  return EBG_OK;
}
} // namespace ebgpu
using ebgpu::ensure_cost_grid;

namespace ebgpu {
// The query pipeline over a workspace of at least n queries. With may_alloc false nothing is
// allocated or created: split packets are used only if the mesh's cost grid already exists, the
// bound network runs on the caller's stream, and the traversal is not timed.
ebg_status query_pipeline(ebg_mesh* m, const float* d_pts, uint64_t n, float* d_dist, uint32_t* d_tri,
                          const ebg_query_opts& o, const WsView& ws, bool may_alloc)
{
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  if (ws.cap < n) return fail(EBG_INVALID_ARG, "query workspace smaller than the batch");
  if (may_alloc && !m->ev0) {
    EBG_HIP(hipEventCreate(&m->ev0));
    EBG_HIP(hipEventCreate(&m->ev1));
  }
  const bool timed = m->ev0 && may_alloc;
  // counting launches run unsplit: the counters describe the plain traversal's own work; so do
  // closest-face queries (each lane decides its face from the band of its own whole walk)
  const bool face        = d_tri != nullptr;
  const bool heavy_first = o.presort && n >= kHeavyMinQ && !o.work && !face && (may_alloc || m->grid_ready);
  m->last_split           = heavy_first;
  m->last_split_cap       = heavy_first ? Workspace::heavy_cap(n) : 0;
  if (heavy_first) {
    const ebg_status gs = ensure_cost_grid(m, s);
    if (gs != EBG_OK) return gs;
  }
  uint32_t* cnt = ws.counters; // [0] recheck, [1] fallback, [2] split packets, [3] early packets
  EBG_HIP(launch_fill_u32(cnt, 4, 0u, s));

  // The bound network needs no presort (it writes the radius by query index), so it runs on a
  // second stream of the handle, in query order, while the presort runs on the caller's stream.
  const float* seeds = o.seed_radius;
  const bool   fork  = o.learned_bound && (may_alloc || m->aux);
  if (o.learned_bound && !fork) {
    // no second stream without creating one: the network runs first on the caller's stream
    EBG_HIP(mlp_forward(m->mlp, d_pts, nullptr, n, true, o.bound_scale, float(o.bound_offset * m->diag), ws.seeds, s));
    seeds = ws.seeds;
  }
  else if (o.learned_bound) {
    if (!m->aux) {
      EBG_HIP(hipStreamCreateWithFlags(&m->aux, hipStreamNonBlocking));
      EBG_HIP(hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming));
      EBG_HIP(hipEventCreateWithFlags(&m->ev_join, hipEventDisableTiming));
    }
    EBG_HIP(hipEventRecord(m->ev_fork, s));
    EBG_HIP(hipStreamWaitEvent(m->aux, m->ev_fork, 0));
    EBG_HIP(mlp_forward(m->mlp, d_pts, nullptr, n, true, o.bound_scale, float(o.bound_offset * m->diag), ws.seeds,
                        m->aux));
    EBG_HIP(hipEventRecord(m->ev_join, m->aux));
    seeds = ws.seeds;
  }

  // the presort keys carry each query's cost flags (cost-grid value against the heavy and early
  // thresholds) to the packet marking
  CostFlags cf{};
  if (heavy_first) {
    cf.grid = m->cost_grid.as<uint32_t>();
    for (int c = 0; c < 3; c++) {
      cf.lo[c]    = m->grid_lo[c];
      cf.scale[c] = m->grid_scale[c];
    }
    cf.g     = kCostGrid;
    cf.heavy = uint32_t(std::max(1.0, heavy_factor(n) * m->grid_mean));
    cf.early = uint32_t(std::max(1.0, kEarlyFactor * m->grid_mean));
  }
  const uint32_t* idx = nullptr;
  if (o.presort && n > 1) {
    EBG_HIP(presort(ws, d_pts, n, s, heavy_first ? &cf : nullptr));
    idx = ws.perm;
  }
  HeavyMarkArgs h{};
  if (heavy_first) {
    h.skeys       = ws.codes_sorted;
    h.n           = n;
    h.cap         = Workspace::heavy_cap(n);
    h.list        = ws.heavy_list;
    h.count       = cnt + 2;
    h.flag        = ws.heavy_flag;
    h.keys        = ws.heavy_keys;
    h.early_cap   = Workspace::early_cap(n);
    h.early_list  = ws.early_list;
    h.early_count = cnt + 3;
    EBG_HIP(launch_heavy_mark(h, s));
  }
  if (fork) EBG_HIP(hipStreamWaitEvent(s, m->ev_join, 0));

  // the FP64 pass also decides the closest face of the lanes whose FP32 band overflowed
  const bool recheck = o.fp64_recheck || o.fp64_all || face;
  QueryArgs a{};
  a.pts            = d_pts;
  a.idx            = idx;
  a.n              = n;
  a.seed_r         = seeds;
  a.out_d          = d_dist;
  a.out_tri        = d_tri;
  a.recheck_abs    = o.fp64_all ? INFINITY : (o.fp64_recheck ? float(o.recheck_eps * m->diag) : 0.0f);
  a.recheck_list   = recheck ? ws.recheck_list : nullptr;
  a.recheck_count  = cnt;
  a.recheck_mask   = recheck ? ws.rmask : nullptr;
  a.fallback_list  = nullptr;
  a.fallback_count = cnt + 1;
  a.work           = reinterpret_cast<unsigned long long*>(o.work);
  a.pair_work      = reinterpret_cast<unsigned long long*>(o.pair_work);
  const bool count = o.work != nullptr;

  if (heavy_first) {
    // split packets run unseeded: exact without a fallback
    a.heavy_flag   = h.flag;
    a.split_roots  = m->split_roots.as<SplitRoot>();
    a.split_list   = h.list;
    a.split_count  = h.count;
    a.split_keys   = h.keys;
    a.split_nroots = m->num_split_roots;
    a.split_cap    = h.cap;
    a.split_blocks = std::min<uint32_t>(kSplitBlocks, (h.cap * m->num_split_roots + 3) / 4);
    if (h.early_list) {
      a.early_list   = h.early_list;
      a.early_count  = h.early_count;
      a.early_cap    = h.early_cap;
      a.early_blocks = (h.early_cap + 3) / 4;
    }
  }

  // events bracket the FP32 traversal (one launch for the bulk and the split items, then the
  // split packets' resolve)
  if (timed) EBG_HIP(hipEventRecord(m->ev0, s));
  EBG_HIP(launch_sdf32(m->dev, m->K, m->banks, a, count, s));
  if (heavy_first) EBG_HIP(launch_sdf32_resolve(m->dev, a, s));
  if (timed) {
    EBG_HIP(hipEventRecord(m->ev1, s));
    m->timed = true;
  }

  if (recheck) {
    // the list follows the presort order, so FP64 packets hold neighbouring queries
    EBG_HIP(launch_recheck_compact(a.recheck_mask, n, idx, ws.recheck_list, cnt, s));
    QueryArgs r = a;
    r.idx       = ws.recheck_list;
    r.count_ptr        = cnt;
    r.seed_r           = nullptr;
    r.recheck_list     = nullptr;
    r.recheck_mask     = nullptr;
    r.fallback_list    = nullptr;
    r.work             = nullptr;
    r.pair_work        = nullptr;
    // bound (|d32| (1 + rel) + abs)^2 always contains the FP64 minimum: FP32 distances are
    // accurate to a few ulp of the coordinates
    r.slack_abs = float(1e-5 * m->diag);
    r.slack_rel = 1e-4f;
    EBG_HIP(launch_sdf64(m->dev, m->K, m->banks, r, s));
  }
  // polygon meshes report the face each closest triangle was clipped from
  if (d_tri && m->tri_face.p) EBG_HIP(launch_map_u32(d_tri, n, m->tri_face.as<uint32_t>(), uint32_t(m->nt), s));
  if (count) EBG_HIP(launch_add_counts(a.work, recheck ? cnt : nullptr, seeds ? cnt + 1 : nullptr, s));
  return EBG_OK;
}

} // namespace ebgpu

extern "C" {

ebg_status ebg_query_sdf(const ebg_mesh* cm, const float* d_pts, uint64_t n, float* d_dist, uint32_t* d_tri,
                         const ebg_query_opts* opts)
{
  if (!cm) return fail(EBG_INVALID_ARG, "mesh is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_dist) return fail(EBG_INVALID_ARG, "null point or output array");
  if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 queries per call");
  ebg_mesh*      m = const_cast<ebg_mesh*>(cm);
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  if (o.learned_bound && !m->mlp.trained) return fail(EBG_NOT_READY, "learned bound requested but no network trained");
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  EBG_HIP(m->ws.ensure(n));
  return ebgpu::query_pipeline(m, d_pts, n, d_dist, d_tri, o, m->ws.view(), true);
}

ebg_status ebg_query_sdf_f64(const ebg_mesh* cm, const double* d_pts, uint64_t n, double* d_dist, uint32_t* d_tri,
                             const ebg_query_opts* opts)
{
  if (!cm) return fail(EBG_INVALID_ARG, "mesh is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_dist) return fail(EBG_INVALID_ARG, "null point or output array");
  if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 queries per call");
  ebg_mesh*      m = const_cast<ebg_mesh*>(cm);
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t                           s = static_cast<hipStream_t>(o.stream);
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  EBG_HIP(m->pts32.reserve(12 * n));
  EBG_HIP(m->dist32.reserve(4 * n));
  EBG_HIP(ebgpu::launch_f64_to_f32(d_pts, m->pts32.as<float>(), 3 * n, s));
  const ebg_status st = ebg_query_sdf(m, m->pts32.as<float>(), n, m->dist32.as<float>(), d_tri, &o);
  if (st != EBG_OK) return st;
  EBG_HIP(ebgpu::launch_f32_to_f64(m->dist32.as<float>(), d_dist, n, s));
  return EBG_OK;
}

} // extern "C"

namespace ebgpu {
// Validate a lattice (origin, spacing, dims) and its point count.
ebg_status grid_spec(const double* origin, const double* spacing, const uint64_t* dims, GridSpec& g, uint64_t& n)
{
  if (!origin || !spacing || !dims) return fail(EBG_INVALID_ARG, "null origin, spacing or dims");
  n = 1;
  for (int c = 0; c < 3; c++) {
    if (!std::isfinite(origin[c]) || !std::isfinite(spacing[c]))
      return fail(EBG_INVALID_ARG, "non-finite lattice origin or spacing");
    if (dims[c] > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "lattice dimension above 2^32-1");
    g.origin[c]  = origin[c];
    g.spacing[c] = spacing[c];
    g.dims[c]    = dims[c];
    n            = dims[c] == 0 ? 0 : n * dims[c];
    if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 lattice points per call");
  }
  return EBG_OK;
}
} // namespace ebgpu

extern "C" {

ebg_status ebg_query_grid(const ebg_mesh* cm, const double* origin, const double* spacing, const uint64_t* dims,
                          float* d_dist, const ebg_query_opts* opts)
{
  if (!cm) return fail(EBG_INVALID_ARG, "mesh is null");
  ebgpu::GridSpec g{};
  uint64_t        n  = 0;
  const ebg_status gs = ebgpu::grid_spec(origin, spacing, dims, g, n);
  if (gs != EBG_OK) return gs;
  if (n == 0) return EBG_OK;
  if (!d_dist) return fail(EBG_INVALID_ARG, "null output array");
  ebg_mesh*      m = const_cast<ebg_mesh*>(cm);
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t                           s = static_cast<hipStream_t>(o.stream);
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  EBG_HIP(m->pts32.reserve(12 * n));
  EBG_HIP(ebgpu::launch_grid_points(g, m->pts32.as<float>(), s));
  return ebg_query_sdf(m, m->pts32.as<float>(), n, d_dist, nullptr, &o);
}

ebg_status ebg_query_grid_host(const ebg_mesh* m, const double* origin, const double* spacing, const uint64_t* dims,
                               float* dist, const ebg_query_opts* opts)
{
  if (!m) return fail(EBG_INVALID_ARG, "mesh is null");
  ebgpu::GridSpec g{};
  uint64_t        n  = 0;
  const ebg_status gs = ebgpu::grid_spec(origin, spacing, dims, g, n);
  if (gs != EBG_OK) return gs;
  if (n == 0) return EBG_OK;
  if (!dist) return fail(EBG_INVALID_ARG, "null output array");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  DevBuf      d;
  EBG_HIP(d.reserve(4 * n));
  const ebg_status st = ebg_query_grid(m, origin, spacing, dims, d.as<float>(), &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(dist, d.p, 4 * n, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_mesh_check(const ebg_mesh* m)
{
  if (!m) return fail(EBG_INVALID_ARG, "mesh is null");
  uint32_t w = 0;
  EBG_HIP(hipDeviceSynchronize());
  EBG_HIP(hipMemcpy(&w, m->status.p, 4, hipMemcpyDeviceToHost));
  if (w & ebgpu::kStatusStackOverflow)
    return fail(EBG_STACK_OVERFLOW, "a traversal of this mesh needed more stack than its build-time worst case");
  return EBG_OK;
}

int64_t ebg_mesh_last_split_packets(const ebg_mesh* cm)
{
  ebg_mesh* m = const_cast<ebg_mesh*>(cm);
  if (!m) return -1;
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  if (!m->ws.cap || !m->last_split) return 0;
  uint32_t c = 0;
  if (hipDeviceSynchronize() != hipSuccess) return -1;
  if (hipMemcpy(&c, m->ws.counters() + 2, 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return std::min<int64_t>(c, m->last_split_cap);
}

double ebg_mesh_last_traversal_ms(const ebg_mesh* cm)
{
  ebg_mesh* m = const_cast<ebg_mesh*>(cm);
  if (!m || !m->timed) return -1.0;
  float ms = -1.0f;
  if (hipEventSynchronize(m->ev1) != hipSuccess || hipEventElapsedTime(&ms, m->ev0, m->ev1) != hipSuccess) {
    (void)hipGetLastError(); // keep the failure out of the caller's next HIP call
    return -1.0;
  }
  return ms;
}

ebg_status ebg_query_sdf_host(const ebg_mesh* m, const float* pts, uint64_t n, float* dist, uint32_t* tri,
                              const ebg_query_opts* opts)
{
  if (!m) return fail(EBG_INVALID_ARG, "mesh is null");
  if (n == 0) return EBG_OK;
  if (!pts || !dist) return fail(EBG_INVALID_ARG, "null point or output array");
  DevBuf p, d, t;
  EBG_HIP(p.reserve(12 * n));
  EBG_HIP(d.reserve(4 * n));
  if (tri) EBG_HIP(t.reserve(4 * n));
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  EBG_HIP(hipMemcpyAsync(p.p, pts, 12 * n, hipMemcpyHostToDevice, s));
  const ebg_status st = ebg_query_sdf(m, p.as<float>(), n, d.as<float>(), tri ? t.as<uint32_t>() : nullptr, &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(dist, d.p, 4 * n, hipMemcpyDeviceToHost, s));
  if (tri) EBG_HIP(hipMemcpyAsync(tri, t.p, 4 * n, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

// ------------------------------------------------------------------------------------------
// SFC
// ------------------------------------------------------------------------------------------

ebg_status ebg_morton_codes_f32(const float* d_pts, uint64_t n, uint64_t* d_codes, void* stream)
{
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_codes) return fail(EBG_INVALID_ARG, "null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevBuf      box;
  EBG_HIP(box.reserve(sizeof(float) * 6 * (ebgpu::kBboxParts + 2)));
  EBG_HIP(ebgpu::launch_bbox_f32(d_pts, n, box.as<float>(), box.as<float>() + 8, s));
  EBG_HIP(ebgpu::launch_codes_f32(d_pts, n, box.as<float>(), d_codes, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_morton_codes_f64(const double* d_pts, uint64_t n, uint64_t* d_codes, void* stream)
{
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_codes) return fail(EBG_INVALID_ARG, "null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevBuf      box;
  EBG_HIP(box.reserve(sizeof(double) * 6 * (ebgpu::kBboxParts + 2)));
  EBG_HIP(ebgpu::launch_bbox_f64(d_pts, n, box.as<double>(), box.as<double>() + 8, s));
  EBG_HIP(ebgpu::launch_codes_f64(d_pts, n, box.as<double>(), d_codes, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_morton_sort(const uint64_t* d_codes, uint64_t n, uint32_t* d_perm, uint64_t* d_sorted, void* stream)
{
  if (n == 0) return EBG_OK;
  if (!d_codes || !d_perm) return fail(EBG_INVALID_ARG, "null pointer");
  if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 keys");
  hipStream_t s = static_cast<hipStream_t>(stream);
  DevBuf      scratch, kout;
  EBG_HIP(scratch.reserve(ebgpu::radix_sort_scratch_bytes(n)));
  uint64_t* ko = d_sorted;
  if (!ko) {
    EBG_HIP(kout.reserve(8 * n));
    ko = kout.as<uint64_t>();
  }
  EBG_HIP(ebgpu::radix_sort_pairs(d_codes, nullptr, ko, d_perm, n, 63, scratch.p, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

} // extern "C"

// ------------------------------------------------------------------------------------------
// learned bound (training lives in mlp_train.hip)
// ------------------------------------------------------------------------------------------
extern "C" {

void ebg_mlp_opts_default(ebg_mlp_opts* o)
{
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->num_samples = 4000000;
  o->epochs      = 8;
  o->batch       = 4096;
  o->lr          = 2e-3f;
  o->seed        = 12345;
  o->near_sigma  = 0.02f;
  o->box_scale   = 1.2f;
  o->under_weight = 16.0f;
}

uint64_t ebg_mlp_param_count(void) { return uint64_t(ebgpu::kMlpParams); }

ebg_status ebg_mlp_check_schedule(uint64_t num_samples, int32_t batch)
{
  if (num_samples < 1 || batch < 32 || batch > 65536) return fail(EBG_INVALID_ARG, "num_samples/batch out of range");
  std::string why;
  if (!ebgpu::mlp_check_train_schedule(num_samples, batch, &why)) return fail(EBG_INVALID_ARG, why);
  return EBG_OK;
}

// Input normalisation of the network: the box_scale-scaled mesh bbox maps to [-1, 1]^3.
static void mlp_set_box(ebg_mesh* m, float box_scale)
{
  ebgpu::Mlp& p = m->mlp;
  for (int c = 0; c < 3; c++) {
    const double ctr = 0.5 * (m->lo[c] + m->hi[c]);
    double       h   = 0.5 * box_scale * (m->hi[c] - m->lo[c]);
    if (!(h > 0.0)) h = 1.0;
    p.lo[c]    = float(ctr - h);
    p.scale[c] = float(1.0 / h);
  }
  p.diag      = float(m->diag);
  p.box_scale = box_scale;
}

ebg_status ebg_mlp_train(ebg_mesh* m, const ebg_mlp_opts* opts, float* loss)
{
  if (!m) return fail(EBG_INVALID_ARG, "mesh is null");
  ebg_mlp_opts o;
  ebg_mlp_opts_default(&o);
  if (opts) o = *opts;
  if (o.num_samples < 64 || o.num_samples > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "num_samples out of range");
  if (o.epochs < 1 || o.batch < 32 || o.batch > 65536) return fail(EBG_INVALID_ARG, "epochs/batch out of range");
  if (!(o.lr > 0.0f) || !(o.box_scale >= 1.0f) || !(o.near_sigma >= 0.0f) || !(o.under_weight >= 1.0f))
    return fail(EBG_INVALID_ARG, "lr, box_scale, near_sigma or under_weight out of range");
  {
    std::string why;
    if (!ebgpu::mlp_check_train_schedule(o.num_samples, o.batch, &why)) return fail(EBG_INVALID_ARG, why);
  }
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  ebgpu::Mlp& mlp = m->mlp;
  EBG_HIP(ebgpu::mlp_alloc(mlp));
  mlp_set_box(m, o.box_scale);
  EBG_HIP(ebgpu::mlp_init(mlp, o.seed));
  mlp.trained      = false;
  mlp.under_weight = o.under_weight;

  // (point, |d|) pairs from the mesh's own BVH
  const uint64_t n = o.num_samples;
  DevBuf         pts, dist;
  EBG_HIP(pts.reserve(12 * n));
  EBG_HIP(dist.reserve(4 * n));
  const float lo[3] = {float(m->lo[0]), float(m->lo[1]), float(m->lo[2])};
  const float hi[3] = {float(m->hi[0]), float(m->hi[1]), float(m->hi[2])};
  EBG_HIP(ebgpu::mlp_samples(m->dev, m->area_cdf.as<float>(), uint32_t(m->nt), lo, hi, o.box_scale, o.near_sigma,
                             o.seed, n, pts.as<float>(), s));
  ebg_query_opts q;
  ebg_query_opts_default(&q);
  q.stream           = o.stream;
  const ebg_status st = ebg_query_sdf(m, pts.as<float>(), n, dist.as<float>(), nullptr, &q);
  if (st != EBG_OK) return st;

  float       last = 0.0f;
  std::string why;
  const hipError_t fe = ebgpu::mlp_fit(mlp, m->train, pts.as<float>(), dist.as<float>(), n, o.epochs, o.batch, o.lr, s,
                                       &last, &why);
  if (fe != hipSuccess)
    return fail(fe == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, why.empty() ? hipGetErrorString(fe) : why);
  mlp.trained = true;
  if (loss) *loss = last;
  return EBG_OK;
}

ebg_status ebg_mlp_predict(const ebg_mesh* cm, const float* d_pts, uint64_t n, float* d_pred, void* stream)
{
  if (!cm) return fail(EBG_INVALID_ARG, "mesh is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_pred) return fail(EBG_INVALID_ARG, "null pointer");
  if (!cm->mlp.master) return fail(EBG_NOT_READY, "no network");
  EBG_HIP(ebgpu::mlp_forward(cm->mlp, d_pts, nullptr, n, false, 1.0f, 0.0f, d_pred, static_cast<hipStream_t>(stream)));
  return EBG_OK;
}

ebg_status ebg_mlp_get_params(const ebg_mesh* m, float* params)
{
  if (!m || !params) return fail(EBG_INVALID_ARG, "null pointer");
  if (!m->mlp.master) return fail(EBG_NOT_READY, "no network");
  EBG_HIP(ebgpu::mlp_get_params_host(m->mlp, params));
  return EBG_OK;
}

ebg_status ebg_mlp_set_params(ebg_mesh* m, const float* params)
{
  if (!m || !params) return fail(EBG_INVALID_ARG, "null pointer");
  for (int i = 0; i < ebgpu::kMlpParams; i++)
    if (!std::isfinite(params[i])) return fail(EBG_INVALID_ARG, "non-finite parameter");
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  mlp_set_box(m, m->mlp.box_scale > 0.0f ? m->mlp.box_scale : 1.2f);
  EBG_HIP(ebgpu::mlp_set_params_host(m->mlp, params));
  m->mlp.trained = true;
  return EBG_OK;
}

// Network file: magic, format version, the mesh hash it was trained for, the input box and
// loss settings, then the kMlpParams parameters in the public order; all little-endian.
namespace {
constexpr char     kMlpMagic[8] = {'E', 'B', 'G', 'M', 'L', 'P', '\0', '\0'};
constexpr uint32_t kMlpVersion  = 1;
struct MlpFileHeader
{
  char     magic[8];
  uint32_t version, params;
  uint64_t mesh_hash;
  float    lo[3], scale[3], diag, box_scale, under_weight, reserved;
};
static_assert(sizeof(MlpFileHeader) == 64, "MlpFileHeader layout");
} // namespace

ebg_status ebg_mlp_save(const ebg_mesh* m, const char* path)
{
  if (!m || !path) return fail(EBG_INVALID_ARG, "null argument");
  if (!m->mlp.master || !m->mlp.trained) return fail(EBG_NOT_READY, "no trained network");
  MlpFileHeader h{};
  std::memcpy(h.magic, kMlpMagic, 8);
  h.version   = kMlpVersion;
  h.params    = uint32_t(ebgpu::kMlpParams);
  h.mesh_hash = m->hash;
  for (int c = 0; c < 3; c++) {
    h.lo[c]    = m->mlp.lo[c];
    h.scale[c] = m->mlp.scale[c];
  }
  h.diag         = m->mlp.diag;
  h.box_scale    = m->mlp.box_scale;
  h.under_weight = m->mlp.under_weight;
  std::vector<float> p(ebgpu::kMlpParams);
  EBG_HIP(ebgpu::mlp_get_params_host(m->mlp, p.data()));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(EBG_INVALID_ARG, std::string("cannot open ") + path + " for writing");
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(p.data(), sizeof(float), p.size(), f) == p.size();
  if (std::fclose(f) != 0 || !ok) return fail(EBG_INVALID_ARG, std::string("write to ") + path + " failed");
  return EBG_OK;
}

ebg_status ebg_mlp_load(ebg_mesh* m, const char* path)
{
  if (!m || !path) return fail(EBG_INVALID_ARG, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(EBG_INVALID_ARG, std::string("cannot open ") + path);
  MlpFileHeader      h{};
  std::vector<float> p(ebgpu::kMlpParams);
  const bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
  std::fclose(f);
  if (!ok || std::memcmp(h.magic, kMlpMagic, 8) != 0 || h.version != kMlpVersion || h.params != uint32_t(ebgpu::kMlpParams))
    return fail(EBG_INVALID_ARG, std::string(path) + " is not a network file of this version");
  if (h.mesh_hash != m->hash) return fail(EBG_INVALID_ARG, "the network in " + std::string(path) + " was trained for another mesh");
  for (float x : p)
    if (!std::isfinite(x)) return fail(EBG_INVALID_ARG, "non-finite parameter in " + std::string(path));
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  EBG_HIP(ebgpu::mlp_set_params_host(m->mlp, p.data()));
  for (int c = 0; c < 3; c++) {
    m->mlp.lo[c]    = h.lo[c];
    m->mlp.scale[c] = h.scale[c];
  }
  m->mlp.diag         = h.diag;
  m->mlp.box_scale    = h.box_scale;
  m->mlp.under_weight = h.under_weight;
  m->mlp.trained      = true;
  return EBG_OK;
}

ebg_status ebg_mlp_loss_grad(ebg_mesh* m, const float* d_pts, const float* d_target, uint64_t n, float* d_grad,
                             float* d_loss, void* stream)
{
  if (!m || !d_pts || !d_target || !d_grad) return fail(EBG_INVALID_ARG, "null pointer");
  if (n < 1 || n > 65536) return fail(EBG_INVALID_ARG, "1 <= n <= 65536");
  if (!m->mlp.master) return fail(EBG_NOT_READY, "no network");
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  EBG_HIP(m->train.ensure(int(n)));
  EBG_HIP(hipMemsetAsync(m->train.loss, 0, sizeof(float), s));
  {
    std::string      why;
    const hipError_t le = ebgpu::mlp_loss_grad(m->mlp, m->train, d_pts, d_target, int(n), s, &why);
    if (le != hipSuccess) return fail(le == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, why.empty() ? hipGetErrorString(le) : why);
  }
  EBG_HIP(ebgpu::mlp_grad_public(m->mlp, d_grad, s));
  if (d_loss) EBG_HIP(hipMemcpyAsync(d_loss, m->train.loss, sizeof(float), hipMemcpyDeviceToDevice, s));
  return EBG_OK;
}

} // extern "C"
