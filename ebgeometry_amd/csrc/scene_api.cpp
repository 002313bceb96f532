// C-ABI for CSG union scenes: build (host SAH over primitive boxes, 4-wide), upload, query.
//
// Query pipeline (caller's stream, no host synchronisation): Morton presort of the queries ->
// FP32 packet traversal of the scene tree with nested mesh traversals -> FP64 pass over the
// queries with |v| < recheck_eps * scene diagonal.
#include "host_common.hpp"
#include "mlp.hpp"
#include "mesh_build.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <string>
#include <vector>

using ebgpu::DevBuf;
using ebgpu::fail;
using ebgpu::upload;
using ebgpu::Workspace;

struct ebg_scene
{
  std::vector<ebg_mesh*> meshes; // retained
  DevBuf                 nodes, prims, prims64, meshdev, status;
  ebgpu::SceneDev        dev{};
  int                    KM = 4, banks = 1, stack_need = 0;
  std::size_t            num_prims = 0, num_nodes = 0;
  double                 lo[3]{}, hi[3]{}, diag = 0;
  std::mutex             mu;
  Workspace              ws;
  std::mutex             grid_mu; // held around the lattice buffer (ebg_scene_query_grid)
  DevBuf                 gridpts;
  hipEvent_t             ev0 = nullptr, ev1 = nullptr;
  bool                   timed = false;
  ebgpu::Mlp             mlp;   // learned cull bound of the union (ebg_scene_mlp_train)
  ebgpu::MlpTrainBuffers train;

  ~ebg_scene()
  {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (ebg_mesh* m : meshes) ebgpu::mesh_release(m);
  }
};

namespace {

float round_down(double x)
{
  float f = float(x);
  if (double(f) > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}
float round_up(double x)
{
  float f = float(x);
  if (double(f) < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}

bool finite_all(const double* p, int n)
{
  for (int i = 0; i < n; i++)
    if (!std::isfinite(p[i])) return false;
  return true;
}

} // namespace

extern "C" {

ebg_status ebg_scene_create(const ebg_prim* prims, uint64_t n, ebg_mesh* const* meshes, uint64_t num_meshes,
                            const ebg_build_opts* opts, ebg_scene** out)
{
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (!prims || n == 0) return fail(EBG_INVALID_ARG, "a scene needs at least one primitive");
  if (n > 0x7FFFFFFFull) return fail(EBG_INVALID_ARG, "too many primitives");
  if (num_meshes && !meshes) return fail(EBG_INVALID_ARG, "meshes is null");
  ebg_build_opts o{};
  if (opts) o = *opts;
  if (o.builder != EBG_BUILD_SAH) return fail(EBG_UNSUPPORTED, "scenes are built with the SAH builder");
  const int leaf = o.leaf_size ? o.leaf_size : 4;
  if (leaf < 1 || leaf > ebgpu::kMaxLeaf) return fail(EBG_INVALID_ARG, "leaf_size must be 1..16");

  int KM = 0, mesh_banks = 1;
  for (uint64_t j = 0; j < num_meshes; j++) {
    if (!meshes[j]) return fail(EBG_INVALID_ARG, "null mesh handle");
    if (KM && meshes[j]->K != KM) return fail(EBG_INVALID_ARG, "all scene meshes must have the same branching");
    KM         = meshes[j]->K;
    mesh_banks = std::max(mesh_banks, meshes[j]->banks);
  }
  if (!KM) KM = 4;

  // primitive records and bounding volumes
  std::vector<ebgpu::PrimRec>   rec(n);
  std::vector<ebgpu::PrimRec64> rec64(n);
  std::vector<float>            blo(3 * n), bhi(3 * n);
  for (uint64_t i = 0; i < n; i++) {
    const ebg_prim& p = prims[i];
    ebgpu::PrimRec   r{};
    ebgpu::PrimRec64 r64{};
    double           lo[3], hi[3];
    const std::string where = "primitive " + std::to_string(i) + ": ";
    if (!finite_all(p.p, 6) || !finite_all(p.translate, 3) || !finite_all(p.matrix, 9) ||
        !std::isfinite(p.scale))
      return fail(EBG_INVALID_ARG, where + "non-finite parameter");
    r.type = uint32_t(p.type);
    r.mesh = 0;
    if (p.type == EBG_PRIM_SPHERE) {
      if (!(p.p[3] >= 0.0)) return fail(EBG_INVALID_ARG, where + "negative sphere radius");
      for (int c = 0; c < 4; c++) {
        r.a[c]   = float(p.p[c]);
        r64.a[c] = p.p[c];
      }
      for (int c = 0; c < 3; c++) {
        lo[c] = p.p[c] - p.p[3];
        hi[c] = p.p[c] + p.p[3];
      }
    }
    else if (p.type == EBG_PRIM_BOX) {
      for (int c = 0; c < 3; c++)
        if (!(p.p[c] <= p.p[3 + c])) return fail(EBG_INVALID_ARG, where + "box low corner above high corner");
      for (int c = 0; c < 6; c++) {
        r.a[c]   = float(p.p[c]);
        r64.a[c] = p.p[c];
      }
      for (int c = 0; c < 3; c++) {
        lo[c] = p.p[c];
        hi[c] = p.p[3 + c];
      }
    }
    else if (p.type == EBG_PRIM_MESH) {
      if (p.mesh < 0 || uint64_t(p.mesh) >= num_meshes) return fail(EBG_INVALID_ARG, where + "mesh index out of range");
      if (!(p.scale > 0.0)) return fail(EBG_INVALID_ARG, where + "mesh scale must be > 0");
      double M[9];
      bool   zero = true;
      for (int c = 0; c < 9; c++) zero = zero && p.matrix[c] == 0.0;
      for (int c = 0; c < 9; c++) M[c] = zero ? ((c % 4 == 0) ? 1.0 : 0.0) : p.matrix[c];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
          const double g = M[i] * M[j] + M[3 + i] * M[3 + j] + M[6 + i] * M[6 + j];
          if (std::fabs(g - (i == j ? 1.0 : 0.0)) > 1e-9) return fail(EBG_INVALID_ARG, where + "matrix is not orthonormal");
        }
      r.mesh = uint32_t(p.mesh);
      for (int c = 0; c < 9; c++) {
        r.A[c]   = float(M[c] / p.scale);
        r64.A[c] = M[c];
      }
      for (int c = 0; c < 3; c++) {
        r.t[c]   = float(p.translate[c]);
        r64.t[c] = p.translate[c];
      }
      r.s   = float(p.scale);
      r64.s = p.scale;
      // world = t + s M^T x over the mesh vertices
      const std::vector<float>& x = meshes[p.mesh]->host_xyz;
      for (int c = 0; c < 3; c++) {
        lo[c] = std::numeric_limits<double>::infinity();
        hi[c] = -lo[c];
      }
      for (std::size_t v = 0; v + 2 < x.size(); v += 3) {
        for (int c = 0; c < 3; c++) {
          const double w = p.translate[c] + p.scale * (M[c] * x[v] + M[3 + c] * x[v + 1] + M[6 + c] * x[v + 2]);
          lo[c]          = std::min(lo[c], w);
          hi[c]          = std::max(hi[c], w);
        }
      }
      const double pad = 1e-6 * (std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                           (hi[2] - lo[2]) * (hi[2] - lo[2])) + 1.0);
      for (int c = 0; c < 3; c++) {
        lo[c] -= pad;
        hi[c] += pad;
      }
    }
    else {
      return fail(EBG_INVALID_ARG, where + "unknown primitive type");
    }
    if (p.bv_lo[0] <= p.bv_hi[0]) { // caller-supplied bounding volume (as BVHUnionIF takes it)
      if (!finite_all(p.bv_lo, 3) || !finite_all(p.bv_hi, 3)) return fail(EBG_INVALID_ARG, where + "non-finite bv");
      for (int c = 0; c < 3; c++) {
        if (!(p.bv_lo[c] <= p.bv_hi[c])) return fail(EBG_INVALID_ARG, where + "empty bounding volume");
        lo[c] = p.bv_lo[c];
        hi[c] = p.bv_hi[c];
      }
    }
    for (int c = 0; c < 3; c++) {
      r.bv[c]          = round_down(lo[c]);
      r.bv[3 + c]      = round_up(hi[c]);
      blo[3 * i + c]   = r.bv[c];
      bhi[3 * i + c]   = r.bv[3 + c];
    }
    rec[i]   = r;
    rec64[i] = r64;
  }

  ebgpu::WideBvh wb;
  std::string    err;
  if (!ebgpu::build_wide_bvh(blo.data(), bhi.data(), n, 4, leaf, wb, err)) return fail(EBG_INVALID_ARG, err);
  std::vector<ebgpu::PrimRec>   rec_o(n);
  std::vector<ebgpu::PrimRec64> rec64_o(n);
  for (uint64_t j = 0; j < n; j++) {
    rec_o[j]   = rec[wb.order[j]];
    rec64_o[j] = rec64[wb.order[j]];
  }
  const int scene_banks = (wb.stack_need + 63) / 64;
  int       banks       = std::max(scene_banks, mesh_banks);
  if (banks > 4)
    return fail(EBG_STACK_OVERFLOW, "scene tree needs a traversal stack of " + std::to_string(wb.stack_need) +
                                      " entries (max 256)");
  if (banks == 3) banks = 4;

  auto* sc       = new ebg_scene;
  sc->KM         = KM;
  sc->banks      = banks;
  sc->stack_need = wb.stack_need;
  sc->num_prims  = n;
  sc->num_nodes  = wb.num_nodes;
  for (uint64_t j = 0; j < num_meshes; j++) {
    ebgpu::mesh_retain(meshes[j]);
    sc->meshes.push_back(meshes[j]);
  }
  std::vector<ebgpu::MeshDev> md(num_meshes);
  for (uint64_t j = 0; j < num_meshes; j++) md[j] = meshes[j]->dev;
  hipError_t e;
  if ((e = upload(sc->nodes, wb.nodes)) != hipSuccess || (e = upload(sc->prims, rec_o)) != hipSuccess ||
      (e = upload(sc->prims64, rec64_o)) != hipSuccess || (e = upload(sc->meshdev, md)) != hipSuccess ||
      (e = upload(sc->status, std::vector<uint32_t>(16, 0u))) != hipSuccess) {
    delete sc;
    return fail(e == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, std::string("scene upload: ") + hipGetErrorString(e));
  }
  for (int c = 0; c < 3; c++) {
    sc->lo[c]          = wb.lo[c];
    sc->hi[c]          = wb.hi[c];
    sc->dev.root_lo[c] = wb.lo[c];
    sc->dev.root_hi[c] = wb.hi[c];
  }
  sc->diag        = wb.diag;
  sc->dev.diag    = float(wb.diag);
  sc->dev.nodes   = sc->nodes.p;
  sc->dev.prims   = sc->prims.as<ebgpu::PrimRec>();
  sc->dev.prims64 = sc->prims64.as<ebgpu::PrimRec64>();
  sc->dev.meshes  = sc->meshdev.as<ebgpu::MeshDev>();
  sc->dev.status  = sc->status.as<uint32_t>();
  *out            = sc;
  return EBG_OK;
}

ebg_status ebg_scene_destroy(ebg_scene* scene)
{
  delete scene;
  return EBG_OK;
}

ebg_status ebg_scene_bounds(const ebg_scene* sc, double* lo3, double* hi3, double* diag)
{
  if (!sc) return fail(EBG_INVALID_ARG, "scene is null");
  for (int c = 0; c < 3; c++) {
    if (lo3) lo3[c] = sc->lo[c];
    if (hi3) hi3[c] = sc->hi[c];
  }
  if (diag) *diag = sc->diag;
  return EBG_OK;
}

ebg_status ebg_scene_query(const ebg_scene* csc, const float* d_pts, uint64_t n, float* d_value,
                           const ebg_query_opts* opts)
{
  if (!csc) return fail(EBG_INVALID_ARG, "scene is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_value) return fail(EBG_INVALID_ARG, "null point or output array");
  if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 queries per call");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  ebg_scene*  sc = const_cast<ebg_scene*>(csc);
  hipStream_t s  = static_cast<hipStream_t>(o.stream);
  if (o.learned_bound && !sc->mlp.trained) return fail(EBG_NOT_READY, "learned bound requested but no network trained");

  std::lock_guard<std::mutex> lock(sc->mu);
  Workspace&                  ws = sc->ws;
  EBG_HIP(ws.ensure(n));
  if (!sc->ev0) {
    EBG_HIP(hipEventCreate(&sc->ev0));
    EBG_HIP(hipEventCreate(&sc->ev1));
  }
  uint32_t* cnt = ws.counters();
  EBG_HIP(ebgpu::launch_fill_u32(cnt, 2, 0u, s));
  const uint32_t* idx = nullptr;
  if (o.presort && n > 1) {
    EBG_HIP(ebgpu::presort(ws.view(), d_pts, n, s));
    idx = ws.perm.as<uint32_t>();
  }
  // seeded walks: the union network's radii (the query order; it needs no presort), or the caller's
  const float* seeds = o.seed_radius;
  if (o.learned_bound) {
    EBG_HIP(ebgpu::mlp_forward(sc->mlp, d_pts, nullptr, n, true, o.bound_scale, float(o.bound_offset * sc->diag),
                               ws.view().seeds, s));
    seeds = ws.view().seeds;
  }
  const bool       recheck = o.fp64_recheck || o.fp64_all;
  ebgpu::QueryArgs a{};
  a.pts            = d_pts;
  a.idx            = idx;
  a.n              = n;
  a.seed_r         = seeds;
  a.fallback_count = cnt + 1;
  a.out_d          = d_value;
  a.recheck_abs   = o.fp64_all ? INFINITY : (recheck ? float(o.recheck_eps * sc->diag) : 0.0f);
  a.recheck_list  = recheck ? ws.recheck_list() : nullptr;
  a.recheck_count = cnt;
  a.work          = reinterpret_cast<unsigned long long*>(o.work);
  const bool count = o.work != nullptr;

  EBG_HIP(hipEventRecord(sc->ev0, s));
  EBG_HIP(ebgpu::launch_scene32(sc->dev, sc->KM, sc->banks, a, count, s));
  EBG_HIP(hipEventRecord(sc->ev1, s));
  sc->timed = true;
  if (recheck) {
    ebgpu::QueryArgs r = a;
    r.idx              = ws.recheck_list();
    r.count_ptr        = cnt;
    r.recheck_list     = nullptr;
    r.work             = nullptr;
    // the FP32 value is within a few ulp of the coordinates of the exact one
    r.slack_abs = float(1e-5 * sc->diag);
    r.slack_rel = 1e-4f;
    EBG_HIP(ebgpu::launch_scene64(sc->dev, sc->KM, sc->banks, r, s));
  }
  if (count) EBG_HIP(ebgpu::launch_add_counts(a.work, recheck ? cnt : nullptr, seeds ? cnt + 1 : nullptr, s));
  return EBG_OK;
}

ebg_status ebg_scene_mlp_train(ebg_scene* sc, const ebg_mlp_opts* opts, float* loss)
{
  if (!sc) return fail(EBG_INVALID_ARG, "scene is null");
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
  ebgpu::Mlp& mlp = sc->mlp;
  EBG_HIP(ebgpu::mlp_alloc(mlp));
  for (int c = 0; c < 3; c++) {
    const double ctr = 0.5 * (sc->lo[c] + sc->hi[c]);
    double       h   = 0.5 * o.box_scale * (sc->hi[c] - sc->lo[c]);
    if (!(h > 0.0)) h = 1.0;
    mlp.lo[c]    = float(ctr - h);
    mlp.scale[c] = float(1.0 / h);
  }
  mlp.diag      = float(sc->diag);
  mlp.box_scale = o.box_scale;
  EBG_HIP(ebgpu::mlp_init(mlp, o.seed));
  mlp.trained      = false;
  mlp.under_weight = o.under_weight;

  // (point, |value|) pairs from the scene's own BVH walk
  const uint64_t n = o.num_samples;
  DevBuf         pts, dist;
  EBG_HIP(pts.reserve(12 * n));
  EBG_HIP(dist.reserve(4 * n));
  const float lo[3] = {float(sc->lo[0]), float(sc->lo[1]), float(sc->lo[2])};
  const float hi[3] = {float(sc->hi[0]), float(sc->hi[1]), float(sc->hi[2])};
  EBG_HIP(ebgpu::mlp_scene_samples(sc->prims.as<ebgpu::PrimRec>(), uint32_t(sc->num_prims), lo, hi, o.box_scale,
                                   float(o.near_sigma * sc->diag), o.seed, n, pts.as<float>(), s));
  ebg_query_opts q;
  ebg_query_opts_default(&q);
  q.stream            = o.stream;
  const ebg_status st = ebg_scene_query(sc, pts.as<float>(), n, dist.as<float>(), &q);
  if (st != EBG_OK) return st;
  float       last = 0.0f;
  std::string why;
  const hipError_t fe = ebgpu::mlp_fit(mlp, sc->train, pts.as<float>(), dist.as<float>(), n, o.epochs, o.batch, o.lr,
                                       s, &last, &why);
  if (fe != hipSuccess)
    return fail(fe == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, why.empty() ? hipGetErrorString(fe) : why);
  mlp.trained = true;
  if (loss) *loss = last;
  return EBG_OK;
}

ebg_status ebg_scene_mlp_predict(const ebg_scene* sc, const float* d_pts, uint64_t n, float* d_pred, void* stream)
{
  if (!sc) return fail(EBG_INVALID_ARG, "scene is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_pred) return fail(EBG_INVALID_ARG, "null pointer");
  if (!sc->mlp.trained) return fail(EBG_NOT_READY, "no network trained");
  EBG_HIP(ebgpu::mlp_forward(sc->mlp, d_pts, nullptr, n, false, 1.0f, 0.0f, d_pred, static_cast<hipStream_t>(stream)));
  return EBG_OK;
}

ebg_status ebg_scene_query_grid(const ebg_scene* csc, const double* origin, const double* spacing,
                                const uint64_t* dims, float* d_value, const ebg_query_opts* opts)
{
  if (!csc) return fail(EBG_INVALID_ARG, "scene is null");
  ebgpu::GridSpec g{};
  uint64_t        n  = 0;
  const ebg_status gs = ebgpu::grid_spec(origin, spacing, dims, g, n);
  if (gs != EBG_OK) return gs;
  if (n == 0) return EBG_OK;
  if (!d_value) return fail(EBG_INVALID_ARG, "null output array");
  ebg_scene*     sc = const_cast<ebg_scene*>(csc);
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t                 s = static_cast<hipStream_t>(o.stream);
  std::lock_guard<std::mutex> lock(sc->grid_mu);
  EBG_HIP(sc->gridpts.reserve(12 * n));
  EBG_HIP(ebgpu::launch_grid_points(g, sc->gridpts.as<float>(), s));
  return ebg_scene_query(sc, sc->gridpts.as<float>(), n, d_value, &o);
}

ebg_status ebg_scene_query_grid_host(const ebg_scene* sc, const double* origin, const double* spacing,
                                     const uint64_t* dims, float* value, const ebg_query_opts* opts)
{
  if (!sc) return fail(EBG_INVALID_ARG, "scene is null");
  ebgpu::GridSpec g{};
  uint64_t        n  = 0;
  const ebg_status gs = ebgpu::grid_spec(origin, spacing, dims, g, n);
  if (gs != EBG_OK) return gs;
  if (n == 0) return EBG_OK;
  if (!value) return fail(EBG_INVALID_ARG, "null output array");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  DevBuf      d;
  EBG_HIP(d.reserve(4 * n));
  const ebg_status st = ebg_scene_query_grid(sc, origin, spacing, dims, d.as<float>(), &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(value, d.p, 4 * n, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_scene_query_host(const ebg_scene* sc, const float* pts, uint64_t n, float* value,
                                const ebg_query_opts* opts)
{
  if (!sc) return fail(EBG_INVALID_ARG, "scene is null");
  if (n == 0) return EBG_OK;
  if (!pts || !value) return fail(EBG_INVALID_ARG, "null point or output array");
  DevBuf p, d;
  EBG_HIP(p.reserve(12 * n));
  EBG_HIP(d.reserve(4 * n));
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  EBG_HIP(hipMemcpyAsync(p.p, pts, 12 * n, hipMemcpyHostToDevice, s));
  const ebg_status st = ebg_scene_query(sc, p.as<float>(), n, d.as<float>(), &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(value, d.p, 4 * n, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_scene_check(const ebg_scene* sc)
{
  if (!sc) return fail(EBG_INVALID_ARG, "scene is null");
  uint32_t w = 0;
  EBG_HIP(hipDeviceSynchronize());
  EBG_HIP(hipMemcpy(&w, sc->status.p, 4, hipMemcpyDeviceToHost));
  if (w & ebgpu::kStatusStackOverflow)
    return fail(EBG_STACK_OVERFLOW, "a walk of this scene's tree needed more stack than its build-time worst case");
  return EBG_OK;
}

double ebg_scene_last_traversal_ms(const ebg_scene* csc)
{
  ebg_scene* sc = const_cast<ebg_scene*>(csc);
  if (!sc || !sc->timed) return -1.0;
  float ms = -1.0f;
  if (hipEventSynchronize(sc->ev1) != hipSuccess || hipEventElapsedTime(&ms, sc->ev0, sc->ev1) != hipSuccess) {
    (void)hipGetLastError(); // keep the failure out of the caller's next HIP call
    return -1.0;
  }
  return ms;
}

} // extern "C"
