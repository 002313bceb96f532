// C-ABI for point clouds: build (host SAH over the points, K-wide, leaves of at most 16 points),
// upload, k-nearest-neighbour queries (external points, presorted along the Hilbert curve; or the
// cloud's own points, which leaf order already keeps coherent).
#include "host_common.hpp"
#include "mesh_build.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using ebgpu::DevBuf;
using ebgpu::fail;
using ebgpu::upload;
using ebgpu::Workspace;

struct ebg_cloud
{
  DevBuf          nodes, pts, status;
  ebgpu::CloudDev dev{};
  int             K = 4, banks = 1;
  uint64_t        n = 0;
  std::mutex      mu;
  Workspace       ws;
};

extern "C" {

ebg_status ebg_cloud_create(const float* xyz, uint64_t n, const ebg_build_opts* opts, ebg_cloud** out)
{
  using namespace ebgpu;
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (!xyz || n == 0) return fail(EBG_INVALID_ARG, "empty point cloud");
  if (n > 0x7FFFFFFFull) return fail(EBG_INVALID_ARG, "point cloud too large");
  for (uint64_t i = 0; i < 3 * n; i++)
    if (!std::isfinite(xyz[i])) return fail(EBG_INVALID_ARG, "non-finite point coordinate");
  ebg_build_opts o{};
  if (opts) o = *opts;
  if (o.builder != EBG_BUILD_SAH) return fail(EBG_UNSUPPORTED, "point clouds are built with the SAH builder");
  const int K    = o.branching ? o.branching : 4;
  const int leaf = o.leaf_size ? o.leaf_size : 8;
  if (K != 4 && K != 8) return fail(EBG_INVALID_ARG, "branching must be 4 or 8");
  if (leaf < 1 || leaf > kMaxLeaf) return fail(EBG_INVALID_ARG, "leaf_size must be 1..16");
  WideBvh     wb;
  std::string err;
  if (!build_wide_bvh(xyz, xyz, n, K, leaf, wb, err)) return fail(EBG_INVALID_ARG, err);
  const int stack_banks = (wb.stack_need + 63) / 64;
  if (stack_banks > 4)
    return fail(EBG_STACK_OVERFLOW, "tree needs a traversal stack of " + std::to_string(wb.stack_need) + " entries");
  std::vector<uint32_t> rec(4 * (n + 16), 0u);
  for (uint64_t j = 0; j < n; j++) {
    const uint32_t i = wb.order[j];
    std::memcpy(&rec[4 * j], xyz + 3 * (size_t)i, 12);
    rec[4 * j + 3] = i;
  }
  auto* c  = new ebg_cloud;
  c->K     = K;
  c->banks = stack_banks == 3 ? 4 : std::max(stack_banks, 1);
  c->n     = n;
  hipError_t e;
  if ((e = upload(c->nodes, wb.nodes)) != hipSuccess || (e = upload(c->pts, rec)) != hipSuccess ||
      (e = upload(c->status, std::vector<uint32_t>(16, 0u))) != hipSuccess) {
    delete c;
    return fail(e == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, std::string("cloud upload: ") + hipGetErrorString(e));
  }
  c->dev.nodes  = c->nodes.p;
  c->dev.pts    = c->pts.as<uint32_t>();
  c->dev.status = c->status.as<uint32_t>();
  for (int k = 0; k < 3; k++) {
    c->dev.root_lo[k] = wb.lo[k];
    c->dev.root_hi[k] = wb.hi[k];
  }
  *out = c;
  return EBG_OK;
}

ebg_status ebg_cloud_destroy(ebg_cloud* c)
{
  delete c;
  return EBG_OK;
}

ebg_status ebg_cloud_knn(const ebg_cloud* cc, const float* d_pts, uint64_t n, uint32_t k, uint32_t* d_index,
                         float* d_dist2, const ebg_query_opts* opts)
{
  if (!cc) return fail(EBG_INVALID_ARG, "cloud is null");
  if (k < 1 || k > ebgpu::kKnnMax) return fail(EBG_INVALID_ARG, "k must be 1..16");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_index) return fail(EBG_INVALID_ARG, "null point or output array");
  if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 queries per call");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  ebg_cloud*                  c = const_cast<ebg_cloud*>(cc);
  hipStream_t                 s = static_cast<hipStream_t>(o.stream);
  std::lock_guard<std::mutex> lock(c->mu);
  const uint32_t*             idx = nullptr;
  if (o.presort && n > 1) {
    EBG_HIP(c->ws.ensure(n));
    EBG_HIP(ebgpu::presort(c->ws.view(), d_pts, n, s));
    idx = c->ws.perm.as<uint32_t>();
  }
  ebgpu::KnnArgs a{};
  a.pts     = d_pts;
  a.idx     = idx;
  a.n       = n;
  a.k       = k;
  a.self    = 0;
  a.out_idx = d_index;
  a.out_d2  = d_dist2;
  EBG_HIP(ebgpu::launch_knn(c->dev, c->K, c->banks, a, s));
  return EBG_OK;
}

ebg_status ebg_cloud_knn_self(const ebg_cloud* cc, uint32_t k, uint32_t* d_index, float* d_dist2,
                              const ebg_query_opts* opts)
{
  if (!cc) return fail(EBG_INVALID_ARG, "cloud is null");
  if (k < 1 || k > ebgpu::kKnnMax) return fail(EBG_INVALID_ARG, "k must be 1..16");
  if (!d_index) return fail(EBG_INVALID_ARG, "null output array");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  ebgpu::KnnArgs a{};
  a.n       = cc->n;
  a.k       = k;
  a.self    = 1;
  a.out_idx = d_index;
  a.out_d2  = d_dist2;
  EBG_HIP(ebgpu::launch_knn(cc->dev, cc->K, cc->banks, a, static_cast<hipStream_t>(o.stream)));
  return EBG_OK;
}

ebg_status ebg_cloud_knn_host(const ebg_cloud* c, const float* pts, uint64_t n, uint32_t k, uint32_t* index,
                              float* dist2, const ebg_query_opts* opts)
{
  if (!c) return fail(EBG_INVALID_ARG, "cloud is null");
  if (k < 1 || k > ebgpu::kKnnMax) return fail(EBG_INVALID_ARG, "k must be 1..16");
  if (n == 0) return EBG_OK;
  if (!pts || !index) return fail(EBG_INVALID_ARG, "null point or output array");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  DevBuf      p, di, dd;
  EBG_HIP(p.reserve(12 * n));
  EBG_HIP(di.reserve(4 * n * k));
  if (dist2) EBG_HIP(dd.reserve(4 * n * k));
  EBG_HIP(hipMemcpyAsync(p.p, pts, 12 * n, hipMemcpyHostToDevice, s));
  const ebg_status st = ebg_cloud_knn(c, p.as<float>(), n, k, di.as<uint32_t>(), dist2 ? dd.as<float>() : nullptr, &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(index, di.p, 4 * n * k, hipMemcpyDeviceToHost, s));
  if (dist2) EBG_HIP(hipMemcpyAsync(dist2, dd.p, 4 * n * k, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_cloud_knn_self_host(const ebg_cloud* c, uint32_t k, uint32_t* index, float* dist2,
                                   const ebg_query_opts* opts)
{
  if (!c) return fail(EBG_INVALID_ARG, "cloud is null");
  if (k < 1 || k > ebgpu::kKnnMax) return fail(EBG_INVALID_ARG, "k must be 1..16");
  if (!index) return fail(EBG_INVALID_ARG, "null output array");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t    s = static_cast<hipStream_t>(o.stream);
  const uint64_t n = c->n;
  DevBuf         di, dd;
  EBG_HIP(di.reserve(4 * n * k));
  if (dist2) EBG_HIP(dd.reserve(4 * n * k));
  const ebg_status st = ebg_cloud_knn_self(c, k, di.as<uint32_t>(), dist2 ? dd.as<float>() : nullptr, &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(index, di.p, 4 * n * k, hipMemcpyDeviceToHost, s));
  if (dist2) EBG_HIP(hipMemcpyAsync(dist2, dd.p, 4 * n * k, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_cloud_check(const ebg_cloud* c)
{
  if (!c) return fail(EBG_INVALID_ARG, "cloud is null");
  uint32_t w = 0;
  EBG_HIP(hipDeviceSynchronize());
  EBG_HIP(hipMemcpy(&w, c->status.p, 4, hipMemcpyDeviceToHost));
  if (w & ebgpu::kStatusStackOverflow)
    return fail(EBG_STACK_OVERFLOW, "a walk of this cloud needed more stack than its build-time worst case");
  return EBG_OK;
}

} // extern "C"
