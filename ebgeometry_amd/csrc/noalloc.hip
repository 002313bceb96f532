// Allocation-free mesh queries: ebg_query_sdf with its workspace in a file-scope device array.
//
// For callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc). The workspace of one call -- presort keys and
// permutation, radix-sort scratch, FP64 recheck list, counters, seeded radii, recheck masks and
// the split-packet schedule -- is carved from `scratch`, which every GPU the library runs on
// holds (its own copy) until the process ends. Its device address is looked up once per device
// with hipGetSymbolAddress; host code uses that address, never the array's name.
#include "host_common.hpp"

#include <hip/hip_runtime.h>

#include <mutex>

namespace {

constexpr uint64_t    kNoAllocMaxQueries = 1ull << 20; // EBG_NOALLOC_MAX_QUERIES
constexpr std::size_t kScratchBytes      = 40ull << 20;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex    g_mu;
uint8_t*      g_scratch[kMaxDevices] = {}; // device address of `scratch` per device

// Address of the current device's scratch array (looked up on the first call on that device).
ebg_status scratch_of_current_device(uint8_t** out)
{
  int dev = 0;
  EBG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ebgpu::fail(EBG_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void*            p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return ebgpu::fail(EBG_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    // the carve-up of the largest batch must fit the array
    if (ebgpu::Workspace::flat_bytes(kNoAllocMaxQueries) > kScratchBytes)
      return ebgpu::fail(EBG_UNSUPPORTED, "allocation-free workspace larger than its scratch array");
    g_scratch[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_scratch[dev];
  return EBG_OK;
}

} // namespace

extern "C" {

ebg_status ebg_query_sdf_noalloc(const ebg_mesh* cm, const float* d_pts, uint64_t n, float* d_dist, uint32_t* d_tri,
                                 const ebg_query_opts* opts)
{
  if (!cm) return ebgpu::fail(EBG_INVALID_ARG, "mesh is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_dist) return ebgpu::fail(EBG_INVALID_ARG, "null point or output array");
  if (n > kNoAllocMaxQueries)
    return ebgpu::fail(EBG_INVALID_ARG, "ebg_query_sdf_noalloc takes at most 1,048,576 queries per call; split the "
                                        "batch or use ebg_query_sdf");
  ebg_mesh*      m = const_cast<ebg_mesh*>(cm);
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  if (o.learned_bound && !m->mlp.trained)
    return ebgpu::fail(EBG_NOT_READY, "learned bound requested but no network trained");
  uint8_t*         base = nullptr;
  const ebg_status st   = scratch_of_current_device(&base);
  if (st != EBG_OK) return st;
  std::lock_guard<std::recursive_mutex> lock(m->mu);
  // sized for this call: the parts of a smaller batch sit closer together
  return ebgpu::query_pipeline(m, d_pts, n, d_dist, d_tri, o, ebgpu::Workspace::carve(base, n), false);
}

} // extern "C"
