// Host-side plumbing shared by the C-ABI translation units: error reporting, device buffers,
// the per-handle query workspace and the mesh handle itself.
#pragma once

#include "../../include/ebgpu.h"
#include "kernels.hpp"
#include "mlp.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

namespace ebgpu {

void set_last_error(const std::string& s);

inline ebg_status fail(ebg_status s, const std::string& msg)
{
  set_last_error(msg);
  return s;
}

#define EBG_HIP(call)                                                                                        \
  do {                                                                                                       \
    const hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess)                                                                                    \
      return ::ebgpu::fail(e_ == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR,                             \
                           std::string(#call) + ": " + hipGetErrorString(e_));                               \
  } while (0)

struct DevBuf
{
  void*       p     = nullptr;
  std::size_t bytes = 0;
  DevBuf()              = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (p) (void)hipFree(p);
    p     = nullptr;
    bytes = 0;
  }
  hipError_t reserve(std::size_t b)
  {
    if (b <= bytes) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    else p = nullptr;
    return e;
  }
  template <class T>
  T* as() const
  {
    return static_cast<T*>(p);
  }
};

template <class T>
hipError_t upload(DevBuf& b, const std::vector<T>& v)
{
  const std::size_t bytes = std::max<std::size_t>(64, v.size() * sizeof(T));
  hipError_t        e     = b.reserve(bytes);
  if (e != hipSuccess) return e;
  if (!v.empty()) e = hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

// Device memory of one query call, carved from a workspace: the presort keys and permutation,
// the radix sort's scratch, the bounding-box partials, the FP64 recheck list, the counters, the
// seeded radii, the recheck lane masks and the split/early packet schedule, for up to `cap`
// queries.
struct WsView
{
  uint64_t            cap = 0;
  uint32_t*           codes = nullptr;        // cap (32-bit presort keys)
  uint32_t*           codes_sorted = nullptr; // cap
  uint32_t*           perm = nullptr;         // cap
  void*               sort_scratch = nullptr; // radix_sort_scratch_bytes(cap)
  void*               box = nullptr;          // 6 doubles x (kBboxParts + 1)
  uint32_t*           recheck_list = nullptr; // cap
  uint32_t*           counters = nullptr;     // 4: recheck, fallback, split packets, early packets
  float*              seeds = nullptr;        // cap
  uint64_t*           rmask = nullptr;        // packets(cap)
  uint8_t*            heavy_flag = nullptr;   // packets(cap)
  uint32_t*           heavy_list = nullptr;   // heavy_cap(cap)
  unsigned long long* heavy_keys = nullptr;   // 64 x heavy_cap(cap)
  uint32_t*           early_list = nullptr;   // early_cap(cap)
};

// Query workspace, grown on demand.
struct Workspace
{
  DevBuf   codes, codes_sorted, perm, sort_scratch, box, lists, seeds, sched, rmask;
  uint64_t cap = 0;

  hipError_t ensure(uint64_t n)
  {
    if (n <= cap && cap) return hipSuccess;
    const uint64_t m = std::max<uint64_t>(n, 1024);
    hipError_t     e;
    if ((e = codes.reserve(8 * m)) != hipSuccess) return e;
    if ((e = codes_sorted.reserve(8 * m)) != hipSuccess) return e;
    if ((e = perm.reserve(4 * m)) != hipSuccess) return e;
    if ((e = sort_scratch.reserve(radix_sort_scratch_bytes(m))) != hipSuccess) return e;
    if ((e = box.reserve(sizeof(double) * 6 * (kBboxParts + 1))) != hipSuccess) return e;
    // recheck list, fallback list, 2 counters (+pad)
    if ((e = lists.reserve(8 * m + 256)) != hipSuccess) return e;
    if ((e = seeds.reserve(4 * m)) != hipSuccess) return e;
    if ((e = rmask.reserve(8 * packets(m))) != hipSuccess) return e; // recheck lane mask per packet
    // split packets: one flag byte per packet, the list of split packets, 64 result keys each
    if ((e = sched.reserve(sched_early_offset(m) + 4 * (size_t)early_cap(m))) != hipSuccess) return e;
    cap = m;
    return hipSuccess;
  }
  static uint64_t packets(uint64_t n) { return (n + 63) / 64; }
  static uint32_t heavy_cap(uint64_t n) { return (uint32_t)(packets(n) / 64 + 64); }
  static size_t   sched_list_offset(uint64_t n) { return (packets(n) + 255) & ~size_t(255); }
  static size_t   sched_keys_offset(uint64_t n) { return sched_list_offset(n) + ((4 * (size_t)heavy_cap(n) + 255) & ~size_t(255)); }
  static uint32_t early_cap(uint64_t n) { return (uint32_t)(packets(n) / 8 + 64); }
  static size_t   sched_early_offset(uint64_t n) { return sched_keys_offset(n) + 512 * (size_t)heavy_cap(n); }
  uint32_t*       early_list() const { return reinterpret_cast<uint32_t*>(sched.as<uint8_t>() + sched_early_offset(cap)); }
  uint8_t*        heavy_flag() const { return sched.as<uint8_t>(); }
  uint32_t*       heavy_list() const { return reinterpret_cast<uint32_t*>(sched.as<uint8_t>() + sched_list_offset(cap)); }
  unsigned long long* heavy_keys() const
  {
    return reinterpret_cast<unsigned long long*>(sched.as<uint8_t>() + sched_keys_offset(cap));
  }
  uint32_t* recheck_list() const { return lists.as<uint32_t>(); }
  uint32_t* fallback_list() const { return lists.as<uint32_t>() + cap; }
  uint32_t* counters() const { return lists.as<uint32_t>() + 2 * cap; } // [0] recheck [1] fallback [2] split [3] early

  WsView view() const
  {
    WsView v;
    v.cap          = cap;
    v.codes        = codes.as<uint32_t>();
    v.codes_sorted = codes_sorted.as<uint32_t>();
    v.perm         = perm.as<uint32_t>();
    v.sort_scratch = sort_scratch.p;
    v.box          = box.p;
    v.recheck_list = recheck_list();
    v.counters     = counters();
    v.seeds        = seeds.as<float>();
    v.rmask        = rmask.as<uint64_t>();
    v.heavy_flag   = heavy_flag();
    v.heavy_list   = heavy_list();
    v.heavy_keys   = heavy_keys();
    v.early_list   = early_list();
    return v;
  }

  // The same arrays carved from one flat block of flat_bytes(cap) bytes (256-byte aligned parts).
  static std::size_t part(std::size_t b) { return (b + 255) & ~std::size_t(255); }
  static std::size_t flat_bytes(uint64_t cap)
  {
    const uint64_t np = packets(cap);
    return 3 * part(4 * cap) + part(radix_sort_scratch_bytes(cap)) + part(sizeof(double) * 6 * (kBboxParts + 1)) +
           part(4 * cap) + part(16) + part(4 * cap) + part(8 * np) + part(np) + part(4 * (size_t)heavy_cap(cap)) +
           part(512 * (size_t)heavy_cap(cap)) + part(4 * (size_t)early_cap(cap));
  }
  static WsView carve(uint8_t* base, uint64_t cap)
  {
    const uint64_t np = packets(cap);
    WsView         v;
    v.cap    = cap;
    auto take = [&](std::size_t b) {
      uint8_t* p = base;
      base += part(b);
      return p;
    };
    v.codes        = reinterpret_cast<uint32_t*>(take(4 * cap));
    v.codes_sorted = reinterpret_cast<uint32_t*>(take(4 * cap));
    v.perm         = reinterpret_cast<uint32_t*>(take(4 * cap));
    v.sort_scratch = take(radix_sort_scratch_bytes(cap));
    v.box          = take(sizeof(double) * 6 * (kBboxParts + 1));
    v.recheck_list = reinterpret_cast<uint32_t*>(take(4 * cap));
    v.counters     = reinterpret_cast<uint32_t*>(take(16));
    v.seeds        = reinterpret_cast<float*>(take(4 * cap));
    v.rmask        = reinterpret_cast<uint64_t*>(take(8 * np));
    v.heavy_flag   = take(np);
    v.heavy_list   = reinterpret_cast<uint32_t*>(take(4 * (size_t)heavy_cap(cap)));
    v.heavy_keys   = reinterpret_cast<unsigned long long*>(take(512 * (size_t)heavy_cap(cap)));
    v.early_list   = reinterpret_cast<uint32_t*>(take(4 * (size_t)early_cap(cap)));
    return v;
  }
};

// Hilbert presort of n device points into a workspace: the permutation lands in w.perm (valid
// until the next call). With cost flags the sorted keys, flags in bits 30 and 31, land in
// w.codes_sorted.
inline hipError_t presort(const WsView& w, const float* d_pts, uint64_t n, hipStream_t s, const CostFlags* cf = nullptr)
{
  hipError_t e;
  float*     box = static_cast<float*>(w.box);
  if ((e = launch_bbox_f32(d_pts, n, box, box + 8, s)) != hipSuccess) return e;
  // 30-bit Hilbert keys on a 1024^3 grid, sorted as 32-bit keys in 4 radix passes
  if ((e = launch_hilbert_f32(d_pts, n, box, w.codes, s, cf)) != hipSuccess) return e;
  return radix_sort_pairs32(w.codes, nullptr, w.codes_sorted, w.perm, n, 30, w.sort_scratch, cf != nullptr, s);
}

} // namespace ebgpu

// The mesh handle. Scenes that instance a mesh hold a reference: the handle (and its device
// memory) lives until ebg_mesh_destroy and every such scene have released it.
struct ebg_mesh
{
  int                    K = 4, leaf = 4, banks = 1, stack_need = 0;
  std::size_t            nv = 0, nt = 0, num_nodes = 0, num_leaves = 0;
  double                 lo[3]{}, hi[3]{}, diag = 0;
  double                 build_seconds = 0;
  ebgpu::DevBuf          nodes, tris, tripos, tsph, feat, vn32, en32, vn64, en64, fn64;
  ebgpu::DevBuf          area_cdf; // normalised cumulative triangle area, leaf order (training samples)
  ebgpu::DevBuf          lthick;   // MeshDev::lthick (one float per triangle, set at leaf starts)
  ebgpu::DevBuf          frames;   // MeshDev::frames (one FrameRec per triangle)
  // cost grid of the split schedule (built at the first large presorted query): per cell, the
  // work of a packet of 64 queries spread over the cell
  ebgpu::DevBuf          cost_grid;
  ebgpu::DevBuf          split_roots; // SplitRoot frontier covering the tree
  ebgpu::DevBuf          status;      // one device word, MeshDev::status
  ebgpu::DevBuf          pts32, dist32; // FP64-point queries: the points rounded to FP32, FP32 results
  uint64_t               hash = 0;      // FNV-1a of the vertex and triangle arrays given at creation
  uint32_t               num_split_roots = 0;
  bool                   grid_ready = false;
  float                  grid_lo[3]{}, grid_scale[3]{};
  double                 grid_mean = 0; // mean packet cost over the grid cells
  bool                   last_split     = false; // the last query listed split packets ...
  int64_t                last_split_cap = 0;     // ... up to this many
  std::vector<float>     host_xyz; // vertex positions (scene instance bounds)
  std::vector<uint32_t>  order;    // original face index of each leaf-order triangle
  // polygon meshes: the polygon each triangle was clipped from (host and device copies); the
  // closest-face outputs are mapped through it
  uint64_t               nf = 0;
  std::vector<uint32_t>  host_tri_face;
  ebgpu::DevBuf          tri_face;
  ebgpu::MeshDev         dev{};
  ebgpu::MlpTrainBuffers train;
  std::size_t            device_bytes = 0;
  std::recursive_mutex   mu; // training runs queries while holding it
  ebgpu::Workspace       ws;
  hipEvent_t             ev0 = nullptr, ev1 = nullptr;
  hipStream_t            aux = nullptr; // bound network, concurrent with the presort
  hipEvent_t             ev_fork = nullptr, ev_join = nullptr;
  bool                   timed = false;
  ebgpu::Mlp             mlp;
  std::atomic<int>       refs{1};

  ~ebg_mesh()
  {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (aux) (void)hipStreamDestroy(aux);
  }
};

namespace ebgpu {
constexpr uint32_t kCostGrid = 32; // cells per axis of a mesh's packet cost grid
// Build the mesh's packet cost grid once (a counting launch of 64 lattice queries per cell on
// stream s, then a wait); defined with the query pipeline (api.cpp).
ebg_status ensure_cost_grid(ebg_mesh* m, hipStream_t s);
// Validate a lattice (origin, spacing, dims; at most 2^32-1 points) into g; n = its point count.
ebg_status grid_spec(const double* origin, const double* spacing, const uint64_t* dims, GridSpec& g, uint64_t& n);
// The mesh query pipeline (presort, [bound network], FP32 traversal with split packets, FP64
// recheck) over a caller-chosen workspace of at least n queries. With may_alloc false nothing is
// allocated or created: split packets only if the cost grid already exists, the bound network
// on the caller's stream, no traversal timing. The caller holds m->mu.
ebg_status query_pipeline(ebg_mesh* m, const float* d_pts, uint64_t n, float* d_dist, uint32_t* d_tri,
                          const ebg_query_opts& o, const WsView& ws, bool may_alloc);
inline void mesh_retain(ebg_mesh* m) { m->refs.fetch_add(1); }
inline void mesh_release(ebg_mesh* m)
{
  if (m->refs.fetch_sub(1) == 1) delete m;
}
} // namespace ebgpu
