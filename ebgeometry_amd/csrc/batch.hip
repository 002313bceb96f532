// Many meshes, one traversal launch (gfx950).
//
// The queries of all items (item = one query array against one mesh) form one flat index
// space. The pipeline:
//   1. keys: item id in the high bits, the query's Morton code (relative to a box around the
//      item's mesh) in the low bits; one stable radix sort then orders every item's queries along
//      the curve and keeps the items apart;
//   2. batch32: a packet = up to 64 consecutive (sorted) queries of one item, one per lane of a
//      wavefront; a packet table (built on the host, most expensive meshes first) maps each
//      wavefront to (item, first query). The wavefront loads its item and mesh records with
//      scalar loads and runs the same packet traversal as the single-mesh kernel, so every
//      wavefront works on one mesh while different wavefronts of a launch work on different
//      meshes;
//   3. batch64: packets that hold near-surface results were listed (one entry per packet with
//      the lane mask); one wavefront per listed packet redoes those lanes in FP64.
#include "kernels.hpp"
#include "sfc_dev.hpp"
#include "traverse_dev.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace ebgpu {

namespace {

using namespace dev;

// Item of flat query g (items[lo].base <= g < items[lo + 1].base): binary search over the item
// bases; with a wave-uniform g the probes are scalar loads.
__device__ __forceinline__ uint32_t item_of(const BatchItemDev* __restrict__ items, uint32_t num_items, uint64_t g)
{
  uint32_t lo = 0, hi = num_items;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (items[mid].base <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Sort key of flat query g: (item << 24) | 24-bit Hilbert index of the query on a 256^3 grid of
// the item's box (the mesh box grown by 25%; 8 bits per axis order 64-query packets well at any
// item size). The item of the wavefront's first and last query is searched once per
// wavefront; lanes of a wavefront that straddles items step forward from the first.
template <class KeyT>
__global__ __launch_bounds__(256) void batch_keys_kernel(const BatchItemDev* __restrict__ items, uint32_t num_items,
                                                         uint64_t total, KeyT* __restrict__ keys)
{
  const uint64_t g0 = (uint64_t)blockIdx.x * 256u + uni(threadIdx.x & ~63u);
  if (g0 >= total) return; // whole wavefront
  const uint64_t g     = g0 + __lane_id();
  const uint64_t glast = min(g0 + 63u, total - 1);
  const uint32_t lo0   = uni(item_of(items, num_items, g0));
  const uint32_t lo1   = uni(item_of(items, num_items, glast));
  uint32_t       lo    = lo0;
  if (lo1 != lo0)
    while (lo < lo1 && items[lo + 1].base <= g) lo++;
  if (g >= total) return;
  const BatchItemDev& it = items[lo];
  const uint64_t      j  = g - it.base;
  uint32_t            x[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    x[c] = (uint32_t)fminf(fmaxf((it.pts[3 * j + c] - it.box_lo[c]) * it.box_inv[c], 0.0f), 255.0f); // NaN -> 0
  keys[g] = KeyT(((uint64_t)lo << 24) | hilbert_index<8>(x[0], x[1], x[2]));
}

// One packet: the wavefront's 64 queries of one item against its mesh.
template <int K, int NB, bool COUNT>
__device__ __forceinline__ void batch_packet(const BatchArgs& a, uint32_t p, PairLds* pl)
{
  const SWords<4>    pk = sload<4>(a.packets + 4 * (size_t)p);
  const uint32_t     item = pk.w[0], first = pk.w[1];
  const BatchItemDev it   = sload_record(a.items + item);
  const MeshDev      m    = sload_record(a.meshes + it.mesh);

  const uint32_t lane   = __lane_id();
  const uint32_t j      = first + lane;
  const bool     active = j < it.n;
  uint32_t       qi     = 0;
  float          px = 0.0f, py = 0.0f, pz = 0.0f, bound = -1.0f;
  if (active) {
    qi    = a.sorted ? (uint32_t)(a.sorted[it.base + j] - it.base) : j;
    px    = it.pts[3 * (size_t)qi];
    py    = it.pts[3 * (size_t)qi + 1];
    pz    = it.pts[3 * (size_t)qi + 2];
    bound = INFINITY;
  }
  float    best = INFINITY;
  uint32_t btri = kInvalid, bfeat = 0, n_inner = 0, n_tris = 0, w_inner = 0, w_tris = 0;
  PairCounts pc{0u, 0u, 0u, 0u};
  traverse<K, NB, float, COUNT, false, false, true>(m, px, py, pz, bound, best, btri, bfeat, n_inner, n_tris, w_inner,
                                                    w_tris, nullptr, pl, &pc);
  const bool found = active && btri != kInvalid;
  float      d     = 0.0f;
  if (found) {
    d            = resolve32(m, btri, bfeat, best, px, py, pz);
    it.out[qi]   = d;
    if (it.tri) it.tri[qi] = m.tris[btri].face;
  }
  if (a.recheck_mask) {
    const uint64_t mask = ballot(found && fabsf(d) < it.recheck_abs);
    if (lane == 0) a.recheck_mask[p] = mask;
  }
  if (COUNT && a.work) {
    const uint64_t si = wave_sum_u64(active ? (uint64_t)n_inner : 0ull);
    const uint64_t st = wave_sum_u64(active ? (uint64_t)n_tris : 0ull);
    if (lane == 0) {
      atomicAdd(a.work, (unsigned long long)si);
      atomicAdd(a.work + 1, (unsigned long long)st);
      atomicAdd(a.work + 4, (unsigned long long)w_inner);
      atomicAdd(a.work + 5, (unsigned long long)w_tris);
      if (a.pair_work) {
        atomicAdd(a.pair_work, (unsigned long long)pc.pairs);
        atomicAdd(a.pair_work + 1, (unsigned long long)pc.batches);
        atomicAdd(a.pair_work + 2, (unsigned long long)pc.groups);
        atomicAdd(a.pair_work + 3, (unsigned long long)pc.leaves);
      }
    }
  }
}

// PERSIST = false: one wavefront per packet (the early list first, then the table without the
// early packets). PERSIST = true: a grid that fits the device at once; each wavefront takes queue
// entries in the same order until the queue is exhausted (no wavefront waits on another).
template <int K, int NB, bool COUNT, bool PERSIST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(K == 4 ? 8 : 1))) void batch32_kernel(const BatchArgs a)
{
  extern __shared__ PairLds pair_lds[]; // one per wavefront of the workgroup
  PairLds* const pl = &pair_lds[uni(threadIdx.x >> 6)];
  if constexpr (PERSIST) {
    const uint32_t ne   = a.early_blocks ? min(*a.early_count, a.early_cap) : 0u;
    const uint32_t skip = a.early_blocks ? a.early_cap : 0u; // queue entries before the table
    for (;;) {
      uint32_t w = 0;
      if (__lane_id() == 0) w = atomicAdd(a.queue, 1u);
      w = uni(rdlane(w, 0));
      if (w >= a.queue_total) return; // whole wavefront
      uint32_t p;
      if (w < skip) {
        if (w >= ne) continue;
        p = uni(a.early_list[w]);
      }
      else {
        p = w - skip;
        if (p >= a.num_packets || (a.early_blocks && a.flag[p])) continue;
      }
      batch_packet<K, NB, COUNT>(a, p, pl);
    }
  }
  else {
    const uint32_t wpb = blockDim.x >> 6;
    // wave-uniform, and known to be: the packet and mesh records below are then scalar loads.
    // Plain dispatch order (no XCD remap): all XCDs work on the same stretch of the LPT-ordered
    // packet table, so the meshes in flight at any time are few and stay in the Infinity Cache;
    // with a contiguous eighth of the table per XCD, eight stretches of meshes were in flight and
    // the launch took 2.5x as long (163 against 66 ms on the 256-mesh workload).
    uint32_t w = uni(blockIdx.x * wpb + (threadIdx.x >> 6));
    uint32_t p;
    if (a.early_blocks) {
      // the early packets first (they run longest), then the table in order without them
      const uint32_t ew = a.early_blocks * wpb;
      if (w < ew) {
        if (w >= min(*a.early_count, a.early_cap)) return; // whole wavefront
        p = uni(a.early_list[w]);
      }
      else {
        p = w - ew;
        if (p >= a.num_packets || a.flag[p]) return; // whole wavefront
      }
    }
    else p = w;
    if (p >= a.num_packets) return; // whole wavefront (p is wave-uniform)
    batch_packet<K, NB, COUNT>(a, p, pl);
  }
}

// One wavefront per 64 consecutive entries of the recheck list (packet << 6 | lane, in packet
// order: neighbouring queries, mostly of one item). The wavefront takes the distinct items of its
// entries in turn; each listed query is redone in FP64, seeded with (|d32| (1 + 1e-4) +
// slack_abs)^2, which contains the FP64 minimum.
constexpr uint32_t kBatch64Lanes = 64; // list entries per wavefront (32 measured the same on 256 meshes, 60M queries)
template <int K, int NB>
__device__ __forceinline__ void batch64_group(const BatchArgs& a, uint32_t count, uint32_t g)
{
  const uint32_t lane   = __lane_id();
  const uint32_t e      = g * kBatch64Lanes + lane;
  const bool     listed = lane < kBatch64Lanes && e < count;
  uint32_t       item = kInvalid, j = 0;
  if (listed) {
    const uint32_t en  = a.recheck_list[e];
    const size_t   pkt = en >> 6;
    item               = a.packets[4 * pkt];
    j                  = a.packets[4 * pkt + 1] + (en & 63u);
  }
  uint64_t pending = ballot(listed);
  while (pending) {
    const uint32_t cur  = rdlane(item, __ffsll((unsigned long long)pending) - 1);
    const bool     mine = listed && item == cur;
    pending &= ~ballot(mine);
    const BatchItemDev it = sload_record(a.items + cur);
    const MeshDev      m  = sload_record(a.meshes + it.mesh);
    uint32_t           qi = 0;
    double             px = 0.0, py = 0.0, pz = 0.0, bound = -1.0;
    if (mine) {
      qi             = a.sorted ? (uint32_t)(a.sorted[it.base + j] - it.base) : j;
      px             = it.pts[3 * (size_t)qi];
      py             = it.pts[3 * (size_t)qi + 1];
      pz             = it.pts[3 * (size_t)qi + 2];
      const double r = fabs((double)it.out[qi]) * (1.0 + 1e-4) + (double)it.slack_abs;
      bound          = r * r;
    }
    double   best = INFINITY;
    uint32_t btri = kInvalid, bfeat = 0, n_inner = 0, n_tris = 0, w_inner = 0, w_tris = 0;
    const double seed2 = bound;
    traverse<K, NB, double, false>(m, px, py, pz, bound, best, btri, bfeat, n_inner, n_tris, w_inner, w_tris);
    if (mine && btri != kInvalid && best <= seed2) {
      it.out[qi] = (float)resolve64(m, btri, bfeat, best, px, py, pz);
      if (it.tri) it.tri[qi] = m.feat[8 * (size_t)btri + 6];
    }
  }
}

// A fixed grid of wavefronts strides over the groups (their number is only known on the device).
template <int K, int NB>
__global__ __launch_bounds__(256) void batch64_kernel(const BatchArgs a)
{
  const uint32_t count  = *a.recheck_count;
  const uint32_t groups = (count + kBatch64Lanes - 1u) / kBatch64Lanes;
  const uint32_t waves  = gridDim.x * 4u;
  for (uint32_t g = uni(blockIdx.x * 4u + (threadIdx.x >> 6)); g < groups; g += waves) batch64_group<K, NB>(a, count, g);
}

// One wavefront per packet: the largest cost-grid value over its queries' cells in its mesh's
// grid decides whether it is listed as early.
__global__ __launch_bounds__(256) void batch_mark_kernel(const BatchArgs a)
{
  const uint32_t p = uni(blockIdx.x * 4u + (threadIdx.x >> 6));
  if (p >= a.num_packets) return; // whole wavefront
  const SWords<4>    pk   = sload<4>(a.packets + 4 * (size_t)p);
  const uint32_t     item = pk.w[0], first = pk.w[1];
  const BatchItemDev it   = sload_record(a.items + item);
  const GridDev      gd   = sload_record(a.grids + it.mesh);
  const uint32_t     lane = __lane_id();
  const uint32_t     j    = first + lane;
  uint32_t           c    = 0;
  if (gd.threshold != kInvalid && j < it.n) {
    const uint32_t qi = a.sorted ? (uint32_t)(a.sorted[it.base + j] - it.base) : j;
    uint32_t       cell[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float x = (it.pts[3 * (size_t)qi + k] - gd.lo[k]) * gd.scale[k];
      cell[k]       = (uint32_t)fminf(fmaxf(x, 0.0f), float(gd.g - 1)); // NaN -> 0
    }
    c = gd.cost[((size_t)cell[2] * gd.g + cell[1]) * gd.g + cell[0]];
  }
  const uint32_t est  = wave_max_u32(c);
  uint32_t       slot = kInvalid;
  if (gd.threshold != kInvalid && est >= gd.threshold && lane == 0) slot = atomicAdd(a.early_count, 1u);
  slot             = rdlane(slot, 0);
  const bool early = slot < a.early_cap;
  if (lane == 0) {
    if (early) a.early_list[slot] = p;
    a.flag[p] = early ? 1 : 0;
  }
}

template <int K, int NB>
hipError_t launch32(const BatchArgs& a, int wg, bool count, hipStream_t s)
{
  const size_t lds = (size_t)(wg / 64) * sizeof(PairLds);
  if (a.persist_blocks) {
    if (!a.queue) return hipErrorInvalidValue;
    if (count) hipLaunchKernelGGL((batch32_kernel<K, NB, true, true>), dim3(a.persist_blocks), dim3(wg), lds, s, a);
    else hipLaunchKernelGGL((batch32_kernel<K, NB, false, true>), dim3(a.persist_blocks), dim3(wg), lds, s, a);
    return hipGetLastError();
  }
  const uint32_t wpb    = (uint32_t)wg / 64u;
  const uint32_t blocks = (a.num_packets + wpb - 1) / wpb + a.early_blocks;
  if (blocks == 0) return hipSuccess;
  if (count) hipLaunchKernelGGL((batch32_kernel<K, NB, true, false>), dim3(blocks), dim3(wg), lds, s, a);
  else hipLaunchKernelGGL((batch32_kernel<K, NB, false, false>), dim3(blocks), dim3(wg), lds, s, a);
  return hipGetLastError();
}

template <int K, int NB>
hipError_t persist_blocks(int wg, bool count, uint32_t* blocks)
{
  int        nb = 0, dev = 0, cus = 0;
  const size_t lds = (size_t)(wg / 64) * sizeof(PairLds);
  hipError_t   e   = count ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch32_kernel<K, NB, true, true>, wg, lds)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch32_kernel<K, NB, false, true>, wg, lds);
  if (e != hipSuccess) return e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  *blocks = (uint32_t)(std::max(nb, 1) * std::max(cus, 1));
  return hipSuccess;
}

template <int K, int NB>
hipError_t launch64(const BatchArgs& a, hipStream_t s)
{
  if (a.num_packets == 0 || !a.recheck_list) return hipSuccess;
  hipLaunchKernelGGL((batch64_kernel<K, NB>), dim3(std::min<uint32_t>((a.num_packets + 3) / 4, 2048u)), dim3(256), 0, s,
                     a);
  return hipGetLastError();
}

} // namespace

hipError_t launch_batch_keys(const BatchItemDev* items, uint32_t num_items, uint64_t total, void* keys, bool narrow,
                             hipStream_t s)
{
  if (total == 0) return hipSuccess;
  if (narrow && num_items > 256) return hipErrorInvalidValue;
  const dim3 grid((uint32_t)((total + 255) / 256));
  if (narrow)
    hipLaunchKernelGGL((batch_keys_kernel<uint32_t>), grid, dim3(256), 0, s, items, num_items, total,
                       static_cast<uint32_t*>(keys));
  else
    hipLaunchKernelGGL((batch_keys_kernel<uint64_t>), grid, dim3(256), 0, s, items, num_items, total,
                       static_cast<uint64_t*>(keys));
  return hipGetLastError();
}

hipError_t launch_batch_mark(const BatchArgs& a, hipStream_t s)
{
  if (a.num_packets == 0) return hipSuccess;
  if (!a.grids || !a.flag || !a.early_list || !a.early_count) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_mark_kernel, dim3((a.num_packets + 3) / 4), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_batch32(const BatchArgs& a, int K, int banks, int wg, bool count, hipStream_t s)
{
  if (wg != 64 && wg != 128 && wg != 256) return hipErrorInvalidValue;
  if (a.early_blocks && (!a.flag || !a.early_list || !a.early_count)) return hipErrorInvalidValue;
  if (K == 4) {
    if (banks <= 1) return launch32<4, 1>(a, wg, count, s);
    if (banks == 2) return launch32<4, 2>(a, wg, count, s);
    return launch32<4, 4>(a, wg, count, s);
  }
  if (banks <= 1) return launch32<8, 1>(a, wg, count, s);
  if (banks == 2) return launch32<8, 2>(a, wg, count, s);
  return launch32<8, 4>(a, wg, count, s);
}

hipError_t batch32_persistent_blocks(int K, int banks, int wg, bool count, uint32_t* blocks)
{
  if (wg != 64 && wg != 128 && wg != 256) return hipErrorInvalidValue;
  if (K == 4) {
    if (banks <= 1) return persist_blocks<4, 1>(wg, count, blocks);
    if (banks == 2) return persist_blocks<4, 2>(wg, count, blocks);
    return persist_blocks<4, 4>(wg, count, blocks);
  }
  if (banks <= 1) return persist_blocks<8, 1>(wg, count, blocks);
  if (banks == 2) return persist_blocks<8, 2>(wg, count, blocks);
  return persist_blocks<8, 4>(wg, count, blocks);
}

hipError_t launch_batch64(const BatchArgs& a, int K, int banks, hipStream_t s)
{
  if (K == 4) {
    if (banks <= 1) return launch64<4, 1>(a, s);
    if (banks == 2) return launch64<4, 2>(a, s);
    return launch64<4, 4>(a, s);
  }
  if (banks <= 1) return launch64<8, 1>(a, s);
  if (banks == 2) return launch64<8, 2>(a, s);
  return launch64<8, 4>(a, s);
}

} // namespace ebgpu
