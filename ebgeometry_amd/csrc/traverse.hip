// Packet (wavefront-coherent) BVH traversal for batched signed-distance queries, gfx950.
//
// One query per lane; the 64 lanes of a wavefront walk the tree together:
//   * the traversal stack is wave-uniform and lives in VGPR lanes (entry j of the stack is
//     lane j%64 of bank j/64), pushed with v_writelane and popped with v_readlane, so it costs
//     no LDS and no scratch;
//   * a node or triangle record is fetched once per wavefront with scalar loads into SGPRs
//     and every lane tests its own query against it;
//   * a child is pushed when any lane's box distance is within that lane's bound (ballot);
//     children are ordered by the box distance seen by the packet's middle lane (nearest
//     popped first), which is the reference's descending-d^2 push order
//     (Source/EBGeometry_BVHImplem.hpp:1178-1194) applied to the packet;
//   * on pop, each lane re-tests the entry box against its current bound (the reference's
//     `entry.dist2 > curBest2` check, :1145-1149); the entry is skipped when no lane needs it.
// Queries arrive presorted along a Hilbert curve, so the 64 queries of a wave are spatial
// neighbours and visit nearly the same nodes.
//
// The per-triangle math follows TriangleSoAT::signedDistanceLane
// (Source/EBGeometry_TriangleSoAImplem.hpp:1199-1293): face region when the three signs of
// (e_i x n).(p - v_i) agree, else the closest of the three vertices and of the edges with
// t in (0,1); the sign comes from the pseudo-normal of the chosen feature. The FP32 pass tracks
// squared distance and the feature; the sign is resolved once per query at the end. The FP64
// pass redoes near-surface queries in double with the reference's exact formulas.
#include "kernels.hpp"
#include "traverse_dev.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace ebgpu {

namespace {

using namespace dev;

// Workgroups of a list-driven launch (8 per CU); each strides over the list.
constexpr uint32_t kListGrid = 2048;

// Launch modes of the FP32 pass (distinct kernels, so that profiles separate them):
//   kPlain     every query, unseeded
//   kSeeded    every query, cull radius seed_r[q] (learned or caller-supplied bound), with the
//              unseeded redo of the queries whose radius held no triangle
enum Mode : int
{
  kPlain  = 0,
  kSeeded = 1
};

// Closest face of a lane whose FP32 walk kept more than one triangle inside the band of its
// minimum (FaceCands): the kept triangles are evaluated with the FP64 lane formulas and the
// first strict FP64 minimum (in FP32 order) wins, as TriMeshSDF::getClosestTriangle keeps the
// first strict minimum (Source/EBGeometry_MeshDistanceFunctionsImplem.hpp:511-549). Returns the
// FP64 signed distance of that face.
__device__ __forceinline__ float decide_face64(const MeshDev& m, const FaceCands& fc, float lim, float px, float py,
                                               float pz, uint32_t& btri)
{
  double   bd = INFINITY;
  uint32_t bt = fc.t[0], bf = 0;
#pragma unroll
  for (int i = 0; i < kFaceCands; i++) {
    if (fc.d[i] <= lim && fc.t[i] != kInvalid) {
      const uint32_t t = fc.t[i];
      double         d2;
      uint32_t       f;
      tri_eval64(m.tripos + 12 * (size_t)t, m.fn64 + 4 * (size_t)t, px, py, pz, d2, f);
      if (d2 < bd) {
        bd = d2;
        bt = t;
        bf = f;
      }
    }
  }
  btri = bt;
  return (float)resolve64(m, bt, bf, bd, px, py, pz);
}

// One packet: the 64 queries pk*64 .. pk*64+63 of the (presorted) query order, one per lane.
// FACE (closest faces requested): each lane decides its face exactly (decide_face64); a lane
// whose band held more triangles than its list goes to the FP64 pass with the near-surface ones.
// Distance-only variants (!FACE) evaluate leaves pair-compacted in the wavefront's PairLds.
template <int K, int NB, bool COUNT, int MODE, bool FACE>
__device__ __forceinline__ void sdf32_packet(const MeshDev& m, const QueryArgs& a, uint64_t n, uint64_t pk, PairLds* pl)
{
  const uint64_t tid    = pk * 64u + (uint64_t)__lane_id();
  const bool     active = tid < n;
  uint32_t   qi     = 0;
  float      px = 0.0f, py = 0.0f, pz = 0.0f, bound = -1.0f;
  if (active) {
    qi = a.idx ? a.idx[tid] : (uint32_t)tid;
    px = a.pts[3 * (size_t)qi];
    py = a.pts[3 * (size_t)qi + 1];
    pz = a.pts[3 * (size_t)qi + 2];
    if (MODE == kSeeded) {
      const float r = a.seed_r[qi];
      bound         = r * r;
    }
    else {
      bound = INFINITY;
    }
  }
  float    best = INFINITY;
  uint32_t btri = kInvalid, bfeat = 0, n_inner = 0, n_tris = 0, w_inner = 0, w_tris = 0;
  const float seed2 = bound;
  FaceCands   fc;
  if constexpr (FACE) fc.init(m.hmax);
  PairCounts pc{0u, 0u, 0u, 0u};
  traverse<K, NB, float, COUNT, false, FACE, !FACE>(m, px, py, pz, bound, best, btri, bfeat, n_inner, n_tris, w_inner,
                                                    w_tris, &fc, pl, &pc);

  // Exact only if some triangle lies within the seeded radius: otherwise nodes between the
  // seeded radius and the true distance were pruned, and the query must be redone unseeded.
  // The redo runs here, in the same wavefront (the other lanes sit it out with bound -1), so a
  // rare expensive query is hidden among the other wavefronts of the launch instead of being
  // serialised in a launch of its own.
  bool found = active && btri != kInvalid && best <= seed2;
  if (MODE == kSeeded) {
    const bool     redo  = active && !found;
    const uint64_t rmask = ballot(redo);
    if (rmask) {
      float     best2 = INFINITY;
      uint32_t  btri2 = kInvalid, bfeat2 = 0, f_inner = 0, f_tris = 0;
      FaceCands fc2;
      if constexpr (FACE) fc2.init(m.hmax);
      traverse<K, NB, float, COUNT, false, FACE, !FACE>(m, px, py, pz, redo ? INFINITY : -1.0f, best2, btri2, bfeat2,
                                                        f_inner, f_tris, w_inner, w_tris, &fc2, pl, &pc);
      if (redo) {
        best  = best2;
        btri  = btri2;
        bfeat = bfeat2;
        if constexpr (FACE) fc = fc2;
      }
      found = active && btri != kInvalid;
      if (__lane_id() == 0) atomicAdd(a.fallback_count, (uint32_t)__popcll(rmask));
      if (COUNT && a.work) {
        const uint64_t fi = wave_sum_u64(redo ? (uint64_t)f_inner : 0ull);
        const uint64_t ft = wave_sum_u64(redo ? (uint64_t)f_tris : 0ull);
        if (__lane_id() == 0) {
          atomicAdd(a.work + 6, (unsigned long long)fi);
          atomicAdd(a.work + 7, (unsigned long long)ft);
        }
      }
    }
  }
  float d   = 0.0f;
  bool  ovf = false; // FACE: the band held more triangles than the list
  if (found) {
    d = resolve32(m, btri, bfeat, best, px, py, pz);
    if constexpr (FACE) {
      // a second triangle inside the band: decide in FP64 among the kept ones, unless the
      // query is a near-surface one (the FP64 pass decides those from scratch) or the band
      // overflowed the list (then the FP64 pass decides it too)
      const float lim = fc.band(best);
      ovf             = fc.evict <= lim;
      if (!ovf && fc.d[1] <= lim && !(fabsf(d) < a.recheck_abs)) d = decide_face64(m, fc, lim, px, py, pz, btri);
    }
    a.out_d[qi] = d;
    if (a.out_tri) a.out_tri[qi] = m.tris[btri].face;
  }
  const bool rchk = found && (fabsf(d) < a.recheck_abs || ovf);
  if (a.recheck_mask) {
    const uint64_t rm = ballot(rchk);
    if (__lane_id() == 0) a.recheck_mask[pk] = rm;
  }
  else if (a.recheck_list) wave_append(rchk, qi, a.recheck_list, a.recheck_count);
  if (COUNT && a.packet_cost && __lane_id() == 0) a.packet_cost[pk] = w_inner + w_tris;
  if (COUNT && a.work) {
    const uint64_t si = wave_sum_u64(active ? (uint64_t)n_inner : 0ull);
    const uint64_t st = wave_sum_u64(active ? (uint64_t)n_tris : 0ull);
    if (__lane_id() == 0) {
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

// One split item: subtree root r for the 64 queries of listed packet j, unseeded; each lane's
// (d^2, triangle) is folded into the packet's keys. d^2 >= 0, so its bits order like the value;
// on equal d^2 the lower triangle index wins.
template <int K, int NB, bool COUNT>
__device__ __forceinline__ void split_item(const MeshDev& m, const QueryArgs& a, uint32_t j, uint32_t r, PairLds* pl)
{
  const uint64_t pk     = a.split_list[j];
  const uint32_t lane   = __lane_id();
  const uint64_t tid    = pk * 64u + lane;
  const bool     active = tid < a.n;
  float          px = 0.0f, py = 0.0f, pz = 0.0f;
  if (active) {
    const uint32_t qi = a.idx ? a.idx[tid] : (uint32_t)tid;
    px                = a.pts[3 * (size_t)qi];
    py                = a.pts[3 * (size_t)qi + 1];
    pz                = a.pts[3 * (size_t)qi + 2];
  }
  const SplitRoot rt = sload_record(a.split_roots + r);
  const Entry     start{rt.ref, rt.cnt, rt.lo[0], rt.lo[1], rt.lo[2], rt.hi[0], rt.hi[1], rt.hi[2]};
  float           best = INFINITY;
  uint32_t        btri = kInvalid, bfeat = 0, n_inner = 0, n_tris = 0, w_inner = 0, w_tris = 0;
  traverse_from<K, NB, float, COUNT, false, false, true>(m, start, px, py, pz, active ? INFINITY : -1.0f, best, btri,
                                                        bfeat, n_inner, n_tris, w_inner, w_tris, nullptr, pl);
  if (active && btri != kInvalid)
    atomicMin(a.split_keys + 64 * (size_t)j + lane, ((unsigned long long)__float_as_uint(best) << 32) | btri);
  if (COUNT && a.work) {
    const uint64_t si = wave_sum_u64(active ? (uint64_t)n_inner : 0ull);
    const uint64_t st = wave_sum_u64(active ? (uint64_t)n_tris : 0ull);
    if (lane == 0) {
      atomicAdd(a.work, (unsigned long long)si);
      atomicAdd(a.work + 1, (unsigned long long)st);
      atomicAdd(a.work + 4, (unsigned long long)w_inner);
      atomicAdd(a.work + 5, (unsigned long long)w_tris);
    }
  }
}

// Grid launch: one wavefront per 64-query packet, 4 packets per workgroup; workgroups are
// mapped XCD-aware so that each XCD works through one contiguous (Morton) range of queries.
// The leading split_blocks workgroups (dispatched first) work through the split items instead,
// beside the packets of the bulk; packets flagged in heavy_flag are skipped.
template <int K, int NB, bool COUNT, int MODE, bool FACE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(K == 4 && !FACE ? 8 : 1))) void sdf32_kernel(const MeshDev m, const QueryArgs a)
{
  const uint32_t wave = uni(threadIdx.x >> 6); // known uniform: list entries load as scalars
  extern __shared__ PairLds pair_lds[];        // one per wavefront (distance-only variants)
  PairLds* const pl = FACE ? nullptr : &pair_lds[wave];
  uint32_t       b = blockIdx.x, nb = gridDim.x;
  if (a.split_blocks) {
    if (b < a.split_blocks) {
      const uint32_t items = min(*a.split_count, a.split_cap) * a.split_nroots;
      for (uint32_t w = b * 4u + wave; w < items; w += a.split_blocks * 4u) {
        const uint32_t j = w / a.split_nroots;
        split_item<K, NB, COUNT>(m, a, j, w - j * a.split_nroots, pl);
      }
      return;
    }
    b -= a.split_blocks;
    nb -= a.split_blocks;
  }
  if (a.early_blocks) {
    if (b < a.early_blocks) {
      const uint32_t j = b * 4u + wave;
      if (j < min(*a.early_count, a.early_cap)) sdf32_packet<K, NB, COUNT, MODE, FACE>(m, a, a.n, a.early_list[j], pl);
      return;
    }
    b -= a.early_blocks;
    nb -= a.early_blocks;
  }
  const uint32_t blk = xcd_remap(b, nb);
  const uint64_t n   = a.n;
  const uint64_t pk  = (uint64_t)blk * 4u + wave;
  if (pk * 64u >= n) return;                    // whole wavefront past the end
  if (a.heavy_flag && a.heavy_flag[pk]) return; // whole wavefront: a split or early packet
  sdf32_packet<K, NB, COUNT, MODE, FACE>(m, a, n, pk, pl);
}

// One wavefront per listed packet: the winning triangle of each query, as the plain pass ends.
__global__ __launch_bounds__(256) void sdf32_resolve_kernel(const MeshDev m, const QueryArgs a)
{
  const uint32_t j = uni(blockIdx.x * 4u + (threadIdx.x >> 6));
  if (j >= min(*a.split_count, a.split_cap)) return; // whole wavefront
  const uint64_t pk     = a.split_list[j];
  const uint64_t tid    = pk * 64u + __lane_id();
  const bool     active = tid < a.n;
  bool           found  = false;
  uint32_t       qi     = 0;
  float          d      = 0.0f;
  if (active) {
    qi                           = a.idx ? a.idx[tid] : (uint32_t)tid;
    const unsigned long long key = a.split_keys[64 * (size_t)j + __lane_id()];
    const uint32_t           t   = (uint32_t)key;
    if (t != kInvalid) {
      const float px = a.pts[3 * (size_t)qi], py = a.pts[3 * (size_t)qi + 1], pz = a.pts[3 * (size_t)qi + 2];
      float       d2;
      uint32_t    feat;
      tri_eval32<true>(m.tris[t], px, py, pz, d2, feat);
      d           = resolve32(m, t, feat, __uint_as_float((uint32_t)(key >> 32)), px, py, pz);
      a.out_d[qi] = d;
      if (a.out_tri) a.out_tri[qi] = m.tris[t].face;
      found = true;
    }
  }
  if (a.recheck_mask) {
    const uint64_t rm = ballot(found && fabsf(d) < a.recheck_abs);
    if (__lane_id() == 0) a.recheck_mask[pk] = rm;
  }
  else if (a.recheck_list) wave_append(found && fabsf(d) < a.recheck_abs, qi, a.recheck_list, a.recheck_count);
}

// one wavefront per 64 consecutive packets
__global__ __launch_bounds__(256) void recheck_compact_kernel(const uint64_t* __restrict__ mask, uint64_t npk,
                                                              const uint32_t* __restrict__ idx, uint32_t* __restrict__ list,
                                                              uint32_t* __restrict__ count)
{
  const uint64_t chunk = (uint64_t)uni(blockIdx.x * 4u + (threadIdx.x >> 6));
  const int      lane  = __lane_id();
  const uint64_t pk    = chunk * 64u + (uint64_t)lane;
  uint64_t       mk    = pk < npk ? mask[pk] : 0ull;
  const uint32_t c     = (uint32_t)__popcll(mk);
  uint32_t       inc   = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  const uint32_t total = rdlane(inc, 63);
  if (total == 0) return; // whole wavefront
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(count, total);
  uint32_t pos = rdlane(base, 0) + inc - c;
  while (mk) {
    const uint64_t tid = pk * 64u + (uint64_t)(__ffsll((unsigned long long)mk) - 1);
    mk &= mk - 1ull;
    list[pos++] = idx ? idx[tid] : (uint32_t)tid;
  }
}

template <int K, int NB>
__device__ __forceinline__ void sdf64_block(const MeshDev& m, const QueryArgs& a, uint64_t n, uint64_t tid)
{
  const bool active = tid < n;
  uint32_t   qi     = 0;
  double     px = 0.0, py = 0.0, pz = 0.0, bound = -1.0;
  if (active) {
    qi = a.idx ? a.idx[tid] : (uint32_t)tid;
    px = a.pts[3 * (size_t)qi];
    py = a.pts[3 * (size_t)qi + 1];
    pz = a.pts[3 * (size_t)qi + 2];
    if (a.slack_abs > 0.0f) {
      const double r = fabs((double)a.out_d[qi]) * (1.0 + (double)a.slack_rel) + (double)a.slack_abs;
      bound          = r * r;
    }
    else {
      bound = INFINITY;
    }
  }
  double   best = INFINITY;
  uint32_t btri = kInvalid, bfeat = 0, n_inner = 0, n_tris = 0, w_inner = 0, w_tris = 0;
  const double seed2 = bound;
  traverse<K, NB, double, false>(m, px, py, pz, bound, best, btri, bfeat, n_inner, n_tris, w_inner, w_tris);
  const bool found = active && btri != kInvalid && best <= seed2;
  if (found) {
    a.out_d[qi] = (float)resolve64(m, btri, bfeat, best, px, py, pz);
    if (a.out_tri) a.out_tri[qi] = m.feat[8 * (size_t)btri + 6];
  }
  if (a.fallback_list) wave_append(active && !found, qi, a.fallback_list, a.fallback_count);
}

// FP64 pass over a device-side list (a.idx, count *a.count_ptr) or over all a.n queries; a fixed
// grid of wavefronts strides over the entries, kSdf64Lanes per wavefront (the other lanes sit
// out). Near-surface queries are sparse even in presort order, so a wavefront walks several
// paths; with 32 entries its slowest wavefronts end earlier (16M queries, 327,680-face displaced sphere: 0.31 -> 0.23 ms).
constexpr uint32_t kSdf64Lanes = 32;
template <int K, int NB>
__global__ __launch_bounds__(256) void sdf64_kernel(const MeshDev m, const QueryArgs a)
{
  const uint64_t n     = a.count_ptr ? (uint64_t)*a.count_ptr : a.n;
  const uint32_t lane  = __lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * 4u;
  for (uint64_t w = uni(blockIdx.x * 4u + (threadIdx.x >> 6)); w * kSdf64Lanes < n; w += waves)
    sdf64_block<K, NB>(m, a, n, lane < kSdf64Lanes ? w * kSdf64Lanes + lane : n);
}

template <int K, int NB, int MODE, bool FACE>
void launch32f(const MeshDev& m, const QueryArgs& a, bool count, uint32_t blocks, hipStream_t s)
{
  const size_t lds = FACE ? 0 : 4 * sizeof(PairLds);
  if (count) hipLaunchKernelGGL((sdf32_kernel<K, NB, true, MODE, FACE>), dim3(blocks), dim3(256), lds, s, m, a);
  else hipLaunchKernelGGL((sdf32_kernel<K, NB, false, MODE, FACE>), dim3(blocks), dim3(256), lds, s, m, a);
}

// closest faces requested: the exact-face variant (its split packets are off: see launch32)
template <int K, int NB, int MODE>
void launch32m(const MeshDev& m, const QueryArgs& a, bool count, uint32_t blocks, hipStream_t s)
{
  if (a.out_tri) launch32f<K, NB, MODE, true>(m, a, count, blocks, s);
  else launch32f<K, NB, MODE, false>(m, a, count, blocks, s);
}

template <int K, int NB>
hipError_t launch32(const MeshDev& m, const QueryArgs& a, bool count, hipStream_t s)
{
  const uint32_t blocks = (uint32_t)((a.n + 255) / 256) + a.split_blocks + a.early_blocks;
  if (blocks == 0) return hipSuccess;
  if (a.count_ptr) return hipErrorInvalidValue;
  if (a.split_blocks && (!a.split_nroots || !a.split_count || !a.heavy_flag)) return hipErrorInvalidValue;
  // exact closest faces need each lane's whole candidate band in one walk: no split packets, and
  // the FP64 pass must follow (lanes whose band overflowed are listed for it)
  if (a.out_tri && (a.split_blocks || (!a.recheck_mask && !a.recheck_list))) return hipErrorInvalidValue;
  if (a.early_blocks && (!a.early_list || !a.early_count || !a.heavy_flag)) return hipErrorInvalidValue;
  if (a.seed_r) {
    if (!a.fallback_count) return hipErrorInvalidValue;
    launch32m<K, NB, kSeeded>(m, a, count, blocks, s);
  }
  else {
    launch32m<K, NB, kPlain>(m, a, count, blocks, s);
  }
  return hipGetLastError();
}

template <int K, int NB>
hipError_t launch64(const MeshDev& m, const QueryArgs& a, hipStream_t s)
{
  const uint32_t blocks = std::min<uint32_t>((uint32_t)((a.n + 255) / 256), kListGrid);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((sdf64_kernel<K, NB>), dim3(blocks), dim3(256), 0, s, m, a);
  return hipGetLastError();
}

} // namespace

hipError_t launch_sdf32(const MeshDev& m, int K, int banks, const QueryArgs& a, bool count, hipStream_t s)
{
  if (K == 4) {
    if (banks <= 1) return launch32<4, 1>(m, a, count, s);
    if (banks == 2) return launch32<4, 2>(m, a, count, s);
    return launch32<4, 4>(m, a, count, s);
  }
  if (banks <= 1) return launch32<8, 1>(m, a, count, s);
  if (banks == 2) return launch32<8, 2>(m, a, count, s);
  return launch32<8, 4>(m, a, count, s);
}

namespace {

// One wavefront per 64 consecutive packets (packet first + j holds the presorted queries
// (first + j) * 64 ..): packet j is heavy (early) when any of its queries carries the heavy
// (early) cost flag in its sorted presort key; lane j collects packet j's flags, and the
// wavefront appends its heavy and early packets with one atomic per list (one atomic per
// packet on the same two counters cost 0.11 ms at 16M queries).
__global__ __launch_bounds__(256) void heavy_mark_kernel(const HeavyMarkArgs h)
{
  const uint64_t packets = (h.n + 63u) / 64u;
  const uint64_t first   = (uint64_t)uni(blockIdx.x * 4u + (threadIdx.x >> 6)) * 64u;
  if (first >= packets) return; // whole wavefront
  const uint32_t lane = __lane_id();
  const uint64_t lt   = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  uint32_t       mine = 0; // bit 0: heavy flag, bit 1: early flag of packet first + lane
#pragma unroll 8
  for (uint32_t j = 0; j < 64u; j++) {
    const uint64_t tid = (first + j) * 64u + lane;
    const uint32_t key = tid < h.n ? h.skeys[tid] : 0u;
    const bool     hv  = ballot((key >> 31) != 0u) != 0ull;
    const bool     er  = ballot(((key >> 30) & 1u) != 0u) != 0ull;
    if (lane == j) mine = (hv ? 1u : 0u) | (er ? 2u : 0u);
  }
  const uint64_t hmask = ballot((mine & 1u) != 0u);
  uint32_t       hbase = 0;
  if (hmask && lane == 0) hbase = atomicAdd(h.count, (uint32_t)__popcll(hmask));
  const uint32_t hslot = rdlane(hbase, 0) + (uint32_t)__popcll(hmask & lt);
  const bool     heavy = (mine & 1u) && hslot < h.cap;
  const bool     ecand = !heavy && h.early_list && (mine & 2u);
  const uint64_t emask = ballot(ecand);
  uint32_t       ebase = 0;
  if (emask && lane == 0) ebase = atomicAdd(h.early_count, (uint32_t)__popcll(emask));
  const uint32_t eslot = rdlane(ebase, 0) + (uint32_t)__popcll(emask & lt);
  const bool     early = ecand && eslot < h.early_cap;
  const uint64_t pk    = first + lane;
  if (pk < packets) {
    if (heavy) h.list[hslot] = (uint32_t)pk;
    if (early) h.early_list[eslot] = (uint32_t)pk;
    h.flag[pk] = heavy ? 1 : (early ? 2 : 0);
  }
  // reset the 64 result keys of each heavy slot
  for (uint64_t hm = ballot(heavy); hm; hm &= hm - 1ull) {
    const uint32_t sl = rdlane(hslot, __ffsll((unsigned long long)hm) - 1);
    h.keys[64 * (size_t)sl + lane] = ~0ull;
  }
}

} // namespace

hipError_t launch_heavy_mark(const HeavyMarkArgs& h, hipStream_t s)
{
  const uint64_t packets = (h.n + 63) / 64;
  if (packets == 0) return hipSuccess;
  const uint64_t waves = (packets + 63) / 64;
  hipLaunchKernelGGL(heavy_mark_kernel, dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, s, h);
  return hipGetLastError();
}

hipError_t launch_recheck_compact(const uint64_t* mask, uint64_t n, const uint32_t* idx, uint32_t* list,
                                  uint32_t* count, hipStream_t s)
{
  const uint64_t npk    = (n + 63) / 64;
  const uint64_t chunks = (npk + 63) / 64;
  if (chunks == 0) return hipSuccess;
  hipLaunchKernelGGL(recheck_compact_kernel, dim3((uint32_t)((chunks + 3) / 4)), dim3(256), 0, s, mask, npk, idx, list,
                     count);
  return hipGetLastError();
}

hipError_t launch_sdf32_resolve(const MeshDev& m, const QueryArgs& a, hipStream_t s)
{
  const uint32_t blocks = (a.split_cap + 3) / 4;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(sdf32_resolve_kernel, dim3(blocks), dim3(256), 0, s, m, a);
  return hipGetLastError();
}

hipError_t launch_sdf64(const MeshDev& m, int K, int banks, const QueryArgs& a, hipStream_t s)
{
  if (K == 4) {
    if (banks <= 1) return launch64<4, 1>(m, a, s);
    if (banks == 2) return launch64<4, 2>(m, a, s);
    return launch64<4, 4>(m, a, s);
  }
  if (banks <= 1) return launch64<8, 1>(m, a, s);
  if (banks == 2) return launch64<8, 2>(m, a, s);
  return launch64<8, 4>(m, a, s);
}

} // namespace ebgpu
