// Batched value() of a CSG union scene on gfx950: spheres, axis-aligned boxes and mesh instances
// under similarity transforms, combined by a union over a 4-wide BVH of their bounding boxes.
//
// Semantics follow BVHUnionIF::value (Source/EBGeometry_CSGImplem.hpp:466-501): the running
// value v starts at +inf, a node is visited when its box distance^2 is within max(0, v)^2, and
// each leaf primitive lowers v to min(v, primitive value). Primitive values:
//   sphere  |p - c| - r                                  (AnalyticDistanceFunctions.hpp:240-247)
//   box     min(0, max_i delta_i) + |max(0, delta)|,
//           delta_i = max(lo_i - p_i, p_i - hi_i)        (AnalyticDistanceFunctions.hpp:377-399)
//   mesh    s * mesh(M (p - t) / s), M = Rx Ry Rz        (TransformImplem.hpp:140-270)
// The packet scheme is the one of the mesh kernel: one query per lane, a wave-uniform stack in
// VGPR lanes, records read with scalar loads. A mesh instance that some lane needs is walked by
// the whole wavefront with the mesh traversal (unseeded for the lanes that need it, since the
// sign of the instance value is only known once its nearest triangle is); lanes that do not
// need it carry a negative bound and take no part.
//
// FP32 pass for all queries; an FP64 pass (same traversal in double, FP64 primitive parameters,
// FP64 triangle math) redoes queries with |v| below the recheck threshold, seeded with the FP32
// value plus slack.
#include "kernels.hpp"
#include "traverse_dev.hpp"

#include <hip/hip_runtime.h>

namespace ebgpu {

namespace {

using namespace dev;

template <class R>
__device__ __forceinline__ R sphere_value(R px, R py, R pz, R cx, R cy, R cz, R r)
{
  const R dx = px - cx, dy = py - cy, dz = pz - cz;
  return sqrt(dx * dx + dy * dy + dz * dz) - r;
}

template <class R>
__device__ __forceinline__ R box_value(R px, R py, R pz, R lx, R ly, R lz, R hx, R hy, R hz)
{
  const R dx = max(lx - px, px - hx), dy = max(ly - py, py - hy), dz = max(lz - pz, pz - hz);
  const R ox = max(R(0), dx), oy = max(R(0), dy), oz = max(R(0), dz);
  return min(R(0), max(dx, max(dy, dz))) + sqrt(ox * ox + oy * oy + oz * oz);
}

struct SceneCounts
{
  uint32_t n_inner = 0, n_prims = 0, n_tris = 0, w_inner = 0, w_prims = 0, w_tris = 0;
};

// Union value of the scene at (px,py,pz) for this lane. v holds the starting value (+inf, or an
// upper bound on the answer when seeded); bound < 0 marks a lane without a query.
template <int KM, int NB, class R, bool COUNT>
__device__ __forceinline__ void scene_traverse(const SceneDev& sc, R px, R py, R pz, R& v, R bound, SceneCounts& c)
{
  const Node<4>* __restrict__ nodes = static_cast<const Node<4>*>(sc.nodes);
  LaneStack<NB>               st;
  st.init();
  const int  lane   = __lane_id();
  const bool active = bound >= R(0);
  const int  rep    = 0; // the scene tree orders children by the wave minimum
  int        top    = 0;
  Entry      root{0u, 0u, sc.root_lo[0], sc.root_lo[1], sc.root_lo[2], sc.root_hi[0], sc.root_hi[1], sc.root_hi[2]};
  st.put(top++, root);

  while (top > 0) {
    const Entry e    = st.get(--top);
    const R     d2e  = box_d2<R>(px, py, pz, e.lx, e.ly, e.lz, e.hx, e.hy, e.hz);
    const bool  need = d2e <= bound;
    if (ballot(need) == 0) continue;

    if (e.cnt == 0u) {
      if (COUNT) { c.n_inner += need ? 1u : 0u; c.w_inner++; }
      push_children<4, NB, R, true>(sload_record(nodes + uni(e.ref)), px, py, pz, bound, st, top, rep);
      if (top > 64 * NB) { // beyond the build-time worst case: record it, keep what is held
        if (lane == 0 && sc.status) atomicOr(sc.status, kStatusStackOverflow);
        top = 64 * NB;
      }
      continue;
    }
    const uint32_t first = uni(e.ref), cnt = uni(e.cnt);
    for (uint32_t i = first; i < first + cnt; i++) {
      const PrimRec  pr = sload_record(sc.prims + i);
      const bool     pn = box_d2<R>(px, py, pz, pr.bv[0], pr.bv[1], pr.bv[2], pr.bv[3], pr.bv[4], pr.bv[5]) <= bound;
      if (ballot(pn) == 0) continue;
      if (COUNT) { c.n_prims += pn ? 1u : 0u; c.w_prims++; }
      const uint32_t type = uni(pr.type);
      R              val;
      if (type == kPrimSphere) {
        if constexpr (sizeof(R) == 4) val = sphere_value<R>(px, py, pz, pr.a[0], pr.a[1], pr.a[2], pr.a[3]);
        else {
          const PrimRec64 q = sload_record(sc.prims64 + i);
          val               = sphere_value<R>(px, py, pz, q.a[0], q.a[1], q.a[2], q.a[3]);
        }
      }
      else if (type == kPrimBox) {
        if constexpr (sizeof(R) == 4)
          val = box_value<R>(px, py, pz, pr.a[0], pr.a[1], pr.a[2], pr.a[3], pr.a[4], pr.a[5]);
        else {
          const PrimRec64 q = sload_record(sc.prims64 + i);
          val               = box_value<R>(px, py, pz, q.a[0], q.a[1], q.a[2], q.a[3], q.a[4], q.a[5]);
        }
      }
      else {
        const MeshDev  md = sload_record(sc.meshes + uni(pr.mesh));
        R              qx, qy, qz, s;
        if constexpr (sizeof(R) == 4) {
          // pr.A = M / s
          const float ux = px - pr.t[0], uy = py - pr.t[1], uz = pz - pr.t[2];
          qx = pr.A[0] * ux + pr.A[1] * uy + pr.A[2] * uz;
          qy = pr.A[3] * ux + pr.A[4] * uy + pr.A[5] * uz;
          qz = pr.A[6] * ux + pr.A[7] * uy + pr.A[8] * uz;
          s  = pr.s;
        }
        else {
          const PrimRec64 q  = sload_record(sc.prims64 + i);
          const double    ux = px - q.t[0], uy = py - q.t[1], uz = pz - q.t[2];
          s  = q.s;
          qx = (q.A[0] * ux + q.A[1] * uy + q.A[2] * uz) / s;
          qy = (q.A[3] * ux + q.A[4] * uy + q.A[5] * uz) / s;
          qz = (q.A[6] * ux + q.A[7] * uy + q.A[8] * uz) / s;
        }
        R        best = INFINITY;
        uint32_t btri = kInvalid, bfeat = 0;
        traverse<KM, NB, R, COUNT, true>(md, qx, qy, qz, pn ? R(INFINITY) : R(-1), best, btri, bfeat, c.n_inner,
                                         c.n_tris, c.w_inner, c.w_tris);
        // Lanes that did not ask for this instance (bound -1) still saw the triangles the
        // wavefront evaluated, so their btri names some triangle but not their nearest one.
        val = R(INFINITY);
        if (pn && btri != kInvalid) {
          if constexpr (sizeof(R) == 4) val = s * resolve32(md, btri, bfeat, best, qx, qy, qz);
          else val = s * resolve64(md, btri, bfeat, best, qx, qy, qz);
        }
      }
      v = min(v, val);
    }
    if (active) {
      const R b = max(R(0), v);
      bound     = b * b;
    }
  }
}

// FP32 pass, one 64-query packet per wavefront, 4 per workgroup, XCD-aware block order.
template <int KM, int NB, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KM == 4 ? 8 : 1))) void scene32_kernel(const SceneDev sc, const QueryArgs a)
{
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const uint64_t n   = a.n;
  if ((uint64_t)blk * 256u >= n) return;
  const uint64_t tid    = (uint64_t)blk * 256u + threadIdx.x;
  const bool     active = tid < n;
  uint32_t       qi     = 0;
  float          px = 0.0f, py = 0.0f, pz = 0.0f;
  if (active) {
    qi = a.idx ? a.idx[tid] : (uint32_t)tid;
    px = a.pts[3 * (size_t)qi];
    py = a.pts[3 * (size_t)qi + 1];
    pz = a.pts[3 * (size_t)qi + 2];
  }
  // Seeded (learned bound or caller radii): the walk starts from v = seed, bound seed^2, and is
  // exact when some primitive comes in below the seed; the other lanes redo the walk unseeded
  // in the same wavefront (the rest sit it out with bound -1).
  const float seed = (active && a.seed_r) ? a.seed_r[qi] : INFINITY;
  float       v    = seed;
  SceneCounts c;
  scene_traverse<KM, NB, float, COUNT>(sc, px, py, pz, v, active ? seed * seed : -1.0f, c);
  if (a.seed_r) {
    const bool     redo  = active && !(v < seed);
    const uint64_t rmask = ballot(redo);
    if (rmask) {
      float       v2 = INFINITY;
      SceneCounts c2;
      scene_traverse<KM, NB, float, COUNT>(sc, px, py, pz, v2, redo ? INFINITY : -1.0f, c2);
      if (redo) v = v2;
      if (__lane_id() == 0 && a.fallback_count) atomicAdd(a.fallback_count, (uint32_t)__popcll(rmask));
      if (COUNT) {
        c.n_inner += c2.n_inner, c.n_prims += c2.n_prims, c.n_tris += c2.n_tris;
        c.w_inner += c2.w_inner, c.w_prims += c2.w_prims, c.w_tris += c2.w_tris;
      }
    }
  }
  if (active) a.out_d[qi] = v;
  if (a.recheck_list) wave_append(active && fabsf(v) < a.recheck_abs, qi, a.recheck_list, a.recheck_count);
  if (COUNT && a.work) {
    const uint64_t si = wave_sum_u64(active ? (uint64_t)c.n_inner : 0ull);
    const uint64_t st = wave_sum_u64(active ? (uint64_t)c.n_tris : 0ull);
    const uint64_t sp = wave_sum_u64(active ? (uint64_t)c.n_prims : 0ull);
    if (__lane_id() == 0) {
      atomicAdd(a.work, (unsigned long long)si);
      atomicAdd(a.work + 1, (unsigned long long)st);
      atomicAdd(a.work + 4, (unsigned long long)c.w_inner);
      atomicAdd(a.work + 5, (unsigned long long)c.w_tris);
      atomicAdd(a.work + 6, (unsigned long long)sp);
      atomicAdd(a.work + 7, (unsigned long long)c.w_prims);
    }
  }
}

// FP64 pass over the recheck list. The seed |v32| (1 + rel) + abs bounds the exact value from
// above, so every primitive that can attain the minimum is visited; the result replaces the
// FP32 value only if some primitive came in at or below the seed.
template <int KM, int NB>
__device__ __forceinline__ void scene64_block(const SceneDev& sc, const QueryArgs& a, uint64_t n, uint64_t tid)
{
  const bool     active = tid < n;
  uint32_t       qi     = 0;
  double         px = 0.0, py = 0.0, pz = 0.0, seed = -1.0;
  if (active) {
    qi   = a.idx ? a.idx[tid] : (uint32_t)tid;
    px   = a.pts[3 * (size_t)qi];
    py   = a.pts[3 * (size_t)qi + 1];
    pz   = a.pts[3 * (size_t)qi + 2];
    seed = a.slack_abs > 0.0f ? fabs((double)a.out_d[qi]) * (1.0 + (double)a.slack_rel) + (double)a.slack_abs
                              : INFINITY;
  }
  double      v = seed;
  SceneCounts c;
  scene_traverse<KM, NB, double, false>(sc, px, py, pz, v, active ? seed * seed : -1.0, c);
  if (active && v < seed) a.out_d[qi] = (float)v;
}

// The list's length is only known on the device: a fixed grid of wavefronts strides over it,
// kScene64Lanes entries per wavefront (the other lanes sit out). The list holds near-surface
// queries from all over the scene, so a wavefront walks about one path per entry; with 32
// entries the slowest wavefront, which ends the launch, ends earlier (32M queries, 4,096 spheres + 1,024 boxes + 8 meshes: 0.98 -> 0.79 ms; 16
// entries 0.81 ms).
constexpr uint32_t kScene64Lanes = 32;
template <int KM, int NB>
__global__ __launch_bounds__(256) void scene64_kernel(const SceneDev sc, const QueryArgs a)
{
  const uint64_t n     = a.count_ptr ? (uint64_t)*a.count_ptr : a.n;
  const uint32_t lane  = __lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * 4u;
  for (uint64_t w = uni(blockIdx.x * 4u + (threadIdx.x >> 6)); w * kScene64Lanes < n; w += waves) {
    const uint64_t tid = lane < kScene64Lanes ? w * kScene64Lanes + lane : n; // n: inactive
    scene64_block<KM, NB>(sc, a, n, tid);
  }
}

template <int KM, int NB>
hipError_t launch32(const SceneDev& sc, const QueryArgs& a, bool count, hipStream_t s)
{
  const uint32_t blocks = (uint32_t)((a.n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  if (count) hipLaunchKernelGGL((scene32_kernel<KM, NB, true>), dim3(blocks), dim3(256), 0, s, sc, a);
  else hipLaunchKernelGGL((scene32_kernel<KM, NB, false>), dim3(blocks), dim3(256), 0, s, sc, a);
  return hipGetLastError();
}

template <int KM, int NB>
hipError_t launch64(const SceneDev& sc, const QueryArgs& a, hipStream_t s)
{
  const uint32_t blocks = std::min<uint32_t>((uint32_t)((a.n + 255) / 256), 2048u);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((scene64_kernel<KM, NB>), dim3(blocks), dim3(256), 0, s, sc, a);
  return hipGetLastError();
}

} // namespace

hipError_t launch_scene32(const SceneDev& sc, int KM, int banks, const QueryArgs& a, bool count, hipStream_t s)
{
  if (KM == 4) {
    if (banks <= 1) return launch32<4, 1>(sc, a, count, s);
    if (banks == 2) return launch32<4, 2>(sc, a, count, s);
    return launch32<4, 4>(sc, a, count, s);
  }
  if (banks <= 1) return launch32<8, 1>(sc, a, count, s);
  if (banks == 2) return launch32<8, 2>(sc, a, count, s);
  return launch32<8, 4>(sc, a, count, s);
}

hipError_t launch_scene64(const SceneDev& sc, int KM, int banks, const QueryArgs& a, hipStream_t s)
{
  if (KM == 4) {
    if (banks <= 1) return launch64<4, 1>(sc, a, s);
    if (banks == 2) return launch64<4, 2>(sc, a, s);
    return launch64<4, 4>(sc, a, s);
  }
  if (banks <= 1) return launch64<8, 1>(sc, a, s);
  if (banks == 2) return launch64<8, 2>(sc, a, s);
  return launch64<8, 4>(sc, a, s);
}

} // namespace ebgpu
