// k-nearest-neighbour queries on point clouds, gfx950: PointCloudBVH::closestPoint(s) and
// nearestNeighbor(s) / allNearestNeighbors (Source/EBGeometry_PointCloudBVH.hpp:115-161,
// PointCloudBVHImplem.hpp:235-470), batched.
//
// The packet traversal of the mesh kernel with a point leaf: one query per lane, the 64 lanes of
// a wavefront walk the K-wide tree together (wave-uniform VGPR-lane stack, node records and leaf
// points read with scalar loads), a child is pushed when some lane's box distance^2 is below its
// bound, and each lane keeps its k best (d^2, cloud index) in registers, sorted, new entries after
// equal ones (the reference's strict-< insertion). The bound is the k-th best d^2 (+inf until k
// points are held). Self queries (the cloud's own points, in leaf order, so packets are spatially
// coherent) exclude the point's own index.
#include "kernels.hpp"
#include "traverse_dev.hpp"

#include <hip/hip_runtime.h>

namespace ebgpu {

namespace {

using namespace dev;

template <int KMAX>
struct KBest
{
  float    d[KMAX];
  uint32_t id[KMAX];

  __device__ __forceinline__ void init()
  {
#pragma unroll
    for (int i = 0; i < KMAX; i++) {
      d[i]  = INFINITY;
      id[i] = kInvalid;
    }
  }
  // sorted insert; equal distances keep their order (the new entry goes after them)
  __device__ __forceinline__ void insert(float x, uint32_t xi)
  {
#pragma unroll
    for (int i = 0; i < KMAX; i++) {
      const bool     sw = x < d[i];
      const float    od = d[i];
      const uint32_t oi = id[i];
      d[i]              = sw ? x : od;
      id[i]             = sw ? xi : oi;
      x                 = sw ? od : x;
      xi                = sw ? oi : xi;
    }
  }
  __device__ __forceinline__ float kth(uint32_t k) const
  {
    float b = INFINITY;
#pragma unroll
    for (int i = 0; i < KMAX; i++) b = (uint32_t)i + 1u == k ? d[i] : b;
    return b;
  }
};

template <int K, int NB, int KMAX>
__device__ __forceinline__ void knn_walk(const CloudDev& c, float px, float py, float pz, bool active, uint32_t k,
                                         uint32_t exclude, KBest<KMAX>& kb)
{
  const Node<K>* __restrict__ nodes = static_cast<const Node<K>*>(c.nodes);
  LaneStack<NB>               st;
  st.init();
  const int lane = __lane_id();
  const int rep  = rep_lane(active ? 0.0f : -1.0f);
  int       top  = 0;
  float     bound = active ? INFINITY : -1.0f;
  st.put(top++, Entry{0u, 0u, c.root_lo[0], c.root_lo[1], c.root_lo[2], c.root_hi[0], c.root_hi[1], c.root_hi[2]});
  while (top > 0) {
    const Entry e = st.get(--top);
    // a box at the bound can hold no new entry (strict <), but a lane without k points yet has
    // bound +inf; inactive lanes have bound -1
    const bool need = box_d2<float>(px, py, pz, e.lx, e.ly, e.lz, e.hx, e.hy, e.hz) < bound;
    if (ballot(need) == 0) continue;
    if (e.cnt == 0u) {
      push_children<K, NB, float, false>(sload_record(nodes + uni(e.ref)), px, py, pz, bound, st, top, rep);
      if (top > 64 * NB) {
        if (lane == 0 && c.status) atomicOr(c.status, kStatusStackOverflow);
        top = 64 * NB;
      }
      continue;
    }
    const uint32_t first = uni(e.ref), cnt = uni(e.cnt);
    for (uint32_t c0 = 0; c0 < cnt; c0 += 4) {
      // four points (x, y, z, cloud index) per scalar round trip; the array is padded by 16
      const SWords<16> w = sload<16>(c.pts + 4 * (size_t)(first + c0));
#pragma unroll
      for (int i = 0; i < 4; i++) {
        if (c0 + i < cnt) {
          const float    dx = px - __uint_as_float(w.w[4 * i]), dy = py - __uint_as_float(w.w[4 * i + 1]),
                         dz = pz - __uint_as_float(w.w[4 * i + 2]);
          const float    d2 = dx * dx + dy * dy + dz * dz;
          const uint32_t pi = w.w[4 * i + 3];
          if (d2 < bound && pi != exclude) {
            kb.insert(d2, pi);
            bound = kb.kth(k);
          }
        }
      }
    }
  }
}

template <int K, int NB, int KMAX>
__global__ __launch_bounds__(256) void knn_kernel(const CloudDev c, const KnnArgs a)
{
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if ((uint64_t)blk * 256u >= a.n) return; // whole workgroup
  const uint64_t tid    = (uint64_t)blk * 256u + threadIdx.x;
  const bool     active = tid < a.n;
  float          px = 0.0f, py = 0.0f, pz = 0.0f;
  uint32_t       row = 0, exclude = kInvalid;
  if (active) {
    if (a.self) { // the cloud's own point at leaf position tid
      px      = __uint_as_float(c.pts[4 * tid]);
      py      = __uint_as_float(c.pts[4 * tid + 1]);
      pz      = __uint_as_float(c.pts[4 * tid + 2]);
      row     = c.pts[4 * tid + 3];
      exclude = row;
    }
    else {
      row = a.idx ? a.idx[tid] : (uint32_t)tid;
      px  = a.pts[3 * (size_t)row];
      py  = a.pts[3 * (size_t)row + 1];
      pz  = a.pts[3 * (size_t)row + 2];
    }
  }
  KBest<KMAX> kb;
  kb.init();
  knn_walk<K, NB, KMAX>(c, px, py, pz, active, a.k, exclude, kb);
  if (!active) return;
#pragma unroll
  for (int i = 0; i < KMAX; i++) {
    if ((uint32_t)i < a.k) {
      a.out_idx[(size_t)row * a.k + i] = kb.id[i];
      if (a.out_d2) a.out_d2[(size_t)row * a.k + i] = kb.d[i];
    }
  }
}

template <int K, int NB, int KMAX>
hipError_t launch_k(const CloudDev& c, const KnnArgs& a, hipStream_t s)
{
  const uint32_t blocks = (uint32_t)((a.n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((knn_kernel<K, NB, KMAX>), dim3(blocks), dim3(256), 0, s, c, a);
  return hipGetLastError();
}

template <int K, int NB>
hipError_t launch_nb(const CloudDev& c, const KnnArgs& a, hipStream_t s)
{
  if (a.k <= 1) return launch_k<K, NB, 1>(c, a, s);
  if (a.k <= 4) return launch_k<K, NB, 4>(c, a, s);
  if (a.k <= 8) return launch_k<K, NB, 8>(c, a, s);
  return launch_k<K, NB, 16>(c, a, s);
}

} // namespace

hipError_t launch_knn(const CloudDev& c, int K, int banks, const KnnArgs& a, hipStream_t s)
{
  if (a.k < 1 || a.k > kKnnMax) return hipErrorInvalidValue;
  if (K == 4) {
    if (banks <= 1) return launch_nb<4, 1>(c, a, s);
    if (banks == 2) return launch_nb<4, 2>(c, a, s);
    return launch_nb<4, 4>(c, a, s);
  }
  if (banks <= 1) return launch_nb<8, 1>(c, a, s);
  if (banks == 2) return launch_nb<8, 2>(c, a, s);
  return launch_nb<8, 4>(c, a, s);
}

} // namespace ebgpu
