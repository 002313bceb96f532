// GPU mesh build: the triangle mesh's pseudo-normals, a binned-SAH BVH and the packed triangle
// records, all computed on the device.
//
//   normals  face normal (x2-x0)x(x2-x1)/|.| per triangle; half-edge pairs found by a radix sort
//            of the directed edge keys a*nv+b and a binary search for b*nv+a, edge normal =
//            normalise(n_f + n_pair); vertex normals by a stable sort of the corners by vertex
//            and one thread per vertex summing acos(e1.e2) n_f over its corners in face order --
//            the formulas and summation order of the host code (compute_pseudo_normals, which
//            follows Source/EBGeometry_DCEL_{Face,Edge,Vertex}Implem.hpp), evaluated without
//            contracted multiply-adds so that each operation rounds as on the host.
//   SAH      the host builder's binary binned SAH (32 bins per axis, cost A_L N_L + A_R N_R, a
//            node of at most leaf_size triangles is a leaf; Source/EBGeometry_BVH.hpp:521-641):
//            nodes of more than kSmall triangles level by level over all workgroups (bins
//            privatised in LDS, merged with atomics; the split of every node of the level by one
//            thread; a stable partition by per-chunk counts), then every smaller node's whole
//            subtree by one workgroup with the triangles in LDS. A child's box and centroid
//            bounds are the unions of its side's bins, so no node re-reads its triangles for
//            bounds; the splits, and hence the tree, are the host builder's (only the order of
//            triangles inside a leaf may differ: the partitions here are stable).
//   packing  TriRec, bounding spheres and planes, positions, feature ids, FP64 face normals in
//            leaf order, the longest edge, and the area CDF the bound's training samples use.
// The host collapses the binary tree into K-wide nodes (O(nodes)).
#include "gpu_build.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace ebgpu {

namespace {

constexpr int      kBins       = 32;
constexpr int      kBinFields  = 13; // count, box lo[3], box hi[3], centroid lo[3], centroid hi[3]
constexpr int      kBinWords   = 3 * kBins * kBinFields;
constexpr uint32_t kSmall      = 1024; // a node of at most this many triangles is built by one workgroup
constexpr uint32_t kChunk      = 4096; // triangles per workgroup in the level-synchronous phase
constexpr uint32_t kMedianDepth = 48;  // from this depth on, nodes are split at their middle
constexpr int      kSmallStack = 64;
constexpr uint32_t kBadFace    = 1u;   // status bits of the normal kernels
constexpr uint32_t kBadStack   = 2u;

// Floats as integers whose order is the floats' order (for integer atomic min / max).
__device__ __forceinline__ int32_t f2o(float f)
{
  const int32_t i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float o2f(int32_t i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__device__ __forceinline__ int32_t bin_init(int w)
{
  const int f = w % kBinFields;
  if (f == 0) return 0;
  const bool lo = (f >= 1 && f <= 3) || (f >= 7 && f <= 9);
  return lo ? f2o(INFINITY) : f2o(-INFINITY);
}

struct OpenNode
{
  float    lo[3], hi[3], clo[3], chi[3];
  uint32_t first, count, depth, id;
};
static_assert(sizeof(OpenNode) == 64, "OpenNode layout");

struct SplitOut
{
  float    llo[3], lhi[3], lclo[3], lchi[3];
  float    rlo[3], rhi[3], rclo[3], rchi[3];
  uint32_t nleft;
  int32_t  axis; // -1: split at the middle, both children keep the parent's bounds
  int32_t  bin;
  uint32_t pad;
};
static_assert(sizeof(SplitOut) == 112, "SplitOut layout");

struct ChunkRec
{
  uint32_t slot, begin, end, pad;
};

// ---------------------------------------------------------------------------------------------
// pseudo-normals
// ---------------------------------------------------------------------------------------------

struct Dv3
{
  double x, y, z;
};
__device__ __forceinline__ Dv3 ld3(const float* p) { return {double(p[0]), double(p[1]), double(p[2])}; }

__global__ __launch_bounds__(256) void face_normals_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                           uint64_t nt, double* __restrict__ fn, uint32_t* status)
{
#pragma clang fp contract(off)
  const uint64_t f = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= nt) return;
  const Dv3 x0 = ld3(xyz + 3 * (size_t)tri[3 * f]), x1 = ld3(xyz + 3 * (size_t)tri[3 * f + 1]),
            x2 = ld3(xyz + 3 * (size_t)tri[3 * f + 2]);
  const Dv3 a = {x2.x - x0.x, x2.y - x0.y, x2.z - x0.z}, u = {x2.x - x1.x, x2.y - x1.y, x2.z - x1.z};
  const Dv3 n = {a.y * u.z - a.z * u.y, a.z * u.x - a.x * u.z, a.x * u.y - a.y * u.x};
  const double l = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
  if (!(l > DBL_EPSILON)) {
    atomicOr(status, kBadFace);
    atomicMin(status + 1, (uint32_t)f);
  }
  const double is = 1.0 / l;
  fn[3 * f]       = is * n.x;
  fn[3 * f + 1]   = is * n.y;
  fn[3 * f + 2]   = is * n.z;
}

__global__ __launch_bounds__(256) void edge_keys_kernel(const uint32_t* __restrict__ tri, uint64_t nt, uint64_t nv,
                                                        uint64_t* __restrict__ keys)
{
  const uint64_t h = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (h >= 3 * nt) return;
  const uint64_t f = h / 3, i = h % 3;
  keys[h]          = (uint64_t)tri[3 * f + i] * nv + tri[3 * f + (i + 1) % 3];
}

__global__ __launch_bounds__(256) void edge_normals_kernel(const uint32_t* __restrict__ tri, uint64_t nt, uint64_t nv,
                                                           const uint64_t* __restrict__ skeys,
                                                           const uint32_t* __restrict__ svals,
                                                           const double* __restrict__ fn, double* __restrict__ en64,
                                                           float* __restrict__ en32)
{
#pragma clang fp contract(off)
  const uint64_t h = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const uint64_t m = 3 * nt;
  if (h >= m) return;
  const uint64_t f = h / 3, i = h % 3;
  const uint64_t q = (uint64_t)tri[3 * f + (i + 1) % 3] * nv + tri[3 * f + i];
  uint64_t       lo = 0, hi = m; // first position with key >= q
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (skeys[mid] < q) lo = mid + 1;
    else hi = mid;
  }
  Dv3 n = {fn[3 * f], fn[3 * f + 1], fn[3 * f + 2]};
  if (lo < m && skeys[lo] == q) {
    const uint64_t g = svals[lo] / 3;
    n                = {n.x + fn[3 * g], n.y + fn[3 * g + 1], n.z + fn[3 * g + 2]};
  }
  const double l = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
  Dv3          e = {0.0, 0.0, 0.0};
  if (l > DBL_EPSILON) {
    const double is = 1.0 / l;
    e               = {is * n.x, is * n.y, is * n.z};
  }
  en64[4 * h] = e.x, en64[4 * h + 1] = e.y, en64[4 * h + 2] = e.z, en64[4 * h + 3] = 0.0;
  en32[4 * h] = float(e.x), en32[4 * h + 1] = float(e.y), en32[4 * h + 2] = float(e.z), en32[4 * h + 3] = 0.0f;
}

__global__ __launch_bounds__(256) void corner_keys_kernel(const uint32_t* __restrict__ tri, uint64_t nt,
                                                          uint32_t* __restrict__ keys)
{
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (c < 3 * nt) keys[c] = tri[c];
}

// One thread per vertex (the first of its run of corners in the sorted order).
__global__ __launch_bounds__(256) void vertex_normals_kernel(const uint32_t* __restrict__ skeys,
                                                             const uint32_t* __restrict__ svals, uint64_t m,
                                                             const uint32_t* __restrict__ tri,
                                                             const float* __restrict__ xyz,
                                                             const double* __restrict__ fn,
                                                             double* __restrict__ vn64, float* __restrict__ vn32)
{
#pragma clang fp contract(off)
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m || (i > 0 && skeys[i - 1] == skeys[i])) return;
  const uint32_t v   = skeys[i];
  const Dv3      x0  = ld3(xyz + 3 * (size_t)v);
  Dv3            acc = {0.0, 0.0, 0.0};
  for (uint64_t j = i; j < m && skeys[j] == v; j++) {
    const uint32_t c = svals[j], f = c / 3, k = c % 3;
    const Dv3      p1 = ld3(xyz + 3 * (size_t)tri[3 * f + (k + 1) % 3]);
    const Dv3      p2 = ld3(xyz + 3 * (size_t)tri[3 * f + (k + 2) % 3]);
    Dv3            e1 = {p1.x - x0.x, p1.y - x0.y, p1.z - x0.z};
    Dv3            e2 = {p2.x - x0.x, p2.y - x0.y, p2.z - x0.z};
    const double   i1 = 1.0 / sqrt(e1.x * e1.x + e1.y * e1.y + e1.z * e1.z);
    const double   i2 = 1.0 / sqrt(e2.x * e2.x + e2.y * e2.y + e2.z * e2.z);
    e1                = {i1 * e1.x, i1 * e1.y, i1 * e1.z};
    e2                = {i2 * e2.x, i2 * e2.y, i2 * e2.z};
    const double d    = fmin(1.0, fmax(-1.0, e1.x * e2.x + e1.y * e2.y + e1.z * e2.z));
    const double al   = acos(d);
    acc = {acc.x + al * fn[3 * f], acc.y + al * fn[3 * f + 1], acc.z + al * fn[3 * f + 2]};
  }
  const double is = 1.0 / sqrt(acc.x * acc.x + acc.y * acc.y + acc.z * acc.z);
  const Dv3    n  = {is * acc.x, is * acc.y, is * acc.z};
  vn64[4 * v] = n.x, vn64[4 * v + 1] = n.y, vn64[4 * v + 2] = n.z, vn64[4 * v + 3] = 0.0;
  vn32[4 * v] = float(n.x), vn32[4 * v + 1] = float(n.y), vn32[4 * v + 2] = float(n.z), vn32[4 * v + 3] = 0.0f;
}

// ---------------------------------------------------------------------------------------------
// binned SAH
// ---------------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void prims_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                    uint64_t nt, float* __restrict__ blo, float* __restrict__ bhi,
                                                    float* __restrict__ cen, uint32_t* __restrict__ idx)
{
  const uint64_t f = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= nt) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float a = xyz[3 * (size_t)tri[3 * f] + c], b = xyz[3 * (size_t)tri[3 * f + 1] + c],
                d = xyz[3 * (size_t)tri[3 * f + 2] + c];
    const float lo = fminf(a, fminf(b, d)), hi = fmaxf(a, fmaxf(b, d));
    blo[3 * f + c] = lo;
    bhi[3 * f + c] = hi;
    cen[3 * f + c] = 0.5f * (lo + hi);
  }
  idx[f] = (uint32_t)f;
}

// Box and centroid bounds of all triangles: 12 ordered integers (lo 3, hi 3, clo 3, chi 3).
__global__ __launch_bounds__(256) void root_bounds_kernel(const float* __restrict__ blo, const float* __restrict__ bhi,
                                                          const float* __restrict__ cen, uint64_t nt, int32_t* out)
{
  float v[12];
  for (int c = 0; c < 3; c++) {
    v[c] = v[6 + c] = INFINITY;
    v[3 + c] = v[9 + c] = -INFINITY;
  }
  for (uint64_t f = (uint64_t)blockIdx.x * 256 + threadIdx.x; f < nt; f += (uint64_t)gridDim.x * 256)
    for (int c = 0; c < 3; c++) {
      v[c]     = fminf(v[c], blo[3 * f + c]);
      v[3 + c] = fmaxf(v[3 + c], bhi[3 * f + c]);
      v[6 + c] = fminf(v[6 + c], cen[3 * f + c]);
      v[9 + c] = fmaxf(v[9 + c], cen[3 * f + c]);
    }
  for (int off = 32; off > 0; off >>= 1)
    for (int c = 0; c < 12; c++) {
      const float o = __shfl_xor(v[c], off);
      v[c]          = (c % 6) < 3 ? fminf(v[c], o) : fmaxf(v[c], o);
    }
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < 12; c++) {
      if ((c % 6) < 3) atomicMin(out + c, f2o(v[c]));
      else atomicMax(out + c, f2o(v[c]));
    }
}

__global__ void bins_init_kernel(int32_t* bins, uint64_t words)
{
  const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (w < words) bins[w] = bin_init(int(w % kBinWords));
}

__device__ __forceinline__ int bin_of(float c, float lo, float scale)
{
  const int k = int((c - lo) * scale);
  return min(kBins - 1, max(0, k));
}

// Accumulate the triangles [i0, i1) of a node (bounds nd) into LDS bins lb.
__device__ __forceinline__ void bin_range(int32_t* lb, const OpenNode& nd, const uint32_t* idx, const float* blo,
                                          const float* bhi, const float* cen, uint32_t i0, uint32_t i1, uint32_t step)
{
  float scale[3];
  bool  ok[3];
  for (int a = 0; a < 3; a++) {
    const float ext = nd.chi[a] - nd.clo[a];
    ok[a]           = ext > 0.0f;
    scale[a]        = ok[a] ? float(kBins) / ext : 0.0f;
  }
  for (uint32_t i = i0; i < i1; i += step) {
    const uint32_t p = idx[i];
    float          bl[3], bh[3], cc[3];
    for (int c = 0; c < 3; c++) {
      bl[c] = blo[3 * (size_t)p + c];
      bh[c] = bhi[3 * (size_t)p + c];
      cc[c] = cen[3 * (size_t)p + c];
    }
    for (int a = 0; a < 3; a++) {
      if (!ok[a]) continue;
      int32_t* b = lb + (a * kBins + bin_of(cc[a], nd.clo[a], scale[a])) * kBinFields;
      atomicAdd(b, 1);
      for (int c = 0; c < 3; c++) {
        atomicMin(b + 1 + c, f2o(bl[c]));
        atomicMax(b + 4 + c, f2o(bh[c]));
        atomicMin(b + 7 + c, f2o(cc[c]));
        atomicMax(b + 10 + c, f2o(cc[c]));
      }
    }
  }
}

__global__ __launch_bounds__(256) void sah_bin_kernel(const ChunkRec* __restrict__ chunks,
                                                      const OpenNode* __restrict__ open,
                                                      const uint32_t* __restrict__ idx, const float* __restrict__ blo,
                                                      const float* __restrict__ bhi, const float* __restrict__ cen,
                                                      int32_t* __restrict__ bins)
{
  __shared__ int32_t lb[kBinWords];
  for (int w = threadIdx.x; w < kBinWords; w += 256) lb[w] = bin_init(w);
  __syncthreads();
  const ChunkRec c  = chunks[blockIdx.x];
  const OpenNode nd = open[c.slot];
  bin_range(lb, nd, idx, blo, bhi, cen, c.begin + threadIdx.x, c.end, 256);
  __syncthreads();
  int32_t* g = bins + (size_t)c.slot * kBinWords;
  for (int w = threadIdx.x; w < kBinWords; w += 256) {
    const int32_t v = lb[w], init = bin_init(w);
    if (v == init) continue;
    const int f = w % kBinFields;
    if (f == 0) atomicAdd(g + w, v);
    else if ((f >= 1 && f <= 3) || (f >= 7 && f <= 9)) atomicMin(g + w, v);
    else atomicMax(g + w, v);
  }
}

__device__ __forceinline__ float box_area(const float* lo, const float* hi)
{
#pragma clang fp contract(off)
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// The host builder's split choice for a node from its bins (b: kBinWords ordered integers).
__device__ void choose_split(const OpenNode& nd, const int32_t* b, SplitOut& o)
{
#pragma clang fp contract(off)
  float bestCost = INFINITY;
  int   bestAxis = -1, bestBin = 0;
  if (nd.depth < kMedianDepth) {
    for (int axis = 0; axis < 3; axis++) {
      if (!(nd.chi[axis] - nd.clo[axis] > 0.0f)) continue;
      float    la[kBins];
      uint32_t lc[kBins];
      float    rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t rc    = 0;
      for (int k = 0; k < kBins - 1; k++) {
        const int32_t* q = b + (axis * kBins + k) * kBinFields;
        for (int c = 0; c < 3; c++) {
          rl[c] = fminf(rl[c], o2f(q[1 + c]));
          rh[c] = fmaxf(rh[c], o2f(q[4 + c]));
        }
        rc += (uint32_t)q[0];
        la[k] = rc ? box_area(rl, rh) : 0.0f;
        lc[k] = rc;
      }
      float    sl[3] = {INFINITY, INFINITY, INFINITY}, sh[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t sc    = 0;
      for (int k = kBins - 1; k >= 1; k--) {
        const int32_t* q = b + (axis * kBins + k) * kBinFields;
        for (int c = 0; c < 3; c++) {
          sl[c] = fminf(sl[c], o2f(q[1 + c]));
          sh[c] = fmaxf(sh[c], o2f(q[4 + c]));
        }
        sc += (uint32_t)q[0];
        if (lc[k - 1] && sc) {
          const float cost = la[k - 1] * float(lc[k - 1]) + box_area(sl, sh) * float(sc);
          if (cost < bestCost) {
            bestCost = cost;
            bestAxis = axis;
            bestBin  = k;
          }
        }
      }
    }
  }
  o.axis = bestAxis;
  o.bin  = bestBin;
  if (bestAxis < 0) {
    o.nleft = nd.count / 2;
    for (int c = 0; c < 3; c++) {
      o.llo[c] = o.rlo[c] = nd.lo[c];
      o.lhi[c] = o.rhi[c] = nd.hi[c];
      o.lclo[c] = o.rclo[c] = nd.clo[c];
      o.lchi[c] = o.rchi[c] = nd.chi[c];
    }
    return;
  }
  uint32_t nl = 0;
  for (int c = 0; c < 3; c++) {
    o.llo[c] = o.lclo[c] = o.rlo[c] = o.rclo[c] = INFINITY;
    o.lhi[c] = o.lchi[c] = o.rhi[c] = o.rchi[c] = -INFINITY;
  }
  for (int k = 0; k < kBins; k++) {
    const int32_t* q    = b + (bestAxis * kBins + k) * kBinFields;
    const bool     left = k < bestBin;
    if (left) nl += (uint32_t)q[0];
    float* lo  = left ? o.llo : o.rlo;
    float* hi  = left ? o.lhi : o.rhi;
    float* clo = left ? o.lclo : o.rclo;
    float* chi = left ? o.lchi : o.rchi;
    for (int c = 0; c < 3; c++) {
      lo[c]  = fminf(lo[c], o2f(q[1 + c]));
      hi[c]  = fmaxf(hi[c], o2f(q[4 + c]));
      clo[c] = fminf(clo[c], o2f(q[7 + c]));
      chi[c] = fmaxf(chi[c], o2f(q[10 + c]));
    }
  }
  o.nleft = nl;
}

__global__ __launch_bounds__(64) void sah_split_kernel(const OpenNode* __restrict__ open, uint32_t nopen,
                                                       const int32_t* __restrict__ bins, SplitOut* __restrict__ out)
{
  const uint32_t j = blockIdx.x * 64 + threadIdx.x;
  if (j >= nopen) return;
  SplitOut o{};
  choose_split(open[j], bins + (size_t)j * kBinWords, o);
  out[j] = o;
}

__device__ __forceinline__ bool goes_left(const OpenNode& nd, const SplitOut& o, const float* cen, uint32_t p,
                                          uint32_t pos)
{
  if (o.axis < 0) return pos - nd.first < o.nleft;
  const int   a     = o.axis;
  const float scale = float(kBins) / (nd.chi[a] - nd.clo[a]);
  return bin_of(cen[3 * (size_t)p + a], nd.clo[a], scale) < o.bin;
}

__global__ __launch_bounds__(256) void sah_count_kernel(const ChunkRec* __restrict__ chunks,
                                                        const OpenNode* __restrict__ open,
                                                        const SplitOut* __restrict__ splits,
                                                        const uint32_t* __restrict__ idx, const float* __restrict__ cen,
                                                        uint32_t* __restrict__ chunk_left)
{
  __shared__ uint32_t tot;
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  const ChunkRec c  = chunks[blockIdx.x];
  const OpenNode nd = open[c.slot];
  const SplitOut o  = splits[c.slot];
  uint32_t       n  = 0;
  for (uint32_t i = c.begin + threadIdx.x; i < c.end; i += 256) n += goes_left(nd, o, cen, idx[i], i) ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(&tot, n);
  __syncthreads();
  if (threadIdx.x == 0) chunk_left[blockIdx.x] = tot;
}

// Stable partition of one chunk: its left triangles go to first + off_left + rank, its right ones
// to first + nleft + off_right + rank (off_*: the node's earlier chunks' counts).
__global__ __launch_bounds__(256) void sah_scatter_kernel(const ChunkRec* __restrict__ chunks,
                                                          const OpenNode* __restrict__ open,
                                                          const SplitOut* __restrict__ splits,
                                                          const uint32_t* __restrict__ idx,
                                                          const float* __restrict__ cen,
                                                          const uint32_t* __restrict__ chunk_off,
                                                          uint32_t* __restrict__ tmp)
{
  __shared__ uint32_t wl[4], wv[4];
  __shared__ uint32_t base_l, base_r;
  const ChunkRec c    = chunks[blockIdx.x];
  const OpenNode nd   = open[c.slot];
  const SplitOut o    = splits[c.slot];
  const int      lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    base_l = nd.first + chunk_off[2 * blockIdx.x];
    base_r = nd.first + o.nleft + chunk_off[2 * blockIdx.x + 1];
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
  for (uint32_t t0 = c.begin; t0 < c.end; t0 += 256) {
    const uint32_t i     = t0 + threadIdx.x;
    const bool     valid = i < c.end;
    const uint32_t p     = valid ? idx[i] : 0u;
    const bool     left  = valid && goes_left(nd, o, cen, p, i);
    const uint64_t ml = __ballot(left), mv = __ballot(valid);
    if (lane == 0) {
      wl[wave] = (uint32_t)__popcll(ml);
      wv[wave] = (uint32_t)__popcll(mv);
    }
    __syncthreads();
    uint32_t offl = 0, offv = 0, totl = 0, totv = 0;
    for (int w = 0; w < 4; w++) {
      if (w < wave) {
        offl += wl[w];
        offv += wv[w];
      }
      totl += wl[w];
      totv += wv[w];
    }
    const uint32_t rl = offl + (uint32_t)__popcll(ml & below);
    const uint32_t rv = offv + (uint32_t)__popcll(mv & below);
    if (valid) {
      if (left) tmp[base_l + rl] = p;
      else tmp[base_r + (rv - rl)] = p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      base_l += totl;
      base_r += totv - totl;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void sah_copy_kernel(const ChunkRec* __restrict__ chunks,
                                                       const uint32_t* __restrict__ tmp, uint32_t* __restrict__ idx)
{
  const ChunkRec c = chunks[blockIdx.x];
  for (uint32_t i = c.begin + threadIdx.x; i < c.end; i += 256) idx[i] = tmp[i];
}

// Small subtrees: one workgroup builds the whole subtree of one node of at most kSmall
// triangles, depth first, with the triangles' boxes and centroids in LDS and a stable partition
// of a local permutation per split.
struct SmallEntry
{
  float    lo[3], hi[3], clo[3], chi[3];
  uint32_t b, e, node, depth;
};
static_assert(sizeof(SmallEntry) == 64, "SmallEntry layout");

struct SmallLds
{
  float      cen[kSmall][3];
  float      blo[kSmall][3];
  float      bhi[kSmall][3];
  uint32_t   pidx[kSmall];
  uint16_t   perm[kSmall], perm2[kSmall];
  int32_t    bins[kBinWords];
  SmallEntry stack[kSmallStack];
  SmallEntry cur;
  SplitOut   split;
  uint32_t   wl[4], wv[4];
  int        sp, done;
};

__global__ __launch_bounds__(256) void sah_small_kernel(const OpenNode* __restrict__ roots, uint32_t* __restrict__ idx,
                                                        const float* __restrict__ blo, const float* __restrict__ bhi,
                                                        const float* __restrict__ cen, BinNode* __restrict__ dn,
                                                        uint32_t* __restrict__ counter, uint32_t leaf_size,
                                                        uint32_t* __restrict__ status)
{
  __shared__ SmallLds s;
  const OpenNode r    = roots[blockIdx.x];
  const int      lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < r.count; i += 256) {
    const uint32_t p = idx[r.first + i];
    s.pidx[i]        = p;
    s.perm[i]        = (uint16_t)i;
    for (int c = 0; c < 3; c++) {
      s.cen[i][c] = cen[3 * (size_t)p + c];
      s.blo[i][c] = blo[3 * (size_t)p + c];
      s.bhi[i][c] = bhi[3 * (size_t)p + c];
    }
  }
  if (threadIdx.x == 0) {
    SmallEntry e;
    for (int c = 0; c < 3; c++) {
      e.lo[c]  = r.lo[c];
      e.hi[c]  = r.hi[c];
      e.clo[c] = r.clo[c];
      e.chi[c] = r.chi[c];
    }
    e.b = 0, e.e = r.count, e.node = r.id, e.depth = r.depth;
    s.stack[0] = e;
    s.sp       = 1;
    s.done     = 0;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      if (s.sp == 0) s.done = 1;
      else s.cur = s.stack[--s.sp];
    }
    __syncthreads();
    if (s.done) break;
    const SmallEntry cur = s.cur;
    const uint32_t   n   = cur.e - cur.b;
    if (n <= leaf_size) {
      if (threadIdx.x == 0) {
        BinNode& d = dn[cur.node];
        for (int c = 0; c < 3; c++) {
          d.lo[c] = cur.lo[c];
          d.hi[c] = cur.hi[c];
        }
        d.left = d.right = -1;
        d.first          = r.first + cur.b;
        d.count          = n;
      }
      continue;
    }
    // bins of this node
    for (int w = threadIdx.x; w < kBinWords; w += 256) s.bins[w] = bin_init(w);
    __syncthreads();
    {
      float scale[3];
      bool  ok[3];
      for (int a = 0; a < 3; a++) {
        const float ext = cur.chi[a] - cur.clo[a];
        ok[a]           = ext > 0.0f;
        scale[a]        = ok[a] ? float(kBins) / ext : 0.0f;
      }
      for (uint32_t i = cur.b + threadIdx.x; i < cur.e; i += 256) {
        const uint32_t l = s.perm[i];
        for (int a = 0; a < 3; a++) {
          if (!ok[a]) continue;
          int32_t* b = s.bins + (a * kBins + bin_of(s.cen[l][a], cur.clo[a], scale[a])) * kBinFields;
          atomicAdd(b, 1);
          for (int c = 0; c < 3; c++) {
            atomicMin(b + 1 + c, f2o(s.blo[l][c]));
            atomicMax(b + 4 + c, f2o(s.bhi[l][c]));
            atomicMin(b + 7 + c, f2o(s.cen[l][c]));
            atomicMax(b + 10 + c, f2o(s.cen[l][c]));
          }
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      OpenNode nd;
      for (int c = 0; c < 3; c++) {
        nd.lo[c]  = cur.lo[c];
        nd.hi[c]  = cur.hi[c];
        nd.clo[c] = cur.clo[c];
        nd.chi[c] = cur.chi[c];
      }
      nd.first = cur.b, nd.count = n, nd.depth = cur.depth, nd.id = cur.node;
      choose_split(nd, s.bins, s.split);
    }
    __syncthreads();
    const SplitOut o = s.split;
    if (o.axis >= 0) {
      // stable partition of perm[b, e) into perm2, then back
      const int   a     = o.axis;
      const float scale = float(kBins) / (cur.chi[a] - cur.clo[a]);
      uint32_t    bl = cur.b, br = cur.b + o.nleft;
      for (uint32_t t0 = cur.b; t0 < cur.e; t0 += 256) {
        const uint32_t i     = t0 + threadIdx.x;
        const bool     valid = i < cur.e;
        const uint16_t l     = valid ? s.perm[i] : (uint16_t)0;
        const bool     left  = valid && bin_of(s.cen[l][a], cur.clo[a], scale) < o.bin;
        const uint64_t ml = __ballot(left), mv = __ballot(valid);
        if (lane == 0) {
          s.wl[wave] = (uint32_t)__popcll(ml);
          s.wv[wave] = (uint32_t)__popcll(mv);
        }
        __syncthreads();
        uint32_t offl = 0, offv = 0, totl = 0, totv = 0;
        for (int w = 0; w < 4; w++) {
          if (w < wave) {
            offl += s.wl[w];
            offv += s.wv[w];
          }
          totl += s.wl[w];
          totv += s.wv[w];
        }
        const uint32_t rl = offl + (uint32_t)__popcll(ml & below);
        const uint32_t rv = offv + (uint32_t)__popcll(mv & below);
        if (valid) {
          if (left) s.perm2[bl + rl] = l;
          else s.perm2[br + (rv - rl)] = l;
        }
        bl += totl;
        br += totv - totl;
        __syncthreads();
      }
      for (uint32_t i = cur.b + threadIdx.x; i < cur.e; i += 256) s.perm[i] = s.perm2[i];
    }
    if (threadIdx.x == 0) {
      const uint32_t ch = atomicAdd(counter, 2u);
      BinNode&       d  = dn[cur.node];
      for (int c = 0; c < 3; c++) {
        d.lo[c] = cur.lo[c];
        d.hi[c] = cur.hi[c];
      }
      d.left  = (int32_t)ch;
      d.right = (int32_t)ch + 1;
      d.first = r.first + cur.b;
      d.count = n;
      if (s.sp + 2 > kSmallStack) {
        atomicOr(status, kBadStack);
        s.sp = 0; // abandon this subtree; the host reports the failure
      }
      else {
        SmallEntry L, R;
        for (int c = 0; c < 3; c++) {
          L.lo[c] = o.llo[c], L.hi[c] = o.lhi[c], L.clo[c] = o.lclo[c], L.chi[c] = o.lchi[c];
          R.lo[c] = o.rlo[c], R.hi[c] = o.rhi[c], R.clo[c] = o.rclo[c], R.chi[c] = o.rchi[c];
        }
        L.b = cur.b, L.e = cur.b + o.nleft, L.node = ch, L.depth = cur.depth + 1;
        R.b = cur.b + o.nleft, R.e = cur.e, R.node = ch + 1, R.depth = cur.depth + 1;
        s.stack[s.sp++] = R;
        s.stack[s.sp++] = L; // left first
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < r.count; i += 256) idx[r.first + i] = s.pidx[s.perm[i]];
}

// ---------------------------------------------------------------------------------------------
// packing
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ Dv3 dsub(Dv3 a, Dv3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dlen(Dv3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ Dv3 dcross(Dv3 a, Dv3 u)
{
  return {a.y * u.z - a.z * u.y, a.z * u.x - a.x * u.z, a.x * u.y - a.y * u.x};
}

// The per-triangle arrays in leaf order (the host's pack_triangles, operation for operation).
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                   const uint32_t* __restrict__ order, const double* __restrict__ fnf,
                                                   uint64_t nt, TriRec* __restrict__ tris, float* __restrict__ tripos,
                                                   uint32_t* __restrict__ feat, double* __restrict__ fn64,
                                                   float* __restrict__ tsph, double* __restrict__ area,
                                                   unsigned long long* __restrict__ hmax_bits)
{
#pragma clang fp contract(off)
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nt) return;
  const uint32_t f  = order[i];
  const float*   p0 = xyz + 3 * (size_t)tri[3 * f];
  const float*   p1 = xyz + 3 * (size_t)tri[3 * f + 1];
  const float*   p2 = xyz + 3 * (size_t)tri[3 * f + 2];
  const Dv3      x0 = ld3(p0), x1 = ld3(p1), x2 = ld3(p2);
  const Dv3      e[3] = {dsub(x1, x0), dsub(x2, x1), dsub(x0, x2)};
  const Dv3      n    = {fnf[3 * f], fnf[3 * f + 1], fnf[3 * f + 2]};
  TriRec         r;
  auto put = [](float* d, Dv3 v) {
    d[0] = float(v.x);
    d[1] = float(v.y);
    d[2] = float(v.z);
  };
  for (int c = 0; c < 3; c++) r.v0[c] = p0[c];
  put(r.e0, e[0]);
  put(r.e1, e[1]);
  put(r.e2, e[2]);
  put(r.n, n);
  // edge normals in the face plane, pointing out of the triangle
  const Dv3    w  = dcross(e[0], e[1]);
  const double sg = w.x * n.x + w.y * n.y + w.z * n.z >= 0.0 ? 1.0 : -1.0;
  auto         sc = [sg](Dv3 a) { return Dv3{sg * a.x, sg * a.y, sg * a.z}; };
  put(r.c0, sc(dcross(e[0], n)));
  put(r.c1, sc(dcross(e[1], n)));
  put(r.c2, sc(dcross(e[2], n)));
  for (int c = 0; c < 3; c++) r.inv[c] = float(1.0 / (e[c].x * e[c].x + e[c].y * e[c].y + e[c].z * e[c].z));
  r.face      = f;
  const Dv3 cen = {double(float((x0.x + x1.x + x2.x) / 3.0)), double(float((x0.y + x1.y + x2.y) / 3.0)),
                   double(float((x0.z + x1.z + x2.z) / 3.0))};
  const double rad = fmax(dlen(dsub(x0, cen)), fmax(dlen(dsub(x1, cen)), dlen(dsub(x2, cen))));
  put(r.sc, cen);
  r.sr    = float(rad * (1.0 + 1e-5) + 1e-7 * (dlen(cen) + rad));
  tris[i] = r;
  float* sp = tsph + 8 * i;
  for (int c = 0; c < 3; c++) sp[c] = r.sc[c];
  sp[3] = r.sr;
  for (int c = 0; c < 3; c++) sp[4 + c] = r.n[c];
  sp[7] = float(-(double(r.n[0]) * r.v0[0] + double(r.n[1]) * r.v0[1] + double(r.n[2]) * r.v0[2]));
  for (int v = 0; v < 3; v++)
    for (int c = 0; c < 3; c++) tripos[12 * i + 3 * v + c] = xyz[3 * (size_t)tri[3 * f + v] + c];
  for (int c = 9; c < 12; c++) tripos[12 * i + c] = 0.0f;
  for (int v = 0; v < 3; v++) {
    feat[8 * i + v]     = tri[3 * f + v];
    feat[8 * i + 3 + v] = 3 * f + v;
  }
  feat[8 * i + 6] = f;
  feat[8 * i + 7] = 0;
  for (int c = 0; c < 3; c++) fn64[4 * i + c] = fnf[3 * f + c];
  fn64[4 * i + 3] = 0.0;
  const double h  = fmax(dlen(e[0]), fmax(dlen(e[1]), dlen(e[2])));
  atomicMax(hmax_bits, (unsigned long long)__double_as_longlong(h));
  // area from the float positions, as the training sampler's CDF takes it
  const float* x  = tripos + 12 * i;
  const double ax = double(x[3]) - x[0], ay = double(x[4]) - x[1], az = double(x[5]) - x[2];
  const double bx = double(x[6]) - x[0], by = double(x[7]) - x[1], bz = double(x[8]) - x[2];
  const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
  area[i]         = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
}

// Normalised inclusive prefix sum of the areas (one workgroup: each thread a contiguous run).
__global__ __launch_bounds__(1024) void area_cdf_kernel(const double* __restrict__ area, uint64_t n,
                                                        float* __restrict__ cdf)
{
  __shared__ double part[1024];
  __shared__ double total;
  const uint64_t per = (n + 1023) / 1024;
  const uint64_t b = min(n, threadIdx.x * per), e = min(n, b + per);
  double         s = 0.0;
  for (uint64_t i = b; i < e; i++) s += area[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double run = 0.0;
    for (int t = 0; t < 1024; t++) {
      const double v = part[t];
      part[t]        = run; // exclusive prefix
      run += v;
    }
    total = run;
  }
  __syncthreads();
  double run = part[threadIdx.x];
  for (uint64_t i = b; i < e; i++) {
    run += area[i];
    cdf[i] = float(run / total);
  }
  if (e == n && e > b) cdf[n - 1] = 1.0f;
}

template <class T>
hipError_t zeros(DevBuf& b, std::size_t count)
{
  hipError_t e = b.reserve(std::max<std::size_t>(64, count * sizeof(T)));
  if (e != hipSuccess) return e;
  return hipMemset(b.p, 0, std::max<std::size_t>(64, count * sizeof(T)));
}

inline uint32_t blocks(uint64_t n, uint32_t per = 256) { return (uint32_t)((n + per - 1) / per); }

int bits_for(uint64_t maxval)
{
  int b = 1;
  while (b < 64 && (maxval >> b) != 0) b++;
  return b;
}

} // namespace

#define EBG_GB(x)                                                                                             \
  do {                                                                                                        \
    const hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) {                                                                                   \
      err = std::string(#x) + ": " + hipGetErrorString(e_);                                                  \
      return e_ == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR;                                             \
    }                                                                                                         \
  } while (0)

ebg_status build_mesh_gpu_sah(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt, int K,
                              int leaf_size, const PseudoNormals* pn, GpuBuiltMesh& out, std::string& err)
{
  using clk = std::chrono::steady_clock;
  if (K != 4 && K != 8) {
    err = "branching must be 4 or 8";
    return EBG_INVALID_ARG;
  }
  if (leaf_size < 1 || leaf_size > kMaxLeaf || nt < 1) {
    err = "leaf_size must be in 1..16 and the mesh must have a triangle";
    return EBG_INVALID_ARG;
  }
  hipStream_t s = nullptr;
  DevBuf      d_xyz, d_tri, d_fn, d_status;
  EBG_GB(upload(d_xyz, std::vector<float>(xyz, xyz + 3 * nv)));
  EBG_GB(upload(d_tri, std::vector<uint32_t>(tri, tri + 3 * nt)));
  EBG_GB(upload(d_status, std::vector<uint32_t>{0u, 0xFFFFFFFFu, 0u, 0u}));

  // ---- pseudo-normals --------------------------------------------------------------------
  const auto t0 = clk::now();
  EBG_GB(zeros<double>(out.vn64, 4 * nv));
  EBG_GB(zeros<float>(out.vn32, 4 * nv));
  EBG_GB(zeros<double>(out.en64, 12 * nt));
  EBG_GB(zeros<float>(out.en32, 12 * nt));
  if (pn) {
    std::vector<double> fn(pn->face), vn64(4 * nv, 0.0), en64(12 * nt, 0.0);
    std::vector<float>  vn32(4 * nv, 0.0f), en32(12 * nt, 0.0f);
    for (std::size_t v = 0; v < nv; v++)
      for (int c = 0; c < 3; c++) vn32[4 * v + c] = float(vn64[4 * v + c] = pn->vert[3 * v + c]);
    for (std::size_t h = 0; h < 3 * nt; h++)
      for (int c = 0; c < 3; c++) en32[4 * h + c] = float(en64[4 * h + c] = pn->edge[3 * h + c]);
    EBG_GB(upload(d_fn, fn));
    EBG_GB(upload(out.vn64, vn64));
    EBG_GB(upload(out.vn32, vn32));
    EBG_GB(upload(out.en64, en64));
    EBG_GB(upload(out.en32, en32));
  }
  else {
    EBG_GB(d_fn.reserve(24 * nt));
    hipLaunchKernelGGL(face_normals_kernel, dim3(blocks(nt)), dim3(256), 0, s, d_xyz.as<float>(),
                       d_tri.as<uint32_t>(), (uint64_t)nt, d_fn.as<double>(), d_status.as<uint32_t>());
    EBG_GB(hipGetLastError());
    uint32_t st[2];
    EBG_GB(hipMemcpy(st, d_status.p, 8, hipMemcpyDeviceToHost));
    if (st[0] & kBadFace) {
      err = "degenerate triangle " + std::to_string(st[1]);
      return EBG_INVALID_ARG;
    }
    const uint64_t m = 3 * nt;
    DevBuf         keys, skeys, svals, scratch;
    EBG_GB(keys.reserve(8 * m));
    EBG_GB(skeys.reserve(8 * m));
    EBG_GB(svals.reserve(4 * m));
    EBG_GB(scratch.reserve(radix_sort_scratch_bytes(m)));
    hipLaunchKernelGGL(edge_keys_kernel, dim3(blocks(m)), dim3(256), 0, s, d_tri.as<uint32_t>(), (uint64_t)nt,
                       (uint64_t)nv, keys.as<uint64_t>());
    EBG_GB(hipGetLastError());
    const uint64_t nvv = (uint64_t)nv;
    EBG_GB(radix_sort_pairs(keys.as<uint64_t>(), nullptr, skeys.as<uint64_t>(), svals.as<uint32_t>(), m,
                            bits_for(nvv * nvv - 1), scratch.p, s));
    hipLaunchKernelGGL(edge_normals_kernel, dim3(blocks(m)), dim3(256), 0, s, d_tri.as<uint32_t>(), (uint64_t)nt,
                       (uint64_t)nv, skeys.as<uint64_t>(), svals.as<uint32_t>(), d_fn.as<double>(),
                       out.en64.as<double>(), out.en32.as<float>());
    EBG_GB(hipGetLastError());
    // corners by vertex: keys in the low half of the (now free) 64-bit key buffers
    uint32_t* ck  = keys.as<uint32_t>();
    uint32_t* sck = skeys.as<uint32_t>();
    hipLaunchKernelGGL(corner_keys_kernel, dim3(blocks(m)), dim3(256), 0, s, d_tri.as<uint32_t>(), (uint64_t)nt, ck);
    EBG_GB(hipGetLastError());
    EBG_GB(radix_sort_pairs32(ck, nullptr, sck, svals.as<uint32_t>(), m, bits_for(nvv - 1), scratch.p, true, s));
    hipLaunchKernelGGL(vertex_normals_kernel, dim3(blocks(m)), dim3(256), 0, s, sck, svals.as<uint32_t>(), m,
                       d_tri.as<uint32_t>(), d_xyz.as<float>(), d_fn.as<double>(), out.vn64.as<double>(),
                       out.vn32.as<float>());
    EBG_GB(hipGetLastError());
  }
  EBG_GB(hipStreamSynchronize(s));
  const auto t1 = clk::now();

  // ---- binned SAH --------------------------------------------------------------------------
  DevBuf d_blo, d_bhi, d_cen, d_idx, d_tmp, d_rootb, d_dn, d_counter;
  EBG_GB(d_blo.reserve(12 * nt));
  EBG_GB(d_bhi.reserve(12 * nt));
  EBG_GB(d_cen.reserve(12 * nt));
  EBG_GB(d_idx.reserve(4 * nt));
  EBG_GB(d_tmp.reserve(4 * nt));
  hipLaunchKernelGGL(prims_kernel, dim3(blocks(nt)), dim3(256), 0, s, d_xyz.as<float>(), d_tri.as<uint32_t>(),
                     (uint64_t)nt, d_blo.as<float>(), d_bhi.as<float>(), d_cen.as<float>(), d_idx.as<uint32_t>());
  EBG_GB(hipGetLastError());
  {
    std::vector<int32_t> init(12);
    for (int c = 0; c < 12; c++) {
      const float v  = (c % 6) < 3 ? INFINITY : -INFINITY;
      int32_t     iv;
      std::memcpy(&iv, &v, 4);
      init[c] = iv >= 0 ? iv : iv ^ 0x7FFFFFFF;
    }
    EBG_GB(upload(d_rootb, init));
  }
  hipLaunchKernelGGL(root_bounds_kernel, dim3(std::min<uint32_t>(1024, blocks(nt))), dim3(256), 0, s,
                     d_blo.as<float>(), d_bhi.as<float>(), d_cen.as<float>(), (uint64_t)nt, d_rootb.as<int32_t>());
  EBG_GB(hipGetLastError());
  int32_t rb[12];
  EBG_GB(hipMemcpy(rb, d_rootb.p, sizeof(rb), hipMemcpyDeviceToHost));
  auto o2f_h = [](int32_t i) {
    const int32_t b = i >= 0 ? i : i ^ 0x7FFFFFFF;
    float         f;
    std::memcpy(&f, &b, 4);
    return f;
  };

  std::vector<BinNode>  bn;
  std::vector<OpenNode> open, small;
  bn.reserve(1024);
  {
    OpenNode r{};
    for (int c = 0; c < 3; c++) {
      r.lo[c]  = o2f_h(rb[c]);
      r.hi[c]  = o2f_h(rb[3 + c]);
      r.clo[c] = o2f_h(rb[6 + c]);
      r.chi[c] = o2f_h(rb[9 + c]);
    }
    r.first = 0, r.count = uint32_t(nt), r.depth = 0, r.id = 0;
    BinNode b{};
    for (int c = 0; c < 3; c++) {
      b.lo[c] = r.lo[c];
      b.hi[c] = r.hi[c];
    }
    b.left = b.right = -1;
    b.first = 0, b.count = uint32_t(nt);
    bn.push_back(b);
    if (nt <= (std::size_t)leaf_size) {
    }
    else if (nt <= kSmall) small.push_back(r);
    else open.push_back(r);
  }

  DevBuf d_open, d_chunks, d_bins, d_splits, d_cleft, d_coff;
  while (!open.empty()) {
    std::vector<ChunkRec> chunks;
    for (uint32_t j = 0; j < open.size(); j++)
      for (uint32_t b = open[j].first, e = open[j].first + open[j].count; b < e; b += kChunk)
        chunks.push_back({j, b, std::min(e, b + kChunk), 0u});
    const uint32_t nopen = uint32_t(open.size()), nch = uint32_t(chunks.size());
    EBG_GB(upload(d_open, open));
    EBG_GB(upload(d_chunks, chunks));
    EBG_GB(d_bins.reserve(4ull * kBinWords * nopen));
    EBG_GB(d_splits.reserve(sizeof(SplitOut) * nopen));
    EBG_GB(d_cleft.reserve(4ull * nch));
    hipLaunchKernelGGL(bins_init_kernel, dim3(blocks((uint64_t)kBinWords * nopen)), dim3(256), 0, s,
                       d_bins.as<int32_t>(), (uint64_t)kBinWords * nopen);
    EBG_GB(hipGetLastError());
    hipLaunchKernelGGL(sah_bin_kernel, dim3(nch), dim3(256), 0, s, d_chunks.as<ChunkRec>(), d_open.as<OpenNode>(),
                       d_idx.as<uint32_t>(), d_blo.as<float>(), d_bhi.as<float>(), d_cen.as<float>(),
                       d_bins.as<int32_t>());
    EBG_GB(hipGetLastError());
    hipLaunchKernelGGL(sah_split_kernel, dim3(blocks(nopen, 64)), dim3(64), 0, s, d_open.as<OpenNode>(), nopen,
                       d_bins.as<int32_t>(), d_splits.as<SplitOut>());
    EBG_GB(hipGetLastError());
    hipLaunchKernelGGL(sah_count_kernel, dim3(nch), dim3(256), 0, s, d_chunks.as<ChunkRec>(), d_open.as<OpenNode>(),
                       d_splits.as<SplitOut>(), d_idx.as<uint32_t>(), d_cen.as<float>(), d_cleft.as<uint32_t>());
    EBG_GB(hipGetLastError());
    std::vector<SplitOut> splits(nopen);
    std::vector<uint32_t> cleft(nch);
    EBG_GB(hipMemcpy(splits.data(), d_splits.p, sizeof(SplitOut) * nopen, hipMemcpyDeviceToHost));
    EBG_GB(hipMemcpy(cleft.data(), d_cleft.p, 4ull * nch, hipMemcpyDeviceToHost));
    // the chunk counts must add up to each node's left count from its bins
    {
      std::vector<uint64_t> sum(nopen, 0);
      for (uint32_t c = 0; c < nch; c++) sum[chunks[c].slot] += cleft[c];
      for (uint32_t j = 0; j < nopen; j++)
        if (sum[j] != splits[j].nleft || splits[j].nleft == 0 || splits[j].nleft >= open[j].count) {
          err = "GPU SAH build: inconsistent split";
          return EBG_HIP_ERROR;
        }
    }
    std::vector<uint32_t> coff(2 * nch);
    {
      uint32_t cur = ~0u, accl = 0, accr = 0;
      for (uint32_t c = 0; c < nch; c++) {
        if (chunks[c].slot != cur) {
          cur  = chunks[c].slot;
          accl = accr = 0;
        }
        coff[2 * c]     = accl;
        coff[2 * c + 1] = accr;
        accl += cleft[c];
        accr += (chunks[c].end - chunks[c].begin) - cleft[c];
      }
    }
    EBG_GB(upload(d_coff, coff));
    hipLaunchKernelGGL(sah_scatter_kernel, dim3(nch), dim3(256), 0, s, d_chunks.as<ChunkRec>(),
                       d_open.as<OpenNode>(), d_splits.as<SplitOut>(), d_idx.as<uint32_t>(), d_cen.as<float>(),
                       d_coff.as<uint32_t>(), d_tmp.as<uint32_t>());
    EBG_GB(hipGetLastError());
    hipLaunchKernelGGL(sah_copy_kernel, dim3(nch), dim3(256), 0, s, d_chunks.as<ChunkRec>(), d_tmp.as<uint32_t>(),
                       d_idx.as<uint32_t>());
    EBG_GB(hipGetLastError());

    std::vector<OpenNode> next;
    for (uint32_t j = 0; j < nopen; j++) {
      const OpenNode& nd = open[j];
      const SplitOut& o  = splits[j];
      const uint32_t  nl = o.nleft;
      const int32_t l = int32_t(bn.size()), r = l + 1;
      bn[nd.id].left  = l;
      bn[nd.id].right = r;
      for (int side = 0; side < 2; side++) {
        OpenNode ch{};
        for (int c = 0; c < 3; c++) {
          ch.lo[c]  = side ? o.rlo[c] : o.llo[c];
          ch.hi[c]  = side ? o.rhi[c] : o.lhi[c];
          ch.clo[c] = side ? o.rclo[c] : o.lclo[c];
          ch.chi[c] = side ? o.rchi[c] : o.lchi[c];
        }
        ch.first = side ? nd.first + nl : nd.first;
        ch.count = side ? nd.count - nl : nl;
        ch.depth = nd.depth + 1;
        ch.id    = uint32_t(side ? r : l);
        BinNode b{};
        for (int c = 0; c < 3; c++) {
          b.lo[c] = ch.lo[c];
          b.hi[c] = ch.hi[c];
        }
        b.left = b.right = -1;
        b.first          = ch.first;
        b.count          = ch.count;
        bn.push_back(b);
        if (ch.count <= uint32_t(leaf_size)) continue;
        if (ch.count <= kSmall) small.push_back(ch);
        else next.push_back(ch);
      }
    }
    open.swap(next);
  }

  // small subtrees: the node array gets the level nodes, the kernel appends the rest
  EBG_GB(d_dn.reserve(sizeof(BinNode) * (2 * nt + 2)));
  EBG_GB(hipMemcpy(d_dn.p, bn.data(), sizeof(BinNode) * bn.size(), hipMemcpyHostToDevice));
  EBG_GB(upload(d_counter, std::vector<uint32_t>{uint32_t(bn.size())}));
  if (!small.empty()) {
    EBG_GB(upload(d_open, small));
    hipLaunchKernelGGL(sah_small_kernel, dim3(uint32_t(small.size())), dim3(256), 0, s, d_open.as<OpenNode>(),
                       d_idx.as<uint32_t>(), d_blo.as<float>(), d_bhi.as<float>(), d_cen.as<float>(),
                       d_dn.as<BinNode>(), d_counter.as<uint32_t>(), uint32_t(leaf_size), d_status.as<uint32_t>());
    EBG_GB(hipGetLastError());
  }
  uint32_t total = 0, st = 0;
  EBG_GB(hipMemcpy(&total, d_counter.p, 4, hipMemcpyDeviceToHost));
  EBG_GB(hipMemcpy(&st, d_status.p, 4, hipMemcpyDeviceToHost));
  if (st & kBadStack) {
    err = "GPU SAH build: subtree stack exhausted";
    return EBG_HIP_ERROR;
  }
  if (total > 2 * nt + 2) {
    err = "GPU SAH build: node count out of range";
    return EBG_HIP_ERROR;
  }
  bn.resize(total);
  EBG_GB(hipMemcpy(bn.data(), d_dn.p, sizeof(BinNode) * total, hipMemcpyDeviceToHost));
  WideBvh w;
  collapse_binary(bn, 0, K, w);
  std::vector<uint32_t> order(nt);
  EBG_GB(hipMemcpy(order.data(), d_idx.p, 4 * nt, hipMemcpyDeviceToHost));
  const auto t2 = clk::now();

  // ---- packing -------------------------------------------------------------------------------
  DevBuf d_area, d_hmax;
  EBG_GB(out.tris.reserve(sizeof(TriRec) * nt));
  EBG_GB(out.tripos.reserve(48 * nt));
  EBG_GB(out.feat.reserve(32 * nt));
  EBG_GB(out.fn64.reserve(32 * nt));
  EBG_GB(zeros<float>(out.tsph, 8 * (nt + kSphPad)));
  EBG_GB(d_area.reserve(8 * nt));
  EBG_GB(out.area_cdf.reserve(4 * (nt + 1)));
  EBG_GB(upload(d_hmax, std::vector<unsigned long long>{0ull}));
  hipLaunchKernelGGL(pack_kernel, dim3(blocks(nt)), dim3(256), 0, s, d_xyz.as<float>(), d_tri.as<uint32_t>(),
                     d_idx.as<uint32_t>(), d_fn.as<double>(), (uint64_t)nt, out.tris.as<TriRec>(),
                     out.tripos.as<float>(), out.feat.as<uint32_t>(), out.fn64.as<double>(), out.tsph.as<float>(),
                     d_area.as<double>(), d_hmax.as<unsigned long long>());
  EBG_GB(hipGetLastError());
  hipLaunchKernelGGL(area_cdf_kernel, dim3(1), dim3(1024), 0, s, d_area.as<double>(), (uint64_t)nt,
                     out.area_cdf.as<float>());
  EBG_GB(hipGetLastError());
  unsigned long long hb = 0;
  EBG_GB(hipMemcpy(&hb, d_hmax.p, 8, hipMemcpyDeviceToHost));
  double hmax;
  std::memcpy(&hmax, &hb, 8);
  const auto t3 = clk::now();

  PackedMesh& pm = out.meta;
  pm.K           = K;
  pm.leaf_size   = leaf_size;
  pm.num_nodes   = w.num_nodes;
  pm.num_leaves  = w.num_leaves;
  pm.stack_need  = w.stack_need;
  pm.nodes       = std::move(w.nodes);
  for (int c = 0; c < 3; c++) {
    pm.lo[c] = w.lo[c];
    pm.hi[c] = w.hi[c];
  }
  pm.diag        = w.diag;
  pm.hmax        = float(hmax * (1.0 + 1e-6));
  pm.order       = std::move(order);
  out.ms_normals = std::chrono::duration<double, std::milli>(t1 - t0).count();
  out.ms_sah     = std::chrono::duration<double, std::milli>(t2 - t1).count();
  out.ms_pack    = std::chrono::duration<double, std::milli>(t3 - t2).count();
  return EBG_OK;
}

} // namespace ebgpu
