// Small device utilities used by the query pipeline. Counters are cleared with a kernel (not
// hipMemsetAsync) so that a captured hipGraph replays them correctly.
#include "kernels.hpp"

#include <hip/hip_runtime.h>

namespace ebgpu {

namespace {

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// v[i] <- map[v[i]] for entries inside the map (kInvalid and other out-of-map values kept)
__global__ void map_u32_kernel(uint32_t* v, uint64_t n, const uint32_t* map, uint32_t nmap)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const uint32_t x = v[i];
    if (x < nmap) v[i] = map[x];
  }
}

__global__ void add_counts_kernel(unsigned long long* work, const uint32_t* a, const uint32_t* b)
{
  if (threadIdx.x == 0) {
    if (a) work[2] += *a;
    if (b) work[3] += *b;
  }
}

// FP64 points rounded to FP32 (round to nearest even, as the host's (float) cast); FP32 results
// widened to FP64.
__global__ void f64_to_f32_kernel(const double* __restrict__ in, float* __restrict__ out, uint64_t n)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = __double2float_rn(in[i]);
}

__global__ void f32_to_f64_kernel(const float* __restrict__ in, double* __restrict__ out, uint64_t n)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (double)in[i];
}

// Lattice points of a box of cells or nodes (x fastest, then y, then z: the layout of an AMReX
// or Chombo FAB): point (i, j, k) = origin + (i, j, k) * spacing, each coordinate a separately
// rounded FP64 product and sum, then rounded to FP32 (round to nearest even).
__global__ void grid_points_kernel(GridSpec g, float* __restrict__ out)
{
  const uint64_t n = g.dims[0] * g.dims[1] * g.dims[2];
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const uint64_t i = p % g.dims[0], r = p / g.dims[0];
  const uint64_t j = r % g.dims[1], k = r / g.dims[1];
  out[3 * p]       = __double2float_rn(__dadd_rn(g.origin[0], __dmul_rn((double)i, g.spacing[0])));
  out[3 * p + 1]   = __double2float_rn(__dadd_rn(g.origin[1], __dmul_rn((double)j, g.spacing[1])));
  out[3 * p + 2]   = __double2float_rn(__dadd_rn(g.origin[2], __dmul_rn((double)k, g.spacing[2])));
}

// Slab of each leaf about its first triangle's plane (n, d of that triangle's culling record):
// every vertex x of the leaf has |n.x + d| <= t, computed in FP64 from the exact vertex
// positions, then widened so that the FP32 test |n.p + d| <= rb + t of the traversal never
// rejects a leaf within rb of p: by 2^-20 (|d| + max |x|_1) for the FP32 evaluation of n.p + d
// and the rounding of rb + t, then rounded up to FP32. (The part of the FP32 error that scales
// with |p - x|, and |n| != 1, are covered by the 1e-6 the traversal adds to rb.)
template <int K>
__global__ __launch_bounds__(256) void leaf_slab_kernel(const Node<K>* __restrict__ nodes, uint32_t num_nodes,
                                                        const float* __restrict__ tsph,
                                                        const float* __restrict__ tripos, float* __restrict__ lthick)
{
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (uint64_t)num_nodes * K) return;
  const Node<K>& nd  = nodes[g / K];
  const uint32_t k   = (uint32_t)(g % K);
  const uint32_t cnt = nd.cnt[k];
  if (cnt == 0u || cnt == kInvalid) return;
  const uint32_t first = nd.ref[k];
  const double   nx = tsph[8 * (size_t)first + 4], ny = tsph[8 * (size_t)first + 5];
  const double   nz = tsph[8 * (size_t)first + 6], d = tsph[8 * (size_t)first + 7];
  double         t = 0.0, r1 = 0.0;
  for (uint32_t j = first; j < first + cnt; j++) {
    const float* x = tripos + 12 * (size_t)j;
    for (int v = 0; v < 3; v++) {
      const double px = x[3 * v], py = x[3 * v + 1], pz = x[3 * v + 2];
      t  = fmax(t, fabs(nx * px + ny * py + nz * pz + d));
      r1 = fmax(r1, fabs(px) + fabs(py) + fabs(pz));
    }
  }
  const double w = t * (1.0 + 0x1p-20) + 0x1p-20 * (fabs(d) + r1);
  lthick[first]  = __double2float_ru(w);
}

} // namespace

hipError_t launch_grid_points(const GridSpec& g, float* out, hipStream_t s)
{
  const uint64_t n = g.dims[0] * g.dims[1] * g.dims[2];
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(grid_points_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, g, out);
  return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double* in, float* out, uint64_t n, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_f32_to_f64(const float* in, double* out, uint64_t n, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_leaf_slabs(const void* nodes, uint32_t num_nodes, int K, const float* tsph, const float* tripos,
                             float* lthick, hipStream_t s)
{
  const uint64_t n = (uint64_t)num_nodes * (uint64_t)K;
  if (n == 0) return hipSuccess;
  const dim3 grid((uint32_t)((n + 255) / 256));
  if (K == 4)
    hipLaunchKernelGGL(leaf_slab_kernel<4>, grid, dim3(256), 0, s, static_cast<const Node<4>*>(nodes), num_nodes,
                       tsph, tripos, lthick);
  else if (K == 8)
    hipLaunchKernelGGL(leaf_slab_kernel<8>, grid, dim3(256), 0, s, static_cast<const Node<8>*>(nodes), num_nodes,
                       tsph, tripos, lthick);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_map_u32(uint32_t* v, uint64_t n, const uint32_t* map, uint32_t nmap, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(map_u32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, v, n, map, nmap);
  return hipGetLastError();
}

hipError_t launch_add_counts(unsigned long long* work, const uint32_t* recheck, const uint32_t* fallback, hipStream_t s)
{
  hipLaunchKernelGGL(add_counts_kernel, dim3(1), dim3(64), 0, s, work, recheck, fallback);
  return hipGetLastError();
}

} // namespace ebgpu
