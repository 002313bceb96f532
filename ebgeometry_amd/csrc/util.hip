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
