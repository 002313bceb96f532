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

// The in-plane frame record of each triangle (FrameRec), in FP64 from the FP32 vertices: the
// longest edge becomes A -> B (a zero-length longest edge, a point triangle, takes the x axis),
// ey is the unit part of C - A normal to ex (when C lies on the line AB, any unit vector normal
// to ex: then c = 0 and no point is inside the face), ez = ex x ey. The edge parameters of a
// zero-length edge are 0, so that edge's closest point is its start.
__global__ __launch_bounds__(256) void frame_kernel(const float* __restrict__ tripos, uint64_t nt,
                                                    FrameRec* __restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nt) return;
  const float* x = tripos + 12 * i;
  double       P[3][3];
  for (int v = 0; v < 3; v++)
    for (int c = 0; c < 3; c++) P[v][c] = x[3 * v + c];
  double L[3];
  for (int j = 0; j < 3; j++) {
    const int k = j == 2 ? 0 : j + 1;
    L[j]        = 0.0;
    for (int c = 0; c < 3; c++) L[j] += (P[k][c] - P[j][c]) * (P[k][c] - P[j][c]);
  }
  int j = 0;
  if (L[1] > L[j]) j = 1;
  if (L[2] > L[j]) j = 2;
  const double* A = P[j];
  const double* B = P[j == 2 ? 0 : j + 1];
  const double* C = P[j == 0 ? 2 : j - 1];
  const double  a = sqrt(L[j]);
  double        ex[3] = {1.0, 0.0, 0.0}, ey[3], ez[3];
  if (a > 0.0)
    for (int c = 0; c < 3; c++) ex[c] = (B[c] - A[c]) / a;
  double b = 0.0;
  for (int c = 0; c < 3; c++) b += (C[c] - A[c]) * ex[c];
  double rr[3], cc = 0.0;
  for (int c = 0; c < 3; c++) {
    rr[c] = (C[c] - A[c]) - b * ex[c];
    cc += rr[c] * rr[c];
  }
  cc = sqrt(cc);
  if (!(cc > 0.0)) {
    // the coordinate axis least aligned with ex, made normal to it
    int m = 0;
    if (fabs(ex[1]) < fabs(ex[m])) m = 1;
    if (fabs(ex[2]) < fabs(ex[m])) m = 2;
    double nr = 0.0;
    for (int c = 0; c < 3; c++) {
      rr[c] = (c == m ? 1.0 : 0.0) - ex[m] * ex[c];
      nr += rr[c] * rr[c];
    }
    nr = sqrt(nr);
    for (int c = 0; c < 3; c++) ey[c] = rr[c] / nr;
    cc = 0.0;
  }
  else
    for (int c = 0; c < 3; c++) ey[c] = rr[c] / cc;
  ez[0] = ex[1] * ey[2] - ex[2] * ey[1];
  ez[1] = ex[2] * ey[0] - ex[0] * ey[2];
  ez[2] = ex[0] * ey[1] - ex[1] * ey[0];
  const double L1 = (b - a) * (b - a) + cc * cc, L2 = b * b + cc * cc;
  FrameRec     r{};
  for (int c = 0; c < 3; c++) {
    r.o[c]  = (float)A[c];
    r.ex[c] = (float)ex[c];
    r.ey[c] = (float)ey[c];
    r.ez[c] = (float)ez[c];
  }
  r.a   = (float)a;
  r.b   = (float)b;
  r.c   = (float)cc;
  r.bma = (float)(b - a);
  r.ca  = (float)(cc * a);
  r.k1u = L1 > 0.0 ? (float)((b - a) / L1) : 0.0f;
  r.k1v = L1 > 0.0 ? (float)(cc / L1) : 0.0f;
  r.k1c = L1 > 0.0 ? (float)(-a * (b - a) / L1) : 0.0f;
  r.k2u = L2 > 0.0 ? (float)(b / L2) : 0.0f;
  r.k2v = L2 > 0.0 ? (float)(cc / L2) : 0.0f;
  out[i] = r;
}

} // namespace

hipError_t launch_frames(const float* tripos, uint64_t nt, FrameRec* out, hipStream_t s)
{
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(frame_kernel, dim3((uint32_t)((nt + 255) / 256)), dim3(256), 0, s, tripos, nt, out);
  return hipGetLastError();
}

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
