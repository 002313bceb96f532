// Morton codes and a stable LSD radix sort on gfx950.
//
// Codes reproduce SFC::computeBins<T> + SFC::Morton::encode bit for bit
// (Source/EBGeometry_SFCImplem.hpp:31-58, 228-276):
//   delta_c = (1/ValidSpan) * (max_c - min_c)  (Vec3T/T multiplies by the rounded reciprocal,
//             Source/EBGeometry_VecImplem.hpp:303-309), delta_c == 0 -> 1,
//   bin_c   = clamp(floor((p_c - min_c) / delta_c), 0, ValidSpan)  (correctly rounded divide),
//   code    = spread(x) | spread(y) << 1 | spread(z) << 2, 21 bits per axis.
// The sort replaces the reference's std::sort on codes (:279-300) with a stable LSD radix sort,
// so equal codes keep increasing index order.
#include "kernels.hpp"
#include "sfc_dev.hpp"

#include <hip/hip_runtime.h>

namespace ebgpu {

namespace {

constexpr uint32_t kValidSpan = 2097151u; // 2^21 - 1

__device__ __forceinline__ uint64_t spread3(uint64_t a)
{
  uint64_t b = a & 0x1fffffull;
  b          = (b | b << 32) & 0x1f00000000ffffull;
  b          = (b | b << 16) & 0x1f0000ff0000ffull;
  b          = (b | b << 8) & 0x100f00f00f00f00full;
  b          = (b | b << 4) & 0x10c30c30c30c30c3ull;
  b          = (b | b << 2) & 0x1249249249249249ull;
  return b;
}

// ------------------------------- bounding box -------------------------------------------

template <class T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <class T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }

template <class T>
__global__ __launch_bounds__(256) void bbox_partial(const T* __restrict__ p, uint64_t n, T* __restrict__ part)
{
  __shared__ T sh[6][256];
  T            lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    for (int c = 0; c < 3; c++) {
      const T v = p[3 * i + c];
      lo[c]     = tmin(lo[c], v);
      hi[c]     = tmax(hi[c], v);
    }
  for (int c = 0; c < 3; c++) {
    sh[c][threadIdx.x]     = lo[c];
    sh[3 + c][threadIdx.x] = hi[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < (unsigned)s)
      for (int c = 0; c < 3; c++) {
        sh[c][threadIdx.x]     = tmin(sh[c][threadIdx.x], sh[c][threadIdx.x + s]);
        sh[3 + c][threadIdx.x] = tmax(sh[3 + c][threadIdx.x], sh[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// One block of 256 threads: strided partial minima / maxima, then a tree in LDS (min and max
// are exact, so the result does not depend on the order).
template <class T>
__global__ __launch_bounds__(256) void bbox_final(const T* __restrict__ part, int nparts, T* __restrict__ box)
{
  __shared__ T sh[6][256];
  T            v[6];
#pragma unroll
  for (int c = 0; c < 6; c++) v[c] = c < 3 ? T(INFINITY) : T(-INFINITY);
  for (int i = threadIdx.x; i < nparts; i += 256)
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = c < 3 ? tmin(v[c], part[6 * i + c]) : tmax(v[c], part[6 * i + c]);
#pragma unroll
  for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o)
#pragma unroll
      for (int c = 0; c < 6; c++)
        sh[c][threadIdx.x] = c < 3 ? tmin(sh[c][threadIdx.x], sh[c][threadIdx.x + o])
                                   : tmax(sh[c][threadIdx.x], sh[c][threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x < 6) box[threadIdx.x] = sh[threadIdx.x][0];
}

template <class T>
hipError_t bbox_impl(const T* pts, uint64_t n, T* box, T* part, hipStream_t s)
{
  const int blocks = (int)std::min<uint64_t>(kBboxParts, std::max<uint64_t>(1, (n + 255) / 256));
  hipLaunchKernelGGL(bbox_partial<T>, dim3(blocks), dim3(256), 0, s, pts, n, part);
  hipLaunchKernelGGL(bbox_final<T>, dim3(1), dim3(256), 0, s, part, blocks, box);
  return hipGetLastError();
}

// ------------------------------- codes ---------------------------------------------------

__global__ __launch_bounds__(256) void codes_f32(const float* __restrict__ p, uint64_t n, const float* __restrict__ box,
                                                 uint64_t* __restrict__ codes, int shift)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float is = 1.0f / (float)kValidSpan;
  uint64_t    code = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float lo = box[c];
    float       dl = is * (box[3 + c] - lo);
    if (dl == 0.0f) dl = 1.0f;
    const float f = floorf((p[3 * i + c] - lo) / dl); // IEEE division (built without fast-math)
    const uint32_t b = (uint32_t)(f < 0.0f ? 0.0f : (f > (float)kValidSpan ? (float)kValidSpan : f));
    code |= spread3(b) << c;
  }
  codes[i] = code >> shift;
}

// Query presort key: position on a 3-D Hilbert curve over a 1024^3 grid of the query box (30
// bits). Consecutive cells of the curve share a face, so a run of 64 sorted queries never jumps
// across the box the way a Morton run can at a high-level boundary (16M queries on a 327,680-face displaced sphere: 10% fewer triangle
// evaluations per wavefront).
// With FLAGS, bits 31 / 30 carry the query's cost flags (see CostFlags).
template <bool FLAGS>
__device__ __forceinline__ uint32_t hilbert_key(const float (&q)[3], const float* __restrict__ box, const CostFlags& cf)
{
  uint32_t x[3];
#pragma unroll
  for (int c = 0; c < 3; c++) x[c] = dev::hilbert_bin<10>(q[c], box[c], box[3 + c] - box[c]);
  uint32_t key = (uint32_t)dev::hilbert_index<10>(x[0], x[1], x[2]);
  if (FLAGS) {
    uint32_t cell[3];
#pragma unroll
    for (int c = 0; c < 3; c++)
      cell[c] = (uint32_t)fminf(fmaxf((q[c] - cf.lo[c]) * cf.scale[c], 0.0f), float(cf.g - 1)); // NaN -> 0
    const uint32_t cost = cf.grid[((size_t)cell[2] * cf.g + cell[1]) * cf.g + cell[0]];
    key |= (cost >= cf.heavy ? 1u << 31 : 0u) | (cost >= cf.early ? 1u << 30 : 0u);
  }
  return key;
}

// Four points per thread: 48 contiguous bytes read as three 16-byte loads and four keys stored
// at once (when both arrays are 16-byte aligned; otherwise point by point).
template <bool FLAGS>
__global__ __launch_bounds__(256) void hilbert_f32(const float* __restrict__ p, uint64_t n,
                                                   const float* __restrict__ box, uint32_t* __restrict__ keys,
                                                   const CostFlags cf)
{
  const uint64_t i0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 4u;
  if (i0 >= n) return;
  if (i0 + 4u <= n && ((reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(keys)) & 15u) == 0) {
    const float4* v = reinterpret_cast<const float4*>(p + 3 * i0);
    const float4  a = v[0], b = v[1], c = v[2];
    const float   q0[3] = {a.x, a.y, a.z}, q1[3] = {a.w, b.x, b.y}, q2[3] = {b.z, b.w, c.x}, q3[3] = {c.y, c.z, c.w};
    *reinterpret_cast<uint4*>(keys + i0) = make_uint4(hilbert_key<FLAGS>(q0, box, cf), hilbert_key<FLAGS>(q1, box, cf),
                                                      hilbert_key<FLAGS>(q2, box, cf), hilbert_key<FLAGS>(q3, box, cf));
    return;
  }
  for (uint64_t i = i0; i < n && i < i0 + 4u; i++) {
    const float q[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    keys[i]          = hilbert_key<FLAGS>(q, box, cf);
  }
}

__global__ __launch_bounds__(256) void codes_f64(const double* __restrict__ p, uint64_t n, const double* __restrict__ box,
                                                 uint64_t* __restrict__ codes)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double is = 1.0 / (double)kValidSpan;
  uint64_t     code = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double lo = box[c];
    double       dl = is * (box[3 + c] - lo);
    if (dl == 0.0) dl = 1.0;
    const double f = floor((p[3 * i + c] - lo) / dl);
    const uint32_t b = (uint32_t)(f < 0.0 ? 0.0 : (f > (double)kValidSpan ? (double)kValidSpan : f));
    code |= spread3(b) << c;
  }
  codes[i] = code;
}

// ------------------------------- radix sort ----------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortItems   = 16;
constexpr int kTile        = kSortThreads * kSortItems; // 4096 keys per workgroup

template <class KeyT>
__global__ __launch_bounds__(kSortThreads) void radix_hist(const KeyT* __restrict__ keys, uint64_t n, int shift,
                                                           uint32_t dmask, uint32_t* __restrict__ hist,
                                                           uint32_t nblocks)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll 4
  for (int r = 0; r < kSortItems; r++) {
    const uint64_t i = base + (uint64_t)r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & dmask], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of m uint32 values in place: three passes over chunks of 1024
constexpr int kScanChunk = 1024;

__global__ __launch_bounds__(256) void scan_reduce(const uint32_t* __restrict__ a, uint64_t m, uint32_t* __restrict__ sums)
{
  __shared__ uint32_t sh[256];
  const uint64_t      base = (uint64_t)blockIdx.x * kScanChunk;
  uint32_t            s    = 0;
  for (int k = 0; k < 4; k++) {
    const uint64_t i = base + (uint64_t)threadIdx.x * 4 + k;
    if (i < m) s += a[i];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(1024) void scan_sums(uint32_t* __restrict__ sums, uint32_t nsum)
{
  __shared__ uint32_t sh[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nsum; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < nsum ? sums[i] : 0u;
    sh[threadIdx.x]  = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const uint32_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nsum) sums[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void scan_apply(uint32_t* __restrict__ a, uint64_t m, const uint32_t* __restrict__ sums)
{
  __shared__ uint32_t sh[256];
  const uint64_t      base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * 4;
  uint32_t            v[4];
  uint32_t            s = 0;
  for (int k = 0; k < 4; k++) {
    v[k] = (base + k < m) ? a[base + k] : 0u;
    s += v[k];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const uint32_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t run = sums[blockIdx.x] + sh[threadIdx.x] - s;
  for (int k = 0; k < 4; k++) {
    if (base + k < m) a[base + k] = run;
    run += v[k];
  }
}

// Stable scatter of one tile of kTile keys, in three phases:
//  1. wavefront w ranks its own contiguous kTile / W keys in index order, 64 at a time (rank
//     among equal digits from 8 ballots), against a per-wavefront digit counter in LDS: no
//     workgroup barrier inside the loop;
//  2. per digit: tile-local start (exclusive scan of the tile's digit counts) plus the counts
//     of the lower wavefronts gives each wavefront's base; keys and values are placed at their
//     tile-local sorted positions in LDS;
//  3. the tile is read back in sorted order and written out, so each digit's run of the tile
//     goes to consecutive addresses (coalesced stores instead of one scattered store per key).
// kout may be null (last pass of a sort whose caller only wants the permutation).
template <class KeyT>
__global__ __launch_bounds__(kSortThreads) void radix_scatter(const KeyT* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin, uint64_t n, int shift,
                                                              uint32_t dmask, const uint32_t* __restrict__ offs,
                                                              uint32_t nblocks, KeyT* __restrict__ kout,
                                                              uint32_t* __restrict__ vout)
{
  constexpr int       W   = kSortThreads / 64;
  constexpr int       PER = kSortItems; // keys per lane
  __shared__ uint32_t cnt[W][256];      // phase 1: running counts; phase 2: wavefront bases
  __shared__ uint32_t lstart[256];      // tile-local start of each digit
  __shared__ uint32_t goff[256];        // global start of each digit's run of this tile
  __shared__ uint32_t wsum[W];
  __shared__ KeyT     sk[kTile];
  __shared__ uint32_t sv[kTile];
  const int           lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int w = 0; w < W; w++) cnt[w][threadIdx.x] = 0;
  goff[threadIdx.x] = offs[(uint64_t)threadIdx.x * nblocks + blockIdx.x];
  __syncthreads();

  const uint64_t tile  = (uint64_t)blockIdx.x * kTile;
  const uint64_t first = tile + (uint64_t)wave * (64 * PER);
  const uint64_t lt    = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  KeyT           key[PER];
  uint32_t       val[PER], rk[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const uint64_t i     = first + (uint64_t)r * 64 + lane;
    const bool     valid = i < n;
    key[r]               = valid ? kin[i] : KeyT(0);
    val[r]               = valid ? (vin ? vin[i] : (uint32_t)i) : 0u;
  }
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const uint64_t i     = first + (uint64_t)r * 64 + lane;
    const bool     valid = i < n;
    const uint32_t d     = (uint32_t)(key[r] >> shift) & dmask;
    uint64_t       peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const uint32_t prev = valid ? cnt[wave][d] : 0u;
    rk[r]               = valid ? prev + rank : 0xFFFFFFFFu;
    if (valid && rank == 0) cnt[wave][d] = prev + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // phase 2: thread t owns digit t
  {
    uint32_t c[W], tot = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      c[w] = cnt[w][threadIdx.x];
      tot += c[w];
    }
    // exclusive scan of tot over the 256 digits: inclusive within each wavefront, then the
    // wavefront totals
    uint32_t inc = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t start = inc - tot;
    for (int w = 0; w < wave; w++) start += wsum[w];
    lstart[threadIdx.x] = start;
#pragma unroll
    for (int w = 0; w < W; w++) {
      cnt[w][threadIdx.x] = start;
      start += c[w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < PER; r++) {
    if (rk[r] == 0xFFFFFFFFu) continue;
    const uint32_t d   = (uint32_t)(key[r] >> shift) & dmask;
    const uint32_t pos = cnt[wave][d] + rk[r];
    sk[pos]            = key[r];
    sv[pos]            = val[r];
  }
  __syncthreads();

  // phase 3: sorted read-out
  const uint32_t valid_n = (uint32_t)((n - tile) < (uint64_t)kTile ? (n - tile) : (uint64_t)kTile);
#pragma unroll
  for (int r = 0; r < PER; r++) {
    const uint32_t j = (uint32_t)r * kSortThreads + threadIdx.x;
    if (j < valid_n) {
      const KeyT     k   = sk[j];
      const uint32_t d   = (uint32_t)(k >> shift) & dmask;
      const uint32_t pos = goff[d] + (j - lstart[d]);
      if (kout) kout[pos] = k;
      vout[pos] = sv[j];
    }
  }
}

} // namespace

hipError_t launch_bbox_f32(const float* pts, uint64_t n, float* d_box6, void* part, hipStream_t s)
{
  return bbox_impl<float>(pts, n, d_box6, static_cast<float*>(part), s);
}
hipError_t launch_bbox_f64(const double* pts, uint64_t n, double* d_box6, void* part, hipStream_t s)
{
  return bbox_impl<double>(pts, n, d_box6, static_cast<double*>(part), s);
}

hipError_t launch_codes_f32(const float* pts, uint64_t n, const float* box, uint64_t* codes, hipStream_t s, int shift)
{
  if (n == 0) return hipSuccess;
  if (shift < 0 || shift > 63) return hipErrorInvalidValue;
  hipLaunchKernelGGL(codes_f32, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, pts, n, box, codes, shift);
  return hipGetLastError();
}
hipError_t launch_hilbert_f32(const float* pts, uint64_t n, const float* box, uint32_t* keys, hipStream_t s,
                              const CostFlags* cf)
{
  if (n == 0) return hipSuccess;
  const dim3 grid((uint32_t)((n + 1023) / 1024));
  if (cf) hipLaunchKernelGGL(hilbert_f32<true>, grid, dim3(256), 0, s, pts, n, box, keys, *cf);
  else hipLaunchKernelGGL(hilbert_f32<false>, grid, dim3(256), 0, s, pts, n, box, keys, CostFlags{});
  return hipGetLastError();
}
hipError_t launch_codes_f64(const double* pts, uint64_t n, const double* box, uint64_t* codes, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(codes_f64, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, pts, n, box, codes);
  return hipGetLastError();
}

static uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

std::size_t radix_sort_scratch_bytes(uint64_t n)
{
  const uint64_t nb   = (n + kTile - 1) / kTile;
  const uint64_t m    = 256 * nb;
  const uint64_t nsum = (m + kScanChunk - 1) / kScanChunk;
  return round_up(n * 8, 256) + round_up(n * 4, 256) + round_up(m * 4, 256) + round_up(nsum * 4, 256) + 256;
}

namespace {
// keys_out is also the ping-pong partner of the scratch keys; with want_keys false the last
// pass writes only the values (the permutation).
template <class KeyT>
hipError_t radix_sort_impl(const KeyT* keys_in, const uint32_t* values_in, KeyT* keys_out, uint32_t* values_out,
                           uint64_t n, int bits, void* scratch, bool want_keys, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  if (n > 0xFFFFFFFFull || bits < 1 || bits > 8 * (int)sizeof(KeyT)) return hipErrorInvalidValue;
  const uint32_t nb    = (uint32_t)((n + kTile - 1) / kTile);
  const uint64_t m     = 256ull * nb;
  const uint32_t nsum  = (uint32_t)((m + kScanChunk - 1) / kScanChunk);
  char*          p     = static_cast<char*>(scratch);
  KeyT*          ktmp  = reinterpret_cast<KeyT*>(p);
  p += round_up(n * 8, 256);
  uint32_t* vtmp = reinterpret_cast<uint32_t*>(p);
  p += round_up(n * 4, 256);
  uint32_t* hist = reinterpret_cast<uint32_t*>(p);
  p += round_up(m * 4, 256);
  uint32_t* sums = reinterpret_cast<uint32_t*>(p);

  const int passes = (bits + 7) / 8;
  const KeyT*     ksrc = keys_in;
  const uint32_t* vsrc = values_in;
  for (int pass = 0; pass < passes; pass++) {
    // the last pass must land in the output buffers
    const bool to_out = ((passes - 1 - pass) % 2) == 0;
    KeyT*      kdst   = to_out ? (pass == passes - 1 && !want_keys ? nullptr : keys_out) : ktmp;
    uint32_t*  vdst   = to_out ? values_out : vtmp;
    const int  shift  = 8 * pass;
    // bits at and above 'bits' are not part of the digit
    const uint32_t dmask = bits - shift >= 8 ? 255u : (1u << (bits - shift)) - 1u;
    hipLaunchKernelGGL((radix_hist<KeyT>), dim3(nb), dim3(kSortThreads), 0, s, ksrc, n, shift, dmask, hist, nb);
    hipLaunchKernelGGL(scan_reduce, dim3(nsum), dim3(256), 0, s, hist, m, sums);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1024), 0, s, sums, nsum);
    hipLaunchKernelGGL(scan_apply, dim3(nsum), dim3(256), 0, s, hist, m, sums);
    hipLaunchKernelGGL((radix_scatter<KeyT>), dim3(nb), dim3(kSortThreads), 0, s, ksrc, vsrc, n, shift, dmask, hist,
                       nb, kdst, vdst);
    ksrc = kdst;
    vsrc = vdst;
  }
  return hipGetLastError();
}
} // namespace

hipError_t radix_sort_pairs(const uint64_t* keys_in, const uint32_t* values_in, uint64_t* keys_out,
                            uint32_t* values_out, uint64_t n, int bits, void* scratch, hipStream_t s)
{
  if (!keys_out) return hipErrorInvalidValue;
  return radix_sort_impl<uint64_t>(keys_in, values_in, keys_out, values_out, n, bits, scratch, true, s);
}

hipError_t radix_sort_pairs32(const uint32_t* keys_in, const uint32_t* values_in, uint32_t* keys_out,
                              uint32_t* values_out, uint64_t n, int bits, void* scratch, bool want_keys,
                              hipStream_t s)
{
  if (!keys_out) return hipErrorInvalidValue;
  return radix_sort_impl<uint32_t>(keys_in, values_in, keys_out, values_out, n, bits, scratch, want_keys, s);
}

} // namespace ebgpu
