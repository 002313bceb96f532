// Learned cull bound: fused MFMA inference, MFMA training GEMMs, Adam, sample generation.
//
// MFMA: v_mfma_f32_16x16x32_bf16 (gfx950). Operand layout per lane l (n = l % 16, g = l / 16):
//   A (16 x 32, M x K): row n, k = 8g .. 8g+7      B (32 x 16, K x N): column n, k = 8g .. 8g+7
//   D (16 x 16):        column n, rows 4g .. 4g+3
// The fused kernel computes h^T = W^T x^T per layer so that queries run along N. Layer L's
// accumulators hold, for query n, features 16t + 4g + e of output tile t; the next layer reads
// its B operand straight from them by taking the K order
//   phi(b, g, e) = 32b + (e < 4 ? 4g + e : 16 + 4g + e - 4)
// and storing the next weight matrix's A fragments in that same K order (no data movement).
#include "mlp.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace ebgpu {

namespace {

typedef __bf16   bf16x8 __attribute__((ext_vector_type(8)));
typedef float    f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// relu(a), relu(b) as two BF16 in one word: one v_cvt_pk_bf16_f32 and one packed 16-bit integer
// max with 0 (a BF16 bit pattern is negative as an int16 exactly when the value is, -0 included)
typedef float  f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short  i16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint32_t relu_pk_bf16(float a, float b)
{
  const bf16x2 h = __builtin_convertvector((f32x2){a, b}, bf16x2);
  const i16x2  r = __builtin_elementwise_max(__builtin_bit_cast(i16x2, h), (i16x2){0, 0});
  return __builtin_bit_cast(uint32_t, r);
}
__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

// Fourier frequencies w_k = pi * 2^(0.4 k), k = 0..15, in revolutions per unit (w_k / 2 pi,
// float32-rounded): v_sin_f32 takes its argument in revolutions.
__constant__ float c_freq[kMlpFreqs];

struct MlpDev
{
  float           lo[3], sc[3], diag;
  const uint16_t* winfer;
  const float*    master;
};

__device__ __forceinline__ float feature(int f, float x0, float x1, float x2)
{
  if (f >= kMlpIn) return 0.0f;
  const int   c = f < 3 ? f : (f - 3) % 3;
  const float x = c == 0 ? x0 : (c == 1 ? x1 : x2);
  if (f < 3) return x;
  return __builtin_amdgcn_sinf(c_freq[(f - 3) / 3] * x);
}

// ------------------------------------------------------------------------------------------
// fused inference: one wavefront = kInferQT tiles of 16 queries, each weight fragment read from
// LDS once per kInferQT MFMAs; kInferWaves waves per workgroup.
// The BF16 weight fragments (116 KB) and the FP32 biases are copied once per workgroup into
// LDS and the workgroup then strides over query tiles (one workgroup per CU), so every A
// fragment comes from LDS instead of an L2 round trip per wavefront.
// ------------------------------------------------------------------------------------------

constexpr int kInferWaves = 12; // 3 per SIMD
constexpr int kInferQT    = 2;  // 16-query tiles per wavefront (32 queries; 148 VGPRs)
constexpr int kInferBias  = 4 * kMlpHidden; // b1..b4 in LDS

// Layer-1 K order of the fused kernel: K index 32 b + 8 g + e holds, for g < 3, the Fourier
// feature sin(w_j x_g) with j = 8 b + e, and for g = 3 the raw coordinate x_e (b = 0, e < 3)
// or zero padding (-1). A lane's frequencies are then compile-time constants and its coordinate
// is one select per query; features are bitwise those of features_kernel.
__host__ __device__ constexpr int infer_feature(int b, int g, int e)
{
  return g < 3 ? 3 + 3 * (8 * b + e) + g : (b == 0 && e < 3 ? e : -1);
}

__device__ __forceinline__ uint32_t pk_bf16(float a, float b)
{
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

// acc[q][t] += A(t, b) B_q(b) over the 8 NB fragments of one layer, K block b outermost so that
// consecutive MFMAs never chain on one accumulator; the A fragments stream from LDS through a
// ring of three registers, two loads ahead of the MFMAs that use them.
template <int NB>
__device__ __forceinline__ void layer_mma(const uint16_t* __restrict__ w, int lane, f32x4 (&acc)[kInferQT][8],
                                          const u16x8 (&bq)[kInferQT][4])
{
  constexpr int F = 8 * NB;
  auto frag = [&](int f) { return *reinterpret_cast<const u16x8*>(w + (((f % 8) * NB + f / 8) * 64 + lane) * 8); };
  u16x8 a[3];
  a[0] = frag(0);
  a[1] = frag(1);
#pragma unroll
  for (int f = 0; f < F; f++) {
    if (f + 2 < F) a[(f + 2) % 3] = frag(f + 2);
    const int t = f % 8, b = f / 8;
#pragma unroll
    for (int q = 0; q < kInferQT; q++) acc[q][t] = mfma(a[f % 3], bq[q][b], acc[q][t]);
  }
}

// One hidden layer (K = 128) from the previous layer's accumulators `prev` into `acc` (separate
// registers): K block b's B operands (relu'd BF16 of prev tiles 2b, 2b+1) are packed right
// before the MFMAs that read them, so that the packing of block b+1 can be scheduled under the
// MFMAs of block b.
__device__ __forceinline__ void dense_layer(const uint16_t* __restrict__ w, const float* __restrict__ bias, int lane,
                                            const f32x4 (&prev)[kInferQT][8], f32x4 (&acc)[kInferQT][8])
{
  const int g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const f32x4 bv = *reinterpret_cast<const f32x4*>(bias + 16 * t + 4 * g);
#pragma unroll
    for (int q = 0; q < kInferQT; q++) acc[q][t] = bv;
  }
  auto  frag = [&](int f) { return *reinterpret_cast<const u16x8*>(w + (((f % 8) * 4 + f / 8) * 64 + lane) * 8); };
  u16x8 a[3], bq[kInferQT];
  a[0] = frag(0);
  a[1] = frag(1);
#pragma unroll
  for (int f = 0; f < 32; f++) {
    const int t = f % 8, b = f / 8;
    if (t == 0) {
#pragma unroll
      for (int q = 0; q < kInferQT; q++) {
        const u32x4v v = {relu_pk_bf16(prev[q][2 * b][0], prev[q][2 * b][1]),
                          relu_pk_bf16(prev[q][2 * b][2], prev[q][2 * b][3]),
                          relu_pk_bf16(prev[q][2 * b + 1][0], prev[q][2 * b + 1][1]),
                          relu_pk_bf16(prev[q][2 * b + 1][2], prev[q][2 * b + 1][3])};
        bq[q]          = __builtin_bit_cast(u16x8, v);
      }
    }
    if (f + 2 < 32) a[(f + 2) % 3] = frag(f + 2);
#pragma unroll
    for (int q = 0; q < kInferQT; q++) acc[q][t] = mfma(a[f % 3], bq[q], acc[q][t]);
  }
}

__device__ __forceinline__ void infer_tile(const uint16_t* __restrict__ sw, const float* __restrict__ sb,
                                           const float* __restrict__ pts, const uint32_t* __restrict__ idx,
                                           uint64_t n, uint64_t base, const MlpDev& p, float b5, float scale,
                                           float offset, int seed_mode, float* __restrict__ out)
{
  const int lane = threadIdx.x & 63;
  const int nn = lane & 15, g = lane >> 4;

  u16x8    bq[kInferQT][4];
  uint32_t qid[kInferQT];
  bool     valid[kInferQT];
#pragma unroll
  for (int q = 0; q < kInferQT; q++) {
    const uint64_t i = base + q * 16 + nn;
    valid[q]         = i < n;
    qid[q]           = valid[q] ? (idx ? idx[i] : (uint32_t)i) : 0u;
    float x[3]       = {0.0f, 0.0f, 0.0f};
    if (valid[q])
      for (int c = 0; c < 3; c++) x[c] = (pts[3 * (size_t)qid[q] + c] - p.lo[c]) * p.sc[c] - 1.0f;
    // g = 3 lanes take 0 here, so their sin() terms are exactly the zero padding
    const float xg = g == 0 ? x[0] : (g == 1 ? x[1] : (g == 2 ? x[2] : 0.0f));
#pragma unroll
    for (int b = 0; b < 2; b++) {
      float v[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        v[e] = __builtin_amdgcn_sinf(c_freq[8 * b + e] * xg);
        if (b == 0 && e < 3) v[e] = g == 3 ? x[e] : v[e];
      }
      const u32x4v w = {pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3]), pk_bf16(v[4], v[5]), pk_bf16(v[6], v[7])};
      bq[q][b] = __builtin_bit_cast(u16x8, w);
    }
  }

  // each layer's accumulators start from its bias (output rows 16 t + 4 g .. + 3 of this lane);
  // layers alternate between two accumulator sets
  f32x4 accA[kInferQT][8], accB[kInferQT][8];
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const f32x4 bv = *reinterpret_cast<const f32x4*>(sb + 16 * t + 4 * g);
#pragma unroll
    for (int q = 0; q < kInferQT; q++) accA[q][t] = bv;
  }
  // layer 1 (K = 64)
  layer_mma<2>(sw + kBfW1, lane, accA, bq);
  // layers 2..4 (K = 128)
  dense_layer(sw + kBfW2, sb + 1 * kMlpHidden, lane, accA, accB);
  dense_layer(sw + kBfW3, sb + 2 * kMlpHidden, lane, accB, accA);
  dense_layer(sw + kBfW4, sb + 3 * kMlpHidden, lane, accA, accB);
  // output layer y = relu(h4) . w5 + b5 on MFMA: h4 rounded to BF16 as in training; rows 0..2 of
  // the W5 fragments are w5's three BF16 parts, so lanes g = 0 receive b5 + hi.h, mid.h, lo.h
  f32x4 y[kInferQT];
#pragma unroll
  for (int q = 0; q < kInferQT; q++) y[q] = (f32x4){b5, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const u16x8 a = *reinterpret_cast<const u16x8*>(sw + kBfW5 + (b * 64 + lane) * 8);
#pragma unroll
    for (int q = 0; q < kInferQT; q++) {
      const u32x4v v = {relu_pk_bf16(accB[q][2 * b][0], accB[q][2 * b][1]),
                        relu_pk_bf16(accB[q][2 * b][2], accB[q][2 * b][3]),
                        relu_pk_bf16(accB[q][2 * b + 1][0], accB[q][2 * b + 1][1]),
                        relu_pk_bf16(accB[q][2 * b + 1][2], accB[q][2 * b + 1][3])};
      y[q]           = mfma(a, __builtin_bit_cast(u16x8, v), y[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < kInferQT; q++) {
    if (g == 0 && valid[q]) {
      const float s = y[q][0] + y[q][1] + y[q][2];
      const float r = (expf(s) - kMlpLogC) * p.diag;
      out[qid[q]]   = seed_mode ? scale * fmaxf(r, 0.0f) + offset : r;
    }
  }
}

__global__ __launch_bounds__(64 * kInferWaves) void infer_kernel(const float* __restrict__ pts,
                                                                 const uint32_t* __restrict__ idx, uint64_t n, MlpDev p,
                                                                 float scale, float offset, int seed_mode,
                                                                 float* __restrict__ out)
{
  __shared__ __attribute__((aligned(16))) uint16_t sw[kBfInfer];
  __shared__ __attribute__((aligned(16))) float    sb[kInferBias];
  for (int i = 8 * threadIdx.x; i < kBfInfer; i += 8 * 64 * kInferWaves)
    *reinterpret_cast<u16x8*>(sw + i) = *reinterpret_cast<const u16x8*>(p.winfer + i);
  {
    const int bofs[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
    for (int i = threadIdx.x; i < kInferBias; i += 64 * kInferWaves) sb[i] = p.master[bofs[i / kMlpHidden] + i % kMlpHidden];
  }
  __syncthreads();
  const float b5 = p.master[kOffB5];
  // every wavefront runs the same number of iterations of this loop shape (no barrier inside)
  const uint64_t per = 16u * kInferQT * kInferWaves;
  for (uint64_t tile = blockIdx.x; tile * per < n; tile += gridDim.x) {
    const uint64_t base = tile * per + 16u * kInferQT * (threadIdx.x >> 6);
    if (base < n) infer_tile(sw, sb, pts, idx, n, base, p, b5, scale, offset, seed_mode, out);
  }
}

// ------------------------------------------------------------------------------------------
// training: generic MFMA GEMM with strided operands and a fused epilogue
// ------------------------------------------------------------------------------------------
struct GemmArgs
{
  const uint16_t* A;
  int64_t         as0, as1; // A(i,k) = A[i*as0 + k*as1]
  const uint16_t* B;
  int64_t         bs0, bs1; // B(k,j) = B[k*bs0 + j*bs1]
  int             M, N, K;
  const float*    bias;  // per column (nullable)
  const uint16_t* mask;  // C(i,j) kept only where mask[i*ms0 + j] > 0 (nullable)
  int64_t         ms0;
  int             relu;
  float*          Cf;    // nullable
  uint16_t*       Cb;    // nullable
  int64_t         ldc;
  int             kchunk; // K per blockIdx.z (set by gemm(); more than one chunk adds into Cf)
};

__global__ __launch_bounds__(64) void gemm_kernel(GemmArgs a)
{
  const int lane = threadIdx.x, nn = lane & 15, g = lane >> 4;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  const int kb = blockIdx.z * a.kchunk, ke = min(a.K, kb + a.kchunk);
  f32x4     acc{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    u16x8 va, vb;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int64_t k = k0 + 8 * g + e;
      va[e]           = a.A[(int64_t)(i0 + nn) * a.as0 + k * a.as1];
      vb[e]           = a.B[k * a.bs0 + (int64_t)(j0 + nn) * a.bs1];
    }
    acc = mfma(va, vb, acc);
  }
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const int i = i0 + 4 * g + e, j = j0 + nn;
    if (gridDim.z > 1) { // split K: plain partial products, summed into the zeroed Cf
      atomicAdd(a.Cf + (int64_t)i * a.ldc + j, acc[e]);
      continue;
    }
    float v = acc[e] + (a.bias ? a.bias[j] : 0.0f);
    if (a.mask) {
      const uint16_t h = a.mask[(int64_t)i * a.ms0 + j];
      if (h == 0 || (h & 0x8000u)) v = 0.0f;
    }
    if (a.relu) v = fmaxf(v, 0.0f);
    if (a.Cf) a.Cf[(int64_t)i * a.ldc + j] = v;
    if (a.Cb) a.Cb[(int64_t)i * a.ldc + j] = f2bf(v);
  }
}

__global__ __launch_bounds__(256) void zero_rows_kernel(float* __restrict__ c, int rows, int cols, int64_t ld)
{
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < (int64_t)rows * cols) c[(i / cols) * ld + i % cols] = 0.0f;
}

// Products with a long K and a plain FP32 output (the weight gradients, K = batch rows) split K
// over blockIdx.z in chunks of kSplitK: a 128 x 128 gradient is only 64 output tiles.
constexpr int kSplitK = 512;

// Element counts of the arrays a GEMM may touch, for the host-side check before its launch.
struct GemmExtent
{
  int64_t a, b, c, mask, bias;
};

// The launch shape and every index gemm_kernel forms, checked on the host: M, N multiples of 16
// (one 16 x 16 tile per workgroup, no edge guards), K and the split chunk multiples of 32 (one
// MFMA per 32), strides >= 0, so the largest index of each operand is reached at its last row
// and column and must lie inside the array.
bool gemm_in_bounds(const GemmArgs& a, const GemmExtent& x, std::string* why)
{
  auto bad = [&](const char* w) {
    if (why) *why = w;
    return false;
  };
  if (a.M <= 0 || a.N <= 0 || a.K <= 0 || a.M % 16 || a.N % 16 || a.K % 32) return bad("gemm: M, N % 16 or K % 32");
  if (a.kchunk <= 0 || a.kchunk % 32 || a.kchunk > a.K) return bad("gemm: split chunk not a multiple of 32");
  if (a.as0 < 0 || a.as1 < 0 || a.bs0 < 0 || a.bs1 < 0 || a.ldc < a.N || (a.mask && a.ms0 < a.N))
    return bad("gemm: negative stride or overlapping rows");
  if (!a.A || !a.B || (!a.Cf && !a.Cb)) return bad("gemm: null operand");
  const int64_t M = a.M, N = a.N, K = a.K;
  if ((M - 1) * a.as0 + (K - 1) * a.as1 >= x.a) return bad("gemm: A read past its end");
  if ((K - 1) * a.bs0 + (N - 1) * a.bs1 >= x.b) return bad("gemm: B read past its end");
  if ((M - 1) * a.ldc + N - 1 >= x.c) return bad("gemm: C written past its end");
  if (a.mask && (M - 1) * a.ms0 + N - 1 >= x.mask) return bad("gemm: mask read past its end");
  if (a.bias && N > x.bias) return bad("gemm: bias read past its end");
  return true;
}

// With check_only nothing is enqueued (and the pointers may be placeholders): only the host
// check runs.
hipError_t gemm(const GemmArgs& a0, const GemmExtent& x, hipStream_t s, bool check_only, std::string* why)
{
  GemmArgs   a     = a0;
  const bool plain = a.Cf && !a.Cb && !a.bias && !a.mask && !a.relu;
  a.kchunk         = (plain && a.K >= 4 * kSplitK) ? kSplitK : a.K;
  if (!gemm_in_bounds(a, x, why)) return hipErrorInvalidValue;
  if (check_only) return hipSuccess;
  const int splits = (a.K + a.kchunk - 1) / a.kchunk;
  if (splits > 1)
    hipLaunchKernelGGL(zero_rows_kernel, dim3((a.M * a.N + 255) / 256), dim3(256), 0, s, a.Cf, a.M, a.N, a.ldc);
  hipLaunchKernelGGL(gemm_kernel, dim3(a.M / 16, a.N / 16, splits), dim3(64), 0, s, a);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void features_kernel(const float* __restrict__ pts, int n, int rows, MlpDev p,
                                                       uint16_t* __restrict__ x0)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= rows) return;
  float x[3] = {0.0f, 0.0f, 0.0f};
  if (i < n)
    for (int c = 0; c < 3; c++) x[c] = (pts[3 * (size_t)i + c] - p.lo[c]) * p.sc[c] - 1.0f;
  for (int f = 0; f < kMlpInPad; f++) x0[(size_t)i * kMlpInPad + f] = f2bf(feature(f, x[0], x[1], x[2]));
}

// Output layer, loss and its gradient. The network predicts y ~ log(|d|/diag + c); the loss is
// an asymmetric squared error in that log space, e = y - log(|d|/diag + c), loss = w e^2 with
// w = under_weight where e < 0 (a too-small cull radius forces the exact fallback) and 1
// otherwise. The prediction is (exp(y) - c) * diag.
// One wavefront per 64 rows: the 64 x 128 BF16 activations come in with 16-byte coalesced loads
// (4 rows per instruction) into an LDS tile whose rows are padded to 65 words, so that each lane
// then reads its own row without bank conflicts; the gradient rows go out the same way. Rows in
// [n, rows) get a zero gradient.
constexpr int kOutLd = kMlpHidden / 2 + 1; // LDS words per tile row

__global__ __launch_bounds__(64) void out_fwd_bwd_kernel(const uint16_t* __restrict__ h4, const float* __restrict__ master,
                                                         const float* __restrict__ target, int n, int rows,
                                                         float inv_diag, float inv_n, float under_weight,
                                                         float* __restrict__ gout, float* __restrict__ loss_sum,
                                                         uint16_t* __restrict__ d4)
{
  __shared__ uint32_t tile[64 * kOutLd];
  __shared__ float    w5[kMlpHidden];
  __shared__ float    gs[64];
  const int lane = threadIdx.x;
  const int row0 = blockIdx.x * 64;
  w5[lane]      = master[kOffW5 + lane];
  w5[lane + 64] = master[kOffW5 + 64 + lane];
  const int sub = lane >> 4, c4 = (lane & 15) * 4; // 16 lanes per row, 4 words (8 BF16) each
  for (int pass = 0; pass < 16; pass++) {
    const int r = pass * 4 + sub, i = row0 + r;
    uint4     v = make_uint4(0u, 0u, 0u, 0u);
    if (i < rows) v = *reinterpret_cast<const uint4*>(h4 + (size_t)i * kMlpHidden + 2 * c4);
    uint32_t* t = tile + r * kOutLd + c4;
    t[0] = v.x, t[1] = v.y, t[2] = v.z, t[3] = v.w;
  }
  __syncthreads();
  const int i  = row0 + lane;
  float     y  = 0.0f;
  for (int k = 0; k < kMlpHidden / 2; k++) {
    const uint32_t hh = tile[lane * kOutLd + k];
    y += bf2f((uint16_t)(hh & 0xFFFFu)) * w5[2 * k] + bf2f((uint16_t)(hh >> 16)) * w5[2 * k + 1];
  }
  y += master[kOffB5];
  float gi = 0.0f, l = 0.0f;
  if (i < n) {
    const float lt = logf(fabsf(target[i]) * inv_diag + kMlpLogC);
    const float e  = y - lt;
    const float w  = e < 0.0f ? under_weight : 1.0f;
    l  = w * e * e;
    gi = 2.0f * w * e * inv_n;
  }
  if (i < rows) gout[i] = gi;
  gs[lane] = gi;
  __syncthreads();
  for (int pass = 0; pass < 16; pass++) {
    const int r = pass * 4 + sub, i2 = row0 + r;
    if (i2 >= rows) continue;
    const float     g = gs[r];
    const uint32_t* t = tile + r * kOutLd + c4;
    uint32_t        o[4];
    for (int q = 0; q < 4; q++) {
      const uint16_t h0 = (uint16_t)(t[q] & 0xFFFFu), h1 = (uint16_t)(t[q] >> 16);
      const int      k  = 2 * (c4 + q);
      const uint16_t d0 = f2bf((h0 == 0 || (h0 & 0x8000u)) ? 0.0f : g * w5[k]);
      const uint16_t d1 = f2bf((h1 == 0 || (h1 & 0x8000u)) ? 0.0f : g * w5[k + 1]);
      o[q]              = (uint32_t)d0 | ((uint32_t)d1 << 16);
    }
    *reinterpret_cast<uint4*>(d4 + (size_t)i2 * kMlpHidden + 2 * c4) = make_uint4(o[0], o[1], o[2], o[3]);
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if (lane == 0 && l != 0.0f) atomicAdd(loss_sum, l);
}

// dw5[k] = sum_i h4[i,k] g[i] (block k < 128), db5 = sum_i g[i] (block 128)
__global__ __launch_bounds__(256) void out_wgrad_kernel(const uint16_t* __restrict__ h4, const float* __restrict__ gin,
                                                        int rows, float* __restrict__ grad)
{
  __shared__ float sh[256];
  const int        k = blockIdx.x;
  float            s = 0.0f;
  for (int i = threadIdx.x; i < rows; i += 256)
    s += (k < kMlpHidden ? bf2f(h4[(size_t)i * kMlpHidden + k]) : 1.0f) * gin[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[k < kMlpHidden ? kOffW5 + k : kOffB5] = sh[0];
}

// db[j] = sum_i D[i, j] (one block per column)
__global__ __launch_bounds__(256) void colsum_kernel(const uint16_t* __restrict__ d, int rows, int ld,
                                                     float* __restrict__ out)
{
  __shared__ float sh[256];
  const int        j = blockIdx.x;
  float            s = 0.0f;
  for (int i = threadIdx.x; i < rows; i += 256) s += bf2f(d[(size_t)i * ld + j]);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < (unsigned)o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sh[0];
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                   float bc1, float bc2)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i]           = mi;
  v[i]           = vi;
  w[i] -= lr * (mi * bc1) / (sqrtf(vi * bc2) + 1e-8f);
}

// BF16 copies: row-major training copy and the permuted inference fragments.
__global__ __launch_bounds__(256) void sync_bf16_kernel(const float* __restrict__ master, uint16_t* __restrict__ wtrain,
                                                        uint16_t* __restrict__ winfer)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kBfInfer) return;
  // inference copy, fragment-major: A fragment (output tile t of 16 rows, K block b of 32) is
  // 64 lanes x 8 BF16 contiguous, lane = row + 16 g holding K elements 8 g .. 8 g + 7, so that a
  // wavefront's 16-byte LDS reads of one fragment fall on consecutive banks
  if (i >= kBfW5) {
    // output layer: row r < 3 of fragment b holds part r of w5 in the K order of layers 2..4
    const int j = i - kBfW5, e = j % 8, lane = (j / 8) % 64, b = j / 512, g = lane / 16, r = lane % 16;
    const int k = 32 * b + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
    const float  w  = master[kOffW5 + k];
    const __bf16 hi = (__bf16)w;
    const float  r1 = w - (float)hi;
    const __bf16 mid = (__bf16)r1;
    const __bf16 lo  = (__bf16)(r1 - (float)mid);
    const __bf16 v   = r == 0 ? hi : (r == 1 ? mid : (r == 2 ? lo : (__bf16)0.0f));
    winfer[i]        = __builtin_bit_cast(uint16_t, v);
    return;
  }
  // training copy: same [in][out] layout as the master
  int src;
  if (i < kBfW2) src = kOffW1 + i;
  else if (i < kBfW3) src = kOffW2 + (i - kBfW2);
  else if (i < kBfW4) src = kOffW3 + (i - kBfW3);
  else src = kOffW4 + (i - kBfW4);
  wtrain[i] = f2bf(master[src]);
  const bool first = i < kBfW2;
  const int  base  = first ? kBfW1 : (i < kBfW3 ? kBfW2 : (i < kBfW4 ? kBfW3 : kBfW4));
  const int  moff  = first ? kOffW1 : (i < kBfW3 ? kOffW2 : (i < kBfW4 ? kOffW3 : kOffW4));
  const int  nb    = first ? kMlpInPad / 32 : kMlpHidden / 32;
  const int  j = i - base, e = j % 8, lane = (j / 8) % 64, frag = j / 512;
  const int  t = frag / nb, b = frag % nb, g = lane / 16;
  const int  r = 16 * t + lane % 16;
  // K order: for layer 1 the feature order of infer_feature (padding -> zero weight); for
  // layers 2..4 the order in which the previous layer's accumulators sit in a lane (rows
  // 4 g .. 4 g + 3 of tiles 2b, 2b+1)
  const int k = first ? infer_feature(b, g, e) : 32 * b + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
  winfer[i]   = k < 0 ? (uint16_t)0 : f2bf(master[moff + k * kMlpHidden + r]); // W^T[out r][in k]
}

__global__ void scale_kernel(float* p, int n, float s)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] *= s;
}

// public order <-> master layout (W1 rows 51..63 are padding)
__global__ void grad_public_kernel(const float* __restrict__ g, float* __restrict__ out)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kMlpParams) return;
  const int w1 = kMlpIn * kMlpHidden;
  out[i]       = i < w1 ? g[i] : g[i + (kMlpInPad - kMlpIn) * kMlpHidden];
}

// ------------------------------------------------------------------------------------------
// training samples
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t key, uint64_t ctr)
{
  return (double)(mix64(key + ctr * 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(256) void samples_kernel(MeshDev m, const float* __restrict__ cdf, uint32_t nt, float lo0,
                                                      float lo1, float lo2, float hi0, float hi1, float hi2,
                                                      float box_scale, float sigma, uint64_t key, uint64_t n,
                                                      float* __restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  double       p[3];
  if ((i & 1u) == 0) { // even: uniform in the box; odd: near the surface (interleaved batches)
    for (int c = 0; c < 3; c++) {
      const double ctr = 0.5 * (lo[c] + hi[c]), h = 0.5 * box_scale * (hi[c] - lo[c]);
      p[c]             = ctr + h * (2.0 * u01(key, 8 * i + c) - 1.0);
    }
  }
  else {
    const float u = (float)u01(key, 8 * i);
    uint32_t    a = 0, b = nt - 1; // first t with cdf[t] > u
    while (a < b) {
      const uint32_t mid = (a + b) / 2;
      if (cdf[mid] > u) b = mid;
      else a = mid + 1;
    }
    const float*  x  = m.tripos + 12 * (size_t)a;
    const double* fn = m.fn64 + 4 * (size_t)a;
    const double  r1 = sqrt(u01(key, 8 * i + 1)), r2 = u01(key, 8 * i + 2);
    const double  u1 = u01(key, 8 * i + 3) + 1e-300, u2 = u01(key, 8 * i + 4);
    const double  z  = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    for (int c = 0; c < 3; c++)
      p[c] = (1.0 - r1) * x[c] + r1 * (1.0 - r2) * x[3 + c] + r1 * r2 * x[6 + c] + sigma * z * fn[c];
  }
  for (int c = 0; c < 3; c++) out[3 * i + c] = (float)p[c];
}

// Training samples of a union scene's bound: even i uniform in the training box, odd i uniform
// in the world box of a random primitive (the PrimRec bounding boxes) grown by sigma on each side.
__global__ __launch_bounds__(256) void scene_samples_kernel(const PrimRec* __restrict__ prims, uint32_t np, float lo0,
                                                            float lo1, float lo2, float hi0, float hi1, float hi2,
                                                            float box_scale, float sigma, uint64_t key, uint64_t n,
                                                            float* __restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  double       p[3];
  if ((i & 1u) == 0) {
    for (int c = 0; c < 3; c++) {
      const double ctr = 0.5 * (lo[c] + hi[c]), h = 0.5 * box_scale * (hi[c] - lo[c]);
      p[c]             = ctr + h * (2.0 * u01(key, 8 * i + c) - 1.0);
    }
  }
  else {
    const uint32_t k  = min(np - 1, (uint32_t)(u01(key, 8 * i) * np));
    const float*   bv = prims[k].bv;
    for (int c = 0; c < 3; c++)
      p[c] = (double)bv[c] - sigma + ((double)bv[3 + c] - bv[c] + 2.0 * sigma) * u01(key, 8 * i + 1 + c);
  }
  for (int c = 0; c < 3; c++) out[3 * i + c] = (float)p[c];
}

MlpDev dev_of(const Mlp& m)
{
  MlpDev d;
  for (int c = 0; c < 3; c++) {
    d.lo[c] = m.lo[c];
    d.sc[c] = m.scale[c];
  }
  d.diag   = m.diag;
  d.winfer = m.winfer;
  d.master = m.master;
  return d;
}

hipError_t upload_freqs()
{
  static bool done = false;
  if (done) return hipSuccess;
  float f[kMlpFreqs];
  for (int k = 0; k < kMlpFreqs; k++) f[k] = (float)(0.5 * std::pow(2.0, 0.4 * k)); // pi 2^(0.4k) / (2 pi)
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_freq), f, sizeof(f));
  if (e == hipSuccess) done = true;
  return e;
}

} // namespace

Mlp::~Mlp()
{
  for (void* p : {(void*)master, (void*)grad, (void*)adam_m, (void*)adam_v, (void*)wtrain, (void*)winfer})
    if (p) (void)hipFree(p);
}

MlpTrainBuffers::~MlpTrainBuffers()
{
  for (void* p : {(void*)x0, (void*)h[0], (void*)h[1], (void*)h[2], (void*)h[3], (void*)dl[0], (void*)dl[1],
                  (void*)dl[2], (void*)dl[3], (void*)g, (void*)loss})
    if (p) (void)hipFree(p);
}

hipError_t MlpTrainBuffers::ensure(int b)
{
  b = (b + 31) / 32 * 32;
  if (b <= batch) return hipSuccess;
  this->~MlpTrainBuffers();
  std::memset(static_cast<void*>(this), 0, sizeof(*this));
  hipError_t e;
  if ((e = hipMalloc(&x0, sizeof(uint16_t) * b * kMlpInPad)) != hipSuccess) return e;
  for (int i = 0; i < 4; i++) {
    if ((e = hipMalloc(&h[i], sizeof(uint16_t) * b * kMlpHidden)) != hipSuccess) return e;
    if ((e = hipMalloc(&dl[i], sizeof(uint16_t) * b * kMlpHidden)) != hipSuccess) return e;
    if ((e = hipMemset(h[i], 0, sizeof(uint16_t) * b * kMlpHidden)) != hipSuccess) return e;
  }
  if ((e = hipMemset(x0, 0, sizeof(uint16_t) * b * kMlpInPad)) != hipSuccess) return e;
  if ((e = hipMalloc(&g, sizeof(float) * b)) != hipSuccess) return e;
  if ((e = hipMalloc(&loss, sizeof(float) * 4)) != hipSuccess) return e;
  batch = b;
  return hipSuccess;
}

hipError_t mlp_alloc(Mlp& m)
{
  hipError_t e = upload_freqs();
  if (e != hipSuccess) return e;
  if (m.master) return hipSuccess;
  if ((e = hipMalloc(&m.master, sizeof(float) * kMaster)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.grad, sizeof(float) * kMaster)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.adam_m, sizeof(float) * kMaster)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.adam_v, sizeof(float) * kMaster)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.wtrain, sizeof(uint16_t) * kBfTotal)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.winfer, sizeof(uint16_t) * kBfInfer)) != hipSuccess) return e;
  if ((e = hipMemset(m.grad, 0, sizeof(float) * kMaster)) != hipSuccess) return e;
  return hipSuccess;
}

static void public_to_master(const float* pub, std::vector<float>& mst)
{
  mst.assign(kMaster, 0.0f);
  const int w1 = kMlpIn * kMlpHidden;
  for (int i = 0; i < kMlpParams; i++) mst[i < w1 ? i : i + (kMlpInPad - kMlpIn) * kMlpHidden] = pub[i];
}

hipError_t mlp_set_params_host(Mlp& m, const float* params)
{
  hipError_t e = mlp_alloc(m);
  if (e != hipSuccess) return e;
  std::vector<float> mst;
  public_to_master(params, mst);
  if ((e = hipMemcpy(m.master, mst.data(), sizeof(float) * kMaster, hipMemcpyHostToDevice)) != hipSuccess) return e;
  if ((e = hipMemset(m.adam_m, 0, sizeof(float) * kMaster)) != hipSuccess) return e;
  if ((e = hipMemset(m.adam_v, 0, sizeof(float) * kMaster)) != hipSuccess) return e;
  m.step = 0;
  if ((e = mlp_sync_bf16(m, nullptr)) != hipSuccess) return e;
  return hipDeviceSynchronize();
}

hipError_t mlp_get_params_host(const Mlp& m, float* params)
{
  if (!m.master) return hipErrorInvalidValue;
  std::vector<float> mst(kMaster);
  hipError_t         e = hipMemcpy(mst.data(), m.master, sizeof(float) * kMaster, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  const int w1 = kMlpIn * kMlpHidden;
  for (int i = 0; i < kMlpParams; i++) params[i] = mst[i < w1 ? i : i + (kMlpInPad - kMlpIn) * kMlpHidden];
  return hipSuccess;
}

hipError_t mlp_init(Mlp& m, uint64_t seed)
{
  // He-uniform weights from SplitMix64, zero biases
  std::vector<float> pub(kMlpParams, 0.0f);
  uint64_t           s    = seed ? seed : 1;
  auto               next = [&s]() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  int o = 0;
  auto fill = [&](int fan_in, int count) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; i++) pub[o + i] = float((2.0 * next() - 1.0) * a);
    o += count;
  };
  fill(kMlpIn, kMlpIn * kMlpHidden);
  o += kMlpHidden;
  for (int l = 0; l < 3; l++) {
    fill(kMlpHidden, kMlpHidden * kMlpHidden);
    o += kMlpHidden;
  }
  // small output layer: the first predictions sit near log(0.05) for every point
  const double a5 = 0.01 * std::sqrt(6.0 / kMlpHidden);
  for (int i = 0; i < kMlpHidden; i++) pub[o + i] = float((2.0 * next() - 1.0) * a5);
  o += kMlpHidden;
  pub[o] = float(std::log(0.05));
  return mlp_set_params_host(m, pub.data());
}

hipError_t mlp_sync_bf16(Mlp& m, hipStream_t s)
{
  hipLaunchKernelGGL(sync_bf16_kernel, dim3((kBfInfer + 255) / 256), dim3(256), 0, s, m.master, m.wtrain, m.winfer);
  return hipGetLastError();
}

hipError_t mlp_forward(const Mlp& m, const float* pts, const uint32_t* idx, uint64_t n, bool seed_mode, float scale,
                       float offset, float* out, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  // one workgroup per CU (the weights take 112 KB of the CU's LDS)
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  const uint64_t tiles  = (n + 16 * kInferQT * kInferWaves - 1) / (16 * kInferQT * kInferWaves);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(tiles, (uint64_t)cus);
  hipLaunchKernelGGL(infer_kernel, dim3(blocks), dim3(64 * kInferWaves), 0, s, pts, idx, n, dev_of(m), scale, offset,
                     seed_mode ? 1 : 0, out);
  return hipGetLastError();
}

// Launch sequence of one training step over n rows (n <= capacity rows of the buffers). Every
// GEMM's shape and extents are checked on the host first; the other kernels index rows < rows
// <= capacity, 128 or 64 columns, guarded by their grid. With check_only nothing is enqueued and
// the buffer pointers may be placeholders.
static hipError_t loss_grad_launches(Mlp& m, MlpTrainBuffers& tb, int capacity, const float* pts, const float* target,
                                     int n, hipStream_t s, bool check_only, std::string* why)
{
  if (n < 1 || n > capacity || capacity % 32) {
    if (why) *why = "training batch larger than its buffers";
    return hipErrorInvalidValue;
  }
  hipError_t    e;
  const int     rows = (n + 31) / 32 * 32;
  const int64_t cap  = capacity;
  const MlpDev  d    = dev_of(m);
  if (!check_only)
    hipLaunchKernelGGL(features_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, pts, n, rows, d, tb.x0);
  // forward
  const uint16_t* in    = tb.x0;
  int             in_ld = kMlpInPad;
  const int       wb[4] = {kBfW1, kBfW2, kBfW3, kBfW4};
  const int       bo[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
  const int       wo[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
  const int64_t   wn[4] = {kMlpInPad * kMlpHidden, kMlpHidden * kMlpHidden, kMlpHidden * kMlpHidden,
                           kMlpHidden * kMlpHidden};
  for (int L = 0; L < 4; L++) {
    GemmArgs a{};
    a.A    = in;
    a.as0  = in_ld;
    a.as1  = 1;
    a.B    = m.wtrain + wb[L];
    a.bs0  = kMlpHidden;
    a.bs1  = 1;
    a.M    = rows;
    a.N    = kMlpHidden;
    a.K    = in_ld;
    a.bias = m.master + bo[L];
    a.relu = 1;
    a.Cb   = tb.h[L];
    a.ldc  = kMlpHidden;
    if ((e = gemm(a, {cap * in_ld, wn[L], cap * kMlpHidden, 0, kMlpHidden}, s, check_only, why)) != hipSuccess) return e;
    in    = tb.h[L];
    in_ld = kMlpHidden;
  }
  // loss (accumulated into tb.loss[0]; the caller clears it) + output layer backward
  if (!check_only) {
    hipLaunchKernelGGL(out_fwd_bwd_kernel, dim3((rows + 63) / 64), dim3(64), 0, s, tb.h[3], m.master, target, n, rows,
                       1.0f / m.diag, 1.0f / (float)n, m.under_weight, tb.g, tb.loss, tb.dl[3]);
    hipLaunchKernelGGL(out_wgrad_kernel, dim3(kMlpHidden + 1), dim3(256), 0, s, tb.h[3], tb.g, rows, m.grad);
  }
  // hidden layers, last to first
  for (int L = 3; L >= 0; L--) {
    const uint16_t* hin    = L > 0 ? tb.h[L - 1] : tb.x0;
    const int       hin_ld = L > 0 ? kMlpHidden : kMlpInPad;
    GemmArgs        w{};
    w.A   = hin;
    w.as0 = 1;
    w.as1 = hin_ld;
    w.B   = tb.dl[L];
    w.bs0 = kMlpHidden;
    w.bs1 = 1;
    w.M   = hin_ld;
    w.N   = kMlpHidden;
    w.K   = rows;
    w.Cf  = m.grad + wo[L];
    w.ldc = kMlpHidden;
    if ((e = gemm(w, {cap * hin_ld, cap * kMlpHidden, kMaster - wo[L], 0, 0}, s, check_only, why)) != hipSuccess)
      return e;
    if (!check_only)
      hipLaunchKernelGGL(colsum_kernel, dim3(kMlpHidden), dim3(256), 0, s, tb.dl[L], rows, kMlpHidden,
                         m.grad + bo[L]);
    if (L > 0) {
      GemmArgs p{};
      p.A    = tb.dl[L];
      p.as0  = kMlpHidden;
      p.as1  = 1;
      p.B    = m.wtrain + wb[L];
      p.bs0  = 1;
      p.bs1  = kMlpHidden;
      p.M    = rows;
      p.N    = kMlpHidden;
      p.K    = kMlpHidden;
      p.mask = tb.h[L - 1];
      p.ms0  = kMlpHidden;
      p.Cb   = tb.dl[L - 1];
      p.ldc  = kMlpHidden;
      if ((e = gemm(p, {cap * kMlpHidden, wn[L], cap * kMlpHidden, cap * kMlpHidden, 0}, s, check_only, why)) !=
          hipSuccess)
        return e;
    }
  }
  return check_only ? hipSuccess : hipGetLastError();
}

hipError_t mlp_loss_grad(Mlp& m, MlpTrainBuffers& tb, const float* pts, const float* target, int n, hipStream_t s,
                         std::string* why)
{
  hipError_t e = tb.ensure(n);
  if (e != hipSuccess) return e;
  return loss_grad_launches(m, tb, tb.batch, pts, target, n, s, false, why);
}

bool mlp_check_train_schedule(uint64_t num_samples, int batch, std::string* why)
{
  // the buffers ensure(batch) allocates, with placeholder addresses: only shapes and extents
  // are examined
  Mlp             m;
  MlpTrainBuffers tb;
  const int       cap  = (batch + 31) / 32 * 32;
  uint16_t*       fake = reinterpret_cast<uint16_t*>(uintptr_t(256));
  m.master             = reinterpret_cast<float*>(fake);
  m.grad               = reinterpret_cast<float*>(fake);
  m.wtrain             = fake;
  tb.x0                = fake;
  for (int i = 0; i < 4; i++) tb.h[i] = tb.dl[i] = fake;
  std::vector<int> sizes{(int)std::min<uint64_t>(num_samples, (uint64_t)batch)};
  if (num_samples > (uint64_t)batch && num_samples % (uint64_t)batch) sizes.push_back(int(num_samples % batch));
  bool ok = true;
  for (int n : sizes)
    if (loss_grad_launches(m, tb, cap, nullptr, nullptr, n, nullptr, true, why) != hipSuccess) ok = false;
  // the placeholders are not device memory: keep the destructors away from them
  m.master = m.grad = nullptr;
  m.wtrain          = nullptr;
  tb.x0             = nullptr;
  for (int i = 0; i < 4; i++) tb.h[i] = tb.dl[i] = nullptr;
  return ok;
}

hipError_t mlp_adam(Mlp& m, float lr, hipStream_t s)
{
  m.step++;
  const float bc1 = 1.0f / (1.0f - std::pow(0.9f, (float)m.step));
  const float bc2 = 1.0f / (1.0f - std::pow(0.999f, (float)m.step));
  hipLaunchKernelGGL(adam_kernel, dim3((kMaster + 255) / 256), dim3(256), 0, s, m.master, m.grad, m.adam_m, m.adam_v,
                     kMaster, lr, bc1, bc2);
  return mlp_sync_bf16(m, s);
}

hipError_t mlp_grad_public(const Mlp& m, float* out, hipStream_t s)
{
  hipLaunchKernelGGL(grad_public_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, m.grad, out);
  return hipGetLastError();
}

hipError_t mlp_samples(const MeshDev& dev, const float* cdf, uint32_t nt, const float* lo, const float* hi,
                       float box_scale, float sigma, uint64_t seed, uint64_t n, float* pts, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  uint64_t key = seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  hipLaunchKernelGGL(samples_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, dev, cdf, nt, lo[0], lo[1],
                     lo[2], hi[0], hi[1], hi[2], box_scale, sigma, key, n, pts);
  return hipGetLastError();
}

hipError_t mlp_scene_samples(const PrimRec* prims, uint32_t num_prims, const float* lo, const float* hi,
                             float box_scale, float sigma, uint64_t seed, uint64_t n, float* pts, hipStream_t s)
{
  if (n == 0 || num_prims == 0) return hipSuccess;
  uint64_t key = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull;
  hipLaunchKernelGGL(scene_samples_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, prims, num_prims, lo[0],
                     lo[1], lo[2], hi[0], hi[1], hi[2], box_scale, sigma, key, n, pts);
  return hipGetLastError();
}

hipError_t mlp_fit(Mlp& m, MlpTrainBuffers& tb, const float* pts, const float* dist, uint64_t n, int epochs, int batch,
                   float lr0, hipStream_t s, float* last_loss, std::string* why)
{
  hipError_t e = tb.ensure(batch);
  if (e != hipSuccess) return e;
  // Adam with a cosine learning-rate decay to 5% over all steps
  const uint64_t spe   = (n + batch - 1) / batch;
  const double   total = double(spe) * epochs;
  uint64_t       k     = 0;
  float          last  = 0.0f;
  for (int ep = 0; ep < epochs; ep++) {
    if ((e = hipMemsetAsync(tb.loss, 0, sizeof(float), s)) != hipSuccess) return e;
    for (uint64_t b = 0; b < n; b += batch, k++) {
      const int   nb = int(std::min<uint64_t>(batch, n - b));
      const float lr = float(lr0 * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * double(k) / total))));
      if ((e = mlp_loss_grad(m, tb, pts + 3 * b, dist + b, nb, s, why)) != hipSuccess) return e;
      if ((e = mlp_adam(m, lr, s)) != hipSuccess) return e;
    }
    float sum = 0.0f;
    if ((e = hipMemcpyAsync(&sum, tb.loss, sizeof(float), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    last = sum / float(n);
  }
  if (last_loss) *last_loss = last;
  return hipSuccess;
}

} // namespace ebgpu
