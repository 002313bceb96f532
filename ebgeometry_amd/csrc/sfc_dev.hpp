// Hilbert index on a 2^B grid (device), shared by the query presort and the batch keys.
//
// Skilling's transpose method (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707,
// 2004): undo the excess work (for each bit from the top down, invert or exchange the lower bits
// of x against each coordinate), Gray-encode across the coordinates, then read the transposed
// bits out most significant first, x before y before z. Consecutive indices are face-adjacent
// cells and every aligned sub-cube is one contiguous index range.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ebgpu {
namespace dev {

__device__ __forceinline__ uint64_t spread3_21(uint64_t a)
{
  uint64_t b = a & 0x1fffffull;
  b          = (b | b << 32) & 0x1f00000000ffffull;
  b          = (b | b << 16) & 0x1f0000ff0000ffull;
  b          = (b | b << 8) & 0x100f00f00f00f00full;
  b          = (b | b << 4) & 0x10c30c30c30c30c3ull;
  b          = (b | b << 2) & 0x1249249249249249ull;
  return b;
}

// the same interleave for values of at most 10 bits, in 32-bit arithmetic
__device__ __forceinline__ uint32_t spread3_10(uint32_t a)
{
  uint32_t b = a & 0x3ffu;
  b          = (b | b << 16) & 0x030000ffu;
  b          = (b | b << 8) & 0x0300f00fu;
  b          = (b | b << 4) & 0x030c30c3u;
  b          = (b | b << 2) & 0x09249249u;
  return b;
}

template <int B>
__device__ __forceinline__ uint64_t hilbert_index(uint32_t x0, uint32_t x1, uint32_t x2)
{
  uint32_t x[3] = {x0, x1, x2};
#pragma unroll
  for (uint32_t q = 1u << (B - 1); q > 1u; q >>= 1) {
    const uint32_t m = q - 1u;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (x[c] & q) x[0] ^= m;
      else {
        const uint32_t t = (x[0] ^ x[c]) & m;
        x[0] ^= t;
        x[c] ^= t;
      }
    }
  }
  x[1] ^= x[0];
  x[2] ^= x[1];
  uint32_t t = 0;
#pragma unroll
  for (uint32_t q = 1u << (B - 1); q > 1u; q >>= 1)
    if (x[2] & q) t ^= q - 1u;
  if constexpr (B <= 10)
    return (uint64_t)((spread3_10(x[0] ^ t) << 2) | (spread3_10(x[1] ^ t) << 1) | spread3_10(x[2] ^ t));
  else
    return (spread3_21(x[0] ^ t) << 2) | (spread3_21(x[1] ^ t) << 1) | spread3_21(x[2] ^ t);
}

// Grid cell of coordinate v in [lo, lo + ext) at 2^B cells per axis (clamped; NaN -> 0). The
// cell size comes from the hardware reciprocal: the presort only needs a deterministic binning.
template <int B>
__device__ __forceinline__ uint32_t hilbert_bin(float v, float lo, float ext)
{
  const float f = ext > 0.0f ? (v - lo) * (float(1u << B) * __builtin_amdgcn_rcpf(ext)) : 0.0f;
  return (uint32_t)fminf(fmaxf(f, 0.0f), float((1u << B) - 1u));
}

} // namespace dev
} // namespace ebgpu
