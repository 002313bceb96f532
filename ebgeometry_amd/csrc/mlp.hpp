// Learned cull bound: a Fourier-feature MLP that predicts the unsigned distance to the mesh.
//
// Network: x_hat = 2 (p - lo) / (hi - lo) - 1 (bbox-normalised); features
// [x_hat, sin(2^k pi x_hat_j) for k = 0..15, j = 0..2] (51 values, zero-padded to 64);
// 4 hidden layers of 128 with ReLU; scalar output (softplus-free: ReLU on the output so the
// prediction is a non-negative distance). BF16 weights and activations on MFMA with FP32
// accumulation; FP32 master weights and Adam moments.
#pragma once

#include "ebgpu_internal.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ebgpu {

constexpr int kMlpIn     = 51; // 3 + 3*16
constexpr int kMlpInPad  = 64;
constexpr int kMlpHidden = 128;
constexpr int kMlpLayers = 4; // hidden layers

// Parameter counts: W1 64x128 (rows >= 51 are zero) + b1, W2..W4 128x128 + b, w5 128 + b5.
constexpr int kMlpW1     = kMlpInPad * kMlpHidden;
constexpr int kMlpWh     = kMlpHidden * kMlpHidden;
constexpr int kMlpParams = kMlpW1 + kMlpHidden + 3 * (kMlpWh + kMlpHidden) + kMlpHidden + 1;

struct Mlp
{
  bool     trained = false;
  float    lo[3]{}, scale[3]{}; // x_hat = (p - lo) * scale - 1
  float*   master  = nullptr;   // kMlpParams FP32 (layout: W1,b1,W2,b2,W3,b3,W4,b4,w5,b5; W row-major [in][out])
  uint16_t* wbf16  = nullptr;   // BF16 copy of the weight matrices in the MFMA fragment layout
  float*   adam_m  = nullptr;
  float*   adam_v  = nullptr;
  float*   grad    = nullptr;
  ~Mlp();
};

// Allocate (if needed) device memory of an Mlp.
hipError_t mlp_alloc(Mlp& m);
// Refresh the BF16 weight copy from the FP32 master weights.
hipError_t mlp_sync_bf16(Mlp& m, hipStream_t s);
// Forward pass: pred[i] = network(pts[idx ? idx[i] : i]) for i < n, written at position
// (idx ? idx[i] : i). With seed mode, writes scale*pred + offset instead.
hipError_t mlp_forward(const Mlp& m, const float* pts, const uint32_t* idx, uint64_t n, float scale, float offset,
                       float* out, hipStream_t s);
// Seeded cull radius for the traversal: scale * prediction + offset, indexed by query.
hipError_t mlp_seed_radius(const Mlp& m, const MeshDev& dev, const float* pts, const uint32_t* idx, uint64_t n,
                           float scale, float offset, float* seeds, hipStream_t s);

} // namespace ebgpu
