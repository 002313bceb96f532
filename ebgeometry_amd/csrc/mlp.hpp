// Learned cull bound: a Fourier-feature MLP that predicts the unsigned distance to the mesh.
//
// Network (56,321 parameters):
//   x_hat = (p - lo) * scale - 1                 (lo/scale from the training box)
//   features[51] = [x_hat_0..2, sin(w_k x_hat_c) for k = 0..15, c = 0..2],
//                  w_k = pi * 2^(0.4 k)  (pi .. 64 pi), zero-padded to 64
//   h1 = relu(features W1 + b1)   W1: 51 x 128
//   h2..h4 = relu(h W + b)        128 x 128 each
//   y = h4 . w5 + b5              prediction of |d| / diag
// Weights and activations are BF16 on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation; master weights and Adam moments are FP32.
//
// Public parameter order (ebg_mlp_get_params): W1 (51x128 row-major, [in][out]), b1, W2, b2,
// W3, b3, W4, b4, w5 (128), b5 (1).
#pragma once

#include "ebgpu_internal.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace ebgpu {

constexpr int kMlpIn     = 51; // 3 + 3*16
constexpr int kMlpInPad  = 64;
constexpr int kMlpHidden = 128;
constexpr int kMlpFreqs  = 16;
constexpr float kMlpLogC = 0.005f; // y ~ log(|d|/diag + kMlpLogC)

constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + 3 * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpHidden + 1; // 56,321

// Device-side FP32 master layout (W1 padded to 64 input rows).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpInPad * kMlpHidden;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden * kMlpHidden;
constexpr int kOffW4 = kOffB3 + kMlpHidden;
constexpr int kOffB4 = kOffW4 + kMlpHidden * kMlpHidden;
constexpr int kOffW5 = kOffB4 + kMlpHidden;
constexpr int kOffB5 = kOffW5 + kMlpHidden;
constexpr int kMaster = kOffB5 + 1; // padded parameter count

// BF16 copies: row-major [in][out] for the training GEMMs (W1..W4), and the A-operand layout
// of the fused inference kernel (W^T with the K order permuted to match the accumulator layout).
constexpr int kBfW1 = 0;
constexpr int kBfW2 = kBfW1 + kMlpInPad * kMlpHidden;
constexpr int kBfW3 = kBfW2 + kMlpHidden * kMlpHidden;
constexpr int kBfW4 = kBfW3 + kMlpHidden * kMlpHidden;
constexpr int kBfTotal = kBfW4 + kMlpHidden * kMlpHidden;
// The inference copy also holds the output layer as four 16 x 32 A fragments whose rows 0..2
// are w5 split into three BF16 parts (hi + mid + lo = w5 to FP32 precision), rows 3..15 zero.
constexpr int kBfW5    = kBfTotal;
constexpr int kBfInfer = kBfW5 + 4 * 512;

struct Mlp
{
  bool      trained = false;
  float     lo[3]{}, scale[3]{}; // x_hat = (p - lo) * scale - 1
  float     diag    = 1.0f;      // output unit
  float     box_scale = 0.0f;    // training box = box_scale x mesh bbox (0: not set)
  float     under_weight = 8.0f; // loss weight of under-predictions
  float*    master  = nullptr;   // kMaster FP32
  float*    grad    = nullptr;   // kMaster FP32
  float*    adam_m  = nullptr;
  float*    adam_v  = nullptr;
  uint16_t* wtrain  = nullptr;   // kBfTotal BF16, row-major [in][out]
  uint16_t* winfer  = nullptr;   // kBfInfer BF16, inference fragment layout
  int       step    = 0;
  ~Mlp();
};

hipError_t mlp_alloc(Mlp& m);
// Upload parameters given in the public order (kMlpParams floats, host pointer).
hipError_t mlp_set_params_host(Mlp& m, const float* params);
hipError_t mlp_get_params_host(const Mlp& m, float* params);
// Deterministic He initialisation from a seed.
hipError_t mlp_init(Mlp& m, uint64_t seed);
// Refresh both BF16 copies from the FP32 master weights.
hipError_t mlp_sync_bf16(Mlp& m, hipStream_t s);

// Fused forward pass over queries q = idx ? idx[i] : i (i < n), written at out[q]:
// the prediction y(q) * diag, or with seed_mode the cull radius scale * max(y, 0) * diag + offset.
hipError_t mlp_forward(const Mlp& m, const float* pts, const uint32_t* idx, uint64_t n, bool seed_mode, float scale,
                       float offset, float* out, hipStream_t s);

// Training state for one batch size.
struct MlpTrainBuffers
{
  int       batch = 0;
  uint16_t* x0    = nullptr; // B x 64
  uint16_t* h[4]  = {};      // B x 128 each
  uint16_t* dl[4] = {};      // B x 128 each (deltas)
  float*    g     = nullptr; // B
  float*    loss  = nullptr; // 1
  ~MlpTrainBuffers();
  hipError_t ensure(int b);
};

// Loss and full parameter gradient (into m.grad, master layout) for n <= buffers.batch points
// with targets |d| (physical units). The summed loss is added to tb.loss[0].
// On failure *why (nullable) names the check that failed.
hipError_t mlp_loss_grad(Mlp& m, MlpTrainBuffers& tb, const float* pts, const float* target, int n,
                         hipStream_t s, std::string* why = nullptr);
// Host-only: run the shape and extent checks of every launch of the training steps that
// ebg_mlp_train enqueues for these sizes (full batches and the last, partial one). No GPU.
bool mlp_check_train_schedule(uint64_t num_samples, int batch, std::string* why);
// One Adam step on m (uses m.grad), then refresh the BF16 copies.
hipError_t mlp_adam(Mlp& m, float lr, hipStream_t s);
// Copy the gradient out in the public parameter order (device pointer, kMlpParams floats).
hipError_t mlp_grad_public(const Mlp& m, float* out, hipStream_t s);

// Training samples: n points, first half uniform in the box (center +- box_scale * half
// extent), second half area-uniform on the surface pushed along the face normal by N(0, sigma).
hipError_t mlp_samples(const MeshDev& dev, const float* area_cdf, uint32_t num_tris, const float* lo, const float* hi,
                       float box_scale, float sigma, uint64_t seed, uint64_t n, float* pts, hipStream_t s);
// Training points of a union scene's bound: half uniform in box_scale x [lo, hi], half uniform in
// a random primitive's world box grown by sigma.
hipError_t mlp_scene_samples(const PrimRec* prims, uint32_t num_prims, const float* lo, const float* hi,
                             float box_scale, float sigma, uint64_t seed, uint64_t n, float* pts, hipStream_t s);
// Train m on n (point, distance) pairs (device arrays): epochs passes in batches of `batch`,
// Adam with a cosine learning-rate decay from lr0 to 5%; *last_loss = the last epoch's mean loss.
hipError_t mlp_fit(Mlp& m, MlpTrainBuffers& tb, const float* pts, const float* dist, uint64_t n, int epochs, int batch,
                   float lr0, hipStream_t s, float* last_loss, std::string* why);

} // namespace ebgpu
