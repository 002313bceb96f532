// Measurement aid, not part of the query library: a kernel of independent FP32 FMA chains that
// keeps every SIMD's VALU busy, so that rocprofv3's VALU counters of the traversal kernels can be
// set against the issue rate the hardware actually sustains (built into libebgcalib.so; bench.py
// --valu-calibration runs it).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kBlocks  = 8192; // 32 wavefronts per CU on 256 CUs
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void fma_chains(float* out, int iters, float b, float c)
{
  const unsigned gid = blockIdx.x * kThreads + threadIdx.x;
  float a0 = gid * 1e-9f, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6, a7 = a0 + 7;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      a0 = fmaf(a0, b, c);
      a1 = fmaf(a1, b, c);
      a2 = fmaf(a2, b, c);
      a3 = fmaf(a3, b, c);
      a4 = fmaf(a4, b, c);
      a5 = fmaf(a5, b, c);
      a6 = fmaf(a6, b, c);
      a7 = fmaf(a7, b, c);
    }
  }
  out[gid] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

} // namespace

extern "C" {

// Runs the kernel `launches` times with `iters` x 64 FMAs per lane; ms[i] = launch i's time.
// Returns 0 on success. *wave_fmas = FMA wavefront-instructions per launch.
int ebgcal_valu_fma(int iters, int launches, float* ms, double* wave_fmas)
{
  float* out = nullptr;
  if (hipMalloc(&out, sizeof(float) * kBlocks * kThreads) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 2;
  int rc = 0;
  for (int r = 0; r < launches && rc == 0; r++) {
    hipEventRecord(e0, nullptr);
    hipLaunchKernelGGL(fma_chains, dim3(kBlocks), dim3(kThreads), 0, nullptr, out, iters, 0.999f, 0.001f);
    hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[r], e0, e1) != hipSuccess) rc = 3;
  }
  *wave_fmas = double(kBlocks) * kThreads / 64.0 * iters * 64.0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc;
}

} // extern "C"
