// Learned cull bound: device memory management and the forward pass.
#include "mlp.hpp"

namespace ebgpu {

Mlp::~Mlp()
{
  if (master) (void)hipFree(master);
  if (wbf16) (void)hipFree(wbf16);
  if (adam_m) (void)hipFree(adam_m);
  if (adam_v) (void)hipFree(adam_v);
  if (grad) (void)hipFree(grad);
}

hipError_t mlp_alloc(Mlp& m)
{
  if (m.master) return hipSuccess;
  hipError_t e;
  if ((e = hipMalloc(&m.master, sizeof(float) * kMlpParams)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.wbf16, sizeof(uint16_t) * kMlpParams)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.adam_m, sizeof(float) * kMlpParams)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.adam_v, sizeof(float) * kMlpParams)) != hipSuccess) return e;
  if ((e = hipMalloc(&m.grad, sizeof(float) * kMlpParams)) != hipSuccess) return e;
  return hipSuccess;
}

hipError_t mlp_sync_bf16(Mlp&, hipStream_t) { return hipErrorNotSupported; }

hipError_t mlp_forward(const Mlp&, const float*, const uint32_t*, uint64_t, float, float, float*, hipStream_t)
{
  return hipErrorNotSupported;
}

hipError_t mlp_seed_radius(const Mlp&, const MeshDev&, const float*, const uint32_t*, uint64_t, float, float, float*,
                           hipStream_t)
{
  return hipErrorNotSupported;
}

} // namespace ebgpu
