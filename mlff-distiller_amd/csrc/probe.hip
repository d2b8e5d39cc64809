// Peak probe for the roofline figures: dense matrix-core throughput measured
// on the device itself (the published peaks assume the full boost clock).
// Each wave keeps kChains independent accumulators in registers and issues
// iters x kChains MFMAs of one shape; nothing is read from memory, and every
// accumulator is written at the end so no product can be dropped.

#include <stdint.h>

#include "common.hpp"
#include "internal.hpp"

namespace mlff {
namespace {

constexpr int kProbeBlock = 256;
constexpr int kChains = 8;

template <int KIND>
__global__ __launch_bounds__(kProbeBlock) void mfma_probe_kernel(int iters, float* __restrict__ out) {
  f4 acc[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) acc[k] = f4{0.f, 0.f, 0.f, 0.f};
  const float a = 1e-3f * (float)(threadIdx.x & 63), b = 1e-3f;
  bf8 ab, bb;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    ab[k] = (__bf16)(1e-3f * (float)((threadIdx.x & 7) + 1));
    bb[k] = (__bf16)1e-3f;
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
      if (KIND == 0)
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[k], 0, 0, 0);
      else
        acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab, bb, acc[k], 0, 0, 0);
    }
  }
  f4 s = acc[0];
#pragma unroll
  for (int k = 1; k < kChains; ++k) s += acc[k];
  out[(size_t)blockIdx.x * kProbeBlock + threadIdx.x] = s[0] + s[1] + s[2] + s[3];
}

}  // namespace
}  // namespace mlff

using namespace mlff;

extern "C" {

int64_t mlff_probe_mfma_flops(int kind, int blocks, int iters) {
  if ((kind != 0 && kind != 1) || blocks < 1 || iters < 1) return -1;
  const int64_t per_mfma = kind == 0 ? 16 * 16 * 4 * 2 : 16 * 16 * 32 * 2;
  return (int64_t)blocks * (kProbeBlock / 64) * iters * kChains * per_mfma;
}

int mlff_probe_mfma(void* stream, int kind, int blocks, int iters, float* out_dev) {
  if (kind != 0 && kind != 1) {
    set_error("mlff_probe_mfma: kind is 0 (f32 16x16x4) or 1 (bf16 16x16x32)");
    return MLFF_ERR_INVALID;
  }
  if (blocks < 1 || iters < 1 || !out_dev) {
    set_error("mlff_probe_mfma: need blocks >= 1, iters >= 1 and an output of blocks x 256 floats");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (kind == 0)
    mfma_probe_kernel<0><<<blocks, kProbeBlock, 0, st>>>(iters, out_dev);
  else
    mfma_probe_kernel<1><<<blocks, kProbeBlock, 0, st>>>(iters, out_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_probe_mfma: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

}  // extern "C"
