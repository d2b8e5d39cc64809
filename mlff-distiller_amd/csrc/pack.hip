// Weight blob packing on the device: every float of the blob in one launch,
// from the parameter tensors where they live (the optimizer updates them in
// place, so a training step repacks without a host round trip or one copy
// per tensor).  The slot table (mlff_weights_layout) says where each tensor
// goes; padding and the gaps between slots are written as zeros.

#include <stdint.h>

#include "common.hpp"
#include "internal.hpp"

namespace mlff {
namespace {

constexpr int kPackBlock = 256;
constexpr int kPackMaxSlots = 256;

// one table row per slot, 8 int64: slot offset, slot rows, slot ld, tensor
// rows, tensor cols, kind (1: the slot holds the transpose), source address,
// unused
struct PackRow {
  int64_t off, srows, ld, rows, cols, kind, src, pad;
};

__global__ __launch_bounds__(kPackBlock) void pack_weights_kernel(
    int n_slots, const PackRow* __restrict__ table, int64_t total, float* __restrict__ blob) {
  __shared__ PackRow t[kPackMaxSlots];
  for (int k = threadIdx.x; k < n_slots; k += kPackBlock) t[k] = table[k];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kPackBlock + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * kPackBlock) {
    // the last slot starting at or before i (offsets ascend)
    int lo = 0, hi = n_slots - 1, s = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      if (t[mid].off <= i) { s = mid; lo = mid + 1; } else { hi = mid - 1; }
    }
    float v = 0.f;
    if (s >= 0) {
      const PackRow& r = t[s];
      const int64_t local = i - r.off;
      if (local < r.srows * r.ld) {
        const int64_t row = local / r.ld, col = local - row * r.ld;
        if (row < r.rows && col < r.cols) {
          const float* src = reinterpret_cast<const float*>(r.src);
          // kind 1: the slot holds src^T, src being [cols][rows]
          v = r.kind ? src[col * r.rows + row] : src[row * r.cols + col];
        }
      }
    }
    blob[i] = v;
  }
}

}  // namespace
}  // namespace mlff

using namespace mlff;

extern "C" {

int mlff_pack_weights(void* stream, int n_slots, const int64_t* table_dev, int64_t total_floats,
                      float* blob_dev) {
  if (n_slots < 0 || n_slots > kPackMaxSlots || total_floats < 0 ||
      (total_floats > 0 && !blob_dev) || (n_slots > 0 && !table_dev)) {
    set_error("mlff_pack_weights: need 0 <= n_slots <= %d, total_floats >= 0 and non-null "
              "buffers", kPackMaxSlots);
    return MLFF_ERR_INVALID;
  }
  if (total_floats == 0) return MLFF_OK;
  const int64_t blocks = (total_floats + kPackBlock - 1) / kPackBlock;
  pack_weights_kernel<<<(unsigned)(blocks < 2048 ? blocks : 2048), kPackBlock, 0,
                        static_cast<hipStream_t>(stream)>>>(
      n_slots, reinterpret_cast<const PackRow*>(table_dev), total_floats, blob_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_pack_weights: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

}  // extern "C"
