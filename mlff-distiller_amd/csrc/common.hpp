// Shared device helpers for the gfx950 PaiNN kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlff {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kWave = 64;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kEdgeEps = 1e-8f;   // edge_vector / (d + 1e-8), student_model.py:715

__host__ __device__ inline int pad16(int x) { return (x + 15) & ~15; }

// Distance rounded exactly as torch.norm(dim=-1) rounds a 3-vector on the
// CPU: x*x first, then two fused multiply-adds, then a correctly rounded
// square root.  Contraction is switched off so the compiler cannot fuse or
// reorder anything else in here.
__device__ __forceinline__ float norm3_exact(float x, float y, float z) {
#pragma clang fp contract(off)
  float acc = x * x;
  acc = __builtin_fmaf(y, y, acc);
  acc = __builtin_fmaf(z, z, acc);
  return __builtin_sqrtf(acc);
}

// SiLU with the library exponential and a correctly rounded division (the
// faster forms tried, hardware exp2 and reciprocal, drift by a few ulp in one
// direction, which adds up over channels and pairs when activations are large)
__device__ __forceinline__ float silu(float x) { return x / (1.0f + expf(-x)); }
// d/dx [x sigmoid(x)] = s (1 + x (1 - s))
__device__ __forceinline__ float silu_grad(float x) {
  const float s = 1.0f / (1.0f + expf(-x));
  return s * (1.0f + x * (1.0f - s));
}

// Gaussian RBF x cosine cutoff (student_model.py:249-253, 285-290, 720-722),
// split so the kernels evaluate the cutoff once per pair and the Gaussian once
// per (pair, basis function):
//   e_k(d) = exp(-gamma_k (d - mu_k)^2) * fc(d),  gamma_k = 1 / width_k^2,
//   fc(d) = (cos(pi d / rc) + 1) / 2 for d < rc, else 0,
//   de_k/dd = -2 gamma_k (d - mu_k) phi_k fc + phi_k dfc/dd.
__device__ __forceinline__ float rbf_gamma(float width) { return 1.0f / (width * width); }

__device__ __forceinline__ void cosine_cutoff(float d, float rc, float& fc, float& dfc) {
  const float arg = kPi * d / rc;
  float s, c;
  sincosf(arg, &s, &c);
  fc = 0.5f * (c + 1.0f);
  dfc = -0.5f * (kPi / rc) * s;
  if (!(d < rc)) { fc = 0.0f; dfc = 0.0f; }
}

__device__ __forceinline__ float edge_feature(float d, float mu, float gamma, float fc) {
  const float diff = d - mu;
  return expf(-gamma * (diff * diff)) * fc;
}

__device__ __forceinline__ void edge_feature_and_grad(float d, float mu, float gamma, float fc,
                                                      float dfc, float& e, float& de) {
  const float diff = d - mu;
  const float phi = expf(-gamma * (diff * diff));
  e = phi * fc;
  de = (-2.0f * gamma * diff * phi) * fc + phi * dfc;
}

// ---------------------------------------------------------------------------
// MFMA tile product on v_mfma_f32_16x16x4f32.
//
// acc[rt][ct] += A[rt*16 .. rt*16+15][0..K) * B[cb[ct] .. cb[ct]+15][0..K)^T
//
// A lives in LDS, row-major with row stride lda floats (lda % 4 == 0).
// B is a row-major [rows][ldb] global matrix (a torch Linear weight, or its
// transpose, zero-padded so K is a multiple of 16).
//
// Within each 16-wide K slab, lane group g = lane/16 feeds k = k0 + 4g + s in
// sub-step s, for both operands.  The MFMA sums over its four k slots, so
// this permutation of k leaves the product unchanged while letting every lane
// fetch its four A values and four B values as single 16-byte loads.
// Result layout (per wave): acc[rt][ct][i] holds
//   D[row = rt*16 + 4*(lane/16) + i][col = cb[ct] + lane%16].
// ---------------------------------------------------------------------------
template <int RT, int NT>
__device__ __forceinline__ void mma_tile(const float* __restrict__ A, int lda,
                                         const float* __restrict__ B, int ldb, int K,
                                         const int (&cb)[NT], f4 (&acc)[RT][NT]) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  const int g = lane >> 4;
  const float* arow = A + r * lda + 4 * g;
  const float* brow[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) brow[ct] = B + (size_t)(cb[ct] + r) * ldb + 4 * g;
  // The global B operand of slab k+1 is fetched before the MFMAs of slab k
  // are issued, so its latency hides behind them.
  f4 b[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) b[ct] = *reinterpret_cast<const f4*>(brow[ct]);
  for (int k0 = 0; k0 < K; k0 += 16) {
    f4 a[RT], bn[NT];
    const int kn = (k0 + 16 < K) ? k0 + 16 : k0;
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) bn[ct] = *reinterpret_cast<const f4*>(brow[ct] + kn);
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) a[rt] = *reinterpret_cast<const f4*>(arow + rt * 16 * lda + k0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
          acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b[ct][s], acc[rt][ct], 0, 0, 0);
      }
    }
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) b[ct] = bn[ct];
  }
}

// Row index of a row-parallel kernel with XCD-aware placement: consecutive
// rows (atoms of one molecule) go to blocks on the same XCD, so the node
// features and filters they share stay in that XCD's L2.
__device__ __forceinline__ int xcd_block() {
  const int G = gridDim.x;
  const int b = blockIdx.x;
  if (G % 8 != 0) return b;
  return (b % 8) * (G / 8) + b / 8;
}

// sched_barrier mask: VALU (0x2), SALU (0x4), all LDS (0x80, 0x100, 0x200)
// and transcendentals (0x400) may cross; global memory and MFMA may not, so
// each slab's MFMAs stay between its own barrier and the next one
constexpr int kSchedSlab = 0x0002 | 0x0004 | 0x0080 | 0x0100 | 0x0200 | 0x0400;

// Same product with K fixed at compile time: the slab loop unrolls fully and
// the weight fetches run PF slabs ahead of the MFMAs that consume them.
// PIN keeps each slab's prefetch above its MFMAs with a scheduling barrier
// (the compiler otherwise may sink the loads next to their use to save
// registers); which is faster depends on the kernel's occupancy.
template <int RT, int NT, int K, int PF = 2, bool PIN = false>
__device__ __forceinline__ void mma_tile_k(const float* __restrict__ A, int lda,
                                           const float* __restrict__ B, int ldb,
                                           const int (&cb)[NT], f4 (&acc)[RT][NT]) {
  static_assert(K % 16 == 0, "K must be a multiple of 16");
  constexpr int S = K / 16;
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  const int g = lane >> 4;
  const float* arow = A + r * lda + 4 * g;
  // uniform weight base + 32-bit per-lane offsets (one register per column
  // tile instead of a 64-bit pointer)
  int boff[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) boff[ct] = (cb[ct] + r) * ldb + 4 * g;
  f4 b[PF + 1][NT];
#pragma unroll
  for (int q = 0; q < PF; ++q)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
      if (q < S) b[q][ct] = *reinterpret_cast<const f4*>(B + boff[ct] + 16 * q);
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    if (sl + PF < S) {
#pragma unroll
      for (int ct = 0; ct < NT; ++ct)
        b[(sl + PF) % (PF + 1)][ct] = *reinterpret_cast<const f4*>(B + boff[ct] + 16 * (sl + PF));
    }
    if (PIN) __builtin_amdgcn_sched_barrier(kSchedSlab);
    f4 a[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
      a[rt] = *reinterpret_cast<const f4*>(arow + rt * 16 * lda + 16 * sl);
    const int cur = sl % (PF + 1);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
          acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b[cur][ct][s], acc[rt][ct],
                                                             0, 0, 0);
  }
}

// ---------------------------------------------------------------------------
// FP32 products on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Each float32 operand x is split into three bfloat16 parts, x = h + m + l
// (each part the round-to-nearest-even bfloat16 of what the previous parts
// leave), which carry 24 significant bits between them.  A product a b is
// then the sum of the six part products whose magnitude is at least 2^-16 of
// |h_a h_b| (l_a h_b, h_a l_b, m_a m_b, m_a h_b, h_a m_b, h_a h_b, small terms
// first); each part product is exact in float32 and the MFMA accumulates in
// float32, so the result carries float32 accuracy at 6/16 of the FP32 MFMA
// cost.  The three parts sit in three planes of the same shape.
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__host__ __device__ inline uint16_t bf16_rne(float x) {
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}
__device__ __forceinline__ void split3_bf16(float x, uint16_t& h, uint16_t& m, uint16_t& l) {
  h = bf16_rne(x);
  const float r = x - bf16_to_f32(h);
  m = bf16_rne(r);
  l = bf16_rne(r - bf16_to_f32(m));
}

// acc[rt][ct] += A[rt*16 .. +16][0..K) * B[cb[ct] .. +16][0..K)^T with A in
// LDS and B in global memory, both as three bf16 planes, row-major with row
// strides lda / ldb (elements, multiples of 8) and plane strides pa / pb.
// Lane (r, g) supplies row r of each operand and 8 consecutive k of slot
// group g; the MFMA sums over all 32 k of the slab, so which k the hardware
// pairs with which slot does not matter as long as A and B agree.
template <int RT, int NT, int K>
__device__ __forceinline__ void mma_tile_x6(const uint16_t* __restrict__ A, int lda, int pa,
                                            const uint16_t* __restrict__ B, int ldb, int pb,
                                            const int (&cb)[NT], f4 (&acc)[RT][NT]) {
  static_assert(K % 32 == 0, "K must be a multiple of 32");
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  const int g = lane >> 4;
  const uint16_t* arow = A + r * lda + 8 * g;
  int boff[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) boff[ct] = (cb[ct] + r) * ldb + 8 * g;
#pragma unroll 1
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf8 bh[NT], bm[NT], bl[NT];
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) {
      bh[ct] = *reinterpret_cast<const bf8*>(B + boff[ct] + k0);
      bm[ct] = *reinterpret_cast<const bf8*>(B + pb + boff[ct] + k0);
      bl[ct] = *reinterpret_cast<const bf8*>(B + 2 * pb + boff[ct] + k0);
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const uint16_t* ar = arow + rt * 16 * lda + k0;
      const bf8 ah = *reinterpret_cast<const bf8*>(ar);
      const bf8 am = *reinterpret_cast<const bf8*>(ar + pa);
      const bf8 al = *reinterpret_cast<const bf8*>(ar + 2 * pa);
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) {
        f4 c = acc[rt][ct];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[ct], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[ct], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[ct], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[ct], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[ct], c, 0, 0, 0);
        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[ct], c, 0, 0, 0);
      }
    }
  }
}

// mma_tile_k with the k sub-steps of every slab split over two accumulator
// sets (even and odd), added at the end: twice the independent MFMA chains
// for tiles with few output blocks per wave.
template <int RT, int NT, int K, int PF = 2, bool PIN = false>
__device__ __forceinline__ void mma_tile_k2(const float* __restrict__ A, int lda,
                                            const float* __restrict__ B, int ldb,
                                            const int (&cb)[NT], f4 (&acc)[RT][NT]) {
  static_assert(K % 16 == 0, "K must be a multiple of 16");
  constexpr int S = K / 16;
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  const int g = lane >> 4;
  const float* arow = A + r * lda + 4 * g;
  int boff[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) boff[ct] = (cb[ct] + r) * ldb + 4 * g;
  f4 odd[RT][NT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) odd[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
  f4 b[PF + 1][NT];
#pragma unroll
  for (int q = 0; q < PF; ++q)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
      if (q < S) b[q][ct] = *reinterpret_cast<const f4*>(B + boff[ct] + 16 * q);
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    if (sl + PF < S) {
#pragma unroll
      for (int ct = 0; ct < NT; ++ct)
        b[(sl + PF) % (PF + 1)][ct] = *reinterpret_cast<const f4*>(B + boff[ct] + 16 * (sl + PF));
    }
    if (PIN) __builtin_amdgcn_sched_barrier(kSchedSlab);
    f4 a[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
      a[rt] = *reinterpret_cast<const f4*>(arow + rt * 16 * lda + 16 * sl);
    const int cur = sl % (PF + 1);
#pragma unroll
    for (int s = 0; s < 4; s += 2)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s], b[cur][ct][s], acc[rt][ct],
                                                             0, 0, 0);
          odd[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt][s + 1], b[cur][ct][s + 1],
                                                             odd[rt][ct], 0, 0, 0);
        }
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) acc[rt][ct] += odd[rt][ct];
}

template <int RT, int NT>
__device__ __forceinline__ void zero_acc(f4 (&acc)[RT][NT]) {
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = f4{0.f, 0.f, 0.f, 0.f};
}

// Work tiles of a persistent block: tiles t0, t0 + step, ... below t1.
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Each XCD owns one contiguous eighth of the tiles, and its blocks
// stride through that eighth together, so at any moment an XCD works on a
// few neighbouring tiles (the same molecules, the same node features) that
// fit its L2, instead of one far-apart range per block.
__device__ __forceinline__ void persistent_range(int ntiles, int& t0, int& t1, int& step) {
  const int G = gridDim.x;
  const int b = blockIdx.x;
  if (G % 8 != 0) {
    t0 = b; t1 = ntiles; step = G;
    return;
  }
  const int per_xcd = (ntiles + 7) / 8;
  const int x = b % 8, k = b / 8;
  t0 = x * per_xcd + k;
  t1 = min(ntiles, (x + 1) * per_xcd);
  step = G / 8;
}

// Sum over aligned groups of `width` lanes (width a power of two <= 64).
template <int width>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = width / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace mlff
