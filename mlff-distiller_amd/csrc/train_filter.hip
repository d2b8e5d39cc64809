// The filter MLP of PaiNNMessage (student_model.py:313-322, 350-355) in the
// force-matching training pass, one fused pair-tile kernel for the forward
// and one for its tangent along a position direction:
//
//   forward:  a = e W1^T + b1,    h = silu(a),          W = h W2^T + b2
//   tangent:  da = de W1^T,       dh = silu'(a) da,     dW = dh W2^T
//
// e (or de) is [P][R] per pair; a, h (da, dh) are [P][H] and W (dW) is
// [P][3H], all row-major float32.
//
// A third kernel runs the reverse of the first layer's product and SiLU for
// the reverse pass (student_model.py:318-322 under autograd), given the
// filter-output adjoint gW [P][3H] and its tangent dgW:
//
//   gh = gW W2,   ga = silu'(a) gh,
//   dgh = dgW W2, dga = silu'(a) dgh + silu''(a) da gh
//
// with the products on the FP32 matrix cores (W2 comes transposed, [H][3H]).  The first product runs on the FP32 matrix
// cores; the second, the large one, on the bf16 matrix cores with both
// operands split into three bf16 parts (six part products, float32
// accuracy), as in the evaluation's filter kernel.  W1 comes padded to RP
// columns (16 or 32, zeros beyond R) and W2 as three bf16 planes
// (mlff_split_bf16x3), so a caller re-splits W2 once per weight update.

#include <stdint.h>

#include "common.hpp"
#include "internal.hpp"

namespace mlff {
namespace {

constexpr int kTfBlock = 512;
constexpr int kTfWaves = kTfBlock / 64;
constexpr int kTfMaxRbf = 32;
// pairs per tile: 128 for the compact hidden size, 64 above (LDS: three bf16
// planes of the tile's hidden activations)
template <int H>
constexpr int kTfPairs = H <= 32 ? 128 : 64;

template <int H, int TM, bool TAN>
__global__ __launch_bounds__(kTfBlock, 4) void train_filter_kernel(
    int P, int R, int RP, const float* __restrict__ x, const float* __restrict__ w1p,
    const float* __restrict__ b1, const uint16_t* __restrict__ w2bf,
    const float* __restrict__ b2, const float* __restrict__ a_in, float* __restrict__ a_out,
    float* __restrict__ h_out, float* __restrict__ w_out) {
  constexpr int RT = TM / 16;
  constexpr int LDE = kTfMaxRbf + 4;
  constexpr int LDHB = H + 8;          // bf16 row stride of the split hidden activations
  constexpr int PHB = TM * LDHB;       // plane stride
  constexpr int CT1 = H / 16, NT1 = (CT1 + kTfWaves - 1) / kTfWaves;
  constexpr int CT2 = 3 * H / 16, NT2 = (CT2 + kTfWaves - 1) / kTfWaves;
  static_assert(CT1 % NT1 == 0 && CT2 % NT2 == 0, "tile split");
  __shared__ __align__(16) float e_lds[TM * LDE];
  __shared__ __align__(16) uint16_t h_bf[3 * PHB];
  const int ntiles = (P + TM - 1) / TM;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int pbase = tile * TM;
    // the tile's basis rows, zero past R and past the last pair
    for (int idx = threadIdx.x; idx < TM * RP; idx += kTfBlock) {
      const int m = idx / RP, k = idx - m * RP, p = pbase + m;
      e_lds[m * LDE + k] = (p < P && k < R) ? x[(size_t)p * R + k] : 0.f;
    }
    __syncthreads();
    if (wave < CT1 / NT1) {
      int cb[NT1];
      f4 acc[RT][NT1];
#pragma unroll
      for (int ct = 0; ct < NT1; ++ct) cb[ct] = (wave * NT1 + ct) * 16;
      zero_acc(acc);
      mma_tile<RT, NT1>(e_lds, LDE, w1p, RP, RP, cb, acc);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT1; ++ct) {
          const int c = cb[ct] + r;
          const float bias = TAN ? 0.f : b1[c];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int m = rt * 16 + 4 * g + i, p = pbase + m;
            const size_t o = (size_t)p * H + c;
            float a = acc[rt][ct][i] + bias, hv = 0.f;
            if (p < P) {
              hv = TAN ? silu_grad(a_in[o]) * a : silu(a);
              a_out[o] = a;
              h_out[o] = hv;
            }
            uint16_t hh, hm, hl;
            split3_bf16(hv, hh, hm, hl);
            h_bf[m * LDHB + c] = hh;
            h_bf[PHB + m * LDHB + c] = hm;
            h_bf[2 * PHB + m * LDHB + c] = hl;
          }
        }
    }
    __syncthreads();
    if (wave < CT2 / NT2) {
      int cb[NT2];
      f4 acc[RT][NT2];
#pragma unroll
      for (int ct = 0; ct < NT2; ++ct) cb[ct] = (wave * NT2 + ct) * 16;
      zero_acc(acc);
      mma_tile_x6<RT, NT2, H>(h_bf, LDHB, PHB, w2bf, H, 3 * H * H, cb, acc);
      float bias[NT2];
#pragma unroll
      for (int ct = 0; ct < NT2; ++ct) bias[ct] = TAN ? 0.f : b2[cb[ct] + r];
      // lane (r, g) holds rows rt*16 + 4g + i, columns cb[ct] + r
      float* out = w_out + (size_t)(pbase + 4 * g) * 3 * H + r;
      if (pbase + TM <= P) {   // full tile: no per-element bounds checks
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int ct = 0; ct < NT2; ++ct)
              out[(size_t)(rt * 16 + i) * 3 * H + cb[ct]] = acc[rt][ct][i] + bias[ct];
      } else {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (pbase + rt * 16 + 4 * g + i >= P) continue;
#pragma unroll
            for (int ct = 0; ct < NT2; ++ct)
              out[(size_t)(rt * 16 + i) * 3 * H + cb[ct]] = acc[rt][ct][i] + bias[ct];
          }
      }
    }
    __syncthreads();
  }
}

// The filter adjoint kernel: 256-thread blocks (four per CU, so one block's
// staging overlaps the others' products), TM pairs per tile, the K = 3H sum
// of the product in three chunks of H columns staged in LDS (for TAN both gW
// and dgW).  The 4 waves split the CT = H/16 output column tiles (NTW each)
// and, when there are fewer than 4, the tile's row tiles (RG groups).
constexpr int kTaBlock = 256;

// mma_tile_k over one or (PAIR) two A operands (the adjoint and its tangent,
// same rows) with one B operand: each B fetch, PF slabs ahead, feeds both
// products.  hook() runs right after the first PF slabs' B fetches are
// issued: loads it issues then complete behind those slabs' MFMAs instead of
// delaying the first B wait (memory loads complete in issue order).
template <int RT, int NT, int K, bool PAIR, typename Hook, int PF = 2>
__device__ __forceinline__ void mma_tile_k_pair(const float* __restrict__ A0,
                                                const float* __restrict__ A1, int lda,
                                                const float* __restrict__ B, int ldb,
                                                const int (&cb)[NT], f4 (&acc0)[RT][NT],
                                                f4 (&acc1)[RT][NT], Hook&& hook) {
  static_assert(K % 16 == 0, "K must be a multiple of 16");
  constexpr int S = K / 16;
  const int lane = threadIdx.x & 63;
  const int r = lane & 15;
  const int g = lane >> 4;
  const int aoff = r * lda + 4 * g;
  int boff[NT];
#pragma unroll
  for (int ct = 0; ct < NT; ++ct) boff[ct] = (cb[ct] + r) * ldb + 4 * g;
  f4 b[PF + 1][NT];
#pragma unroll
  for (int q = 0; q < PF; ++q)
#pragma unroll
    for (int ct = 0; ct < NT; ++ct)
      if (q < S) b[q][ct] = *reinterpret_cast<const f4*>(B + boff[ct] + 16 * q);
  hook();
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    if (sl + PF < S) {
#pragma unroll
      for (int ct = 0; ct < NT; ++ct)
        b[(sl + PF) % (PF + 1)][ct] = *reinterpret_cast<const f4*>(B + boff[ct] + 16 * (sl + PF));
    }
    f4 a0[RT], a1[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      a0[rt] = *reinterpret_cast<const f4*>(A0 + aoff + rt * 16 * lda + 16 * sl);
      if (PAIR) a1[rt] = *reinterpret_cast<const f4*>(A1 + aoff + rt * 16 * lda + 16 * sl);
    }
    const int cur = sl % (PF + 1);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) {
          acc0[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[rt][s], b[cur][ct][s],
                                                              acc0[rt][ct], 0, 0, 0);
          if (PAIR)
            acc1[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[rt][s], b[cur][ct][s],
                                                                acc1[rt][ct], 0, 0, 0);
        }
  }
}
constexpr int kTaWaves = kTaBlock / 64;
template <int H>
constexpr int kTaPairs = H >= 128 ? 32 : 64;

template <int H, int TM, bool TAN>
__global__ __launch_bounds__(kTaBlock, 4) void train_filter_adjoint_kernel(
    int P, const float* __restrict__ gw, const float* __restrict__ dgw,
    const float* __restrict__ w2t, const float* __restrict__ a, const float* __restrict__ da,
    float* __restrict__ ga, float* __restrict__ dga) {
  constexpr int KC = H;                 // columns per chunk
  constexpr int LDA = KC + 4;
  constexpr int CT = H / 16;
  constexpr int NTW = CT >= kTaWaves ? CT / kTaWaves : 1;
  constexpr int RG = CT >= kTaWaves ? 1 : kTaWaves / CT;
  constexpr int RTW = TM / 16 / RG;
  static_assert(NTW * kTaWaves == CT * RG && (TM / 16) % RG == 0, "wave split");
  constexpr int NSTAGE = TAN ? 2 : 1;
  __shared__ __align__(16) float a_lds[NSTAGE][TM * LDA];
  const int ntiles = (P + TM - 1) / TM;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const int cw = wave % (CT / NTW), rg = wave / (CT / NTW);
  int cb[NTW];
#pragma unroll
  for (int ct = 0; ct < NTW; ++ct) cb[ct] = (cw * NTW + ct) * 16;
  // Each thread's share of a chunk (NV float4s per operand) is fetched into
  // registers while the previous chunk's products run, then stored to LDS
  // (PRE, the tangent kernels; 136 -> 132 µs at H = 128 on 50,132 pairs).  The
  // tangent kernel at H = 64 would spill with those registers live across the
  // products, and the forward-only kernels (one operand, half the products)
  // gain nothing: those stage each chunk straight from memory.
  constexpr bool PRE = TAN && H != 64;
  constexpr int NV = TM * (KC / 4) / kTaBlock;
  static_assert(NV * kTaBlock == TM * (KC / 4), "chunk split");
  f4 st_g[NV], st_dg[TAN ? NV : 1];
  auto fetch = [&](int t, int ch) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int idx = threadIdx.x + v * kTaBlock;
      const int m = idx / (KC / 4), k4 = idx - m * (KC / 4), p = t * TM + m;
      const f4 zero = {0.f, 0.f, 0.f, 0.f};
      const bool on = t < ntiles && p < P;
      const size_t o = on ? (size_t)p * 3 * H + ch * KC + 4 * k4 : 0;
      st_g[v] = on ? *reinterpret_cast<const f4*>(gw + o) : zero;
      if (TAN) st_dg[v] = on ? *reinterpret_cast<const f4*>(dgw + o) : zero;
    }
  };
  if (PRE) fetch(blockIdx.x, 0);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int pbase = tile * TM;
    f4 acc[RTW][NTW], dacc[RTW][NTW];
    zero_acc(acc);
    if (TAN) zero_acc(dacc);
#pragma unroll 1
    for (int ch = 0; ch < 3; ++ch) {
      __syncthreads();   // the previous chunk's products are done with the LDS
      if (PRE) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int idx = threadIdx.x + v * kTaBlock;
          const int m = idx / (KC / 4), k4 = idx - m * (KC / 4);
          *reinterpret_cast<f4*>(&a_lds[0][m * LDA + 4 * k4]) = st_g[v];
          if (TAN) *reinterpret_cast<f4*>(&a_lds[NSTAGE - 1][m * LDA + 4 * k4]) = st_dg[v];
        }
      } else {
        for (int idx = threadIdx.x; idx < TM * (KC / 4); idx += kTaBlock) {
          const int m = idx / (KC / 4), k4 = idx - m * (KC / 4), p = pbase + m;
          const size_t o = (size_t)p * 3 * H + ch * KC + 4 * k4;
          const f4 zero = {0.f, 0.f, 0.f, 0.f};
          *reinterpret_cast<f4*>(&a_lds[0][m * LDA + 4 * k4]) =
              p < P ? *reinterpret_cast<const f4*>(gw + o) : zero;
          if (TAN)
            *reinterpret_cast<f4*>(&a_lds[NSTAGE - 1][m * LDA + 4 * k4]) =
                p < P ? *reinterpret_cast<const f4*>(dgw + o) : zero;
        }
      }
      __syncthreads();
      if (TAN) {
        // the next chunk (this tile's, or the next tile's first), in flight
        // during this chunk's products
        const int nt = ch < 2 ? tile : tile + gridDim.x, nch = ch < 2 ? ch + 1 : 0;
        mma_tile_k_pair<RTW, NTW, KC, true>(&a_lds[0][rg * RTW * 16 * LDA],
                                            &a_lds[NSTAGE - 1][rg * RTW * 16 * LDA], LDA,
                                            w2t + ch * KC, 3 * H, cb, acc, dacc,
                                            [&] {
                                              if (PRE) fetch(nt, nch);
                                            });
      } else {
        mma_tile_k<RTW, NTW, KC>(&a_lds[0][rg * RTW * 16 * LDA], LDA, w2t + ch * KC, 3 * H,
                                 cb, acc);
      }
    }
    // lane (r, g) holds rows (rg RTW + rt) 16 + 4g + i of columns cb[ct] + r
#pragma unroll
    for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int p = pbase + (rg * RTW + rt) * 16 + 4 * g + i;
        if (p >= P) continue;
#pragma unroll
        for (int ct = 0; ct < NTW; ++ct) {
          const size_t o = (size_t)p * H + cb[ct] + r;
          const float x = a[o];
          // s = sigmoid(x); 1 - s as 1 / (1 + exp(x)) (exact for large x)
          const float sg = 1.0f / (1.0f + expf(-x));
          const float sm = 1.0f / (1.0f + expf(x));
          const float d1 = sg * (1.0f + x * sm);
          const float gh = acc[rt][ct][i];
          ga[o] = gh * d1;
          if (TAN) {
            const float d2 = sg * sm * (2.0f + x * (sm - sg));
            dga[o] = dacc[rt][ct][i] * d1 + gh * d2 * da[o];
          }
        }
      }
  }
}

__global__ void split_bf16x3_kernel(int64_t n, const float* __restrict__ src,
                                    uint16_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint16_t h, m, l;
    split3_bf16(src[i], h, m, l);
    dst[i] = h;
    dst[n + i] = m;
    dst[2 * n + i] = l;
  }
}

template <int H, bool TAN>
hipError_t launch_train_filter(hipStream_t s, int P, int R, int RP, const float* x,
                               const float* w1p, const float* b1, const uint16_t* w2bf,
                               const float* b2, const float* a_in, float* a_out, float* h_out,
                               float* w_out) {
  constexpr int TM = kTfPairs<H>;
  const int tiles = (P + TM - 1) / TM;
  // two blocks per CU fit the LDS; a few rounds of them
  const int grid = tiles < 2048 ? tiles : 2048;
  train_filter_kernel<H, TM, TAN><<<grid, kTfBlock, 0, s>>>(P, R, RP, x, w1p, b1, w2bf, b2,
                                                            a_in, a_out, h_out, w_out);
  return hipGetLastError();
}

template <int H, bool TAN>
hipError_t launch_train_filter_adjoint(hipStream_t s, int P, const float* gw, const float* dgw,
                                       const float* w2t, const float* a, const float* da,
                                       float* ga, float* dga) {
  constexpr int TM = kTaPairs<H>;
  const int tiles = (P + TM - 1) / TM;
  // one tile per block (the dispatcher balances the CUs), strided past 8192
  const int grid = tiles < 8192 ? tiles : 8192;
  train_filter_adjoint_kernel<H, TM, TAN><<<grid, kTaBlock, 0, s>>>(P, gw, dgw, w2t, a, da, ga,
                                                                    dga);
  return hipGetLastError();
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace
}  // namespace mlff

using namespace mlff;

extern "C" {

int mlff_split_bf16x3(void* stream, int64_t n, const float* src_dev, uint16_t* planes_dev) {
  if (n < 0 || (n > 0 && (!src_dev || !planes_dev))) {
    set_error("mlff_split_bf16x3: need n >= 0 and non-null buffers");
    return MLFF_ERR_INVALID;
  }
  if (n == 0) return MLFF_OK;
  const int64_t blocks = (n + 255) / 256;
  split_bf16x3_kernel<<<(unsigned)(blocks < 1024 ? blocks : 1024), 256, 0,
                        static_cast<hipStream_t>(stream)>>>(n, src_dev, planes_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_split_bf16x3: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_train_filter(void* stream, int tangent, int hidden, int n_pairs, int num_rbf,
                      const float* x_dev, const float* w1_padded_dev, const float* b1_dev,
                      const uint16_t* w2_planes_dev, const float* b2_dev, const float* a_dev,
                      float* a_out_dev, float* h_out_dev, float* w_out_dev) {
  const char* nm = "mlff_train_filter";
  if (hidden != 32 && hidden != 64 && hidden != 128) {
    set_error("%s: hidden_dim %d not supported (32, 64, 128)", nm, hidden);
    return MLFF_ERR_UNSUPPORTED;
  }
  if (n_pairs < 0 || num_rbf < 1 || num_rbf > 32) {
    set_error("%s: need n_pairs >= 0 and 1 <= num_rbf <= 32", nm);
    return MLFF_ERR_INVALID;
  }
  const bool tan = tangent != 0;
  if (!x_dev || !w1_padded_dev || !w2_planes_dev || !a_out_dev || !h_out_dev || !w_out_dev ||
      (tan ? !a_dev : (!b1_dev || !b2_dev))) {
    set_error("%s: null buffer (the tangent needs a, the forward b1 and b2)", nm);
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(w1_padded_dev) || !aligned16(w2_planes_dev)) {
    set_error("%s: padded W1 and the W2 planes must be 16-byte aligned", nm);
    return MLFF_ERR_INVALID;
  }
  if (n_pairs == 0) return MLFF_OK;
  const int RP = num_rbf <= 16 ? 16 : 32;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden * 2 + (tan ? 1 : 0)) {
#define MLFF_TF(HH, TT)                                                                      \
  case HH * 2 + TT:                                                                          \
    e = launch_train_filter<HH, TT>(s, n_pairs, num_rbf, RP, x_dev, w1_padded_dev, b1_dev, \
                                    w2_planes_dev, b2_dev, a_dev, a_out_dev, h_out_dev,     \
                                    w_out_dev);                                              \
    break;
    MLFF_TF(32, false)
    MLFF_TF(32, true)
    MLFF_TF(64, false)
    MLFF_TF(64, true)
    MLFF_TF(128, false)
    MLFF_TF(128, true)
#undef MLFF_TF
    default:
      e = hipErrorInvalidValue;
  }
  if (e != hipSuccess) {
    set_error("%s: %s", nm, hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_train_filter_adjoint(void* stream, int tangent, int hidden, int n_pairs,
                              const float* gw_dev, const float* dgw_dev, const float* w2t_dev,
                              const float* a_dev, const float* da_dev, float* ga_out_dev,
                              float* dga_out_dev) {
  const char* nm = "mlff_train_filter_adjoint";
  if (hidden != 32 && hidden != 64 && hidden != 128) {
    set_error("%s: hidden_dim %d not supported (32, 64, 128)", nm, hidden);
    return MLFF_ERR_UNSUPPORTED;
  }
  if (n_pairs < 0) {
    set_error("%s: need n_pairs >= 0", nm);
    return MLFF_ERR_INVALID;
  }
  const bool tan = tangent != 0;
  if (!gw_dev || !w2t_dev || !a_dev || !ga_out_dev ||
      (tan && (!dgw_dev || !da_dev || !dga_out_dev))) {
    set_error("%s: null buffer (the tangent needs dgw, da and dga_out)", nm);
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(gw_dev) || !aligned16(w2t_dev) || (tan && !aligned16(dgw_dev))) {
    set_error("%s: gw, dgw and w2t must be 16-byte aligned", nm);
    return MLFF_ERR_INVALID;
  }
  if (n_pairs == 0) return MLFF_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden * 2 + (tan ? 1 : 0)) {
#define MLFF_TA(HH, TT)                                                                    \
  case HH * 2 + TT:                                                                        \
    e = launch_train_filter_adjoint<HH, TT>(s, n_pairs, gw_dev, dgw_dev, w2t_dev, a_dev,   \
                                            da_dev, ga_out_dev, dga_out_dev);              \
    break;
    MLFF_TA(32, false)
    MLFF_TA(32, true)
    MLFF_TA(64, false)
    MLFF_TA(64, true)
    MLFF_TA(128, false)
    MLFF_TA(128, true)
#undef MLFF_TA
    default:
      e = hipErrorInvalidValue;
  }
  if (e != hipSuccess) {
    set_error("%s: %s", nm, hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

}  // extern "C"
