// PaiNN student energy and hand-written adjoint (forces, virial) on gfx950.
//
// Forward follows StudentForceField.forward (student_model.py:634-757):
//   embedding -> L x [message (PaiNNMessage :346-387), update (PaiNNUpdate
//   :434-470)] -> energy_head (:599-605) -> per-structure sum (FP64 here).
// Forces are -dE/dx from the adjoint below instead of torch.autograd
// (student_model.py:788-793).
//
// Data layout (all float32, row-major): node scalars s [N][H], node vectors
// v [N][3][H] (xyz in the middle, as the reference); per unordered pair p
// (lo < hi): geometry {vec = p_hi - p_lo (+ image shift), d} and the filter
// output W_p [3H] = (W_s | W_v1 | W_v2) of each block.
//
// Kernels and what bounds them:
//   filter_fwd   pairs, MFMA (R->H, SiLU, H->3H; the H->3H product on the bf16
//                matrix cores with three-way split operands); 2(RH + 3H^2) flop/pair
//   message      receiver rows, gather of s_j, v_j, W_p (HBM/L2)
//   update_fwd   atoms, MFMA (2H->H, SiLU, H->3H) + gated vector epilogue
//   readout      atoms, MFMA head and its input gradient in one pass
//   update_bwd   atoms, MFMA (3H->H, H->2H) transposed products
//   filter_bwd   pairs, gathers both endpoints' features and gradients, MFMA
//                (3H->H, H->R) back to dE/dd and dE/du
//   force_gather receiver rows, forces and per-atom virial from dE/dvec
//   reduce       per structure FP64 sums (energy, virial)

#include "common.hpp"
#include "internal.hpp"

namespace mlff {

namespace {

constexpr int kBlock = 256;
constexpr int kPairBlock = 512;       // 8 waves: each owns few output tiles, many in flight
constexpr int kPairWaves = kPairBlock / 64;
constexpr int kMaxRbf = 32;          // radial basis size limit (padded to 16: 16 or 32)
// pairs per tile, filter forward: 128 for the compact hidden size, 64 above
template <int H>
constexpr int kFwdPairs = H <= 32 ? 128 : 64;
// pairs per tile, filter adjoint: 64 for the compact hidden size (its tiles are
// short, so fewer, longer ones), 32 above (two blocks per CU fit the LDS)
template <int H>
constexpr int kBwdPairs = H <= 32 ? 64 : 32;
constexpr int kUpdWaves = 8;          // update kernels: waves per 32-atom tile
constexpr int kUpdBlock = 64 * kUpdWaves;
// edges in flight per row in the message kernels (forward, adjoint)
constexpr int kMsgUnrollFwd = 5;
constexpr int kMsgUnrollBwd = 4;
// The other launch shape of each kernel, pre-instantiated so the launch-shape
// tuner can choose per workload (mlff_plan_set_tiles).  Results are
// bit-identical across shapes: every output element sums in the same order.
template <int H>
constexpr int kFwdPairsAlt = H <= 32 ? 64 : 32;
template <int H>
constexpr int kBwdPairsAlt = H <= 32 ? 32 : 64;
constexpr int kMsgUnrollFwdAlt = 3;
constexpr int kMsgUnrollBwdAlt = 2;

// Grid of a pair kernel: the plan's persistent grid, capped at the number of
// tiles the pair capacity allows (a multiple of 8 for the XCD-interleaved tile
// order), so small batches do not dispatch thousands of idle workgroups.
inline int pair_grid(const Buffers& b, int pairs_per_tile) {
  const int tiles = (b.pcap + pairs_per_tile - 1) / pairs_per_tile;
  return std::max(8, std::min(b.blocks, (tiles + 7) / 8 * 8));
}

// S0 = emb[Z] and V0 = 0, four channels per thread: thread t of an atom's
// H/4 threads writes S0 channels 4t..4t+3 and the same channels of the
// three V0 components (16-byte stores; rows are 16-byte aligned since H % 4 == 0)
template <int H>
__global__ void embed_kernel(int N, const int64_t* __restrict__ Z,
                             const float* __restrict__ emb, float* __restrict__ S0,
                             float* __restrict__ V0) {
  constexpr int Q = H / 4;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * Q) return;
  const int i = idx / Q, c = idx % Q;
  const float4 e = reinterpret_cast<const float4*>(emb + Z[i] * H)[c];
  reinterpret_cast<float4*>(S0)[idx] = e;
  float4* v = reinterpret_cast<float4*>(V0) + (size_t)i * 3 * Q + c;
  const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
  v[0] = z;
  v[Q] = z;
  v[2 * Q] = z;
}

// ---------------------------------------------------------------------------
// filter forward: W_p = W2 silu(W1 e(d_p) + b1) + b2 for kFwdPairs<H> pairs per tile
// ---------------------------------------------------------------------------
template <int H, int TM>
__global__ __launch_bounds__(kPairBlock, 4) void filter_fwd_kernel(
    const float4* __restrict__ geom, const int32_t* __restrict__ counts, int32_t pcap,
    const float* __restrict__ w, LayerW lw, const float* __restrict__ centers,
    const float* __restrict__ widths, int R, int RP, float rc, float* __restrict__ Wout,
    const uint16_t* __restrict__ w2bf) {
  constexpr int RT = TM / 16;
  constexpr int LDE = kMaxRbf + 4;
  constexpr int LDHB = H + 8;          // bf16 row stride of the split hidden activations
  constexpr int PHB = TM * LDHB;       // plane stride
  constexpr int CT1 = H / 16, NT1 = (CT1 + kPairWaves - 1) / kPairWaves;
  constexpr int CT2 = 3 * H / 16, NT2 = (CT2 + kPairWaves - 1) / kPairWaves;
  static_assert(CT1 % NT1 == 0 && CT2 % NT2 == 0, "tile split");
  __shared__ __align__(16) float e_lds[TM * LDE];
  __shared__ __align__(16) uint16_t h_bf[3 * PHB];   // silu(W1 e + b1) as three bf16 planes
  __shared__ float mu_lds[kMaxRbf], gam_lds[kMaxRbf];
  __shared__ float d_lds[TM], fc_lds[TM];
  if (threadIdx.x < kMaxRbf) {
    const int k = threadIdx.x;
    mu_lds[k] = k < R ? centers[k] : 0.f;
    gam_lds[k] = k < R ? rbf_gamma(widths[k]) : 0.f;
  }
  const int P = min(counts[MLFF_COUNT_PAIRS], pcap);
  const int ntiles = (P + TM - 1) / TM;
  int t0, t1;
  int tstep;
  persistent_range(ntiles, t0, t1, tstep);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const float* f1w = w + lw.f1w;
  const float* f1b = w + lw.f1b;
  const float* f2b = w + lw.f2b;
  for (int tile = t0; tile < t1; tile += tstep) {
    const int pbase = tile * TM;
    if (threadIdx.x < TM) {   // the cutoff once per pair
      const int p = pbase + threadIdx.x;
      float d = 0.f, fc = 0.f, dfc;
      if (p < P) {
        d = geom[p].w;
        cosine_cutoff(d, rc, fc, dfc);
      }
      d_lds[threadIdx.x] = d;
      fc_lds[threadIdx.x] = fc;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < TM * RP; idx += kPairBlock) {
      const int m = idx / RP, k = idx - m * RP;
      const int p = pbase + m;
      float val = 0.f;
      if (p < P && k < R) val = edge_feature(d_lds[m], mu_lds[k], gam_lds[k], fc_lds[m]);
      e_lds[m * LDE + k] = val;
    }
    __syncthreads();
    if (wave < CT1 / NT1) {
      int cb[NT1];
      f4 acc[RT][NT1];
#pragma unroll
      for (int ct = 0; ct < NT1; ++ct) cb[ct] = (wave * NT1 + ct) * 16;
      zero_acc(acc);
      mma_tile<RT, NT1>(e_lds, LDE, f1w, RP, RP, cb, acc);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT1; ++ct) {
          const int c = cb[ct] + r;
          const float bias = f1b[c];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int m = rt * 16 + 4 * g + i;
            uint16_t hh, hm, hl;
            split3_bf16(silu(acc[rt][ct][i] + bias), hh, hm, hl);
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
      // W_p = h W2^T + b2 on the bf16 matrix cores at float32 accuracy
      mma_tile_x6<RT, NT2, H>(h_bf, LDHB, PHB, w2bf, H, (int)wbf_plane(H), cb, acc);
      float bias[NT2];
#pragma unroll
      for (int ct = 0; ct < NT2; ++ct) bias[ct] = f2b[cb[ct] + r];
      // lane (r, g) holds rows rt*16 + 4g + i, columns cb[ct] + r.  W_p is
      // written once and read by a later kernel: non-temporal stores keep it
      // from displacing L2 lines.
      float* out = Wout + (size_t)(pbase + 4 * g) * 3 * H + r;
      if (pbase + TM <= P) {   // full tile: no per-element bounds checks
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int ct = 0; ct < NT2; ++ct)
              __builtin_nontemporal_store(acc[rt][ct][i] + bias[ct],
                                          &out[(size_t)(rt * 16 + i) * 3 * H + cb[ct]]);
      } else {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if (pbase + rt * 16 + 4 * g + i >= P) continue;
#pragma unroll
            for (int ct = 0; ct < NT2; ++ct)
              __builtin_nontemporal_store(acc[rt][ct][i] + bias[ct],
                                          &out[(size_t)(rt * 16 + i) * 3 * H + cb[ct]]);
          }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// message (forward) and its node-gradient transpose (backward).
// FWD:  s'_i = s_i + sum_j s_j * Ws_p ;  v'_i = v_i + sum_j v_j * Wv1_p + u_ji (x) Wv2_p
// BWD:  gs_i = gs'_i + sum_j gs'_j * Ws_p ; gv_i = gv'_i + sum_j gv'_j * Wv1_p
// One row per H/4 lanes, four channels (one float4) per lane; neighbours in
// ascending j, a few edges' operands loaded before any is added.
// ---------------------------------------------------------------------------
template <int H, bool FWD, int UNROLL>
__global__ __launch_bounds__(kBlock) void message_kernel(
    int N, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ edge_pair, int32_t ecap, const float4* __restrict__ geom,
    const float* __restrict__ Wf, const float* __restrict__ s_in, const float* __restrict__ v_in,
    float* __restrict__ s_out, float* __restrict__ v_out) {
  constexpr int LPR = H / 4;
  constexpr int RPB = kBlock / LPR;
  const int sub = threadIdx.x % LPR;
  const int i = xcd_block() * RPB + threadIdx.x / LPR;
  if (i >= N) return;
  const int c = 4 * sub;
  f4 as = {0.f, 0.f, 0.f, 0.f}, a0 = as, a1 = as, a2 = as;
  const int e0 = min(row_ptr[i], ecap), e1 = min(row_ptr[i + 1], ecap);
  // operands of one edge; kUnroll edges are loaded before any is accumulated
  // (more loads in flight per row), and they are accumulated in edge order,
  // so the sums round exactly as a one-edge-at-a-time loop
  struct Edge {
    f4 ws, wv1, wv2, sj, v0, v1, v2;
    float ux, uy, uz;
    bool on;
  };
  auto load = [&](int j, int p) {
    Edge q;
    q.on = p >= 0;
    const int pc = q.on ? p : 0;
    const float* wp = Wf + (size_t)pc * 3 * H + c;
    const float* vj = v_in + (size_t)j * 3 * H + c;
    q.ws = *reinterpret_cast<const f4*>(wp);
    q.wv1 = *reinterpret_cast<const f4*>(wp + H);
    q.sj = *reinterpret_cast<const f4*>(s_in + (size_t)j * H + c);
    q.v0 = *reinterpret_cast<const f4*>(vj);
    q.v1 = *reinterpret_cast<const f4*>(vj + H);
    q.v2 = *reinterpret_cast<const f4*>(vj + 2 * H);
    if (FWD) {
      q.wv2 = *reinterpret_cast<const f4*>(wp + 2 * H);
      const float4 gm = geom[pc];
      const float den = gm.w + kEdgeEps;
      // u_ji = (p_j - p_i)/(d + eps): +pair vector when j is the pair's hi atom
      q.ux = gm.x / den; q.uy = gm.y / den; q.uz = gm.z / den;
      if (j < i) { q.ux = -q.ux; q.uy = -q.uy; q.uz = -q.uz; }
    }
    return q;
  };
  auto accumulate = [&](const Edge& q) {
    if (!q.on) return;
    as += q.sj * q.ws;
    a0 += q.v0 * q.wv1;
    a1 += q.v1 * q.wv1;
    a2 += q.v2 * q.wv1;
    if (FWD) {
      a0 += q.ux * q.wv2;
      a1 += q.uy * q.wv2;
      a2 += q.uz * q.wv2;
    }
  };
  constexpr int kUnroll = UNROLL;
  // The row's lanes load the neighbour and pair indices of a batch of
  // kBatch edges together (one edge per lane), the next batch's while this
  // one is processed; each group of kUnroll edges takes its indices from the
  // row's lanes by shuffle, so its operand loads wait for no index load.
  // Past the row's end, edges are switched off (p = -1).
  constexpr int kBatch = (LPR / kUnroll) * kUnroll;
  static_assert(kBatch >= kUnroll, "a batch holds at least one group");
  const int rbase = (threadIdx.x & 63) - sub;   // the row's first lane in the wave
  int jn = i, pn = -1;
  if (sub < kBatch && e0 + sub < e1) { jn = col[e0 + sub]; pn = edge_pair[e0 + sub]; }
  for (int bs = e0; bs < e1; bs += kBatch) {
    const int jc = jn, pcur = pn;
    jn = i; pn = -1;
    if (sub < kBatch && bs + kBatch + sub < e1) {
      jn = col[bs + kBatch + sub];
      pn = edge_pair[bs + kBatch + sub];
    }
    const int nb = min(kBatch, e1 - bs);
    for (int k = 0; k < nb; k += kUnroll) {
      Edge q[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        q[u] = load(__shfl(jc, rbase + k + u, 64), __shfl(pcur, rbase + k + u, 64));
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) accumulate(q[u]);
    }
  }
  const float* vi = v_in + (size_t)i * 3 * H + c;
  float* vo = v_out + (size_t)i * 3 * H + c;
  *reinterpret_cast<f4*>(s_out + (size_t)i * H + c) =
      *reinterpret_cast<const f4*>(s_in + (size_t)i * H + c) + as;
  *reinterpret_cast<f4*>(vo) = *reinterpret_cast<const f4*>(vi) + a0;
  *reinterpret_cast<f4*>(vo + H) = *reinterpret_cast<const f4*>(vi + H) + a1;
  *reinterpret_cast<f4*>(vo + 2 * H) = *reinterpret_cast<const f4*>(vi + 2 * H) + a2;
}

// ---------------------------------------------------------------------------
// update forward, 32 atoms per tile
// ---------------------------------------------------------------------------
template <int H>
__global__ __launch_bounds__(kUpdBlock) void update_fwd_kernel(
    int N, const float* __restrict__ w, LayerW lw, const float* __restrict__ s_mid,
    const float* __restrict__ v_mid, float* __restrict__ s_out, float* __restrict__ v_out,
    float* __restrict__ a_upd, float* __restrict__ g12) {
  constexpr int TM = 32, RT = 2;
  constexpr int LDX = 2 * H + 4, LDH = H + 4;
  constexpr int CT1 = H / 16, NT1 = (CT1 + kUpdWaves - 1) / kUpdWaves;
  constexpr int NTC = NT1, NT2 = 3 * NTC;
  __shared__ __align__(16) float x_lds[TM * LDX];
  __shared__ __align__(16) float h_lds[TM * LDH];
  constexpr int PER = TM * H / kUpdBlock;   // tile elements per thread
  static_assert(PER * kUpdBlock == TM * H, "tile split");
  const int base = blockIdx.x * TM;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  {
    // every load of the tile in flight at once (rows past N read atom N-1
    // and are zeroed), then the norms
    float sv[PER], x0[PER], x1[PER], x2[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = threadIdx.x + k * kUpdBlock, m = idx / H, c = idx % H;
      const int i = min(base + m, N - 1);
      sv[k] = s_mid[(size_t)i * H + c];
      const float* vv = v_mid + (size_t)i * 3 * H;
      x0[k] = vv[c];
      x1[k] = vv[H + c];
      x2[k] = vv[2 * H + c];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = threadIdx.x + k * kUpdBlock, m = idx / H, c = idx % H;
      const bool in = base + m < N;
      x_lds[m * LDX + c] = in ? sv[k] : 0.f;
      x_lds[m * LDX + H + c] = in ? norm3_exact(x0[k], x1[k], x2[k]) : 0.f;
    }
  }
  __syncthreads();
  if (wave < CT1 / NT1) {
    int cb[NT1];
    f4 acc[RT][NT1];
#pragma unroll
    for (int ct = 0; ct < NT1; ++ct) cb[ct] = (wave * NT1 + ct) * 16;
    zero_acc(acc);
    mma_tile_k<RT, NT1, 2 * H>(x_lds, LDX, w + lw.u1w, 2 * H, cb, acc);
    const float* b1 = w + lw.u1b;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int ct = 0; ct < NT1; ++ct) {
        const int c = cb[ct] + r;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int m = rt * 16 + 4 * g + k, i = base + m;
          const float a = acc[rt][ct][k] + b1[c];
          if (i < N) a_upd[(size_t)i * H + c] = a;
          h_lds[m * LDH + c] = silu(a);
        }
      }
  }
  __syncthreads();
  if (wave < CT1 / NTC) {
    int cb[NT2];
    f4 acc[RT][NT2];
#pragma unroll
    for (int q = 0; q < NTC; ++q)
#pragma unroll
      for (int sp = 0; sp < 3; ++sp) cb[sp * NTC + q] = sp * H + (wave * NTC + q) * 16;
    // the epilogue's operands (biases, mixing matrix, the tile's vectors)
    // are loaded before the product, so their latency hides behind it
    const float* b2 = w + lw.u2b;
    const float* U = w + lw.mix;
    float u[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) u[k] = U[k];
    float bias[NTC][3];
    float xv[RT][NTC][4][3];
#pragma unroll
    for (int q = 0; q < NTC; ++q) {
      const int c = (wave * NTC + q) * 16 + r;
#pragma unroll
      for (int sp = 0; sp < 3; ++sp) bias[q][sp] = b2[sp * H + c];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int i = min(base + rt * 16 + 4 * g + k, N - 1);
          const float* vv = v_mid + (size_t)i * 3 * H;
#pragma unroll
          for (int sp = 0; sp < 3; ++sp) xv[rt][q][k][sp] = vv[sp * H + c];
        }
    }
    zero_acc(acc);
    mma_tile_k<RT, NT2, H>(h_lds, LDH, w + lw.u2w, H, cb, acc);
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int q = 0; q < NTC; ++q) {
        const int c = (wave * NTC + q) * 16 + r;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int m = rt * 16 + 4 * g + k, i = base + m;
          if (i >= N) continue;
          const float ds = acc[rt][q][k] + bias[q][0];
          const float gt1 = acc[rt][NTC + q][k] + bias[q][1];
          const float gt2 = acc[rt][2 * NTC + q][k] + bias[q][2];
          s_out[(size_t)i * H + c] = x_lds[m * LDX + c] + ds;
          const float x0 = xv[rt][q][k][0], x1 = xv[rt][q][k][1], x2 = xv[rt][q][k][2];
          float* vo = v_out + (size_t)i * 3 * H;
          vo[c] = x0 * gt1 + (u[0] * x0 + u[1] * x1 + u[2] * x2) * gt2;
          vo[H + c] = x1 * gt1 + (u[3] * x0 + u[4] * x1 + u[5] * x2) * gt2;
          vo[2 * H + c] = x2 * gt1 + (u[6] * x0 + u[7] * x1 + u[8] * x2) * gt2;
          g12[(size_t)i * 2 * H + c] = gt1;
          g12[(size_t)i * 2 * H + H + c] = gt2;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// readout: eps_i = head(s_i) and gs_i = d eps_i / d s_i, 32 atoms per tile
// ---------------------------------------------------------------------------
template <int H>
__global__ __launch_bounds__(kBlock) void readout_kernel(int N, const float* __restrict__ w,
                                                         ModelW mw, const float* __restrict__ sL,
                                                         float* __restrict__ eps,
                                                         float* __restrict__ gS) {
  constexpr int TM = 32, RT = 2;
  constexpr int HP2 = ((H / 2) + 15) & ~15, HP4 = ((H / 4) + 15) & ~15;
  constexpr int LDX = H + 4, LD2 = HP2 + 4, LD4 = HP4 + 4;
  __shared__ __align__(16) float x_lds[TM * LDX];
  __shared__ __align__(16) float a1_lds[TM * LD2];
  __shared__ __align__(16) float h1_lds[TM * LD2];
  __shared__ __align__(16) float a2_lds[TM * LD4];
  __shared__ __align__(16) float g2_lds[TM * LD4];
  const int base = blockIdx.x * TM;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  {
    constexpr int PER = TM * H / kBlock;
    static_assert(PER * kBlock == TM * H, "tile split");
    float xs[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = threadIdx.x + k * kBlock, m = idx / H, c = idx % H;
      xs[k] = sL[(size_t)min(base + m, N - 1) * H + c];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int idx = threadIdx.x + k * kBlock, m = idx / H, c = idx % H;
      x_lds[m * LDX + c] = (base + m < N) ? xs[k] : 0.f;
    }
  }
  __syncthreads();
  if (wave < HP2 / 16) {   // Linear(H, H/2) + SiLU, one column tile per wave
    int cb[1] = {wave * 16};
    f4 acc[RT][1];
    zero_acc(acc);
    mma_tile_k<RT, 1, H>(x_lds, LDX, w + mw.r1w, H, cb, acc);
    const int c = cb[0] + r;
    const float bias = w[mw.r1b + c];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int m = rt * 16 + 4 * g + k;
        const float a = acc[rt][0][k] + bias;
        a1_lds[m * LD2 + c] = a;
        h1_lds[m * LD2 + c] = silu(a);
      }
  }
  __syncthreads();
  if (wave < HP4 / 16) {   // Linear(H/2, H/4)
    int cb[1] = {wave * 16};
    f4 acc[RT][1];
    zero_acc(acc);
    mma_tile_k<RT, 1, HP2>(h1_lds, LD2, w + mw.r2w, HP2, cb, acc);
    const int c = cb[0] + r;
    const float bias = w[mw.r2b + c];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int k = 0; k < 4; ++k) a2_lds[(rt * 16 + 4 * g + k) * LD4 + c] = acc[rt][0][k] + bias;
  }
  __syncthreads();
  const float* w3 = w + mw.r3w;
  if (threadIdx.x < TM) {   // Linear(H/4, 1), its input gradient, and SiLU'
    const int m = threadIdx.x, i = base + m;
    float e = 0.f;
    for (int c = 0; c < HP4; ++c) {
      const float a = a2_lds[m * LD4 + c];
      e += silu(a) * w3[c];
      g2_lds[m * LD4 + c] = w3[c] * silu_grad(a);
    }
    if (i < N) eps[i] = e + w[mw.r3b];
  }
  __syncthreads();
  if (wave < HP2 / 16) {   // d/dh1 = g2 W2, then SiLU'
    int cb[1] = {wave * 16};
    f4 acc[RT][1];
    zero_acc(acc);
    mma_tile_k<RT, 1, HP4>(g2_lds, LD4, w + mw.r2t, HP4, cb, acc);
    const int c = cb[0] + r;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int m = rt * 16 + 4 * g + k;
        h1_lds[m * LD2 + c] = acc[rt][0][k] * silu_grad(a1_lds[m * LD2 + c]);
      }
  }
  __syncthreads();
  {   // d/ds = g1 W1 over H output columns
    constexpr int CT = H / 16, NT = (CT + 3) / 4;
    if (wave < CT / NT) {
      int cb[NT];
      f4 acc[RT][NT];
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) cb[ct] = (wave * NT + ct) * 16;
      zero_acc(acc);
      mma_tile_k<RT, NT, HP2>(h1_lds, LD2, w + mw.r1t, HP2, cb, acc);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int i = base + rt * 16 + 4 * g + k;
            if (i < N) gS[(size_t)i * H + cb[ct] + r] = acc[rt][ct][k];
          }
    }
  }
}

// ---------------------------------------------------------------------------
// update backward, 32 atoms per tile.  gv2 == nullptr means a zero gradient.
// ---------------------------------------------------------------------------
template <int H>
__global__ __launch_bounds__(kUpdBlock) void update_bwd_kernel(
    int N, const float* __restrict__ w, LayerW lw, const float* __restrict__ gs2,
    const float* __restrict__ gv2, const float* __restrict__ v_mid,
    const float* __restrict__ a_upd, const float* __restrict__ g12, float* __restrict__ gs1,
    float* __restrict__ gv1) {
  constexpr int TM = 32, RT = 2;
  constexpr int LDY = 3 * H + 4, LDA = H + 4, LDX = 2 * H + 4;
  __shared__ __align__(16) float y_lds[TM * LDY];
  __shared__ __align__(16) float a_lds[TM * LDA];
  const int base = blockIdx.x * TM;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  const float* U = w + lw.mix;
  float u[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) u[k] = U[k];
  // loads in batches of kB elements per thread, all of a batch in flight at
  // once (rows past N read atom N-1 and are zeroed)
  constexpr int PER = TM * H / kUpdBlock;
  constexpr int kB = PER < 4 ? PER : 4;
  static_assert(PER * kUpdBlock == TM * H && PER % kB == 0, "tile split");
#pragma unroll
  for (int kb = 0; kb < PER; kb += kB) {
    float y0[kB], x0[kB], x1[kB], x2[kB], q0[kB], q1[kB], q2[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      const int idx = threadIdx.x + (kb + k) * kUpdBlock, m = idx / H, c = idx % H;
      const int i = min(base + m, N - 1);
      y0[k] = gs2[(size_t)i * H + c];
      if (gv2 != nullptr) {
        const float* vv = v_mid + (size_t)i * 3 * H;
        const float* gg = gv2 + (size_t)i * 3 * H;
        x0[k] = vv[c]; x1[k] = vv[H + c]; x2[k] = vv[2 * H + c];
        q0[k] = gg[c]; q1[k] = gg[H + c]; q2[k] = gg[2 * H + c];
      }
    }
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      const int idx = threadIdx.x + (kb + k) * kUpdBlock, m = idx / H, c = idx % H;
      const bool in = base + m < N;
      float y1 = 0.f, y2 = 0.f;
      if (gv2 != nullptr) {
        y1 = q0[k] * x0[k] + q1[k] * x1[k] + q2[k] * x2[k];
        y2 = q0[k] * (u[0] * x0[k] + u[1] * x1[k] + u[2] * x2[k]) +
             q1[k] * (u[3] * x0[k] + u[4] * x1[k] + u[5] * x2[k]) +
             q2[k] * (u[6] * x0[k] + u[7] * x1[k] + u[8] * x2[k]);
      }
      y_lds[m * LDY + c] = in ? y0[k] : 0.f;
      y_lds[m * LDY + H + c] = in ? y1 : 0.f;
      y_lds[m * LDY + 2 * H + c] = in ? y2 : 0.f;
    }
  }
  __syncthreads();
  {
    constexpr int CT = H / 16, NT = (CT + kUpdWaves - 1) / kUpdWaves;
    if (wave < CT / NT) {
      int cb[NT];
      f4 acc[RT][NT];
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) cb[ct] = (wave * NT + ct) * 16;
      float av[RT][NT][4];   // pre-activations, loaded before the product
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int i = min(base + rt * 16 + 4 * g + k, N - 1);
            av[rt][ct][k] = a_upd[(size_t)i * H + cb[ct] + r];
          }
      zero_acc(acc);
      mma_tile_k<RT, NT, 3 * H, 2, true>(y_lds, LDY, w + lw.u2t, 3 * H, cb, acc);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int m = rt * 16 + 4 * g + k, i = base + m, c = cb[ct] + r;
            const float a = (i < N) ? av[rt][ct][k] : 0.f;
            a_lds[m * LDA + c] = acc[rt][ct][k] * silu_grad(a);
          }
    }
  }
  __syncthreads();
  float* x_lds = y_lds;   // y is dead after the barrier above
  {
    constexpr int CT = 2 * H / 16, NT = (CT + kUpdWaves - 1) / kUpdWaves;
    if (wave < CT / NT) {
      int cb[NT];
      f4 acc[RT][NT];
#pragma unroll
      for (int ct = 0; ct < NT; ++ct) cb[ct] = (wave * NT + ct) * 16;
      zero_acc(acc);
      mma_tile_k<RT, NT, H, 2, true>(a_lds, LDA, w + lw.u1t, H, cb, acc);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < NT; ++ct)
#pragma unroll
          for (int k = 0; k < 4; ++k) x_lds[(rt * 16 + 4 * g + k) * LDX + cb[ct] + r] = acc[rt][ct][k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int kb = 0; kb < PER; kb += kB) {
    float s2[kB], x0[kB], x1[kB], x2[kB], t1[kB], t2[kB], q0[kB], q1[kB], q2[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      const int idx = threadIdx.x + (kb + k) * kUpdBlock, m = idx / H, c = idx % H;
      const int i = min(base + m, N - 1);
      s2[k] = gs2[(size_t)i * H + c];
      const float* vv = v_mid + (size_t)i * 3 * H;
      x0[k] = vv[c]; x1[k] = vv[H + c]; x2[k] = vv[2 * H + c];
      if (gv2 != nullptr) {
        t1[k] = g12[(size_t)i * 2 * H + c];
        t2[k] = g12[(size_t)i * 2 * H + H + c];
        const float* gg = gv2 + (size_t)i * 3 * H;
        q0[k] = gg[c]; q1[k] = gg[H + c]; q2[k] = gg[2 * H + c];
      }
    }
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      const int idx = threadIdx.x + (kb + k) * kUpdBlock, m = idx / H, c = idx % H;
      const int i = base + m;
      if (i >= N) continue;
      gs1[(size_t)i * H + c] = s2[k] + x_lds[m * LDX + c];
      const float gn = x_lds[m * LDX + H + c];
      const float n = norm3_exact(x0[k], x1[k], x2[k]);
      const float sc = (n > 0.f) ? gn / n : 0.f;
      float o0 = sc * x0[k], o1 = sc * x1[k], o2 = sc * x2[k];
      if (gv2 != nullptr) {
        o0 += q0[k] * t1[k] + (u[0] * q0[k] + u[3] * q1[k] + u[6] * q2[k]) * t2[k];
        o1 += q1[k] * t1[k] + (u[1] * q0[k] + u[4] * q1[k] + u[7] * q2[k]) * t2[k];
        o2 += q2[k] * t1[k] + (u[2] * q0[k] + u[5] * q1[k] + u[8] * q2[k]) * t2[k];
      }
      float* go = gv1 + (size_t)i * 3 * H;
      go[c] = o0;
      go[H + c] = o1;
      go[2 * H + c] = o2;
    }
  }
}

// Partial dL/dd of a tile: (ga[:, k0:k0+KC] W1t[c0:c0+16, k0:k0+KC]^T) times
// de/dd for the 16 basis columns c0 + r, summed over those columns.
template <int RT, int KC, int TM>
__device__ __forceinline__ void gd_partial(const float* ga, int lda, const float* w1t, int ldb,
                                           int c0, int k0, const float* ed, float* out) {
  const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  int cb[1] = {c0};
  f4 acc[RT][1];
  zero_acc(acc);
  mma_tile_k<RT, 1, KC>(ga + k0, lda, w1t + k0, ldb, cb, acc);
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = rt * 16 + 4 * g + i;
      float t = acc[rt][0][i] * ed[m * kMaxRbf + c0 + r];
      t = group_sum<16>(t);
      if (r == 0) out[m] = t;
    }
}

// ---------------------------------------------------------------------------
// filter backward, kBwdPairs<H> pairs per tile.
// For pair p = (a < b): edge b->a carries s_b*Ws, v_b*Wv1 + u (x) Wv2 to a
// and edge a->b carries s_a*Ws, v_a*Wv1 - u (x) Wv2 to b, with
// u = (p_b - p_a)/(d + 1e-8).  The tile forms dL/dW_p from both directions,
// pushes it through W2^T, SiLU', W1^T to dL/de, then to dL/dd, and adds
// dL/du.  The running (dL/du, dL/dd) of all blocks is kept per pair; the
// last block processed turns it into dE/dvec.
// ---------------------------------------------------------------------------
// Pair indices of this thread's gather items of tile t (0 past the last tile
// or the last pair: those items are switched off by p >= P in the gather).
template <int TM, int C4, int NIT>
__device__ __forceinline__ void load_pair_indices(int t, int t1, int P,
                                                  const int32_t* __restrict__ pair_lo,
                                                  const int32_t* __restrict__ pair_hi,
                                                  int (&pa)[NIT], int (&pb)[NIT]) {
#pragma unroll
  for (int k = 0; k < NIT; ++k) {
    const int idx = threadIdx.x + k * kPairBlock;
    const int p = t * TM + idx / C4;
    const bool on = t < t1 && idx < TM * C4 && p < P;
    pa[k] = on ? pair_lo[p] : 0;
    pb[k] = on ? pair_hi[p] : 0;
  }
}

// PF: the pair indices of a thread's gather items of its next tile are loaded
// before the current tile's matrix products, so the next gather starts with
// its operand loads instead of a dependent index load (same arithmetic,
// bit-identical results).
template <int H, int TM, bool PF>
__global__ __launch_bounds__(kPairBlock, 4) void filter_bwd_kernel(
    const int32_t* __restrict__ counts, int32_t pcap, const int32_t* __restrict__ pair_lo,
    const int32_t* __restrict__ pair_hi, const float4* __restrict__ geom,
    const float* __restrict__ w, LayerW lw, const float* __restrict__ centers,
    const float* __restrict__ widths, int R, int RP, float rc, const float* __restrict__ Wf,
    const float* __restrict__ s_in, const float* __restrict__ v_in,
    const float* __restrict__ gs1, const float* __restrict__ gv1, float4* __restrict__ pacc,
    float4* __restrict__ gvec, int first, int last) {
  constexpr int RT = TM / 16;
  constexpr int LDE = kMaxRbf + 4, LDG = 3 * H + 4, LDA = H + 4;
  constexpr int C4 = H / 4;                    // lanes per pair in the gather (4 channels each)
  constexpr int NPART = 1;                     // one dE/du partial per pair
  __shared__ __align__(16) float e_lds[TM * LDE];
  __shared__ __align__(16) float gw_lds[TM * LDG];
  __shared__ __align__(16) float ga_lds[TM * LDA];
  __shared__ float gu_lds[TM * NPART * 3];
  __shared__ float ed_lds[TM * kMaxRbf];       // d e_k / d d of each pair
  __shared__ float gd_lds[kPairWaves * TM];    // dL/dd partials, one set per wave
  __shared__ float mu_lds[kMaxRbf], gam_lds[kMaxRbf];
  __shared__ float d_lds[TM], fc_lds[TM], dfc_lds[TM];
  if (threadIdx.x < kMaxRbf) {
    const int k = threadIdx.x;
    mu_lds[k] = k < R ? centers[k] : 0.f;
    gam_lds[k] = k < R ? rbf_gamma(widths[k]) : 0.f;
  }
  const int P = min(counts[MLFF_COUNT_PAIRS], pcap);
  const int ntiles = (P + TM - 1) / TM;
  int t0, t1;
  int tstep;
  persistent_range(ntiles, t0, t1, tstep);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
  constexpr int NIT = (TM * C4 + kPairBlock - 1) / kPairBlock;   // gather items per thread
  int pa[NIT], pb[NIT];
  if (PF) load_pair_indices<TM, C4, NIT>(t0, t1, P, pair_lo, pair_hi, pa, pb);
  for (int tile = t0; tile < t1; tile += tstep) {
    const int pbase = tile * TM;
    if (threadIdx.x < TM) {   // the cutoff once per pair
      const int p = pbase + threadIdx.x;
      float d = 0.f, fc = 0.f, dfc = 0.f;
      if (p < P) {
        d = geom[p].w;
        cosine_cutoff(d, rc, fc, dfc);
      }
      d_lds[threadIdx.x] = d;
      fc_lds[threadIdx.x] = fc;
      dfc_lds[threadIdx.x] = dfc;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < TM * RP; idx += kPairBlock) {
      const int m = idx / RP, k = idx - m * RP, p = pbase + m;
      float val = 0.f, der = 0.f;
      if (p < P && k < R)
        edge_feature_and_grad(d_lds[m], mu_lds[k], gam_lds[k], fc_lds[m], dfc_lds[m], val, der);
      e_lds[m * LDE + k] = val;
      ed_lds[m * kMaxRbf + k] = der;
    }
    // gather both endpoints, four channels per lane, every load of one item
    // issued up front (items one at a time: two blocks per CU fit the registers)
#pragma unroll 1
    for (int k = 0; k < NIT; ++k) {
      const int idx = threadIdx.x + k * kPairBlock;
      if (idx >= TM * C4) break;
      const int m = idx / C4, c = 4 * (idx - m * C4), p = pbase + m;
      f4 gws = {0.f, 0.f, 0.f, 0.f}, gwv1 = gws, gwv2 = gws;
      float gu0 = 0.f, gu1 = 0.f, gu2 = 0.f;
      if (p < P) {
        const int a = PF ? pa[k] : pair_lo[p], b = PF ? pb[k] : pair_hi[p];
        const float4 gm = geom[p];
        const float* va = v_in + (size_t)a * 3 * H + c;
        const float* vb = v_in + (size_t)b * 3 * H + c;
        const float* qa = gv1 + (size_t)a * 3 * H + c;
        const float* qb = gv1 + (size_t)b * 3 * H + c;
        const f4 sa = *reinterpret_cast<const f4*>(s_in + (size_t)a * H + c);
        const f4 sb = *reinterpret_cast<const f4*>(s_in + (size_t)b * H + c);
        const f4 ga = *reinterpret_cast<const f4*>(gs1 + (size_t)a * H + c);
        const f4 gb = *reinterpret_cast<const f4*>(gs1 + (size_t)b * H + c);
        const f4 va0 = *reinterpret_cast<const f4*>(va);
        const f4 va1 = *reinterpret_cast<const f4*>(va + H);
        const f4 va2 = *reinterpret_cast<const f4*>(va + 2 * H);
        const f4 vb0 = *reinterpret_cast<const f4*>(vb);
        const f4 vb1 = *reinterpret_cast<const f4*>(vb + H);
        const f4 vb2 = *reinterpret_cast<const f4*>(vb + 2 * H);
        const f4 qa0 = *reinterpret_cast<const f4*>(qa);
        const f4 qa1 = *reinterpret_cast<const f4*>(qa + H);
        const f4 qa2 = *reinterpret_cast<const f4*>(qa + 2 * H);
        const f4 qb0 = *reinterpret_cast<const f4*>(qb);
        const f4 qb1 = *reinterpret_cast<const f4*>(qb + H);
        const f4 qb2 = *reinterpret_cast<const f4*>(qb + 2 * H);
        const f4 wv2 = *reinterpret_cast<const f4*>(Wf + (size_t)p * 3 * H + 2 * H + c);
        const float den = gm.w + kEdgeEps;
        const float ux = gm.x / den, uy = gm.y / den, uz = gm.z / den;
        gws = ga * sb + gb * sa;
        gwv1 = qa0 * vb0 + qa1 * vb1 + qa2 * vb2 + qb0 * va0 + qb1 * va1 + qb2 * va2;
        const f4 d0 = qa0 - qb0, d1 = qa1 - qb1, d2 = qa2 - qb2;
        gwv2 = ux * d0 + uy * d1 + uz * d2;
        const f4 w0 = wv2 * d0, w1 = wv2 * d1, w2 = wv2 * d2;
        gu0 = (w0[0] + w0[1]) + (w0[2] + w0[3]);
        gu1 = (w1[0] + w1[1]) + (w1[2] + w1[3]);
        gu2 = (w2[0] + w2[1]) + (w2[2] + w2[3]);
      }
      *reinterpret_cast<f4*>(gw_lds + m * LDG + c) = gws;
      *reinterpret_cast<f4*>(gw_lds + m * LDG + H + c) = gwv1;
      *reinterpret_cast<f4*>(gw_lds + m * LDG + 2 * H + c) = gwv2;
      gu0 = group_sum<C4>(gu0);
      gu1 = group_sum<C4>(gu1);
      gu2 = group_sum<C4>(gu2);
      if (c == 0) {
        gu_lds[m * 3 + 0] = gu0;
        gu_lds[m * 3 + 1] = gu1;
        gu_lds[m * 3 + 2] = gu2;
      }
    }
    __syncthreads();
    if (PF) {   // the next tile's pair indices, in flight during the products below
      load_pair_indices<TM, C4, NIT>(tile + tstep, t1, P, pair_lo, pair_hi, pa, pb);
      __builtin_amdgcn_sched_barrier(0);
    }
    {
      // When the tile has at most one 16x16 output block per wave, each wave
      // takes one (row tile, column tile); otherwise every wave takes all row
      // tiles of its column tiles.  Each output element sums over K in the
      // same order either way.
      constexpr int CT = H / 16;
      constexpr bool kOneBlock = CT * RT <= kPairWaves;
      constexpr int RTW = kOneBlock ? 1 : RT;
      constexpr int NT = kOneBlock ? 1 : (CT + kPairWaves - 1) / kPairWaves;
      constexpr int NWB = kOneBlock ? CT * RT : CT / NT;
      if (wave < NWB) {
        const int r0 = kOneBlock ? wave / CT : 0;
        const int c0 = kOneBlock ? wave % CT : wave * NT;
        int cb[NT];
        f4 pre[RTW][NT], acc[RTW][NT];
#pragma unroll
        for (int ct = 0; ct < NT; ++ct) cb[ct] = (c0 + ct) * 16;
        zero_acc(pre);
        zero_acc(acc);
        mma_tile<RTW, NT>(e_lds + r0 * 16 * LDE, LDE, w + lw.f1w, RP, RP, cb, pre);
        mma_tile_k2<RTW, NT, 3 * H, 2, true>(gw_lds + r0 * 16 * LDG, LDG, w + lw.f2t, 3 * H, cb,
                                             acc);
        const float* b1 = w + lw.f1b;
#pragma unroll
        for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
          for (int ct = 0; ct < NT; ++ct) {
            const int c = cb[ct] + r;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int m = (r0 + rt) * 16 + 4 * g + k;
              ga_lds[m * LDA + c] = acc[rt][ct][k] * silu_grad(pre[rt][ct][k] + b1[c]);
            }
          }
      }
    }
    __syncthreads();
    // dL/dd = sum_k (ga W1)_k de_k/dd.  All waves share the W1 product: wave
    // w takes basis columns 16 (w % ctiles) + r and the K (hidden) slice
    // w / ctiles; each multiplies its partial dL/de by de/dd, sums its 16
    // columns by shuffles, and the partials of all waves are added in wave
    // order below (the sum is linear in dL/de).
    const int ctiles = RP / 16;
    if (ctiles == 2) {
      constexpr int NP = (H / 16 < 4) ? H / 16 : 4, KC = H / NP;
      if (wave < 2 * NP) gd_partial<RT, KC, TM>(ga_lds, LDA, w + lw.f1t, H, (wave % 2) * 16,
                                                (wave / 2) * KC, ed_lds, gd_lds + wave * TM);
    } else {
      constexpr int NP = (H / 16 < 8) ? H / 16 : 8, KC = H / NP;
      if (wave < NP) gd_partial<RT, KC, TM>(ga_lds, LDA, w + lw.f1t, H, 0, wave * KC, ed_lds,
                                            gd_lds + wave * TM);
    }
    __syncthreads();
    if (threadIdx.x < TM) {
      const int m = threadIdx.x, p = pbase + m;
      if (p < P) {
        const float4 gm = geom[p];
        const float d = gm.w;
        const int nparts = (RP > 16) ? 2 * ((H / 16 < 4) ? H / 16 : 4) : ((H / 16 < 8) ? H / 16 : 8);
        float gd = 0.f;
        for (int q = 0; q < nparts; ++q) gd += gd_lds[q * TM + m];
        float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int part = 0; part < NPART; ++part) {
          gx += gu_lds[(m * NPART + part) * 3 + 0];
          gy += gu_lds[(m * NPART + part) * 3 + 1];
          gz += gu_lds[(m * NPART + part) * 3 + 2];
        }
        float4 acc = first ? make_float4(0.f, 0.f, 0.f, 0.f) : pacc[p];
        acc.x += gx; acc.y += gy; acc.z += gz; acc.w += gd;
        if (last) {
          // dE/dvec = gd vec/d + gu/(d+eps) - vec (vec.gu)/(d (d+eps)^2)
          const float den = d + kEdgeEps;
          const float vdot = gm.x * acc.x + gm.y * acc.y + gm.z * acc.z;
          const float a1 = (d > 0.f) ? acc.w / d : 0.f;
          const float a2 = (d > 0.f) ? vdot / (d * den * den) : 0.f;
          gvec[p] = make_float4(a1 * gm.x + acc.x / den - a2 * gm.x,
                                a1 * gm.y + acc.y / den - a2 * gm.y,
                                a1 * gm.z + acc.z / den - a2 * gm.z, 0.f);
        } else {
          pacc[p] = acc;
        }
      }
    }
    __syncthreads();
  }
}

// forces F_i = sum_{j>i} G_p - sum_{j<i} G_p; per-atom virial sum_{j>i} vec (x) G
__global__ void force_gather_kernel(int N, const int32_t* __restrict__ row_ptr,
                                    const int32_t* __restrict__ col,
                                    const int32_t* __restrict__ edge_pair, int32_t ecap,
                                    const float4* __restrict__ gvec,
                                    const float4* __restrict__ geom, float* __restrict__ forces,
                                    float* __restrict__ vir_atom) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  float vr[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int e0 = min(row_ptr[i], ecap), e1 = min(row_ptr[i + 1], ecap);
  // one edge's contribution, in edge order
  auto add = [&](int j, int p, const float4& G, const float4& v) {
    if (p < 0) return;
    if (j > i) {
      fx += G.x; fy += G.y; fz += G.z;
      if (vir_atom != nullptr) {
        vr[0] += v.x * G.x; vr[1] += v.x * G.y; vr[2] += v.x * G.z;
        vr[3] += v.y * G.x; vr[4] += v.y * G.y; vr[5] += v.y * G.z;
        vr[6] += v.z * G.x; vr[7] += v.z * G.y; vr[8] += v.z * G.z;
      }
    } else {
      fx -= G.x; fy -= G.y; fz -= G.z;
    }
  };
  // four edges' loads in flight before any is added (one thread per atom
  // leaves about one wave per CU, so the loop is latency bound)
  constexpr int U = 4;
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  int e = e0;
  for (; e + U - 1 < e1; e += U) {
    int jj[U], pp[U];
    float4 GG[U], VV[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      jj[u] = col[e + u];
      pp[u] = edge_pair[e + u];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int pc = pp[u] >= 0 ? pp[u] : 0;
      GG[u] = gvec[pc];
      VV[u] = vir_atom != nullptr ? geom[pc] : zero;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) add(jj[u], pp[u], GG[u], VV[u]);
  }
  for (; e < e1; ++e) {
    const int p = edge_pair[e];
    const int pc = p >= 0 ? p : 0;
    add(col[e], p, gvec[pc], vir_atom != nullptr ? geom[pc] : zero);
  }
  if (forces != nullptr) {
    forces[3 * i] = fx;
    forces[3 * i + 1] = fy;
    forces[3 * i + 2] = fz;
  }
  if (vir_atom != nullptr)
    for (int k = 0; k < 9; ++k) vir_atom[9 * i + k] = vr[k];
}

// per-structure FP64 sums in a fixed order (one workgroup per structure)
__global__ __launch_bounds__(kBlock) void reduce_kernel(int B, const int32_t* __restrict__ mol_ptr,
                                                        const float* __restrict__ eps,
                                                        const float* __restrict__ vir_atom,
                                                        double* __restrict__ energy,
                                                        double* __restrict__ virial) {
  __shared__ double sh[kBlock][10];
  const int b = blockIdx.x;
  if (b >= B) return;
  const int i0 = mol_ptr[b], i1 = mol_ptr[b + 1];
  double acc[10];
  for (int k = 0; k < 10; ++k) acc[k] = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += kBlock) {
    acc[0] += (double)eps[i];
    if (vir_atom != nullptr)
      for (int k = 0; k < 9; ++k) acc[1 + k] += (double)vir_atom[9 * i + k];
  }
  for (int k = 0; k < 10; ++k) sh[threadIdx.x][k] = acc[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      for (int k = 0; k < 10; ++k) sh[threadIdx.x][k] += sh[threadIdx.x + off][k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (energy != nullptr) energy[b] = sh[0][0];
    if (virial != nullptr)
      for (int k = 0; k < 9; ++k) virial[9 * b + k] = sh[0][1 + k];
  }
}

__global__ void copy_f32_kernel(int n, const float* __restrict__ src, float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// bf16 planes of W2 [3H][H] of every block (dense in the blob)
__global__ void split_filter_weights_kernel(const float* __restrict__ w, ModelW mw,
                                            uint16_t* __restrict__ wbf) {
  const int H = mw.H;
  const int64_t S = wbf_plane(H);
  const int64_t total = (int64_t)mw.L * S;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(idx / S);
    const int64_t e = idx % S;
    uint16_t h, m, lo;
    split3_bf16(w[mw.layer[l].f2w + e], h, m, lo);
    uint16_t* dst = wbf + wbf_f2w(H, l) + e;
    dst[0] = h;
    dst[S] = m;
    dst[2 * S] = lo;
  }
}

inline size_t nh(const Buffers& b) { return (size_t)b.N * b.H; }

// Launches in the plan's chosen shapes (Buffers::tile_fwd, tile_bwd,
// unroll_fwd, unroll_bwd; each is the default or its alternative).
template <int H>
void launch_filter_fwd(const Buffers& b, hipStream_t st, int l) {
  const ModelW& mw = *b.mw;
  const float* centers = b.w + mw.centers;
  const float* widths = b.w + mw.widths;
  float* Wl = b.Wf + (size_t)l * b.pcap * 3 * H;
  const uint16_t* w2 = b.wbf + wbf_f2w(H, l);
  if (b.tile_fwd == kFwdPairsAlt<H>) {
    constexpr int TM = kFwdPairsAlt<H>;
    filter_fwd_kernel<H, TM><<<pair_grid(b, TM), kPairBlock, 0, st>>>(
        b.geom, b.counts, b.pcap, b.w, mw.layer[l], centers, widths, b.R, mw.RP, b.rc, Wl, w2);
  } else {
    constexpr int TM = kFwdPairs<H>;
    filter_fwd_kernel<H, TM><<<pair_grid(b, TM), kPairBlock, 0, st>>>(
        b.geom, b.counts, b.pcap, b.w, mw.layer[l], centers, widths, b.R, mw.RP, b.rc, Wl, w2);
  }
}

template <int H>
void launch_filter_bwd(const Buffers& b, hipStream_t st, int l, const float* gS1,
                       const float* gV1, int first, int last) {
  const ModelW& mw = *b.mw;
  const float* centers = b.w + mw.centers;
  const float* widths = b.w + mw.widths;
  const float* Wl = b.Wf + (size_t)l * b.pcap * 3 * H;
  const float* Sl = b.S + (size_t)l * nh(b);
  const float* Vl = b.V + (size_t)l * 3 * nh(b);
  if (b.tile_bwd == kBwdPairsAlt<H>) {
    constexpr int TM = kBwdPairsAlt<H>;
    filter_bwd_kernel<H, TM, false><<<pair_grid(b, TM), kPairBlock, 0, st>>>(
        b.counts, b.pcap, b.pair_lo, b.pair_hi, b.geom, b.w, mw.layer[l], centers, widths, b.R,
        mw.RP, b.rc, Wl, Sl, Vl, gS1, gV1, b.pacc, b.gvec, first, last);
  } else if (b.bwd_variant == 1) {
    constexpr int TM = kBwdPairs<H>;
    filter_bwd_kernel<H, TM, true><<<pair_grid(b, TM), kPairBlock, 0, st>>>(
        b.counts, b.pcap, b.pair_lo, b.pair_hi, b.geom, b.w, mw.layer[l], centers, widths, b.R,
        mw.RP, b.rc, Wl, Sl, Vl, gS1, gV1, b.pacc, b.gvec, first, last);
  } else {
    constexpr int TM = kBwdPairs<H>;
    filter_bwd_kernel<H, TM, false><<<pair_grid(b, TM), kPairBlock, 0, st>>>(
        b.counts, b.pcap, b.pair_lo, b.pair_hi, b.geom, b.w, mw.layer[l], centers, widths, b.R,
        mw.RP, b.rc, Wl, Sl, Vl, gS1, gV1, b.pacc, b.gvec, first, last);
  }
}

template <int H, bool FWD>
void launch_message(const Buffers& b, hipStream_t st, int grid, const float* Wl,
                    const float* s_in, const float* v_in, float* s_out, float* v_out) {
  constexpr int U0 = FWD ? kMsgUnrollFwd : kMsgUnrollBwd;
  constexpr int U1 = FWD ? kMsgUnrollFwdAlt : kMsgUnrollBwdAlt;
  if ((FWD ? b.unroll_fwd : b.unroll_bwd) == U1)
    message_kernel<H, FWD, U1><<<grid, kBlock, 0, st>>>(b.N, b.row_ptr, b.col, b.edge_pair,
                                                        b.ecap, b.geom, Wl, s_in, v_in, s_out,
                                                        v_out);
  else
    message_kernel<H, FWD, U0><<<grid, kBlock, 0, st>>>(b.N, b.row_ptr, b.col, b.edge_pair,
                                                        b.ecap, b.geom, Wl, s_in, v_in, s_out,
                                                        v_out);
}

// Fork/join between the caller's stream and the filter stream.  Events order
// the two streams both eagerly and under graph capture (a wait on an event
// recorded in a capturing stream pulls the other stream into the capture).
struct Streams {
  hipStream_t main, filt;
  hipEvent_t* ev;
  hipError_t err = hipSuccess;
  Streams(const Buffers& b, hipStream_t s) : main(s), filt(b.aux ? b.aux : s), ev(b.ev) {}
  bool two() const { return filt != main; }
  // `to` waits for everything issued so far on `from`
  void order(hipStream_t from, hipStream_t to, int k) {
    if (!two() || err != hipSuccess) return;
    err = hipEventRecord(ev[k], from);
    if (err == hipSuccess) err = hipStreamWaitEvent(to, ev[k], 0);
  }
};

template <int H>
hipError_t forward_impl(const Buffers& b, hipStream_t s) {
  const ModelW& mw = *b.mw;
  const int N = b.N;
  const float* centers = b.w + mw.centers;
  const float* widths = b.w + mw.widths;
  Streams st(b, s);
  // the filters depend on the pair geometry only: all L of them run on the
  // filter stream while the caller's stream walks the atom-side chain
  st.order(s, st.filt, 0);
  for (int l = 0; l < b.L; ++l) {
    launch_filter_fwd<H>(b, st.filt, l);
    if (st.two() && st.err == hipSuccess) st.err = hipEventRecord(b.ev[1 + l], st.filt);
  }
  {
    const int total = N * (H / 4);
    embed_kernel<H><<<(unsigned)((total + kBlock - 1) / kBlock), kBlock, 0, s>>>(
        N, b.numbers, b.w + mw.emb, b.S, b.V);
  }
  const int rows_per_block = kBlock / (H / 4);
  const int msg_grid = ((N + rows_per_block - 1) / rows_per_block + 7) / 8 * 8;
  const int atom_grid = (N + 31) / 32;
  for (int l = 0; l < b.L; ++l) {
    const LayerW& lw = mw.layer[l];
    float* Wl = b.Wf + (size_t)l * b.pcap * 3 * H;
    if (st.two() && st.err == hipSuccess) st.err = hipStreamWaitEvent(s, b.ev[1 + l], 0);
    float* Sl = b.S + (size_t)l * nh(b);
    float* Vl = b.V + (size_t)l * 3 * nh(b);
    float* Sm = b.Smid + (size_t)l * nh(b);
    float* Vm = b.Vmid + (size_t)l * 3 * nh(b);
    launch_message<H, true>(b, s, msg_grid, Wl, Sl, Vl, Sm, Vm);
    update_fwd_kernel<H><<<atom_grid, kUpdBlock, 0, s>>>(
        N, b.w, lw, Sm, Vm, b.S + (size_t)(l + 1) * nh(b), b.V + (size_t)(l + 1) * 3 * nh(b),
        b.Aupd + (size_t)l * nh(b), b.G12 + (size_t)l * 2 * nh(b));
  }
  readout_kernel<H><<<atom_grid, kBlock, 0, s>>>(N, b.w, mw, b.S + (size_t)b.L * nh(b), b.eps,
                                                 b.gS0);
  if (st.err != hipSuccess) return st.err;
  return hipGetLastError();
}

template <int H>
hipError_t backward_impl(const Buffers& b, hipStream_t s, float* forces, bool want_virial) {
  const ModelW& mw = *b.mw;
  const int N = b.N;
  const float* centers = b.w + mw.centers;
  const float* widths = b.w + mw.widths;
  const int rows_per_block = kBlock / (H / 4);
  const int msg_grid = ((N + rows_per_block - 1) / rows_per_block + 7) / 8 * 8;
  const int atom_grid = (N + 31) / 32;
  Streams st(b, s);
  // gradient w.r.t. the update outputs of the last block: (gS0, 0)
  float* gs2 = b.gS0;
  float* gv2 = nullptr;
  for (int l = b.L - 1; l >= 0; --l) {
    const LayerW& lw = mw.layer[l];
    const float* Wl = b.Wf + (size_t)l * b.pcap * 3 * H;
    // block l's message gradient has a slice of its own, so the filter
    // adjoint of block l (filter stream) and the message adjoint of block l
    // and update adjoint of block l-1 (caller's stream) run side by side
    float* gS1 = b.gS1 + (size_t)l * nh(b);
    float* gV1 = b.gV1 + (size_t)l * 3 * nh(b);
    update_bwd_kernel<H><<<atom_grid, kUpdBlock, 0, s>>>(
        N, b.w, lw, gs2, gv2, b.Vmid + (size_t)l * 3 * nh(b), b.Aupd + (size_t)l * nh(b),
        b.G12 + (size_t)l * 2 * nh(b), gS1, gV1);
    st.order(s, st.filt, b.L + 1 + l);
    launch_filter_bwd<H>(b, st.filt, l, gS1, gV1, l == b.L - 1, l == 0);
    if (l > 0) {
      launch_message<H, false>(b, s, msg_grid, Wl, gS1, gV1, b.gS0, b.gV0);
      gs2 = b.gS0;
      gv2 = b.gV0;
    }
  }
  st.order(st.filt, s, 2 * b.L + 1);
  force_gather_kernel<<<(N + 127) / 128, 128, 0, s>>>(N, b.row_ptr, b.col, b.edge_pair, b.ecap,
                                                      b.gvec, b.geom, forces,
                                                      want_virial ? b.vir_atom : nullptr);
  if (st.err != hipSuccess) return st.err;
  return hipGetLastError();
}

template <int H>
hipError_t stage_impl(const Buffers& b, hipStream_t s, int stage, int l) {
  const ModelW& mw = *b.mw;
  const int N = b.N;
  const float* centers = b.w + mw.centers;
  const float* widths = b.w + mw.widths;
  const int rows_per_block = kBlock / (H / 4);
  const int msg_grid = ((N + rows_per_block - 1) / rows_per_block + 7) / 8 * 8;
  const int atom_grid = (N + 31) / 32;
  const LayerW& lw = mw.layer[l];
  float* Wl = b.Wf + (size_t)l * b.pcap * 3 * H;
  switch (stage) {
    case MLFF_STAGE_FILTER_FWD:
      launch_filter_fwd<H>(b, s, l);
      break;
    case MLFF_STAGE_MESSAGE_FWD:
      launch_message<H, true>(b, s, msg_grid, Wl, b.S + (size_t)l * nh(b),
                              b.V + (size_t)l * 3 * nh(b), b.Smid + (size_t)l * nh(b),
                              b.Vmid + (size_t)l * 3 * nh(b));
      break;
    case MLFF_STAGE_UPDATE_FWD:
      update_fwd_kernel<H><<<atom_grid, kUpdBlock, 0, s>>>(
          N, b.w, lw, b.Smid + (size_t)l * nh(b), b.Vmid + (size_t)l * 3 * nh(b),
          b.S + (size_t)(l + 1) * nh(b), b.V + (size_t)(l + 1) * 3 * nh(b),
          b.Aupd + (size_t)l * nh(b), b.G12 + (size_t)l * 2 * nh(b));
      break;
    case MLFF_STAGE_READOUT:
      readout_kernel<H><<<atom_grid, kBlock, 0, s>>>(N, b.w, mw, b.S + (size_t)b.L * nh(b), b.eps,
                                                     b.gS0);
      break;
    case MLFF_STAGE_UPDATE_BWD:
      // reads gS0/gV0, writes block l's slice of gS1/gV1
      update_bwd_kernel<H><<<atom_grid, kUpdBlock, 0, s>>>(
          N, b.w, lw, b.gS0, (l == b.L - 1) ? nullptr : b.gV0, b.Vmid + (size_t)l * 3 * nh(b),
          b.Aupd + (size_t)l * nh(b), b.G12 + (size_t)l * 2 * nh(b), b.gS1 + (size_t)l * nh(b),
          b.gV1 + (size_t)l * 3 * nh(b));
      break;
    case MLFF_STAGE_FILTER_BWD:
      // first=1 overwrites the pair accumulator, so repeated launches agree
      launch_filter_bwd<H>(b, s, l, b.gS1 + (size_t)l * nh(b), b.gV1 + (size_t)l * 3 * nh(b),
                           1, 0);
      break;
    case MLFF_STAGE_MESSAGE_BWD:
      launch_message<H, false>(b, s, msg_grid, Wl, b.gS1 + (size_t)l * nh(b),
                               b.gV1 + (size_t)l * 3 * nh(b), b.gS0, b.gV0);
      break;
    case MLFF_STAGE_FORCES:
      force_gather_kernel<<<(N + 127) / 128, 128, 0, s>>>(N, b.row_ptr, b.col, b.edge_pair,
                                                          b.ecap, b.gvec, b.geom, nullptr,
                                                          b.vir_atom);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

int default_shape(int H, int which) {
  switch (which) {
    case 0: return H <= 32 ? kFwdPairs<32> : kFwdPairs<128>;
    case 1: return H <= 32 ? kBwdPairs<32> : kBwdPairs<128>;
    case 2: return kMsgUnrollFwd;
    case 3: return kMsgUnrollBwd;
    case 4: return 0;
    default: return 0;
  }
}

int alternative_shape(int H, int which) {
  switch (which) {
    case 0: return H <= 32 ? kFwdPairsAlt<32> : kFwdPairsAlt<128>;
    case 1: return H <= 32 ? kBwdPairsAlt<32> : kBwdPairsAlt<128>;
    case 2: return kMsgUnrollFwdAlt;
    case 3: return kMsgUnrollBwdAlt;
    case 4: return 1;
    default: return 0;
  }
}

hipError_t split_filter_weights(const Buffers& b) {
  if (b.wbf == nullptr || b.L < 1) return hipSuccess;
  // the weights may have just been written on another stream
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return e;
  split_filter_weights_kernel<<<256, 256, 0, nullptr>>>(b.w, *b.mw, b.wbf);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipDeviceSynchronize();
}

hipError_t split_filter_weights_async(const Buffers& b, hipStream_t s) {
  if (b.wbf == nullptr || b.L < 1) return hipSuccess;
  split_filter_weights_kernel<<<256, 256, 0, s>>>(b.w, *b.mw, b.wbf);
  return hipGetLastError();
}

hipError_t launch_stage(const Buffers& b, hipStream_t s, int stage, int layer) {
  if (layer < 0 || layer >= b.L) return hipErrorInvalidValue;
  switch (b.H) {
    case 32: return stage_impl<32>(b, s, stage, layer);
    case 64: return stage_impl<64>(b, s, stage, layer);
    case 128: return stage_impl<128>(b, s, stage, layer);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_forward(const Buffers& b, hipStream_t s) {
  switch (b.H) {
    case 32: return forward_impl<32>(b, s);
    case 64: return forward_impl<64>(b, s);
    case 128: return forward_impl<128>(b, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_backward(const Buffers& b, hipStream_t s, float* forces, bool want_virial) {
  switch (b.H) {
    case 32: return backward_impl<32>(b, s, forces, want_virial);
    case 64: return backward_impl<64>(b, s, forces, want_virial);
    case 128: return backward_impl<128>(b, s, forces, want_virial);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_reduce(const Buffers& b, hipStream_t s, double* energy, double* virial,
                         float* atom_energy) {
  if (b.B > 0)
    reduce_kernel<<<b.B, kBlock, 0, s>>>(b.B, b.mol_ptr, b.eps, virial ? b.vir_atom : nullptr,
                                         energy, virial);
  if (atom_energy != nullptr && b.N > 0)
    copy_f32_kernel<<<(b.N + 255) / 256, 256, 0, s>>>(b.N, b.eps, atom_energy);
  return hipGetLastError();
}

}  // namespace mlff
