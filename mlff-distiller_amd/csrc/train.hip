// Kernels of the force-matching training pass (include/mlff_amd.h,
// mlff_train_*): the edge- and pair-sized parts of the forward pass, its
// tangent along a position direction, the reverse pass and the reverse
// pass's tangent.  The reference builds these by autograd double backward
// (src/mlff_distiller/models/distillation_wrapper.py:86-91,
// training/trainer.py:325-328); here they are explicit derivatives of
// PaiNNMessage.forward (student_model.py:346-387).
//
// Every sum runs in a fixed order -- a receiver row adds its edges in
// ascending neighbour order, a segment adds its rows in index order -- so
// results are bit-reproducible; there are no atomics.
//
// Row kernels: a group of H/4 lanes owns one receiver row, each lane four
// channels (one f4) of every [.][H] plane; a 256-thread block holds 1024/H
// rows.  Pair kernel: a group of H/4 lanes per unordered pair.

#include <stdint.h>

#include "common.hpp"
#include "internal.hpp"

namespace mlff {
namespace {

constexpr int kTrainBlock = 256;
constexpr int kTrainUnroll = 2;   // edges whose operands are in flight together

__device__ __forceinline__ f4 ld4(const float* p) { return *reinterpret_cast<const f4*>(p); }
__device__ __forceinline__ void st4(float* p, f4 v) { *reinterpret_cast<f4*>(p) = v; }

// Message forward and its tangent, receiver row i (student_model.py:357-387):
//   s_out[i] = s[i] + sum_e s[j] Ws[p]
//   v_out[i] = v[i] + sum_e (v[j] Wv1[p] + sg u[p] (x) Wv2[p])
// with sg = +1 when i is the pair's lower atom (u points lo -> hi, the edge
// vector p[src] - p[dst] of edge hi -> lo) and -1 otherwise.  TAN adds
//   ds_out[i] = ds[i] + sum_e (ds[j] Ws + s[j] dWs)
//   dv_out[i] = dv[i] + sum_e (dv[j] Wv1 + v[j] dWv1 + sg (du (x) Wv2 + u (x) dWv2)).
template <int H, bool TAN>
__global__ __launch_bounds__(kTrainBlock) void train_message_kernel(
    int N, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ edge_pair, const float* __restrict__ s,
    const float* __restrict__ v, const float* __restrict__ W, const float* __restrict__ u,
    const float* __restrict__ ds, const float* __restrict__ dv, const float* __restrict__ dW,
    const float* __restrict__ du, float* __restrict__ s_out, float* __restrict__ v_out,
    float* __restrict__ ds_out, float* __restrict__ dv_out) {
  constexpr int G = H / 4, RPB = kTrainBlock / G;
  const int i = blockIdx.x * RPB + threadIdx.x / G;
  const int c = 4 * (threadIdx.x % G);
  if (i >= N) return;
  f4 as = ld4(s + (size_t)i * H + c);
  f4 av[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) av[x] = ld4(v + ((size_t)i * 3 + x) * H + c);
  f4 ads = {0.f, 0.f, 0.f, 0.f}, adv[3] = {ads, ads, ads};
  if (TAN) {
    ads = ld4(ds + (size_t)i * H + c);
#pragma unroll
    for (int x = 0; x < 3; ++x) adv[x] = ld4(dv + ((size_t)i * 3 + x) * H + c);
  }
  const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
  for (int e = e0; e < e1; e += kTrainUnroll) {
    int jj[kTrainUnroll], pp[kTrainUnroll];
#pragma unroll
    for (int k = 0; k < kTrainUnroll; ++k) {
      const int ek = e + k < e1 ? e + k : e;
      jj[k] = col[ek];
      pp[k] = edge_pair[ek];
    }
#pragma unroll
    for (int k = 0; k < kTrainUnroll; ++k) {
      if (e + k >= e1) break;
      const int j = jj[k], p = pp[k];
      const float sg = j > i ? 1.f : -1.f;
      const float* w = W + (size_t)p * 3 * H + c;
      const f4 ws = ld4(w), wv1 = ld4(w + H), wv2 = ld4(w + 2 * H);
      const f4 sj = ld4(s + (size_t)j * H + c);
      const float ux = sg * u[3 * p], uy = sg * u[3 * p + 1], uz = sg * u[3 * p + 2];
      const float uu[3] = {ux, uy, uz};
      f4 vj[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) vj[x] = ld4(v + ((size_t)j * 3 + x) * H + c);
      as += sj * ws;
#pragma unroll
      for (int x = 0; x < 3; ++x) av[x] += vj[x] * wv1 + uu[x] * wv2;
      if (TAN) {
        const float* dw = dW + (size_t)p * 3 * H + c;
        const f4 dws = ld4(dw), dwv1 = ld4(dw + H), dwv2 = ld4(dw + 2 * H);
        const f4 dsj = ld4(ds + (size_t)j * H + c);
        const float dd3[3] = {sg * du[3 * p], sg * du[3 * p + 1], sg * du[3 * p + 2]};
        ads += dsj * ws + sj * dws;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const f4 dvj = ld4(dv + ((size_t)j * 3 + x) * H + c);
          adv[x] += dvj * wv1 + vj[x] * dwv1 + (dd3[x] * wv2 + uu[x] * dwv2);
        }
      }
    }
  }
  st4(s_out + (size_t)i * H + c, as);
#pragma unroll
  for (int x = 0; x < 3; ++x) st4(v_out + ((size_t)i * 3 + x) * H + c, av[x]);
  if (TAN) {
    st4(ds_out + (size_t)i * H + c, ads);
#pragma unroll
    for (int x = 0; x < 3; ++x) st4(dv_out + ((size_t)i * 3 + x) * H + c, adv[x]);
  }
}

// Message adjoint (the transpose of the gathers above) and its tangent:
//   gs_out[i] = gs[i] + sum_e gs[j] Ws[p]
//   gv_out[i] = gv[i] + sum_e gv[j] Wv1[p]
//   dgs_out[i] = dgs[i] + sum_e (dgs[j] Ws + gs[j] dWs)
//   dgv_out[i] = dgv[i] + sum_e (dgv[j] Wv1 + gv[j] dWv1)
template <int H, bool TAN>
__global__ __launch_bounds__(kTrainBlock) void train_message_adjoint_kernel(
    int N, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ edge_pair, const float* __restrict__ gs,
    const float* __restrict__ gv, const float* __restrict__ W, const float* __restrict__ dgs,
    const float* __restrict__ dgv, const float* __restrict__ dW, float* __restrict__ gs_out,
    float* __restrict__ gv_out, float* __restrict__ dgs_out, float* __restrict__ dgv_out) {
  constexpr int G = H / 4, RPB = kTrainBlock / G;
  const int i = blockIdx.x * RPB + threadIdx.x / G;
  const int c = 4 * (threadIdx.x % G);
  if (i >= N) return;
  f4 as = ld4(gs + (size_t)i * H + c);
  f4 av[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) av[x] = ld4(gv + ((size_t)i * 3 + x) * H + c);
  f4 ads = {0.f, 0.f, 0.f, 0.f}, adv[3] = {ads, ads, ads};
  if (TAN) {
    ads = ld4(dgs + (size_t)i * H + c);
#pragma unroll
    for (int x = 0; x < 3; ++x) adv[x] = ld4(dgv + ((size_t)i * 3 + x) * H + c);
  }
  const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
  for (int e = e0; e < e1; e += kTrainUnroll) {
    int jj[kTrainUnroll], pp[kTrainUnroll];
#pragma unroll
    for (int k = 0; k < kTrainUnroll; ++k) {
      const int ek = e + k < e1 ? e + k : e;
      jj[k] = col[ek];
      pp[k] = edge_pair[ek];
    }
#pragma unroll
    for (int k = 0; k < kTrainUnroll; ++k) {
      if (e + k >= e1) break;
      const int j = jj[k], p = pp[k];
      const float* w = W + (size_t)p * 3 * H + c;
      const f4 ws = ld4(w), wv1 = ld4(w + H);
      const f4 gsj = ld4(gs + (size_t)j * H + c);
      f4 gvj[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) gvj[x] = ld4(gv + ((size_t)j * 3 + x) * H + c);
      as += gsj * ws;
#pragma unroll
      for (int x = 0; x < 3; ++x) av[x] += gvj[x] * wv1;
      if (TAN) {
        const float* dw = dW + (size_t)p * 3 * H + c;
        const f4 dws = ld4(dw), dwv1 = ld4(dw + H);
        ads += ld4(dgs + (size_t)j * H + c) * ws + gsj * dws;
#pragma unroll
        for (int x = 0; x < 3; ++x)
          adv[x] += ld4(dgv + ((size_t)j * 3 + x) * H + c) * wv1 + gvj[x] * dwv1;
      }
    }
  }
  st4(gs_out + (size_t)i * H + c, as);
#pragma unroll
  for (int x = 0; x < 3; ++x) st4(gv_out + ((size_t)i * 3 + x) * H + c, av[x]);
  if (TAN) {
    st4(dgs_out + (size_t)i * H + c, ads);
#pragma unroll
    for (int x = 0; x < 3; ++x) st4(dgv_out + ((size_t)i * 3 + x) * H + c, adv[x]);
  }
}

// Gradient of the loss with respect to a pair's filter output W_p = (Ws, Wv1,
// Wv2), given the gradients (gs, gv) of the message outputs, and its tangent.
// Pair p = (a < b): edge b -> a carries s_b Ws, v_b Wv1 + u (x) Wv2 to a, edge
// a -> b carries s_a Ws, v_a Wv1 - u (x) Wv2 to b, so
//   gWs  = gs_a s_b + gs_b s_a
//   gWv1 = sum_x (gv_a,x v_b,x + gv_b,x v_a,x)
//   gWv2 = sum_x u_x (gv_a,x - gv_b,x)
// and TAN differentiates each product along (ds, dv, du, dgs, dgv).
template <int H, bool TAN>
__global__ __launch_bounds__(kTrainBlock) void train_pair_grad_kernel(
    int P, const int32_t* __restrict__ pair_lo, const int32_t* __restrict__ pair_hi,
    const float* __restrict__ u, const float* __restrict__ du, const float* __restrict__ s,
    const float* __restrict__ v, const float* __restrict__ ds, const float* __restrict__ dv,
    const float* __restrict__ gs, const float* __restrict__ gv, const float* __restrict__ dgs,
    const float* __restrict__ dgv, float* __restrict__ gW, float* __restrict__ dgW) {
  constexpr int G = H / 4, PPB = kTrainBlock / G;
  const int p = blockIdx.x * PPB + threadIdx.x / G;
  const int c = 4 * (threadIdx.x % G);
  if (p >= P) return;
  const int a = pair_lo[p], b = pair_hi[p];
  const float uu[3] = {u[3 * p], u[3 * p + 1], u[3 * p + 2]};
  const f4 sa = ld4(s + (size_t)a * H + c), sb = ld4(s + (size_t)b * H + c);
  const f4 ga = ld4(gs + (size_t)a * H + c), gb = ld4(gs + (size_t)b * H + c);
  f4 gws = ga * sb + gb * sa;
  f4 gwv1 = {0.f, 0.f, 0.f, 0.f}, gwv2 = gwv1;
  f4 dgws = gwv1, dgwv1 = gwv1, dgwv2 = gwv1;
  float dd3[3] = {0.f, 0.f, 0.f};
  if (TAN) {
    dd3[0] = du[3 * p]; dd3[1] = du[3 * p + 1]; dd3[2] = du[3 * p + 2];
    const f4 dsa = ld4(ds + (size_t)a * H + c), dsb = ld4(ds + (size_t)b * H + c);
    const f4 dga = ld4(dgs + (size_t)a * H + c), dgb = ld4(dgs + (size_t)b * H + c);
    dgws = (dga * sb + ga * dsb) + (dgb * sa + gb * dsa);
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const size_t oa = ((size_t)a * 3 + x) * H + c, ob = ((size_t)b * 3 + x) * H + c;
    const f4 va = ld4(v + oa), vb = ld4(v + ob);
    const f4 qa = ld4(gv + oa), qb = ld4(gv + ob);
    gwv1 += qa * vb + qb * va;
    gwv2 += uu[x] * (qa - qb);
    if (TAN) {
      const f4 dva = ld4(dv + oa), dvb = ld4(dv + ob);
      const f4 dqa = ld4(dgv + oa), dqb = ld4(dgv + ob);
      dgwv1 += (dqa * vb + qa * dvb) + (dqb * va + qb * dva);
      dgwv2 += dd3[x] * (qa - qb) + uu[x] * (dqa - dqb);
    }
  }
  float* o = gW + (size_t)p * 3 * H + c;
  st4(o, gws);
  st4(o + H, gwv1);
  st4(o + 2 * H, gwv2);
  if (TAN) {
    float* q = dgW + (size_t)p * 3 * H + c;
    st4(q, dgws);
    st4(q + H, dgwv1);
    st4(q + 2 * H, dgwv2);
  }
}

// Message block on a caller-given directed edge list (PaiNNMessage.forward,
// student_model.py:346-387): each edge e has its own filter row W[e] and
// edge vector ev[e]; edges are grouped by receiver (row_ptr over dst, col =
// src, edge_id = the edge's index in the caller's list), added in that order:
//   s_out[i] = s[i] + sum_e s[src] Ws[e],
//   v_out[i] = v[i] + sum_e (v[src] Wv1[e] + ev[e] (x) Wv2[e]).
template <int H>
__global__ __launch_bounds__(kTrainBlock) void edge_message_kernel(
    int N, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const int32_t* __restrict__ edge_id, const float* __restrict__ s,
    const float* __restrict__ v, const float* __restrict__ W, const float* __restrict__ ev,
    float* __restrict__ s_out, float* __restrict__ v_out) {
  constexpr int G = H / 4, RPB = kTrainBlock / G;
  const int i = blockIdx.x * RPB + threadIdx.x / G;
  const int c = 4 * (threadIdx.x % G);
  if (i >= N) return;
  f4 as = ld4(s + (size_t)i * H + c);
  f4 av[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) av[x] = ld4(v + ((size_t)i * 3 + x) * H + c);
  const int e0 = row_ptr[i], e1 = row_ptr[i + 1];
  for (int e = e0; e < e1; ++e) {
    const int j = col[e], q = edge_id[e];
    const float* w = W + (size_t)q * 3 * H + c;
    const f4 ws = ld4(w), wv1 = ld4(w + H), wv2 = ld4(w + 2 * H);
    const float uu[3] = {ev[3 * q], ev[3 * q + 1], ev[3 * q + 2]};
    as += ld4(s + (size_t)j * H + c) * ws;
#pragma unroll
    for (int x = 0; x < 3; ++x) av[x] += ld4(v + ((size_t)j * 3 + x) * H + c) * wv1 + uu[x] * wv2;
  }
  st4(s_out + (size_t)i * H + c, as);
#pragma unroll
  for (int x = 0; x < 3; ++x) st4(v_out + ((size_t)i * 3 + x) * H + c, av[x]);
}

// Gradients per directed edge e = (src -> dst) of the message above, given the
// gradients (gs, gv) of its outputs:
//   gW[e]  = (gs[dst] s[src],  sum_x gv[dst]_x v[src]_x,  sum_x ev_x gv[dst]_x)
//   gev[e]_x = sum_h gv[dst]_x,h Wv2[e]_h   (a shuffle sum over the group)
template <int H>
__global__ __launch_bounds__(kTrainBlock) void edge_grad_kernel(
    int E, const int32_t* __restrict__ src, const int32_t* __restrict__ dst,
    const float* __restrict__ s, const float* __restrict__ v, const float* __restrict__ W,
    const float* __restrict__ ev, const float* __restrict__ gs, const float* __restrict__ gv,
    float* __restrict__ gW, float* __restrict__ gev) {
  constexpr int G = H / 4, EPB = kTrainBlock / G;
  const int e = blockIdx.x * EPB + threadIdx.x / G;
  const int c = 4 * (threadIdx.x % G);
  const bool on = e < E;
  const int ec = on ? e : 0;
  const int a = src[ec], b = dst[ec];
  const float uu[3] = {ev[3 * ec], ev[3 * ec + 1], ev[3 * ec + 2]};
  const f4 wv2 = ld4(W + (size_t)ec * 3 * H + 2 * H + c);
  f4 gws = ld4(gs + (size_t)b * H + c) * ld4(s + (size_t)a * H + c);
  f4 gwv1 = {0.f, 0.f, 0.f, 0.f}, gwv2 = gwv1;
  float ge[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const f4 q = ld4(gv + ((size_t)b * 3 + x) * H + c);
    gwv1 += q * ld4(v + ((size_t)a * 3 + x) * H + c);
    gwv2 += uu[x] * q;
    const f4 t = q * wv2;
    ge[x] = group_sum<G>((t[0] + t[1]) + (t[2] + t[3]));
  }
  if (!on) return;
  float* o = gW + (size_t)e * 3 * H + c;
  st4(o, gws);
  st4(o + H, gwv1);
  st4(o + 2 * H, gwv2);
  if (c == 0) {
    gev[3 * e] = ge[0];
    gev[3 * e + 1] = ge[1];
    gev[3 * e + 2] = ge[2];
  }
}

// out[k][c] = sum over rows r in [seg_ptr[k], seg_ptr[k+1]) of
// src[perm ? perm[r] : r][c].  One workgroup per (segment, chunk of CW
// columns): its 256 threads are CW columns x G = 256/CW row groups; group g
// adds rows g, g + G, ... of the segment in order, then the G partials of a
// column are added in group order (a fixed partition and order, so the sum
// is bit-reproducible; float64 or float32 accumulation).
template <typename T>
__global__ __launch_bounds__(kTrainBlock) void segment_sum_kernel(
    int width, int cw, const int32_t* __restrict__ seg_ptr, const int32_t* __restrict__ perm,
    const float* __restrict__ src, T* __restrict__ out) {
  __shared__ T part[kTrainBlock];
  const int k = blockIdx.x;
  const int G = kTrainBlock / cw;
  const int col = threadIdx.x % cw, grp = threadIdx.x / cw;
  const int c = blockIdx.y * cw + col;
  T acc = 0;
  if (c < width) {
    const int r1 = seg_ptr[k + 1];
#pragma unroll 4
    for (int r = seg_ptr[k] + grp; r < r1; r += G) {
      const int row = perm ? perm[r] : r;
      acc += (T)src[(size_t)row * width + c];
    }
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  if (grp == 0 && c < width) {
    T sum = 0;
    for (int q = 0; q < G; ++q) sum += part[q * cw + col];
    out[(size_t)k * width + c] = sum;
  }
}

template <int H>
hipError_t message_impl(hipStream_t st, int N, const int32_t* row_ptr, const int32_t* col,
                        const int32_t* edge_pair, const float* s, const float* v, const float* W,
                        const float* u, const float* ds, const float* dv, const float* dW,
                        const float* du, float* s_out, float* v_out, float* ds_out,
                        float* dv_out) {
  constexpr int RPB = kTrainBlock / (H / 4);
  const int grid = (N + RPB - 1) / RPB;
  if (ds)
    train_message_kernel<H, true><<<grid, kTrainBlock, 0, st>>>(
        N, row_ptr, col, edge_pair, s, v, W, u, ds, dv, dW, du, s_out, v_out, ds_out, dv_out);
  else
    train_message_kernel<H, false><<<grid, kTrainBlock, 0, st>>>(
        N, row_ptr, col, edge_pair, s, v, W, u, ds, dv, dW, du, s_out, v_out, ds_out, dv_out);
  return hipGetLastError();
}

template <int H>
hipError_t adjoint_impl(hipStream_t st, int N, const int32_t* row_ptr, const int32_t* col,
                        const int32_t* edge_pair, const float* gs, const float* gv, const float* W,
                        const float* dgs, const float* dgv, const float* dW, float* gs_out,
                        float* gv_out, float* dgs_out, float* dgv_out) {
  constexpr int RPB = kTrainBlock / (H / 4);
  const int grid = (N + RPB - 1) / RPB;
  if (dgs)
    train_message_adjoint_kernel<H, true><<<grid, kTrainBlock, 0, st>>>(
        N, row_ptr, col, edge_pair, gs, gv, W, dgs, dgv, dW, gs_out, gv_out, dgs_out, dgv_out);
  else
    train_message_adjoint_kernel<H, false><<<grid, kTrainBlock, 0, st>>>(
        N, row_ptr, col, edge_pair, gs, gv, W, dgs, dgv, dW, gs_out, gv_out, dgs_out, dgv_out);
  return hipGetLastError();
}

template <int H>
hipError_t pair_grad_impl(hipStream_t st, int P, const int32_t* lo, const int32_t* hi,
                          const float* u, const float* du, const float* s, const float* v,
                          const float* ds, const float* dv, const float* gs, const float* gv,
                          const float* dgs, const float* dgv, float* gW, float* dgW) {
  constexpr int PPB = kTrainBlock / (H / 4);
  const int grid = (P + PPB - 1) / PPB;
  if (dgW)
    train_pair_grad_kernel<H, true><<<grid, kTrainBlock, 0, st>>>(
        P, lo, hi, u, du, s, v, ds, dv, gs, gv, dgs, dgv, gW, dgW);
  else
    train_pair_grad_kernel<H, false><<<grid, kTrainBlock, 0, st>>>(
        P, lo, hi, u, du, s, v, ds, dv, gs, gv, dgs, dgv, gW, dgW);
  return hipGetLastError();
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <int H>
hipError_t edge_message_impl(hipStream_t st, int N, const int32_t* row_ptr, const int32_t* col,
                             const int32_t* edge_id, const float* s, const float* v,
                             const float* W, const float* ev, float* s_out, float* v_out) {
  constexpr int RPB = kTrainBlock / (H / 4);
  edge_message_kernel<H><<<(N + RPB - 1) / RPB, kTrainBlock, 0, st>>>(
      N, row_ptr, col, edge_id, s, v, W, ev, s_out, v_out);
  return hipGetLastError();
}

template <int H>
hipError_t edge_grad_impl(hipStream_t st, int E, const int32_t* src, const int32_t* dst,
                          const float* s, const float* v, const float* W, const float* ev,
                          const float* gs, const float* gv, float* gW, float* gev) {
  constexpr int EPB = kTrainBlock / (H / 4);
  edge_grad_kernel<H><<<(E + EPB - 1) / EPB, kTrainBlock, 0, st>>>(E, src, dst, s, v, W, ev, gs,
                                                                   gv, gW, gev);
  return hipGetLastError();
}

}  // namespace
}  // namespace mlff

using namespace mlff;

extern "C" {

int mlff_train_message(void* stream, int hidden_dim, int n_atoms, const int32_t* row_ptr_dev,
                       const int32_t* col_dev, const int32_t* edge_pair_dev, const float* s_dev,
                       const float* v_dev, const float* w_dev, const float* u_dev,
                       const float* ds_dev, const float* dv_dev, const float* dw_dev,
                       const float* du_dev, float* s_out_dev, float* v_out_dev,
                       float* ds_out_dev, float* dv_out_dev) {
  if (n_atoms < 0) { set_error("mlff_train_message: negative size"); return MLFF_ERR_INVALID; }
  if (n_atoms == 0) return MLFF_OK;
  const bool tan = ds_dev != nullptr;
  if (!row_ptr_dev || !col_dev || !edge_pair_dev || !s_dev || !v_dev || !w_dev || !u_dev ||
      !s_out_dev || !v_out_dev ||
      (tan && (!dv_dev || !dw_dev || !du_dev || !ds_out_dev || !dv_out_dev))) {
    set_error("mlff_train_message: null pointer");
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(s_dev) || !aligned16(v_dev) || !aligned16(w_dev) || !aligned16(s_out_dev) ||
      !aligned16(v_out_dev) || (tan && (!aligned16(ds_dev) || !aligned16(dv_dev) ||
                                        !aligned16(dw_dev) || !aligned16(ds_out_dev) ||
                                        !aligned16(dv_out_dev)))) {
    set_error("mlff_train_message: feature buffers must be 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden_dim) {
    case 32: e = message_impl<32>(st, n_atoms, row_ptr_dev, col_dev, edge_pair_dev, s_dev, v_dev, w_dev, u_dev, ds_dev, dv_dev, dw_dev, du_dev, s_out_dev, v_out_dev, ds_out_dev, dv_out_dev); break;
    case 64: e = message_impl<64>(st, n_atoms, row_ptr_dev, col_dev, edge_pair_dev, s_dev, v_dev, w_dev, u_dev, ds_dev, dv_dev, dw_dev, du_dev, s_out_dev, v_out_dev, ds_out_dev, dv_out_dev); break;
    case 128: e = message_impl<128>(st, n_atoms, row_ptr_dev, col_dev, edge_pair_dev, s_dev, v_dev, w_dev, u_dev, ds_dev, dv_dev, dw_dev, du_dev, s_out_dev, v_out_dev, ds_out_dev, dv_out_dev); break;
    default:
      set_error("mlff_train_message: hidden_dim %d not supported (32, 64, 128)", hidden_dim);
      return MLFF_ERR_UNSUPPORTED;
  }
  if (e != hipSuccess) {
    set_error("mlff_train_message: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_train_message_adjoint(void* stream, int hidden_dim, int n_atoms,
                               const int32_t* row_ptr_dev, const int32_t* col_dev,
                               const int32_t* edge_pair_dev, const float* gs_dev,
                               const float* gv_dev, const float* w_dev, const float* dgs_dev,
                               const float* dgv_dev, const float* dw_dev, float* gs_out_dev,
                               float* gv_out_dev, float* dgs_out_dev, float* dgv_out_dev) {
  if (n_atoms < 0) {
    set_error("mlff_train_message_adjoint: negative size");
    return MLFF_ERR_INVALID;
  }
  if (n_atoms == 0) return MLFF_OK;
  const bool tan = dgs_dev != nullptr;
  if (!row_ptr_dev || !col_dev || !edge_pair_dev || !gs_dev || !gv_dev || !w_dev ||
      !gs_out_dev || !gv_out_dev ||
      (tan && (!dgv_dev || !dw_dev || !dgs_out_dev || !dgv_out_dev))) {
    set_error("mlff_train_message_adjoint: null pointer");
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(gs_dev) || !aligned16(gv_dev) || !aligned16(w_dev) || !aligned16(gs_out_dev) ||
      !aligned16(gv_out_dev) || (tan && (!aligned16(dgs_dev) || !aligned16(dgv_dev) ||
                                         !aligned16(dw_dev) || !aligned16(dgs_out_dev) ||
                                         !aligned16(dgv_out_dev)))) {
    set_error("mlff_train_message_adjoint: feature buffers must be 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden_dim) {
    case 32: e = adjoint_impl<32>(st, n_atoms, row_ptr_dev, col_dev, edge_pair_dev, gs_dev, gv_dev, w_dev, dgs_dev, dgv_dev, dw_dev, gs_out_dev, gv_out_dev, dgs_out_dev, dgv_out_dev); break;
    case 64: e = adjoint_impl<64>(st, n_atoms, row_ptr_dev, col_dev, edge_pair_dev, gs_dev, gv_dev, w_dev, dgs_dev, dgv_dev, dw_dev, gs_out_dev, gv_out_dev, dgs_out_dev, dgv_out_dev); break;
    case 128: e = adjoint_impl<128>(st, n_atoms, row_ptr_dev, col_dev, edge_pair_dev, gs_dev, gv_dev, w_dev, dgs_dev, dgv_dev, dw_dev, gs_out_dev, gv_out_dev, dgs_out_dev, dgv_out_dev); break;
    default:
      set_error("mlff_train_message_adjoint: hidden_dim %d not supported (32, 64, 128)",
                hidden_dim);
      return MLFF_ERR_UNSUPPORTED;
  }
  if (e != hipSuccess) {
    set_error("mlff_train_message_adjoint: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_train_pair_grad(void* stream, int hidden_dim, int n_pairs, const int32_t* pair_lo_dev,
                         const int32_t* pair_hi_dev, const float* u_dev, const float* du_dev,
                         const float* s_dev, const float* v_dev, const float* ds_dev,
                         const float* dv_dev, const float* gs_dev, const float* gv_dev,
                         const float* dgs_dev, const float* dgv_dev, float* gw_dev,
                         float* dgw_dev) {
  if (n_pairs < 0) { set_error("mlff_train_pair_grad: negative size"); return MLFF_ERR_INVALID; }
  if (n_pairs == 0) return MLFF_OK;
  const bool tan = dgw_dev != nullptr;
  if (!pair_lo_dev || !pair_hi_dev || !u_dev || !s_dev || !v_dev || !gs_dev || !gv_dev ||
      !gw_dev || (tan && (!du_dev || !ds_dev || !dv_dev || !dgs_dev || !dgv_dev))) {
    set_error("mlff_train_pair_grad: null pointer");
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(s_dev) || !aligned16(v_dev) || !aligned16(gs_dev) || !aligned16(gv_dev) ||
      !aligned16(gw_dev) || (tan && (!aligned16(ds_dev) || !aligned16(dv_dev) ||
                                     !aligned16(dgs_dev) || !aligned16(dgv_dev) ||
                                     !aligned16(dgw_dev)))) {
    set_error("mlff_train_pair_grad: feature buffers must be 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden_dim) {
    case 32: e = pair_grad_impl<32>(st, n_pairs, pair_lo_dev, pair_hi_dev, u_dev, du_dev, s_dev, v_dev, ds_dev, dv_dev, gs_dev, gv_dev, dgs_dev, dgv_dev, gw_dev, dgw_dev); break;
    case 64: e = pair_grad_impl<64>(st, n_pairs, pair_lo_dev, pair_hi_dev, u_dev, du_dev, s_dev, v_dev, ds_dev, dv_dev, gs_dev, gv_dev, dgs_dev, dgv_dev, gw_dev, dgw_dev); break;
    case 128: e = pair_grad_impl<128>(st, n_pairs, pair_lo_dev, pair_hi_dev, u_dev, du_dev, s_dev, v_dev, ds_dev, dv_dev, gs_dev, gv_dev, dgs_dev, dgv_dev, gw_dev, dgw_dev); break;
    default:
      set_error("mlff_train_pair_grad: hidden_dim %d not supported (32, 64, 128)", hidden_dim);
      return MLFF_ERR_UNSUPPORTED;
  }
  if (e != hipSuccess) {
    set_error("mlff_train_pair_grad: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_edge_message(void* stream, int hidden_dim, int n_atoms, const int32_t* row_ptr_dev,
                      const int32_t* col_dev, const int32_t* edge_id_dev, const float* s_dev,
                      const float* v_dev, const float* w_dev, const float* ev_dev,
                      float* s_out_dev, float* v_out_dev) {
  if (n_atoms < 0) { set_error("mlff_edge_message: negative size"); return MLFF_ERR_INVALID; }
  if (n_atoms == 0) return MLFF_OK;
  if (!row_ptr_dev || !s_dev || !v_dev || !s_out_dev || !v_out_dev) {
    set_error("mlff_edge_message: null pointer");
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(s_dev) || !aligned16(v_dev) || (w_dev && !aligned16(w_dev)) ||
      !aligned16(s_out_dev) || !aligned16(v_out_dev)) {
    set_error("mlff_edge_message: feature buffers must be 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden_dim) {
    case 32: e = edge_message_impl<32>(st, n_atoms, row_ptr_dev, col_dev, edge_id_dev, s_dev, v_dev, w_dev, ev_dev, s_out_dev, v_out_dev); break;
    case 64: e = edge_message_impl<64>(st, n_atoms, row_ptr_dev, col_dev, edge_id_dev, s_dev, v_dev, w_dev, ev_dev, s_out_dev, v_out_dev); break;
    case 128: e = edge_message_impl<128>(st, n_atoms, row_ptr_dev, col_dev, edge_id_dev, s_dev, v_dev, w_dev, ev_dev, s_out_dev, v_out_dev); break;
    default:
      set_error("mlff_edge_message: hidden_dim %d not supported (32, 64, 128)", hidden_dim);
      return MLFF_ERR_UNSUPPORTED;
  }
  if (e != hipSuccess) {
    set_error("mlff_edge_message: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_edge_grad(void* stream, int hidden_dim, int n_edges, const int32_t* src_dev,
                   const int32_t* dst_dev, const float* s_dev, const float* v_dev,
                   const float* w_dev, const float* ev_dev, const float* gs_dev,
                   const float* gv_dev, float* gw_dev, float* gev_dev) {
  if (n_edges < 0) { set_error("mlff_edge_grad: negative size"); return MLFF_ERR_INVALID; }
  if (n_edges == 0) return MLFF_OK;
  if (!src_dev || !dst_dev || !s_dev || !v_dev || !w_dev || !ev_dev || !gs_dev || !gv_dev ||
      !gw_dev || !gev_dev) {
    set_error("mlff_edge_grad: null pointer");
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(s_dev) || !aligned16(v_dev) || !aligned16(w_dev) || !aligned16(gs_dev) ||
      !aligned16(gv_dev) || !aligned16(gw_dev)) {
    set_error("mlff_edge_grad: feature buffers must be 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  switch (hidden_dim) {
    case 32: e = edge_grad_impl<32>(st, n_edges, src_dev, dst_dev, s_dev, v_dev, w_dev, ev_dev, gs_dev, gv_dev, gw_dev, gev_dev); break;
    case 64: e = edge_grad_impl<64>(st, n_edges, src_dev, dst_dev, s_dev, v_dev, w_dev, ev_dev, gs_dev, gv_dev, gw_dev, gev_dev); break;
    case 128: e = edge_grad_impl<128>(st, n_edges, src_dev, dst_dev, s_dev, v_dev, w_dev, ev_dev, gs_dev, gv_dev, gw_dev, gev_dev); break;
    default:
      set_error("mlff_edge_grad: hidden_dim %d not supported (32, 64, 128)", hidden_dim);
      return MLFF_ERR_UNSUPPORTED;
  }
  if (e != hipSuccess) {
    set_error("mlff_edge_grad: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_segment_sum(void* stream, int n_segments, int width, const int32_t* seg_ptr_dev,
                     const int32_t* perm_dev, const float* src_dev, double* out64_dev,
                     float* out32_dev) {
  if (n_segments < 0 || width < 1) {
    set_error("mlff_segment_sum: invalid sizes");
    return MLFF_ERR_INVALID;
  }
  if (n_segments == 0) return MLFF_OK;
  if (!seg_ptr_dev || !src_dev || (out64_dev == nullptr) == (out32_dev == nullptr)) {
    set_error("mlff_segment_sum: need seg_ptr, src and exactly one output");
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  // narrow column slabs: a long segment (one species' atoms) is then read by
  // many row groups at once rather than by a few long dependent chains
  int cw = 1;
  while (cw < width && cw < 8) cw *= 2;
  const dim3 grid((unsigned)n_segments, (unsigned)((width + cw - 1) / cw));
  if (out64_dev)
    segment_sum_kernel<double><<<grid, kTrainBlock, 0, st>>>(width, cw, seg_ptr_dev, perm_dev,
                                                             src_dev, out64_dev);
  else
    segment_sum_kernel<float><<<grid, kTrainBlock, 0, st>>>(width, cw, seg_ptr_dev, perm_dev,
                                                            src_dev, out32_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_segment_sum: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

}  // extern "C"
