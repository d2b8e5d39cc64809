// Kernels of the force-matching training pass (include/mlff_amd.h,
// mlff_train_*): the edge- and pair-sized parts of the forward pass, its
// tangent along a position direction, the reverse pass and the reverse
// pass's tangent.  The reference builds these by autograd double backward
// (src/mlff_distiller/models/distillation_wrapper.py:86-91,
// training/trainer.py:325-328); here they are explicit derivatives of
// PaiNNMessage.forward (student_model.py:346-387).
//
// Every sum runs in a fixed order -- a receiver row adds its edges in
// ascending neighbour order, segment and weight-gradient sums use a fixed
// partition of their rows added in a fixed order -- so results are
// bit-reproducible; there are no atomics.
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
// edges of a row handled per iteration.  The adjoint kernel loads all of an
// iteration's operands before using any (one row per H/4 lanes leaves few
// waves per SIMD at training batch sizes: 94 -> 83 µs for the tangent
// adjoint on 64 molecules); the same form made the forward kernel slower
// (61 -> 98 µs), which keeps its per-edge loads.
template <bool TAN> constexpr int kTrainUnroll = 2;
// lane groups per receiver row in the message kernels: group k sums the k-th
// of kTrainSegs consecutive parts of the row's edges (group 0 onto the row's
// own features), and group 0 adds the other groups' sums in group order (a
// fixed partition and order).  At training batch sizes (a few thousand rows)
// one group per row leaves about one wave per SIMD.
constexpr int kTrainSegs = 8;

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
  constexpr int G = H / 4, RPB = kTrainBlock / (G * kTrainSegs);
  __shared__ f4 part[kTrainSegs - 1][RPB * G][TAN ? 8 : 4];   // groups 1.. sums, per (row, lane)
  const int grp = threadIdx.x / G, seg = grp % kTrainSegs;
  const int i = blockIdx.x * RPB + grp / kTrainSegs;
  const int c = 4 * (threadIdx.x % G);
  const int slot = (grp / kTrainSegs) * G + threadIdx.x % G;
  const bool row_ok = i < N;
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  f4 as = zero, av[3] = {zero, zero, zero}, ads = zero, adv[3] = {zero, zero, zero};
  int e0 = 0, e1 = 0;
  if (row_ok) {
    if (seg == 0) {
      as = ld4(s + (size_t)i * H + c);
#pragma unroll
      for (int x = 0; x < 3; ++x) av[x] = ld4(v + ((size_t)i * 3 + x) * H + c);
      if (TAN) {
        ads = ld4(ds + (size_t)i * H + c);
#pragma unroll
        for (int x = 0; x < 3; ++x) adv[x] = ld4(dv + ((size_t)i * 3 + x) * H + c);
      }
    }
    const int r0 = row_ptr[i], len = row_ptr[i + 1] - r0;
    e0 = r0 + len * seg / kTrainSegs;
    e1 = r0 + len * (seg + 1) / kTrainSegs;
  }
  for (int e = e0; e < e1; e += kTrainUnroll<TAN>) {
    int jj[kTrainUnroll<TAN>], pp[kTrainUnroll<TAN>];
#pragma unroll
    for (int k = 0; k < kTrainUnroll<TAN>; ++k) {
      const int ek = e + k < e1 ? e + k : e;
      jj[k] = col[ek];
      pp[k] = edge_pair[ek];
    }
#pragma unroll
    for (int k = 0; k < kTrainUnroll<TAN>; ++k) {
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
  if (seg > 0) {
    f4* q = part[seg - 1][slot];
    q[0] = as;
#pragma unroll
    for (int x = 0; x < 3; ++x) q[1 + x] = av[x];
    if (TAN) {
      q[TAN ? 4 : 0] = ads;
#pragma unroll
      for (int x = 0; x < 3; ++x) q[TAN ? 5 + x : 0] = adv[x];
    }
  }
  __syncthreads();
  if (seg != 0 || !row_ok) return;
#pragma unroll
  for (int k = 0; k < kTrainSegs - 1; ++k) {
    as += part[k][slot][0];
#pragma unroll
    for (int x = 0; x < 3; ++x) av[x] += part[k][slot][1 + x];
    if (TAN) {
      ads += part[k][slot][TAN ? 4 : 0];
#pragma unroll
      for (int x = 0; x < 3; ++x) adv[x] += part[k][slot][TAN ? 5 + x : 0];
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
  constexpr int G = H / 4, RPB = kTrainBlock / (G * kTrainSegs);
  __shared__ f4 part[kTrainSegs - 1][RPB * G][TAN ? 8 : 4];   // groups 1.. sums, per (row, lane)
  const int grp = threadIdx.x / G, seg = grp % kTrainSegs;
  const int i = blockIdx.x * RPB + grp / kTrainSegs;
  const int c = 4 * (threadIdx.x % G);
  const int slot = (grp / kTrainSegs) * G + threadIdx.x % G;
  const bool row_ok = i < N;
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  f4 as = zero, av[3] = {zero, zero, zero}, ads = zero, adv[3] = {zero, zero, zero};
  int e0 = 0, e1 = 0;
  if (row_ok) {
    if (seg == 0) {
      as = ld4(gs + (size_t)i * H + c);
#pragma unroll
      for (int x = 0; x < 3; ++x) av[x] = ld4(gv + ((size_t)i * 3 + x) * H + c);
      if (TAN) {
        ads = ld4(dgs + (size_t)i * H + c);
#pragma unroll
        for (int x = 0; x < 3; ++x) adv[x] = ld4(dgv + ((size_t)i * 3 + x) * H + c);
      }
    }
    const int r0 = row_ptr[i], len = row_ptr[i + 1] - r0;
    e0 = r0 + len * seg / kTrainSegs;
    e1 = r0 + len * (seg + 1) / kTrainSegs;
  }
  constexpr int U = kTrainUnroll<TAN>;
  constexpr int UT = TAN ? U : 1;
  for (int e = e0; e < e1; e += U) {
    int jj[U], pp[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int ek = e + k < e1 ? e + k : e;   // past the row's end: a valid edge, unused
      jj[k] = col[ek];
      pp[k] = edge_pair[ek];
    }
    f4 ws[U], wv1[U], gsj[U], gvj[U][3];
    f4 dws[UT], dwv1[UT], dgsj[UT], dgvj[UT][3];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int j = jj[k], p = pp[k];
      const float* w = W + (size_t)p * 3 * H + c;
      ws[k] = ld4(w);
      wv1[k] = ld4(w + H);
      gsj[k] = ld4(gs + (size_t)j * H + c);
#pragma unroll
      for (int x = 0; x < 3; ++x) gvj[k][x] = ld4(gv + ((size_t)j * 3 + x) * H + c);
      if (TAN) {
        const float* dw = dW + (size_t)p * 3 * H + c;
        dws[k] = ld4(dw);
        dwv1[k] = ld4(dw + H);
        dgsj[k] = ld4(dgs + (size_t)j * H + c);
#pragma unroll
        for (int x = 0; x < 3; ++x) dgvj[k][x] = ld4(dgv + ((size_t)j * 3 + x) * H + c);
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      if (e + k >= e1) break;
      as += gsj[k] * ws[k];
#pragma unroll
      for (int x = 0; x < 3; ++x) av[x] += gvj[k][x] * wv1[k];
      if (TAN) {
        ads += dgsj[k] * ws[k] + gsj[k] * dws[k];
#pragma unroll
        for (int x = 0; x < 3; ++x) adv[x] += dgvj[k][x] * wv1[k] + gvj[k][x] * dwv1[k];
      }
    }
  }
  if (seg > 0) {
    f4* q = part[seg - 1][slot];
    q[0] = as;
#pragma unroll
    for (int x = 0; x < 3; ++x) q[1 + x] = av[x];
    if (TAN) {
      q[TAN ? 4 : 0] = ads;
#pragma unroll
      for (int x = 0; x < 3; ++x) q[TAN ? 5 + x : 0] = adv[x];
    }
  }
  __syncthreads();
  if (seg != 0 || !row_ok) return;
#pragma unroll
  for (int k = 0; k < kTrainSegs - 1; ++k) {
    as += part[k][slot][0];
#pragma unroll
    for (int x = 0; x < 3; ++x) av[x] += part[k][slot][1 + x];
    if (TAN) {
      ads += part[k][slot][TAN ? 4 : 0];
#pragma unroll
      for (int x = 0; x < 3; ++x) adv[x] += part[k][slot][TAN ? 5 + x : 0];
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

// ---------------------------------------------------------------------------
// Weight gradients over K rows (pairs or atoms):
//   out[m][n] = sum_k w_k (a0[k][m] b0[k][n] + a1[k][m] b1[k][n]),
//   bias[m]   = sum_k w_k a0[k][m].
//
// One wave owns a 64 x 64 output tile over RW consecutive rows; the four waves
// of a block take four consecutive row ranges of the same tile, and waves 1-3
// hand their tiles to wave 0 through LDS, which adds them in wave order.
// Block (tile, s) writes partial s; weight_grad_reduce_kernel adds the
// partials in a fixed order.  The split depends on (K, M, N) only, so the
// sums are bit-reproducible.  Blocks are numbered so that the tiles of one
// row range run on one XCD: the rows they share are fetched into that XCD's
// L2 once.
//
// Operands (v_mfma_f32_16x16x4f32): a k-step is four rows, lane group
// g = lane/16 feeding row k0 + g.  Lane r = lane%16 loads the 16-byte runs
// a[k][m0 + 4r .. 4r+3] and b[k][n0 + 4r .. 4r+3]; component q of the A run
// is row i = r of 16x16 tile q (tile q holds columns m0 + 4i + q), and the
// same for B.  A k-step is then two coalesced 16-byte loads per term and
// 4 x 4 MFMAs, and
//   acc[q][t][j] = out[m0 + 4 (4g + j) + q][n0 + 4r + t].
// The operands of the next kWgUnroll k-steps are loaded before the MFMAs of
// the current ones.
// ---------------------------------------------------------------------------
constexpr int kWgWaves = 4;
constexpr int kWgMaxSplits = 256;
// k-steps per operand batch (two terms: half as many).  With one batch in
// flight beside the current one the kernel fits 168 VGPRs, three waves per
// SIMD (three blocks per CU, also the LDS limit: 52 KB per block).
template <bool TWO> constexpr int kWgUnroll = TWO ? 1 : 2;

__device__ __forceinline__ void wg_mfma(f4 (&acc)[4][4], const f4& a, const f4& b) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[q][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q], b[t], acc[q][t], 0, 0, 0);
  }
}

template <bool TWO, bool ROWW>
struct WgBatch {
  // raw operands of rows min(k0 + 4u + g, k_end - 1); the masks for rows at
  // or past k_end and columns past M or N are applied when they are used.
  // The loads are unconditional (column 0 where masked), so the compiler
  // keeps count of them and the next batch stays in flight behind the MFMAs.
  static constexpr int U = kWgUnroll<TWO>;
  f4 a[U], b[U], a2[TWO ? U : 1], b2[TWO ? U : 1];
  float w[ROWW ? U : 1];

  __device__ __forceinline__ void load(int k0, int k_end, int g, int ma, int nb,
                                       const float* __restrict__ a0, int lda0,
                                       const float* __restrict__ b0, int ldb0,
                                       const float* __restrict__ a1, int lda1,
                                       const float* __restrict__ b1, int ldb1,
                                       const float* __restrict__ roww) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t kc = (size_t)min(k0 + 4 * u + g, k_end - 1);
      a[u] = ld4(a0 + kc * lda0 + ma);
      b[u] = ld4(b0 + kc * ldb0 + nb);
      if (TWO) {
        a2[u] = ld4(a1 + kc * lda1 + ma);
        b2[u] = ld4(b1 + kc * ldb1 + nb);
      }
      if (ROWW) w[u] = roww[kc];
    }
  }
};

template <bool TWO, bool ROWW, bool BIAS>
__global__ __launch_bounds__(64 * kWgWaves, 3) void weight_grad_kernel(
    int K, int M, int N, int RW, int tiles_m, int tiles_n, int S,
    const float* __restrict__ a0, int lda0, const float* __restrict__ b0, int ldb0,
    const float* __restrict__ a1, int lda1, const float* __restrict__ b1, int ldb1,
    const float* __restrict__ roww, float* __restrict__ part, float* __restrict__ part_bias) {
  __shared__ __align__(16) f4 red[kWgWaves - 1][16 * 64];
  __shared__ f4 bred[kWgWaves][64];
  const int tiles = tiles_m * tiles_n;
  const int logical = xcd_block();          // consecutive: the tiles of one row range
  if (logical >= tiles * S) return;         // grid padding (whole block)
  const int s = logical / tiles, tile = logical - s * tiles;
  const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  const int m0 = tm * 64, n0 = tn * 64;
  const int k_begin = (s * kWgWaves + wave) * RW;
  const int k_end = min(k_begin + RW, K);
  const int mc = m0 + 4 * r, nc = n0 + 4 * r;
  const bool mok = mc < M, nok = nc < N;
  const int ma = mok ? mc : 0, nb = nok ? nc : 0;
  const f4 zero = {0.f, 0.f, 0.f, 0.f};
  f4 acc[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[q][t] = zero;
  }
  f4 cs = zero;   // column sums of w a0 (bias)
  if (k_begin < k_end) {
    // the last batch's successor is loaded too (clamped rows, never used)
    WgBatch<TWO, ROWW> cur, nxt;
    cur.load(k_begin, k_end, g, ma, nb, a0, lda0, b0, ldb0, a1, lda1, b1, ldb1, roww);
    constexpr int U = kWgUnroll<TWO>;
    for (int k0 = k_begin; k0 < k_end; k0 += 4 * U) {
      nxt.load(k0 + 4 * U, k_end, g, ma, nb, a0, lda0, b0, ldb0, a1, lda1, b1, ldb1,
               roww);
      __builtin_amdgcn_sched_barrier(0);   // the loads above stay ahead of the MFMAs
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool kok = k0 + 4 * u + g < k_end;
        f4 a = (kok && mok) ? cur.a[u] : zero;
        const f4 b = (kok && nok) ? cur.b[u] : zero;
        f4 a2 = zero, b2 = zero;
        if (TWO) {
          a2 = (kok && mok) ? cur.a2[u] : zero;
          b2 = (kok && nok) ? cur.b2[u] : zero;
        }
        if (ROWW) {
          const float w = kok ? cur.w[u] : 0.f;
          a = a * w;
          if (TWO) a2 = a2 * w;
        }
        if (BIAS) cs += a;
        wg_mfma(acc, a, b);
        if (TWO) wg_mfma(acc, a2, b2);
      }
      __builtin_amdgcn_sched_barrier(0);
      cur = nxt;
    }
  }
  // waves 1.. hand their tiles to wave 0 (f4 (q, j) of lane: columns t = 0..3)
  if (wave > 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const f4 v = {acc[q][0][j], acc[q][1][j], acc[q][2][j], acc[q][3][j]};
        red[wave - 1][(q * 4 + j) * 64 + lane] = v;
      }
    }
  }
  if (BIAS) bred[wave][lane] = cs;
  __syncthreads();
  if (wave != 0) return;
  float* dst = part + (size_t)s * M * N;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      f4 v = {acc[q][0][j], acc[q][1][j], acc[q][2][j], acc[q][3][j]};
#pragma unroll
      for (int w = 0; w < kWgWaves - 1; ++w) v += red[w][(q * 4 + j) * 64 + lane];
      const int m = m0 + 4 * (4 * g + j) + q;
      if (m < M && nok) st4(dst + (size_t)m * N + nc, v);
    }
  }
  if (BIAS && tn == 0) {
    // column m0 + lane: component q = lane % 4 of lane group r = lane / 4,
    // added over waves then lane groups g in order
    const int rr = lane >> 2, q = lane & 3;
    float sum = 0.f;
#pragma unroll
    for (int w = 0; w < kWgWaves; ++w) {
#pragma unroll
      for (int gg = 0; gg < 4; ++gg) sum += bred[w][gg * 16 + rr][q];
    }
    if (m0 + lane < M) part_bias[(size_t)s * M + m0 + lane] = sum;
  }
}

// One float4 of a weight gradient from its S partials: the partials added in
// order in chunks of kWgChunk (loaded together), then the chunk sums added in
// order.  The partition and order depend on S only (bit-reproducible).
constexpr int kWgChunk = 8;

__device__ __forceinline__ f4 wg_column_sum(int S, int64_t n4, const f4* __restrict__ part,
                                            int64_t i) {
  f4 tot = {0.f, 0.f, 0.f, 0.f};
  for (int c = 0; c < S; c += kWgChunk) {
    f4 v[kWgChunk];
#pragma unroll
    for (int k = 0; k < kWgChunk; ++k)
      if (c + k < S) v[k] = part[(size_t)(c + k) * n4 + i];
    f4 cs = v[0];
#pragma unroll
    for (int k = 1; k < kWgChunk; ++k)
      if (c + k < S) cs += v[k];
    tot = c == 0 ? cs : tot + cs;
  }
  return tot;
}

// out[i] = sum_s part[s][i] over S partials of n4 float4 elements, one thread
// per element, and the same for the bias partials (nb4 elements; the blocks
// after the weight's take them, so one launch reduces both).
__global__ __launch_bounds__(kTrainBlock) void weight_grad_reduce_kernel(
    int S, int n4, const f4* __restrict__ part, f4* __restrict__ out, int nb4,
    const f4* __restrict__ part_b, f4* __restrict__ out_b) {
  const int blocks0 = (n4 + kTrainBlock - 1) / kTrainBlock;
  int blk = blockIdx.x;
  if (blk >= blocks0) {   // the bias partials (uniform per block)
    blk -= blocks0;
    part = part_b;
    out = out_b;
    n4 = nb4;
  }
  const int i = blk * kTrainBlock + threadIdx.x;
  if (i < n4) out[i] = wg_column_sum(S, n4, part, i);
}

// The same sums for several weight gradients in one launch.  Row r of the
// table is {S, n4, partials address, output address} (int64 each); row r takes
// ceil(n4 / kTrainBlock) consecutive virtual blocks, and the grid walks the
// virtual blocks with a stride of its size, so any grid covers every row.
// Each element is added by wg_column_sum, as in weight_grad_reduce_kernel,
// so the results are bit-identical to it.
constexpr int kWgGroupMaxRows = kTrainBlock;
struct WgRedRow {
  int64_t S, n4, part, out;
};

__global__ __launch_bounds__(kTrainBlock) void weight_grad_reduce_group_kernel(
    int n_rows, const WgRedRow* __restrict__ table) {
  __shared__ int64_t scan[kTrainBlock];
  __shared__ int64_t first[kWgGroupMaxRows + 1];   // first virtual block of each row
  const int t = threadIdx.x;
  // blocks per row, then an inclusive scan over the rows
  scan[t] = t < n_rows ? (table[t].n4 + kTrainBlock - 1) / kTrainBlock : 0;
  __syncthreads();
  for (int off = 1; off < kTrainBlock; off <<= 1) {
    const int64_t add = t >= off ? scan[t - off] : 0;
    __syncthreads();
    scan[t] += add;
    __syncthreads();
  }
  if (t == 0) first[0] = 0;
  if (t < n_rows) first[t + 1] = scan[t];
  __syncthreads();
  for (int64_t vb = blockIdx.x; vb < first[n_rows]; vb += gridDim.x) {
    int lo = 0, hi = n_rows - 1;   // the last row whose first block is <= vb
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (first[mid] <= vb) lo = mid; else hi = mid - 1;
    }
    const WgRedRow row = table[lo];
    const int64_t i = (vb - first[lo]) * kTrainBlock + t;
    if (i < row.n4)
      reinterpret_cast<f4*>(row.out)[i] =
          wg_column_sum((int)row.S, row.n4, reinterpret_cast<const f4*>(row.part), i);
  }
}

// Number of partials S and rows per wave RW (a multiple of 4: whole k-steps):
// at most kWgSlots blocks, the ones resident at once on MI355X (256 CUs, three
// blocks of this kernel each), so the grid runs as a single round instead of
// leaving a few blocks for a second one; at most kWgMaxSplits partials.
// A wave takes at least kWgMinRows rows: with fewer (a few thousand atoms
// over hundreds of blocks) the partial sums written and re-read cost more
// than the rows they split.
constexpr long kWgSlots = 768;
constexpr long kWgMinRows = 32;
void weight_grad_split(int K, int M, int N, int& RW, int& S) {
  const long tiles = (long)((M + 63) / 64) * ((N + 63) / 64);
  long s = kWgSlots / tiles;
  if (s > kWgMaxSplits) s = kWgMaxSplits;
  if (s < 1) s = 1;
  long rw = ((long)K + 4 * s - 1) / (4 * s);
  rw = (rw + 3) / 4 * 4;
  if (rw < kWgMinRows) rw = kWgMinRows;
  RW = (int)rw;
  S = K > 0 ? (int)(((long)K + 4 * rw - 1) / (4 * rw)) : 1;
}

template <bool TWO, bool ROWW>
void launch_weight_grad(hipStream_t st, bool bias, int blocks, int K, int M, int N, int RW,
                        int tiles_m, int tiles_n, int S, const float* a0, int lda0,
                        const float* b0, int ldb0, const float* a1, int lda1, const float* b1,
                        int ldb1, const float* roww, float* part, float* part_bias) {
  if (bias)
    weight_grad_kernel<TWO, ROWW, true><<<blocks, 64 * kWgWaves, 0, st>>>(
        K, M, N, RW, tiles_m, tiles_n, S, a0, lda0, b0, ldb0, a1, lda1, b1, ldb1, roww, part,
        part_bias);
  else
    weight_grad_kernel<TWO, ROWW, false><<<blocks, 64 * kWgWaves, 0, st>>>(
        K, M, N, RW, tiles_m, tiles_n, S, a0, lda0, b0, ldb0, a1, lda1, b1, ldb1, roww, part,
        part_bias);
}

// Pair geometry and radial basis of the training pass, one thread per pair
// p = (lo, hi) (student_model.py:706-722: edge vectors, distances and unit
// vectors; 249-253 and 285-290: the Gaussian basis times the cosine cutoff):
//   vec = (x[hi] - x[lo]) + shift,  d = |vec|,  den = d + eps,  u = vec / den,
//   e_k = phi_k fc,  de_k = (-2 gamma_k (d - mu_k) phi_k) fc + phi_k dfc,
// phi_k = exp(-gamma_k (d - mu_k)^2), gamma_k = 1 / w_k^2,
// fc = (cos(pi d / rc) + 1) / 2 and dfc = -(pi / rc) sin(pi d / rc) / 2 for
// d < rc, both 0 otherwise.  With a position direction lam (TAN) it writes
// the tangent instead: dvec = lam[hi] - lam[lo], dd = (vec . dvec) / d,
// du = dvec / den - vec (dd / den^2), de_t = de dd.
template <bool TAN>
__global__ __launch_bounds__(kTrainBlock) void pair_geometry_kernel(
    int P, int R, const float* __restrict__ x, const float* __restrict__ lam,
    const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
    const float* __restrict__ shift, const float* __restrict__ centers,
    const float* __restrict__ widths, float rc, float eps, float* __restrict__ vec_out,
    float* __restrict__ d_out, float* __restrict__ den_out, float* __restrict__ u_out,
    float* __restrict__ e_out, float* __restrict__ de_out, float* __restrict__ dd_out,
    float* __restrict__ du_out, float* __restrict__ de_t_out) {
  const int p = blockIdx.x * kTrainBlock + threadIdx.x;
  if (p >= P) return;
  const int a = lo[p], b = hi[p];
  float v[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v[c] = x[3 * b + c] - x[3 * a + c];
    if (shift) v[c] += shift[3 * p + c];
  }
  const float d = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const float den = d + eps;
  const bool inside = d < rc;
  const float arg = (float)M_PI * d / rc;
  const float fc = inside ? 0.5f * (cosf(arg) + 1.0f) : 0.f;
  const float dfc = inside ? (float)(-0.5 * M_PI / rc) * sinf(arg) : 0.f;
  float dd = 0.f;
  if (TAN) {
    float dv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dv[c] = lam[3 * b + c] - lam[3 * a + c];
    dd = (v[0] * dv[0] + v[1] * dv[1] + v[2] * dv[2]) / d;
    dd_out[p] = dd;
    const float q = dd / (den * den);
#pragma unroll
    for (int c = 0; c < 3; ++c) du_out[3 * p + c] = dv[c] / den - v[c] * q;
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      vec_out[3 * p + c] = v[c];
      u_out[3 * p + c] = v[c] / den;
    }
    d_out[p] = d;
    den_out[p] = den;
  }
  for (int k = 0; k < R; ++k) {
    const float w = widths[k];
    const float gamma = 1.0f / (w * w);
    const float diff = d - centers[k];
    const float phi = expf(-gamma * (diff * diff));
    const float de = (-2.0f * gamma * diff * phi) * fc + phi * dfc;
    if (TAN) {
      de_t_out[(size_t)p * R + k] = de * dd;
    } else {
      e_out[(size_t)p * R + k] = phi * fc;
      de_out[(size_t)p * R + k] = de;
    }
  }
}

// SiLU and its derivatives along a forward tangent and a reverse adjoint
// (student_model.py:318-322, 407-411, 599-605: the SiLU of every MLP),
// elementwise over n values, four per thread:
//   h = a s,  dh = silu'(a) da,  ga = g silu'(a),
//   dga = dg silu'(a) + g silu''(a) da,
// s = 1 / (1 + exp(-a)), silu' = s (1 + a (1 - s)),
// silu'' = s (1 - s) (2 + a (1 - 2 s)).  1 - s is evaluated as
// 1 / (1 + exp(a)), not by subtraction, which would lose its low bits for
// large positive a.  Outputs with a NULL pointer are skipped; a NULL dg is
// zero.
__global__ __launch_bounds__(kTrainBlock) void silu_chain_kernel(
    int n4, const f4* __restrict__ a, const f4* __restrict__ da, const f4* __restrict__ g,
    const f4* __restrict__ dg, f4* __restrict__ h, f4* __restrict__ dh, f4* __restrict__ ga,
    f4* __restrict__ dga) {
  const int i = blockIdx.x * kTrainBlock + threadIdx.x;
  if (i >= n4) return;
  const f4 x = a[i];
  f4 sg, sm, d1;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    sg[c] = 1.0f / (1.0f + expf(-x[c]));
    sm[c] = 1.0f / (1.0f + expf(x[c]));
    d1[c] = sg[c] * (1.0f + x[c] * sm[c]);
  }
  if (h) h[i] = x * sg;
  const f4 dx = da ? da[i] : f4{0.f, 0.f, 0.f, 0.f};
  if (dh) dh[i] = d1 * dx;
  if (ga || dga) {
    const f4 gx = g[i];
    if (ga) ga[i] = gx * d1;
    if (dga) {
      f4 d2;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        d2[c] = sg[c] * sm[c] * (2.0f + x[c] * (sm[c] - sg[c]));
      dga[i] = dg ? dg[i] * d1 + gx * d2 * dx : gx * d2 * dx;
    }
  }
}

template <int H>
hipError_t message_impl(hipStream_t st, int N, const int32_t* row_ptr, const int32_t* col,
                        const int32_t* edge_pair, const float* s, const float* v, const float* W,
                        const float* u, const float* ds, const float* dv, const float* dW,
                        const float* du, float* s_out, float* v_out, float* ds_out,
                        float* dv_out) {
  constexpr int RPB = kTrainBlock / (H / 4 * kTrainSegs);
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
  constexpr int RPB = kTrainBlock / (H / 4 * kTrainSegs);
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

size_t mlff_weight_grad_workspace_bytes(int K, int M, int N) {
  if (K < 0 || M < 1 || N < 1) return 0;
  int RW, S;
  weight_grad_split(K, M, N, RW, S);
  return (size_t)S * ((size_t)M * N + M) * sizeof(float);
}

// Checks the arguments and launches the partial-sum kernel (weight partials
// [S][M][N] at ws, bias partials [S][M] after them); returns S through s_out.
static int weight_grad_partials(const char* who, void* stream, int K, int M, int N,
                                const float* a0_dev, int lda0, const float* b0_dev, int ldb0,
                                const float* a1_dev, int lda1, const float* b1_dev, int ldb1,
                                const float* row_w_dev, bool bias, void* ws_dev,
                                size_t ws_bytes, int* s_out) {
  if (K < 0 || M < 4 || N < 4 || M % 4 || N % 4) {
    set_error("%s: need K >= 0 and M, N positive multiples of 4 (K=%d M=%d N=%d)", who, K, M,
              N);
    return MLFF_ERR_INVALID;
  }
  const bool two = a1_dev != nullptr;
  if ((K > 0 && (!a0_dev || !b0_dev)) || !ws_dev || two != (b1_dev != nullptr)) {
    set_error("%s: need a0, b0 (K > 0), workspace, and a1 and b1 together", who);
    return MLFF_ERR_INVALID;
  }
  if (lda0 < M || ldb0 < N || lda0 % 4 || ldb0 % 4 ||
      (two && (lda1 < M || ldb1 < N || lda1 % 4 || ldb1 % 4))) {
    set_error("%s: row strides must be multiples of 4 and at least M, N", who);
    return MLFF_ERR_INVALID;
  }
  if (!aligned16(a0_dev) || !aligned16(b0_dev) || !aligned16(ws_dev) ||
      (two && (!aligned16(a1_dev) || !aligned16(b1_dev)))) {
    set_error("%s: buffers must be 16-byte aligned", who);
    return MLFF_ERR_INVALID;
  }
  const size_t need = mlff_weight_grad_workspace_bytes(K, M, N);
  if (ws_bytes < need) {
    set_error("%s: workspace of %zu bytes, need %zu", who, ws_bytes, need);
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  int RW, S;
  weight_grad_split(K, M, N, RW, S);
  const int tiles_m = (M + 63) / 64, tiles_n = (N + 63) / 64;
  // a multiple of 8 blocks, so that xcd_block() groups them by XCD
  const int blocks = (tiles_m * tiles_n * S + 7) / 8 * 8;
  float* part = static_cast<float*>(ws_dev);
  float* part_bias = part + (size_t)S * M * N;
  if (two && row_w_dev)
    launch_weight_grad<true, true>(st, bias, blocks, K, M, N, RW, tiles_m, tiles_n, S, a0_dev,
                                   lda0, b0_dev, ldb0, a1_dev, lda1, b1_dev, ldb1, row_w_dev,
                                   part, part_bias);
  else if (two)
    launch_weight_grad<true, false>(st, bias, blocks, K, M, N, RW, tiles_m, tiles_n, S, a0_dev,
                                    lda0, b0_dev, ldb0, a1_dev, lda1, b1_dev, ldb1, row_w_dev,
                                    part, part_bias);
  else if (row_w_dev)
    launch_weight_grad<false, true>(st, bias, blocks, K, M, N, RW, tiles_m, tiles_n, S, a0_dev,
                                    lda0, b0_dev, ldb0, a1_dev, lda1, b1_dev, ldb1, row_w_dev,
                                    part, part_bias);
  else
    launch_weight_grad<false, false>(st, bias, blocks, K, M, N, RW, tiles_m, tiles_n, S,
                                     a0_dev, lda0, b0_dev, ldb0, a1_dev, lda1, b1_dev, ldb1,
                                     row_w_dev, part, part_bias);
  *s_out = S;
  return MLFF_OK;
}

int mlff_weight_grad(void* stream, int K, int M, int N, const float* a0_dev, int lda0,
                     const float* b0_dev, int ldb0, const float* a1_dev, int lda1,
                     const float* b1_dev, int ldb1, const float* row_w_dev, void* ws_dev,
                     size_t ws_bytes, float* out_dev, float* bias_out_dev) {
  if (!out_dev || !aligned16(out_dev) || (bias_out_dev && !aligned16(bias_out_dev))) {
    set_error("mlff_weight_grad: need out, and out and bias_out 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  const bool bias = bias_out_dev != nullptr;
  int S = 0;
  const int rc = weight_grad_partials("mlff_weight_grad", stream, K, M, N, a0_dev, lda0, b0_dev,
                                      ldb0, a1_dev, lda1, b1_dev, ldb1, row_w_dev, bias, ws_dev,
                                      ws_bytes, &S);
  if (rc != MLFF_OK) return rc;
  const float* part = static_cast<const float*>(ws_dev);
  const float* part_bias = part + (size_t)S * M * N;
  const int n4 = M * N / 4, nb4 = bias ? M / 4 : 0;
  const int red_blocks = (n4 + kTrainBlock - 1) / kTrainBlock + (nb4 + kTrainBlock - 1) / kTrainBlock;
  weight_grad_reduce_kernel<<<red_blocks, kTrainBlock, 0, static_cast<hipStream_t>(stream)>>>(
      S, n4, reinterpret_cast<const f4*>(part), reinterpret_cast<f4*>(out_dev), nb4,
      reinterpret_cast<const f4*>(part_bias), reinterpret_cast<f4*>(bias_out_dev));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_weight_grad: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_weight_grad_partial_count(int K, int M, int N) {
  if (K < 0 || M < 1 || N < 1) return 0;
  int RW, S;
  weight_grad_split(K, M, N, RW, S);
  return S;
}

int mlff_weight_grad_partials(void* stream, int K, int M, int N, const float* a0_dev, int lda0,
                              const float* b0_dev, int ldb0, const float* a1_dev, int lda1,
                              const float* b1_dev, int ldb1, const float* row_w_dev,
                              int with_bias, void* ws_dev, size_t ws_bytes) {
  int S = 0;
  const int rc = weight_grad_partials("mlff_weight_grad_partials", stream, K, M, N, a0_dev,
                                      lda0, b0_dev, ldb0, a1_dev, lda1, b1_dev, ldb1, row_w_dev,
                                      with_bias != 0, ws_dev, ws_bytes, &S);
  if (rc != MLFF_OK) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_weight_grad_partials: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_weight_grad_reduce_group(void* stream, int n_rows, const int64_t* table_dev,
                                  int n_blocks) {
  if (n_rows < 0 || n_rows > kWgGroupMaxRows || (n_rows > 0 && (!table_dev || n_blocks < 1)) ||
      !aligned16(table_dev)) {
    set_error("mlff_weight_grad_reduce_group: need 0 <= n_rows <= %d, a 16-byte aligned table "
              "and n_blocks >= 1", kWgGroupMaxRows);
    return MLFF_ERR_INVALID;
  }
  if (n_rows == 0) return MLFF_OK;
  weight_grad_reduce_group_kernel<<<n_blocks, kTrainBlock, 0, static_cast<hipStream_t>(stream)>>>(
      n_rows, reinterpret_cast<const WgRedRow*>(table_dev));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_weight_grad_reduce_group: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_train_geometry(void* stream, int n_pairs, int num_rbf, const float* x_dev,
                        const float* lam_dev, const int32_t* lo_dev, const int32_t* hi_dev,
                        const float* shift_dev, const float* centers_dev,
                        const float* widths_dev, float cutoff, float eps, float* vec_dev,
                        float* d_dev, float* den_dev, float* u_dev, float* e_dev,
                        float* de_dev, float* dd_dev, float* du_dev, float* de_t_dev) {
  const char* nm = "mlff_train_geometry";
  const bool tan = lam_dev != nullptr;
  if (n_pairs < 0 || num_rbf < 1 || !(cutoff > 0.f)) {
    set_error("%s: need n_pairs >= 0, num_rbf >= 1 and a positive cutoff", nm);
    return MLFF_ERR_INVALID;
  }
  if (n_pairs == 0) return MLFF_OK;
  if (!x_dev || !lo_dev || !hi_dev || !centers_dev || !widths_dev ||
      (tan ? (!dd_dev || !du_dev || !de_t_dev)
           : (!vec_dev || !d_dev || !den_dev || !u_dev || !e_dev || !de_dev))) {
    set_error("%s: null buffer (the forward writes vec, d, den, u, e and de; the tangent "
              "dd, du and de_t)", nm);
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int grid = (n_pairs + kTrainBlock - 1) / kTrainBlock;
  if (tan)
    pair_geometry_kernel<true><<<grid, kTrainBlock, 0, st>>>(
        n_pairs, num_rbf, x_dev, lam_dev, lo_dev, hi_dev, shift_dev, centers_dev, widths_dev,
        cutoff, eps, vec_dev, d_dev, den_dev, u_dev, e_dev, de_dev, dd_dev, du_dev, de_t_dev);
  else
    pair_geometry_kernel<false><<<grid, kTrainBlock, 0, st>>>(
        n_pairs, num_rbf, x_dev, lam_dev, lo_dev, hi_dev, shift_dev, centers_dev, widths_dev,
        cutoff, eps, vec_dev, d_dev, den_dev, u_dev, e_dev, de_dev, dd_dev, du_dev, de_t_dev);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", nm, hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

int mlff_silu_chain(void* stream, int64_t n, const float* a_dev, const float* da_dev,
                    const float* g_dev, const float* dg_dev, float* h_dev, float* dh_dev,
                    float* ga_dev, float* dga_dev) {
  if (n < 0 || n % 4 || n / 4 > INT32_MAX) {
    set_error("mlff_silu_chain: n must be a non-negative multiple of 4");
    return MLFF_ERR_INVALID;
  }
  if (!a_dev || (dh_dev && !da_dev) || (ga_dev && !g_dev) ||
      (dga_dev && (!g_dev || !da_dev))) {
    set_error("mlff_silu_chain: a missing input for a requested output");
    return MLFF_ERR_INVALID;
  }
  const void* bufs[8] = {a_dev, da_dev, g_dev, dg_dev, h_dev, dh_dev, ga_dev, dga_dev};
  for (const void* b : bufs)
    if (!aligned16(b)) {
      set_error("mlff_silu_chain: buffers must be 16-byte aligned");
      return MLFF_ERR_INVALID;
    }
  if (n == 0) return MLFF_OK;
  const int n4 = (int)(n / 4);
  silu_chain_kernel<<<(n4 + kTrainBlock - 1) / kTrainBlock, kTrainBlock, 0,
                      static_cast<hipStream_t>(stream)>>>(
      n4, reinterpret_cast<const f4*>(a_dev), reinterpret_cast<const f4*>(da_dev),
      reinterpret_cast<const f4*>(g_dev), reinterpret_cast<const f4*>(dg_dev),
      reinterpret_cast<f4*>(h_dev), reinterpret_cast<f4*>(dh_dev), reinterpret_cast<f4*>(ga_dev),
      reinterpret_cast<f4*>(dga_dev));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("mlff_silu_chain: %s", hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

}  // extern "C"
