// Elementwise parts of PaiNNUpdate (student_model.py:430-470) in the
// force-matching training pass: the forward, its tangent along a position
// direction, and the reverse pass with its tangent, around the update MLP's
// two products (which run as library GEMMs between these kernels).
//
// Forward of the reference block:
//   nrm = |v|_xyz,  xin = (s, nrm),  y = MLP(xin) = (ds, g1, g2),
//   s' = s + ds,  v' = v g1 + (M v) g2.
// Every kernel here is elementwise over (atom, four channels): one thread
// per f4 of an [N][H] plane, so H is a multiple of 4 and every buffer is
// 16-byte aligned float32; sums over xyz run in x order.

#include <stdint.h>

#include <initializer_list>

#include "common.hpp"
#include "internal.hpp"

namespace mlff {
namespace {

constexpr int kGlueBlock = 256;

__device__ __forceinline__ f4 ld4(const float* p) { return *reinterpret_cast<const f4*>(p); }
__device__ __forceinline__ void st4(float* p, f4 v) { *reinterpret_cast<f4*>(p) = v; }

struct Mix {
  float m[9];
};

__device__ __forceinline__ Mix load_mix(const float* __restrict__ M) {
  Mix r;
#pragma unroll
  for (int k = 0; k < 9; ++k) r.m[k] = M[k];
  return r;
}

// (M v)_x = sum_z M[x][z] v_z, or (M^T v)_x = sum_z M[z][x] v_z
template <bool TR>
__device__ __forceinline__ f4 mix_row(const Mix& M, int x, const f4 (&v)[3]) {
  const float c0 = TR ? M.m[x] : M.m[3 * x], c1 = TR ? M.m[3 + x] : M.m[3 * x + 1],
              c2 = TR ? M.m[6 + x] : M.m[3 * x + 2];
  return c0 * v[0] + c1 * v[1] + c2 * v[2];
}

__device__ __forceinline__ void load3(const float* __restrict__ p, int n, int H, int c,
                                      f4 (&v)[3]) {
#pragma unroll
  for (int x = 0; x < 3; ++x) v[x] = ld4(p + ((size_t)n * 3 + x) * H + c);
}

__device__ __forceinline__ void store3(float* __restrict__ p, int n, int H, int c,
                                       const f4 (&v)[3]) {
#pragma unroll
  for (int x = 0; x < 3; ++x) st4(p + ((size_t)n * 3 + x) * H + c, v[x]);
}

// xin = (s, |v|_xyz)  [N][2H]
__global__ __launch_bounds__(kGlueBlock) void update_norm_kernel(
    int N, int H, const float* __restrict__ s, const float* __restrict__ v,
    float* __restrict__ xin) {
  const int h4 = H / 4;
  const int i = blockIdx.x * kGlueBlock + threadIdx.x;
  if (i >= N * h4) return;
  const int n = i / h4, c = 4 * (i - n * h4);
  f4 vv[3];
  load3(v, n, H, c, vv);
  const f4 sq = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
  f4 nrm;
#pragma unroll
  for (int k = 0; k < 4; ++k) nrm[k] = sqrtf(sq[k]);
  st4(xin + (size_t)n * 2 * H + c, ld4(s + (size_t)n * H + c));
  st4(xin + (size_t)n * 2 * H + H + c, nrm);
}

// s' = s + ds, Mv = M v, v' = v g1 + Mv g2, with y = (ds, g1, g2) [N][3H]
__global__ __launch_bounds__(kGlueBlock) void update_apply_kernel(
    int N, int H, const float* __restrict__ Mm, const float* __restrict__ y,
    const float* __restrict__ s, const float* __restrict__ v, float* __restrict__ s_out,
    float* __restrict__ v_out, float* __restrict__ mv_out) {
  const int h4 = H / 4;
  const int i = blockIdx.x * kGlueBlock + threadIdx.x;
  if (i >= N * h4) return;
  const int n = i / h4, c = 4 * (i - n * h4);
  const Mix M = load_mix(Mm);
  const float* yr = y + (size_t)n * 3 * H + c;
  const f4 ds = ld4(yr), g1 = ld4(yr + H), g2 = ld4(yr + 2 * H);
  st4(s_out + (size_t)n * H + c, ld4(s + (size_t)n * H + c) + ds);
  f4 vv[3], mv[3], vo[3];
  load3(v, n, H, c, vv);
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    mv[x] = mix_row<false>(M, x, vv);
    vo[x] = vv[x] * g1 + mv[x] * g2;
  }
  store3(mv_out, n, H, c, mv);
  store3(v_out, n, H, c, vo);
}

// dxin = (ds, dnrm), dnrm = (sum_x v dv) / nrm where nrm > 0, else 0
__global__ __launch_bounds__(kGlueBlock) void update_norm_tangent_kernel(
    int N, int H, const float* __restrict__ v, const float* __restrict__ xin,
    const float* __restrict__ ds, const float* __restrict__ dv, float* __restrict__ dxin) {
  const int h4 = H / 4;
  const int i = blockIdx.x * kGlueBlock + threadIdx.x;
  if (i >= N * h4) return;
  const int n = i / h4, c = 4 * (i - n * h4);
  f4 vv[3], dd[3];
  load3(v, n, H, c, vv);
  load3(dv, n, H, c, dd);
  const f4 nrm = ld4(xin + (size_t)n * 2 * H + H + c);
  const f4 dot = vv[0] * dd[0] + vv[1] * dd[1] + vv[2] * dd[2];
  f4 dn;
#pragma unroll
  for (int k = 0; k < 4; ++k) dn[k] = nrm[k] > 0.f ? dot[k] / nrm[k] : 0.f;
  st4(dxin + (size_t)n * 2 * H + c, ld4(ds + (size_t)n * H + c));
  st4(dxin + (size_t)n * 2 * H + H + c, dn);
}

// tangent of update_apply along (ds_m, dv_m, dy):
//   ds' = ds_m + dds,  dMv = M dv_m,
//   dv' = dv_m g1 + v dg1 + dMv g2 + Mv dg2
__global__ __launch_bounds__(kGlueBlock) void update_apply_tangent_kernel(
    int N, int H, const float* __restrict__ Mm, const float* __restrict__ y,
    const float* __restrict__ dy, const float* __restrict__ ds, const float* __restrict__ v,
    const float* __restrict__ dv, const float* __restrict__ mv, float* __restrict__ ds_out,
    float* __restrict__ dv_out, float* __restrict__ dmv_out) {
  const int h4 = H / 4;
  const int i = blockIdx.x * kGlueBlock + threadIdx.x;
  if (i >= N * h4) return;
  const int n = i / h4, c = 4 * (i - n * h4);
  const Mix M = load_mix(Mm);
  const float* yr = y + (size_t)n * 3 * H + c;
  const float* dyr = dy + (size_t)n * 3 * H + c;
  const f4 g1 = ld4(yr + H), g2 = ld4(yr + 2 * H);
  const f4 dds = ld4(dyr), dg1 = ld4(dyr + H), dg2 = ld4(dyr + 2 * H);
  st4(ds_out + (size_t)n * H + c, ld4(ds + (size_t)n * H + c) + dds);
  f4 vv[3], dd[3], mvv[3], dmv[3], out[3];
  load3(v, n, H, c, vv);
  load3(dv, n, H, c, dd);
  load3(mv, n, H, c, mvv);
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    dmv[x] = mix_row<false>(M, x, dd);
    out[x] = dd[x] * g1 + vv[x] * dg1 + dmv[x] * g2 + mvv[x] * dg2;
  }
  store3(dmv_out, n, H, c, dmv);
  store3(dv_out, n, H, c, out);
}

// adjoint of the gates: gy = (gs, sum_x gv v, sum_x gv Mv) [N][3H], and with
// TAN its tangent dgy = (dgs, sum_x (dgv v + gv dv), sum_x (dgv Mv + gv dMv))
template <bool TAN>
__global__ __launch_bounds__(kGlueBlock) void update_gate_adjoint_kernel(
    int N, int H, const float* __restrict__ gs, const float* __restrict__ gv,
    const float* __restrict__ v, const float* __restrict__ mv, const float* __restrict__ dgs,
    const float* __restrict__ dgv, const float* __restrict__ dv, const float* __restrict__ dmv,
    float* __restrict__ gy, float* __restrict__ dgy) {
  const int h4 = H / 4;
  const int i = blockIdx.x * kGlueBlock + threadIdx.x;
  if (i >= N * h4) return;
  const int n = i / h4, c = 4 * (i - n * h4);
  f4 g[3], vv[3], mvv[3];
  load3(gv, n, H, c, g);
  load3(v, n, H, c, vv);
  load3(mv, n, H, c, mvv);
  float* o = gy + (size_t)n * 3 * H + c;
  st4(o, ld4(gs + (size_t)n * H + c));
  st4(o + H, g[0] * vv[0] + g[1] * vv[1] + g[2] * vv[2]);
  st4(o + 2 * H, g[0] * mvv[0] + g[1] * mvv[1] + g[2] * mvv[2]);
  if (TAN) {
    f4 dg[3], dd[3], dm[3];
    load3(dgv, n, H, c, dg);
    load3(dv, n, H, c, dd);
    load3(dmv, n, H, c, dm);
    float* d = dgy + (size_t)n * 3 * H + c;
    st4(d, ld4(dgs + (size_t)n * H + c));
    st4(d + H, (dg[0] * vv[0] + g[0] * dd[0]) + (dg[1] * vv[1] + g[1] * dd[1]) +
                   (dg[2] * vv[2] + g[2] * dd[2]));
    st4(d + 2 * H, (dg[0] * mvv[0] + g[0] * dm[0]) + (dg[1] * mvv[1] + g[1] * dm[1]) +
                       (dg[2] * mvv[2] + g[2] * dm[2]));
  }
}

// adjoint into the block's inputs, given gx = (gs_x, gn) = dL/dxin [N][2H]:
//   gvg2 = gv g2,  unit = v / nrm (0 where nrm = 0),
//   gs_m = gs + gs_x,  gv_m = gv g1 + M^T gvg2 + gn unit,
// and with TAN their tangents along (dgx, dgs, dgv, dy, dv, dnrm):
//   dgvg2 = dgv g2 + gv dg2,  dunit = dv / nrm - v dnrm / nrm^2 (0 where nrm = 0),
//   dgs_m = dgs + dgs_x,
//   dgv_m = dgv g1 + gv dg1 + M^T dgvg2 + dgn unit + gn dunit.
template <bool TAN>
__global__ __launch_bounds__(kGlueBlock) void update_input_adjoint_kernel(
    int N, int H, const float* __restrict__ Mm, const float* __restrict__ gx,
    const float* __restrict__ gs, const float* __restrict__ gv, const float* __restrict__ y,
    const float* __restrict__ v, const float* __restrict__ xin, const float* __restrict__ dgx,
    const float* __restrict__ dgs, const float* __restrict__ dgv, const float* __restrict__ dy,
    const float* __restrict__ dv, const float* __restrict__ dxin, float* __restrict__ gs_out,
    float* __restrict__ gv_out, float* __restrict__ gvg2_out, float* __restrict__ dgs_out,
    float* __restrict__ dgv_out, float* __restrict__ dgvg2_out) {
  const int h4 = H / 4;
  const int i = blockIdx.x * kGlueBlock + threadIdx.x;
  if (i >= N * h4) return;
  const int n = i / h4, c = 4 * (i - n * h4);
  const Mix M = load_mix(Mm);
  const float* yr = y + (size_t)n * 3 * H + c;
  const f4 g1 = ld4(yr + H), g2 = ld4(yr + 2 * H);
  const float* gxr = gx + (size_t)n * 2 * H + c;
  const f4 gsx = ld4(gxr), gn = ld4(gxr + H);
  const f4 nrm = ld4(xin + (size_t)n * 2 * H + H + c);
  f4 inv;
#pragma unroll
  for (int k = 0; k < 4; ++k) inv[k] = nrm[k] > 0.f ? 1.0f / nrm[k] : 0.f;
  f4 g[3], vv[3], gvg2[3], unit[3], out[3];
  load3(gv, n, H, c, g);
  load3(v, n, H, c, vv);
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    gvg2[x] = g[x] * g2;
    unit[x] = vv[x] * inv;
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) out[x] = g[x] * g1 + mix_row<true>(M, x, gvg2) + gn * unit[x];
  st4(gs_out + (size_t)n * H + c, ld4(gs + (size_t)n * H + c) + gsx);
  store3(gv_out, n, H, c, out);
  store3(gvg2_out, n, H, c, gvg2);
  if (TAN) {
    const float* dyr = dy + (size_t)n * 3 * H + c;
    const f4 dg1 = ld4(dyr + H), dg2 = ld4(dyr + 2 * H);
    const float* dgxr = dgx + (size_t)n * 2 * H + c;
    const f4 dgsx = ld4(dgxr), dgn = ld4(dgxr + H);
    const f4 dn = ld4(dxin + (size_t)n * 2 * H + H + c);
    const f4 dninv2 = dn * inv * inv;
    f4 dg[3], dd[3], dgvg2[3], dout[3];
    load3(dgv, n, H, c, dg);
    load3(dv, n, H, c, dd);
#pragma unroll
    for (int x = 0; x < 3; ++x) dgvg2[x] = dg[x] * g2 + g[x] * dg2;
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const f4 dunit = dd[x] * inv - vv[x] * dninv2;
      dout[x] = dg[x] * g1 + g[x] * dg1 + mix_row<true>(M, x, dgvg2) + dgn * unit[x] +
                gn * dunit;
    }
    st4(dgs_out + (size_t)n * H + c, ld4(dgs + (size_t)n * H + c) + dgsx);
    store3(dgv_out, n, H, c, dout);
    store3(dgvg2_out, n, H, c, dgvg2);
  }
}

// Gradient of the 3x3 mixing matrix (student_model.py:414 and 460 under
// autograd):
//   out[x][z] = sum_n w_n sum_h (a0[n][x][h] b0[n][z][h] + a1[n][x][h] b1[n][z][h])
// over n atoms ([n][3][H] operands), as partial sums: block k takes the atoms
// [k A, (k + 1) A) with A = xyz_atoms(n) (a partition that depends on n only),
// each thread the items (atom, float4 of channels) t, t + 256, ... in order,
// and the block adds its threads' sums in a fixed tree.  Partial k is 12
// floats: the 9 sums row-major, then zeros (three float4 outputs, added by the
// weight-gradient reduce).
constexpr int kXyzMinAtoms = 32;
constexpr int kXyzMaxParts = 256;
int xyz_atoms(int n) {
  const int a = (n + kXyzMaxParts - 1) / kXyzMaxParts;
  return a > kXyzMinAtoms ? a : kXyzMinAtoms;
}
int xyz_parts(int n) { return n > 0 ? (n + xyz_atoms(n) - 1) / xyz_atoms(n) : 1; }

template <bool TWO, bool ROWW>
__global__ __launch_bounds__(kGlueBlock) void xyz_outer_kernel(
    int n, int H, int A, const float* __restrict__ a0, const float* __restrict__ b0,
    const float* __restrict__ a1, const float* __restrict__ b1, const float* __restrict__ w,
    float* __restrict__ part) {
  __shared__ float red[9][kGlueBlock];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * A, i1 = min(n, i0 + A);
  const int G = H / 4;
  float acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = 0.f;
  for (int item = t; item < (i1 - i0) * G; item += kGlueBlock) {
    const int i = i0 + item / G, c = 4 * (item - (item / G) * G);
    f4 av[3], bv[3];
    load3(a0, i, H, c, av);
    load3(b0, i, H, c, bv);
    f4 av1[3], bv1[3];
    if (TWO) {
      load3(a1, i, H, c, av1);
      load3(b1, i, H, c, bv1);
    }
    const float wi = ROWW ? w[i] : 1.f;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int z = 0; z < 3; ++z) {
        const f4 q = av[x] * bv[z];
        float d = (q[0] + q[1]) + (q[2] + q[3]);
        if (TWO) {
          const f4 q1 = av1[x] * bv1[z];
          d += (q1[0] + q1[1]) + (q1[2] + q1[3]);
        }
        acc[3 * x + z] += ROWW ? wi * d : d;
      }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) red[k][t] = acc[k];
  __syncthreads();
  for (int h = kGlueBlock / 2; h > 0; h >>= 1) {
    if (t < h) {
#pragma unroll
      for (int k = 0; k < 9; ++k) red[k][t] += red[k][t + h];
    }
    __syncthreads();
  }
  if (t < 12) part[(size_t)blockIdx.x * 12 + t] = t < 9 ? red[t][0] : 0.f;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// shared checks: sizes, required pointers non-null, every pointer aligned
int check_glue(const char* name, int N, int H, std::initializer_list<const void*> need,
               std::initializer_list<const void*> any) {
  if (N < 0 || H < 4 || H % 4 || (int64_t)N * (H / 4) > INT32_MAX) {
    set_error("%s: need n_atoms >= 0 and hidden a positive multiple of 4", name);
    return MLFF_ERR_INVALID;
  }
  for (const void* p : need)
    if (!p) {
      set_error("%s: null buffer", name);
      return MLFF_ERR_INVALID;
    }
  for (const void* p : any)
    if (!aligned16(p)) {
      set_error("%s: buffers must be 16-byte aligned", name);
      return MLFF_ERR_INVALID;
    }
  return MLFF_OK;
}

int finish(const char* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", name, hipGetErrorString(e));
    return MLFF_ERR_HIP;
  }
  return MLFF_OK;
}

unsigned glue_grid(int N, int H) { return (unsigned)(((int64_t)N * (H / 4) + kGlueBlock - 1) / kGlueBlock); }

}  // namespace
}  // namespace mlff

using namespace mlff;

extern "C" {

int mlff_update_norm(void* stream, int n_atoms, int hidden, const float* s_dev,
                     const float* v_dev, float* xin_dev) {
  const char* nm = "mlff_update_norm";
  if (int rc = check_glue(nm, n_atoms, hidden, {s_dev, v_dev, xin_dev}, {s_dev, v_dev, xin_dev}))
    return rc;
  if (n_atoms == 0) return MLFF_OK;
  update_norm_kernel<<<glue_grid(n_atoms, hidden), kGlueBlock, 0,
                       static_cast<hipStream_t>(stream)>>>(n_atoms, hidden, s_dev, v_dev,
                                                           xin_dev);
  return finish(nm);
}

int mlff_update_apply(void* stream, int n_atoms, int hidden, const float* m_dev,
                      const float* y_dev, const float* s_dev, const float* v_dev,
                      float* s_out_dev, float* v_out_dev, float* mv_out_dev) {
  const char* nm = "mlff_update_apply";
  if (int rc = check_glue(nm, n_atoms, hidden,
                          {m_dev, y_dev, s_dev, v_dev, s_out_dev, v_out_dev, mv_out_dev},
                          {y_dev, s_dev, v_dev, s_out_dev, v_out_dev, mv_out_dev}))
    return rc;
  if (n_atoms == 0) return MLFF_OK;
  update_apply_kernel<<<glue_grid(n_atoms, hidden), kGlueBlock, 0,
                        static_cast<hipStream_t>(stream)>>>(
      n_atoms, hidden, m_dev, y_dev, s_dev, v_dev, s_out_dev, v_out_dev, mv_out_dev);
  return finish(nm);
}

int mlff_update_norm_tangent(void* stream, int n_atoms, int hidden, const float* v_dev,
                             const float* xin_dev, const float* ds_dev, const float* dv_dev,
                             float* dxin_dev) {
  const char* nm = "mlff_update_norm_tangent";
  if (int rc = check_glue(nm, n_atoms, hidden, {v_dev, xin_dev, ds_dev, dv_dev, dxin_dev},
                          {v_dev, xin_dev, ds_dev, dv_dev, dxin_dev}))
    return rc;
  if (n_atoms == 0) return MLFF_OK;
  update_norm_tangent_kernel<<<glue_grid(n_atoms, hidden), kGlueBlock, 0,
                               static_cast<hipStream_t>(stream)>>>(n_atoms, hidden, v_dev,
                                                                   xin_dev, ds_dev, dv_dev,
                                                                   dxin_dev);
  return finish(nm);
}

int mlff_update_apply_tangent(void* stream, int n_atoms, int hidden, const float* m_dev,
                              const float* y_dev, const float* dy_dev, const float* ds_dev,
                              const float* v_dev, const float* dv_dev, const float* mv_dev,
                              float* ds_out_dev, float* dv_out_dev, float* dmv_out_dev) {
  const char* nm = "mlff_update_apply_tangent";
  if (int rc = check_glue(nm, n_atoms, hidden,
                          {m_dev, y_dev, dy_dev, ds_dev, v_dev, dv_dev, mv_dev, ds_out_dev,
                           dv_out_dev, dmv_out_dev},
                          {y_dev, dy_dev, ds_dev, v_dev, dv_dev, mv_dev, ds_out_dev, dv_out_dev,
                           dmv_out_dev}))
    return rc;
  if (n_atoms == 0) return MLFF_OK;
  update_apply_tangent_kernel<<<glue_grid(n_atoms, hidden), kGlueBlock, 0,
                                static_cast<hipStream_t>(stream)>>>(
      n_atoms, hidden, m_dev, y_dev, dy_dev, ds_dev, v_dev, dv_dev, mv_dev, ds_out_dev,
      dv_out_dev, dmv_out_dev);
  return finish(nm);
}

int mlff_update_gate_adjoint(void* stream, int n_atoms, int hidden, const float* gs_dev,
                             const float* gv_dev, const float* v_dev, const float* mv_dev,
                             const float* dgs_dev, const float* dgv_dev, const float* dv_dev,
                             const float* dmv_dev, float* gy_dev, float* dgy_dev) {
  const char* nm = "mlff_update_gate_adjoint";
  const bool tan = dgy_dev != nullptr;
  if (int rc = check_glue(nm, n_atoms, hidden, {gs_dev, gv_dev, v_dev, mv_dev, gy_dev},
                          {gs_dev, gv_dev, v_dev, mv_dev, gy_dev, dgs_dev, dgv_dev, dv_dev,
                           dmv_dev, dgy_dev}))
    return rc;
  if (tan && (!dgs_dev || !dgv_dev || !dv_dev || !dmv_dev)) {
    set_error("%s: the tangent needs dgs, dgv, dv and dmv", nm);
    return MLFF_ERR_INVALID;
  }
  if (n_atoms == 0) return MLFF_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (tan)
    update_gate_adjoint_kernel<true><<<glue_grid(n_atoms, hidden), kGlueBlock, 0, st>>>(
        n_atoms, hidden, gs_dev, gv_dev, v_dev, mv_dev, dgs_dev, dgv_dev, dv_dev, dmv_dev,
        gy_dev, dgy_dev);
  else
    update_gate_adjoint_kernel<false><<<glue_grid(n_atoms, hidden), kGlueBlock, 0, st>>>(
        n_atoms, hidden, gs_dev, gv_dev, v_dev, mv_dev, dgs_dev, dgv_dev, dv_dev, dmv_dev,
        gy_dev, dgy_dev);
  return finish(nm);
}

int mlff_update_input_adjoint(void* stream, int n_atoms, int hidden, const float* m_dev,
                              const float* gx_dev, const float* gs_dev, const float* gv_dev,
                              const float* y_dev, const float* v_dev, const float* xin_dev,
                              const float* dgx_dev, const float* dgs_dev, const float* dgv_dev,
                              const float* dy_dev, const float* dv_dev, const float* dxin_dev,
                              float* gs_out_dev, float* gv_out_dev, float* gvg2_out_dev,
                              float* dgs_out_dev, float* dgv_out_dev, float* dgvg2_out_dev) {
  const char* nm = "mlff_update_input_adjoint";
  const bool tan = dgs_out_dev != nullptr;
  if (int rc = check_glue(
          nm, n_atoms, hidden,
          {m_dev, gx_dev, gs_dev, gv_dev, y_dev, v_dev, xin_dev, gs_out_dev, gv_out_dev,
           gvg2_out_dev},
          {gx_dev, gs_dev, gv_dev, y_dev, v_dev, xin_dev, dgx_dev, dgs_dev, dgv_dev, dy_dev,
           dv_dev, dxin_dev, gs_out_dev, gv_out_dev, gvg2_out_dev, dgs_out_dev, dgv_out_dev,
           dgvg2_out_dev}))
    return rc;
  if (tan && (!dgx_dev || !dgs_dev || !dgv_dev || !dy_dev || !dv_dev || !dxin_dev ||
              !dgv_out_dev || !dgvg2_out_dev)) {
    set_error("%s: the tangent needs dgx, dgs, dgv, dy, dv, dxin and its three outputs", nm);
    return MLFF_ERR_INVALID;
  }
  if (n_atoms == 0) return MLFF_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (tan)
    update_input_adjoint_kernel<true><<<glue_grid(n_atoms, hidden), kGlueBlock, 0, st>>>(
        n_atoms, hidden, m_dev, gx_dev, gs_dev, gv_dev, y_dev, v_dev, xin_dev, dgx_dev, dgs_dev,
        dgv_dev, dy_dev, dv_dev, dxin_dev, gs_out_dev, gv_out_dev, gvg2_out_dev, dgs_out_dev,
        dgv_out_dev, dgvg2_out_dev);
  else
    update_input_adjoint_kernel<false><<<glue_grid(n_atoms, hidden), kGlueBlock, 0, st>>>(
        n_atoms, hidden, m_dev, gx_dev, gs_dev, gv_dev, y_dev, v_dev, xin_dev, dgx_dev, dgs_dev,
        dgv_dev, dy_dev, dv_dev, dxin_dev, gs_out_dev, gv_out_dev, gvg2_out_dev, dgs_out_dev,
        dgv_out_dev, dgvg2_out_dev);
  return finish(nm);
}

int mlff_xyz_outer_partial_count(int n_atoms) { return n_atoms < 0 ? 0 : xyz_parts(n_atoms); }

int mlff_xyz_outer_partials(void* stream, int n_atoms, int hidden, const float* a0_dev,
                            const float* b0_dev, const float* a1_dev, const float* b1_dev,
                            const float* row_w_dev, float* part_dev) {
  const char* nm = "mlff_xyz_outer_partials";
  const bool two = a1_dev != nullptr;
  if (int rc = check_glue(nm, n_atoms, hidden, {part_dev},
                          {a0_dev, b0_dev, a1_dev, b1_dev, part_dev}))
    return rc;
  if ((n_atoms > 0 && (!a0_dev || !b0_dev)) || two != (b1_dev != nullptr)) {
    set_error("%s: need a0 and b0 (n_atoms > 0), and a1 and b1 together", nm);
    return MLFF_ERR_INVALID;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int A = xyz_atoms(n_atoms), S = xyz_parts(n_atoms);
#define MLFF_XYZ(T, W)                                                                        \
  xyz_outer_kernel<T, W><<<S, kGlueBlock, 0, st>>>(n_atoms, hidden, A, a0_dev, b0_dev, a1_dev, \
                                                   b1_dev, row_w_dev, part_dev)
  if (two && row_w_dev) MLFF_XYZ(true, true);
  else if (two) MLFF_XYZ(true, false);
  else if (row_w_dev) MLFF_XYZ(false, true);
  else MLFF_XYZ(false, false);
#undef MLFF_XYZ
  return finish(nm);
}

}  // extern "C"
