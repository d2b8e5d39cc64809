// Velocity Verlet on the device, as ASE's VelocityVerlet.step
// (ase/md/verlet.py; driven by nve_harness.py:214-235 in the reference):
//   p += dt/2 F ;  x += dt p / m ;  F = F(x) ;  p += dt/2 F
// State (x, p) is float64 like ASE's; the model reads a float32 copy of x.
// The second half-kick also sums the kinetic energy of each structure in
// FP64 (sum p^2 / 2m, as Atoms.get_kinetic_energy) and appends
// (E_pot, E_kin) to the ledger row of the current step.  When frame buffers
// are bound it also stores (x, p, F) of the step in the frame row of the
// same step counter, the per-step data the reference harness copies through
// the host (nve_harness.py:263-287).

#include "common.hpp"
#include "internal.hpp"

namespace mlff {

namespace {

__global__ void vv_first_kernel(int N, double* __restrict__ x, double* __restrict__ p,
                                const double* __restrict__ inv_mass, const float* __restrict__ F,
                                float* __restrict__ pos32, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double im = inv_mass[i];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double pa = p[3 * i + a] + 0.5 * dt * (double)F[3 * i + a];
    const double xa = x[3 * i + a] + dt * pa * im;
    p[3 * i + a] = pa;
    x[3 * i + a] = xa;
    pos32[3 * i + a] = (float)xa;
  }
}

__global__ __launch_bounds__(256) void vv_second_kernel(
    int B, const int32_t* __restrict__ mol_ptr, double* __restrict__ p,
    const double* __restrict__ inv_mass, const float* __restrict__ F, double dt,
    const double* __restrict__ epot, double* __restrict__ ledger, int ledger_len,
    const int32_t* __restrict__ counts, int N, const double* __restrict__ x,
    double* __restrict__ fx, double* __restrict__ fp, float* __restrict__ fF, int n_frames) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  if (b >= B) return;
  const int i0 = mol_ptr[b], i1 = mol_ptr[b + 1];
  // frame row of this step (all three frame buffers are bound together)
  const size_t fbase = n_frames > 0
      ? (size_t)(counts[MLFF_COUNT_STEP] % n_frames) * (size_t)N * 3 : 0;
  double ke = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double pa = p[3 * i + a] + 0.5 * dt * (double)F[3 * i + a];
      p[3 * i + a] = pa;
      s += pa * pa;
      if (n_frames > 0) {
        fx[fbase + 3 * i + a] = x[3 * i + a];
        fp[fbase + 3 * i + a] = pa;
        fF[fbase + 3 * i + a] = F[3 * i + a];
      }
    }
    ke += 0.5 * s * inv_mass[i];
  }
  sh[threadIdx.x] = ke;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && ledger != nullptr && ledger_len > 0) {
    const int row = counts[MLFF_COUNT_STEP] % ledger_len;
    ledger[((size_t)row * B + b) * 2 + 0] = epot[b];
    ledger[((size_t)row * B + b) * 2 + 1] = sh[0];
  }
}

__global__ void step_advance_kernel(int32_t* counts) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counts[MLFF_COUNT_STEP] += 1;
}

// ---- multiple time step (r-RESPA) ------------------------------------------

// p += coef * (Fa - Fb), or p += coef * Fa when Fb is null
__global__ void mts_kick_kernel(int N, double* __restrict__ p, const float* __restrict__ Fa,
                                const float* __restrict__ Fb, double coef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * N) return;
  const double f = (double)Fa[i] - (Fb ? (double)Fb[i] : 0.0);
  p[i] += coef * f;
}

// inner force: the full model for flagged structures, the student elsewhere
__global__ void mts_select_kernel(int N, const int32_t* __restrict__ atom_mol,
                                  const uint8_t* __restrict__ flag,
                                  const float* __restrict__ F_full,
                                  const float* __restrict__ F_fast, float* __restrict__ F_in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool f = flag[atom_mol[i]] != 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) F_in[3 * i + a] = f ? F_full[3 * i + a] : F_fast[3 * i + a];
}

// per-structure running max of |F_full - F_fast| (non-negative floats order as
// their bit patterns, so an integer max is exact and order independent)
__global__ void mts_disagree_kernel(int N, const int32_t* __restrict__ atom_mol,
                                    const float* __restrict__ F_full,
                                    const float* __restrict__ F_fast,
                                    float* __restrict__ dis_max) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float dx = F_full[3 * i] - F_fast[3 * i];
  const float dy = F_full[3 * i + 1] - F_fast[3 * i + 1];
  const float dz = F_full[3 * i + 2] - F_fast[3 * i + 2];
  const float d = sqrtf(dx * dx + dy * dy + dz * dz);
  atomicMax(reinterpret_cast<int*>(dis_max) + atom_mol[i], __float_as_int(d));
}

// (E_full, E_kin) of every structure into the ledger row of the current step
__global__ __launch_bounds__(256) void mts_ledger_kernel(
    int B, const int32_t* __restrict__ mol_ptr, const double* __restrict__ p,
    const double* __restrict__ inv_mass, const double* __restrict__ epot,
    double* __restrict__ ledger, int ledger_len, const int32_t* __restrict__ counts) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  if (b >= B) return;
  double ke = 0.0;
  for (int i = mol_ptr[b] + threadIdx.x; i < mol_ptr[b + 1]; i += 256) {
    const double s = p[3 * i] * p[3 * i] + p[3 * i + 1] * p[3 * i + 1] + p[3 * i + 2] * p[3 * i + 2];
    ke += 0.5 * s * inv_mass[i];
  }
  sh[threadIdx.x] = ke;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && ledger_len > 0) {
    const int row = counts[MLFF_COUNT_STEP] % ledger_len;
    ledger[((size_t)row * B + b) * 2 + 0] = epot[b];
    ledger[((size_t)row * B + b) * 2 + 1] = sh[0];
  }
}

}  // namespace

hipError_t launch_mts_kick(int N, double* p, const float* Fa, const float* Fb, double coef,
                           hipStream_t s) {
  if (N > 0) mts_kick_kernel<<<(3 * N + 255) / 256, 256, 0, s>>>(N, p, Fa, Fb, coef);
  return hipGetLastError();
}

hipError_t launch_mts_select(int N, const int32_t* atom_mol, const uint8_t* flag,
                             const float* F_full, const float* F_fast, float* F_in,
                             hipStream_t s) {
  if (N > 0)
    mts_select_kernel<<<(N + 255) / 256, 256, 0, s>>>(N, atom_mol, flag, F_full, F_fast, F_in);
  return hipGetLastError();
}

hipError_t launch_mts_disagree(int N, const int32_t* atom_mol, const float* F_full,
                               const float* F_fast, float* dis_max, hipStream_t s) {
  if (N > 0)
    mts_disagree_kernel<<<(N + 255) / 256, 256, 0, s>>>(N, atom_mol, F_full, F_fast, dis_max);
  return hipGetLastError();
}

hipError_t launch_mts_ledger(int B, const int32_t* mol_ptr, const double* p,
                             const double* inv_mass, const double* epot, double* ledger,
                             int ledger_len, int32_t* counts, hipStream_t s) {
  if (B > 0)
    mts_ledger_kernel<<<B, 256, 0, s>>>(B, mol_ptr, p, inv_mass, epot, ledger, ledger_len, counts);
  step_advance_kernel<<<1, 64, 0, s>>>(counts);
  return hipGetLastError();
}

hipError_t launch_vv_first(int N, double* x, double* p, const double* inv_mass, const float* F,
                           float* pos32, double dt, hipStream_t s) {
  if (N > 0) vv_first_kernel<<<(N + 255) / 256, 256, 0, s>>>(N, x, p, inv_mass, F, pos32, dt);
  return hipGetLastError();
}

hipError_t launch_vv_second(int N, int B, const int32_t* mol_ptr, double* p,
                            const double* inv_mass, const float* F, double dt,
                            const double* epot, double* ledger, int ledger_len,
                            int32_t* counts, const double* x, double* fx, double* fp, float* fF,
                            int n_frames, hipStream_t s) {
  if (B > 0)
    vv_second_kernel<<<B, 256, 0, s>>>(B, mol_ptr, p, inv_mass, F, dt, epot, ledger, ledger_len,
                                       counts, N, x, fx, fp, fF, n_frames);
  step_advance_kernel<<<1, 64, 0, s>>>(counts);
  return hipGetLastError();
}

}  // namespace mlff
