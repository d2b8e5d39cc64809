// Velocity Verlet on the device, as ASE's VelocityVerlet.step
// (ase/md/verlet.py; driven by nve_harness.py:214-235 in the reference):
//   p += dt/2 F ;  x += dt p / m ;  F = F(x) ;  p += dt/2 F
// State (x, p) is float64 like ASE's; the model reads a float32 copy of x.
// The second half-kick also sums the kinetic energy of each structure in
// FP64 (sum p^2 / 2m, as Atoms.get_kinetic_energy) and appends
// (E_pot, E_kin) to the ledger row of the current step.

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
    const int32_t* __restrict__ counts) {
  __shared__ double sh[256];
  const int b = blockIdx.x;
  if (b >= B) return;
  const int i0 = mol_ptr[b], i1 = mol_ptr[b + 1];
  double ke = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += 256) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double pa = p[3 * i + a] + 0.5 * dt * (double)F[3 * i + a];
      p[3 * i + a] = pa;
      s += pa * pa;
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

}  // namespace

hipError_t launch_vv_first(int N, double* x, double* p, const double* inv_mass, const float* F,
                           float* pos32, double dt, hipStream_t s) {
  if (N > 0) vv_first_kernel<<<(N + 255) / 256, 256, 0, s>>>(N, x, p, inv_mass, F, pos32, dt);
  return hipGetLastError();
}

hipError_t launch_vv_second(int N, int B, const int32_t* mol_ptr, double* p,
                            const double* inv_mass, const float* F, double dt,
                            const double* epot, double* ledger, int ledger_len,
                            int32_t* counts, hipStream_t s) {
  (void)N;
  if (B > 0)
    vv_second_kernel<<<B, 256, 0, s>>>(B, mol_ptr, p, inv_mass, F, dt, epot, ledger, ledger_len,
                                       counts);
  step_advance_kernel<<<1, 64, 0, s>>>(counts);
  return hipGetLastError();
}

}  // namespace mlff
