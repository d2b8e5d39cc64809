// Internal declarations shared by the kernel translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mlff_amd.h"

namespace mlff {

// Per-interaction-block weight offsets (floats into the packed blob).
struct LayerW {
  int64_t f1w, f1b, f1t;   // filter Linear(R,H): W [H][RP], b [H], W^T [RP][H]
  int64_t f2w, f2b, f2t;   // filter Linear(H,3H): W [3H][H], b [3H], W^T [H][3H]
  int64_t u1w, u1b, u1t;   // update Linear(2H,H): W [H][2H], b [H], W^T [2H][H]
  int64_t u2w, u2b, u2t;   // update Linear(H,3H): W [3H][H], b [3H], W^T [H][3H]
  int64_t mix;             // mixing matrix [3][3]
};

struct ModelW {
  int H, L, R, RP, HP2, HP4, max_z;
  int64_t emb, centers, widths;
  int64_t r1w, r1b, r1t;   // Linear(H, H/2): [HP2][H], b [HP2], W^T [H][HP2]
  int64_t r2w, r2b, r2t;   // Linear(H/2, H/4): [HP4][HP2], b [HP4], W^T [HP2][HP4]
  int64_t r3w, r3b;        // Linear(H/4, 1): w [HP4], b [1] (stored padded to 16)
  LayerW layer[8];
  int64_t total;
};

int build_layout(int H, int L, int R, int max_z, ModelW* mw, mlff_slot* slots, int max_slots,
                 int* n_slots);

// Neighbour-list launches (nl.hip).
struct NLArgs {
  const float* pos; int N; int B;
  const int32_t* mol_ptr; const int32_t* atom_mol;
  const float* cells; const uint8_t* pbc; float rc;
  int32_t* deg; int32_t* udeg; int32_t* row_ptr; int32_t* pair_ptr;
  int32_t* col; int32_t* edge_shift; int32_t* edge_pair; int32_t ecap;
  int32_t* pair_lo; int32_t* pair_hi; int32_t* pair_shift; float4* geom; int32_t pcap;
  int32_t* counts;
  // cell list (periodic structures): grid[b] = {nc0, nc1, nc2, base}, nc0 == 0 -> brute force
  const int4* grid; int32_t n_cells;
  int32_t* cell_count; int32_t* cell_start; int32_t* cell_atoms; int32_t* atom_cell;
  int32_t* edge_tmp_shift;   // [ecap*3] image shifts of cell-list rows before sorting
};
hipError_t launch_neighbor_list(const NLArgs& a, hipStream_t s, bool build_pairs);

// Model launches (painn_fwd.hip / painn_bwd.hip).
struct Buffers {
  int N, B, H, L, R;
  float rc;
  int32_t ecap, pcap;
  int blocks;                       // persistent grid for pair kernels
  // launch shapes: pairs per filter tile (forward, adjoint) and edges in
  // flight per message row (forward, adjoint); default_shape() or
  // alternative_shape() of each
  int tile_fwd, tile_bwd, unroll_fwd, unroll_bwd;
  int bwd_variant;    // filter adjoint: 0 plain, 1 next-tile pair-index prefetch
  const float* w;                   // weight blob
  const ModelW* mw;
  // bf16 three-plane copies of the filter's second layer, per block l:
  // W2 [3H][H] at wbf + wbf_f2w(l); plane stride 3H^2
  uint16_t* wbf;
  const int64_t* numbers;
  const int32_t* mol_ptr;
  const int32_t* atom_mol;
  const float* cells;
  int32_t* counts;
  const int32_t* row_ptr; const int32_t* col; const int32_t* edge_pair;
  const int32_t* pair_lo; const int32_t* pair_hi; const int32_t* pair_shift;
  const float4* geom;
  float4* pacc;                     // per pair (gu.xyz, gd)
  float4* gvec;                     // per pair dE/dvec
  float* Wf;                        // [L][pcap][3H]
  float* S; float* V;               // [L+1][N][H], [L+1][N][3][H]
  float* Smid; float* Vmid;         // [L][N][H], [L][N][3][H]
  float* Aupd; float* G12;          // [L][N][H], [L][N][2H]
  float* gS0; float* gV0;           // gradient w.r.t. the update outputs of the block below
  float* gS1; float* gV1;           // [L][N][H], [L][N][3][H]: gradient w.r.t. message outputs
  float* eps;                       // [N]
  float* vir_atom;                  // [N][9]
  // second stream for the filter kernels (overlaps the memory-bound message
  // kernels); ev holds 3L+2 events.  aux == nullptr: one stream.
  hipStream_t aux;
  hipEvent_t* ev;
};
__host__ __device__ inline int64_t wbf_plane(int H) { return 3 * (int64_t)H * H; }
__host__ __device__ inline int64_t wbf_f2w(int H, int l) { return (int64_t)l * 3 * wbf_plane(H); }
__host__ __device__ inline int64_t wbf_elems(int H, int L) { return (int64_t)L * 3 * wbf_plane(H); }
// Launch shapes (which: 0 filter-forward pairs per tile, 1 filter-adjoint
// pairs per tile, 2 message-forward edges in flight, 3 message-adjoint edges
// in flight, 4 filter-adjoint variant, default tile only): the default for
// hidden size H and the pre-instantiated alternative (kind 4: the highest
// variant number; every value up to it exists).
int default_shape(int H, int which);
int alternative_shape(int H, int which);
// Split the filter weights into their bf16 planes (once per bind; synchronous).
hipError_t split_filter_weights(const Buffers& b);
// The same split enqueued on stream s (no synchronisation).
hipError_t split_filter_weights_async(const Buffers& b, hipStream_t s);
hipError_t launch_forward(const Buffers& b, hipStream_t s);
hipError_t launch_backward(const Buffers& b, hipStream_t s, float* forces, bool want_virial);
// One model stage (mlff_stage 1..8) of block `layer`, on the workspace as the
// last full evaluation left it.
hipError_t launch_stage(const Buffers& b, hipStream_t s, int stage, int layer);
hipError_t launch_reduce(const Buffers& b, hipStream_t s, double* energy, double* virial,
                         float* atom_energy);

// MD launches (md.hip).
hipError_t launch_vv_first(int N, double* x, double* p, const double* inv_mass, const float* F,
                           float* pos32, double dt, hipStream_t s);
hipError_t launch_vv_second(int N, int B, const int32_t* mol_ptr, double* p,
                            const double* inv_mass, const float* F, double dt,
                            const double* epot, double* ledger, int ledger_len,
                            int32_t* counts, const double* x, double* fx, double* fp, float* fF,
                            int n_frames, hipStream_t s);

hipError_t launch_mts_kick(int N, double* p, const float* Fa, const float* Fb, double coef,
                           hipStream_t s);
hipError_t launch_mts_select(int N, const int32_t* atom_mol, const uint8_t* flag,
                             const float* F_full, const float* F_fast, float* F_in,
                             hipStream_t s);
hipError_t launch_mts_disagree(int N, const int32_t* atom_mol, const float* F_full,
                               const float* F_fast, float* dis_max, hipStream_t s);
hipError_t launch_mts_ledger(int B, const int32_t* mol_ptr, const double* p,
                             const double* inv_mass, const double* epot, double* ledger,
                             int ledger_len, int32_t* counts, hipStream_t s);

void set_error(const char* fmt, ...);

}  // namespace mlff
