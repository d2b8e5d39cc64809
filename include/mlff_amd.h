/*
 * mlff_amd.h -- C ABI of the MI355X (gfx950) PaiNN student force field.
 *
 * The reference (atfrank/MLFF-Distiller) is pure Python: its hot path is
 * StudentForceField.forward / predict_energy_and_forces
 * (src/mlff_distiller/models/student_model.py:634-795) called per MD step by
 * StudentForceFieldCalculator.calculate / calculate_batch
 * (src/mlff_distiller/inference/ase_calculator.py:279-412, 590-817) and by
 * ASE VelocityVerlet in NVEMDHarness
 * (src/mlff_distiller/testing/nve_harness.py:173-261).  It has no FFI of its
 * own; the entry points below are what its model/calculator would bind
 * through ctypes to replace those calls (the package's own ctypes binding is
 * mlff-distiller_amd/_capi.py).
 *
 * Conventions
 *  - Every pointer argument named *_dev is device memory owned by the
 *    caller.  The library never allocates device memory: all scratch lives
 *    in one caller-provided workspace whose size mlff_plan_workspace_bytes
 *    reports.
 *  - Functions return 0 on success, a negative mlff_status on error; they
 *    never throw.  mlff_last_error() returns a message for the calling
 *    thread's last failure.
 *  - Work is enqueued on the hipStream_t passed in (as void*); no call
 *    synchronises the device except where stated.
 *  - Atoms of one structure are contiguous; structure b owns atoms
 *    [mol_ptr[b], mol_ptr[b+1]).
 *  - Results are bit-reproducible run to run: every reduction runs in a
 *    fixed order (no floating-point atomics).
 */
#ifndef MLFF_AMD_H
#define MLFF_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum mlff_status {
  MLFF_OK = 0,
  MLFF_ERR_INVALID = -1,      /* bad argument (maps to ValueError) */
  MLFF_ERR_UNSUPPORTED = -2,  /* configuration the kernels do not cover */
  MLFF_ERR_HIP = -3,          /* HIP runtime error */
  MLFF_ERR_WORKSPACE = -4,    /* workspace too small / not bound */
  MLFF_ERR_STATE = -5         /* call order (e.g. MD not bound) */
};

/* counts[] slots written by the device (int32) */
enum mlff_count_slot {
  MLFF_COUNT_EDGES = 0,       /* directed edges found (may exceed capacity) */
  MLFF_COUNT_PAIRS = 1,       /* unordered pairs found */
  MLFF_COUNT_OVERFLOW = 2,    /* sticky: non-zero when a capacity was exceeded */
  MLFF_COUNT_CELL_SMALL = 3,  /* sticky: periodic cell narrower than 2*cutoff */
  MLFF_COUNT_STEP = 4,        /* MD step counter */
  MLFF_COUNT_EDGES_MAX = 5,   /* sticky: largest edge count any neighbour list found */
  MLFF_COUNT_SLOTS = 8
};

/* Library version (major*10000 + minor*100 + patch). */
int mlff_version(void);

/* Message of the calling thread's last error ("" if none). */
const char* mlff_last_error(void);

/* ---------------------------------------------------------------------------
 * Weights.  The network parameters are packed into one float32 device blob.
 * Each slot is a row-major [rows][ld] matrix whose padding is zero; the
 * Python side fills the blob from the StudentForceField state dict
 * (student_model.py:585-605 names the tensors).  kind: 0 = copy of the
 * state-dict tensor, 1 = its transpose.
 * ------------------------------------------------------------------------- */
typedef struct mlff_slot {
  int32_t tensor;   /* mlff_tensor_id */
  int32_t layer;    /* interaction block, -1 for model-level tensors */
  int32_t kind;     /* 0 = as stored, 1 = transposed */
  int32_t rows, cols, ld;
  int64_t offset;   /* in floats from the blob start */
} mlff_slot;

enum mlff_tensor_id {
  MLFF_T_EMBEDDING = 0,   /* embedding.weight [max_z+1, H] */
  MLFF_T_RBF_CENTERS = 1, /* rbf.centers [R] */
  MLFF_T_RBF_WIDTHS = 2,  /* rbf.widths [R] */
  MLFF_T_F1_W = 3,        /* interactions.k.message.rbf_to_scalar.0.weight [H, R] */
  MLFF_T_F1_B = 4,
  MLFF_T_F2_W = 5,        /* interactions.k.message.rbf_to_scalar.2.weight [3H, H] */
  MLFF_T_F2_B = 6,
  MLFF_T_U1_W = 7,        /* interactions.k.update.update_mlp.0.weight [H, 2H] */
  MLFF_T_U1_B = 8,
  MLFF_T_U2_W = 9,        /* interactions.k.update.update_mlp.2.weight [3H, H] */
  MLFF_T_U2_B = 10,
  MLFF_T_MIX = 11,        /* interactions.k.update.mixing_matrix [3, 3] */
  MLFF_T_R1_W = 12,       /* energy_head.0.weight [H/2, H] */
  MLFF_T_R1_B = 13,
  MLFF_T_R2_W = 14,       /* energy_head.2.weight [H/4, H/2] */
  MLFF_T_R2_B = 15,
  MLFF_T_R3_W = 16,       /* energy_head.4.weight [1, H/4] */
  MLFF_T_R3_B = 17
};

/* Fill up to max_slots slot descriptors for a (H, L, R, max_z) student;
 * *n_slots and *total_floats receive the slot count and blob size. */
int mlff_weights_layout(int hidden_dim, int num_interactions, int num_rbf, int max_z,
                        mlff_slot* slots, int max_slots, int* n_slots, int64_t* total_floats);

/* Write the whole blob (total_floats, as mlff_weights_layout reports) on the
 * device in one launch: table_dev holds n_slots rows of 8 int64, ascending
 * slot offset: {slot offset, slot rows, slot ld, tensor rows, tensor cols,
 * kind, address of the tensor's float32 data on the device, 0}; a kind-1
 * slot holds the transpose of a [cols][rows] tensor.  Padding and the gaps
 * between slots are written as zeros.  n_slots <= 256.  Replaces reading
 * the parameters in place (student_model.py:585-605) after an optimizer
 * step. */
int mlff_pack_weights(void* stream, int n_slots, const int64_t* table_dev, int64_t total_floats,
                      float* blob_dev);

/* ---------------------------------------------------------------------------
 * Neighbour list.  Replaces radius_graph / radius_graph_native
 * (student_model.py:63-109, 165-191): all ordered pairs i != j of the same
 * structure with ||p_j - p_i|| <= cutoff, the distance rounded exactly as
 * torch.norm does on the CPU (sqrtf(fmaf(z,z,fmaf(y,y,x*x)))), so the edge
 * set is bit-identical.  Periodic structures (pbc all true) use the minimum
 * image; cutoff must be below half the cell's smallest perpendicular width.
 *
 * Output is a receiver CSR: row i lists its neighbours j in ascending order
 * (col_dev[row_ptr[i] .. row_ptr[i+1]) ), which for this symmetric relation
 * is exactly torch.where's (src, dst) order with src = row, dst = col.
 * shift_dev (may be NULL) receives the integer lattice shift of each edge:
 * edge vector p[col] - p[row] + shift @ cell.  counts_dev[MLFF_COUNT_EDGES]
 * receives the edge count; entries beyond edge_capacity are not written and
 * MLFF_COUNT_OVERFLOW is set; edge_capacity = 0 with col_dev = NULL only
 * counts.  workspace_dev must hold mlff_neighbor_workspace_bytes(n_atoms)
 * bytes; its contents are scratch.
 * ------------------------------------------------------------------------- */
int64_t mlff_neighbor_workspace_bytes(int n_atoms);
int mlff_neighbor_list(void* stream, const float* pos_dev, int n_atoms, int n_structures,
                       const int32_t* mol_ptr_dev, const int32_t* atom_mol_dev,
                       const float* cells_dev, const uint8_t* pbc_dev, float cutoff,
                       int32_t* row_ptr_dev, int32_t* col_dev, int32_t* shift_dev,
                       int32_t edge_capacity, int32_t* counts_dev, void* workspace_dev);

/* The same list for isolated structures (no cell), with its unordered pairs:
 * pair p = (pair_lo[p] < pair_hi[p]) numbered in CSR order of the upper edges
 * (row < col), geom_dev [P][4] = (p_hi - p_lo, distance) as float4 (16-byte
 * aligned), and edge_pair_dev [E] the pair of every directed edge.
 * counts_dev[MLFF_COUNT_PAIRS] receives the pair count; pairs beyond
 * pair_capacity are not written (edge_pair then -1) and MLFF_COUNT_OVERFLOW is
 * set.  The pair list the reference's message passing iterates as directed
 * edges (student_model.py:706-722), built once for both directions. */
int mlff_neighbor_pairs(void* stream, const float* pos_dev, int n_atoms, int n_structures,
                        const int32_t* mol_ptr_dev, const int32_t* atom_mol_dev, float cutoff,
                        int32_t* row_ptr_dev, int32_t* col_dev, int32_t edge_capacity,
                        int32_t* edge_pair_dev, int32_t* pair_lo_dev, int32_t* pair_hi_dev,
                        float* geom_dev, int32_t pair_capacity, int32_t* counts_dev,
                        void* workspace_dev);

/* ---------------------------------------------------------------------------
 * Plan: host-side descriptor of one batch shape.  Holds no device memory of
 * its own; mlff_plan_bind points it at caller buffers.
 * ------------------------------------------------------------------------- */
typedef struct mlff_plan mlff_plan;

typedef struct mlff_config {
  int32_t hidden_dim, num_interactions, num_rbf, max_z;
  float cutoff;
  int32_t n_atoms, n_structures;
  int32_t edge_capacity;      /* directed edges */
  int32_t pair_capacity;      /* unordered pairs */
  int32_t any_periodic;       /* non-zero when some structure is periodic */
  int32_t persistent_blocks;  /* 0 = library default (32 workgroups per CU) */
  int32_t cell_capacity;      /* cells reserved for cell-list neighbour search (0 = none) */
} mlff_config;

int mlff_plan_create(mlff_plan** plan, const mlff_config* cfg);
int mlff_plan_destroy(mlff_plan* plan);
int64_t mlff_plan_workspace_bytes(const mlff_plan* plan);

/* Bind the workspace and the per-batch constants.  numbers_dev [N] int64,
 * mol_ptr_dev [B+1], atom_mol_dev [N] (structure index of each atom),
 * cells_dev [B,3,3] (rows = lattice vectors) and pbc_dev [B,3] may be NULL
 * for non-periodic batches.  weights_dev is the packed blob.  The workspace
 * and the weights must be 16-byte aligned (MLFF_ERR_INVALID otherwise).  Binding is
 * synchronous: it clears the plan's counters and splits the filter weights
 * into their bf16 planes inside the workspace, then waits for the device;
 * it also creates the plan's second stream on the current device. */
int mlff_plan_bind(mlff_plan* plan, void* workspace_dev, int64_t workspace_bytes,
                   const float* weights_dev, const int64_t* numbers_dev,
                   const int32_t* mol_ptr_dev, const int32_t* atom_mol_dev,
                   const float* cells_dev, const uint8_t* pbc_dev);

/* Cell-list neighbour search for periodic structures (replaces the O(N^2)
 * scan of radius_graph_native, student_model.py:89-109, for large boxes).
 * grid_host [n_structures][4] = {nc0, nc1, nc2, first_cell} per structure;
 * nc0 = 0 keeps the brute-force scan.  Every used axis needs nc >= 3 and
 * cells at least cutoff wide (the 27-cell stencil then holds every neighbour);
 * the cells must fit cell_capacity.  The edge set is identical to the brute
 * force one.  Synchronous (copies the grid to the device); not for capture. */
int mlff_plan_set_cell_grid(mlff_plan* plan, const int32_t* grid_host, int n_structures);

/* Rebind a bound plan to a new weight blob of the same layout (a training
 * step's updated weights) without the host synchronisation of
 * mlff_plan_bind: the bf16 planes of the filter weights are re-split on
 * `stream`, ordered after the blob's writes when those are on the same
 * stream.  Evaluations on another stream must be ordered after it by the
 * caller.  Drops the plan's captured graphs.  (The reference's modules read
 * their parameters in place, student_model.py:585-605.) */
int mlff_plan_set_weights(mlff_plan* plan, void* stream, const float* weights_dev);

/* Device pointer to the plan's int32 counts[MLFF_COUNT_SLOTS]. */
int32_t* mlff_plan_counts(mlff_plan* plan);

/* Energy (float64 per structure), forces (-dE/dx, float32 [N,3]) and the
 * virial dE/d(strain) (float64 [B,9], row-major 3x3) of positions pos_dev
 * [N,3].  forces_dev / virial_dev may be NULL.  atom_energy_dev (may be
 * NULL) receives the per-atom readout [N] float32.  Replaces
 * StudentForceField.forward + torch.autograd.grad (student_model.py:634-795,
 * ase_calculator.py:708-770); the forces come from the hand-written adjoint. */
int mlff_energy_forces(mlff_plan* plan, void* stream, const float* pos_dev,
                       double* energy_dev, float* forces_dev, double* virial_dev,
                       float* atom_energy_dev);

/* ---------------------------------------------------------------------------
 * Evaluation that allocates nothing.  No counterpart in the reference: this
 * is for callers that cannot allocate device memory or pass a workspace --
 * code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same result as mlff_plan_create + mlff_plan_bind +
 * mlff_energy_forces with the same config (one stream, bit-identical), but
 * the workspace is a static device array of MLFF_STATIC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends; a batch whose
 * workspace (mlff_plan_workspace_bytes of the same config) exceeds it is
 * rejected with MLFF_ERR_WORKSPACE.  No hipMalloc/hipFree, no stream or event
 * creation, no host synchronisation: everything is enqueued on `stream`.
 *
 * Every call on one GPU shares that GPU's scratch, which keeps the last
 * call's contents: calls that overlap on one GPU (different streams, or
 * different threads) race and must be ordered by the caller.  Periodic
 * structures use the brute-force minimum-image search (cfg->cell_capacity is
 * ignored).  counts_dev [MLFF_COUNT_SLOTS] int32 (caller-owned output) is
 * cleared by the call and receives the edge and pair counts and the overflow
 * flag of this evaluation.  The weights' bf16 planes are re-split from
 * weights_dev on every call.
 * ------------------------------------------------------------------------- */
#define MLFF_STATIC_SCRATCH_BYTES (64ll << 20)
int mlff_energy_forces_static(void* stream, const mlff_config* cfg, const float* weights_dev,
                              const int64_t* numbers_dev, const int32_t* mol_ptr_dev,
                              const int32_t* atom_mol_dev, const float* cells_dev,
                              const uint8_t* pbc_dev, const float* pos_dev, double* energy_dev,
                              float* forces_dev, double* virial_dev, float* atom_energy_dev,
                              int32_t* counts_dev);

/* Re-launch one stage of the last mlff_energy_forces call on its workspace
 * contents (for timing individual kernels and for the launch-shape tuner;
 * no counterpart in the reference).  Stages are idempotent: they recompute
 * the same outputs from the same inputs.  layer selects the interaction
 * block for per-block stages. */
enum mlff_stage {
  MLFF_STAGE_NEIGHBOR = 0,
  MLFF_STAGE_FILTER_FWD = 1,
  MLFF_STAGE_MESSAGE_FWD = 2,
  MLFF_STAGE_UPDATE_FWD = 3,
  MLFF_STAGE_READOUT = 4,
  MLFF_STAGE_UPDATE_BWD = 5,
  MLFF_STAGE_FILTER_BWD = 6,
  MLFF_STAGE_MESSAGE_BWD = 7,
  MLFF_STAGE_FORCES = 8
};
int mlff_launch_stage(mlff_plan* plan, void* stream, int stage, int layer);

/* Set the persistent grid (workgroups) used by the pair kernels. */
int mlff_plan_set_blocks(mlff_plan* plan, int blocks);

/* Launch shapes of the pair and message kernels (for the launch-shape
 * tuner; no counterpart in the reference, whose Triton kernels fix their
 * block sizes, kernels/fused_rbf_cutoff.py:121-132): pairs per filter-forward
 * tile, pairs per filter-adjoint tile, and edges in flight per row in the
 * message forward and adjoint.  0 keeps the default; each kind has one
 * pre-instantiated alternative (forward tile 64 or 32 for H >= 64, 128 or 64
 * for H = 32; adjoint tile 32 or 64 for H >= 64, 64 or 32 for H = 32;
 * message 5 or 3 and 4 or 2 edges), anything else is MLFF_ERR_UNSUPPORTED.
 * bwd_variant (default adjoint tile only) picks the filter adjoint's form:
 * 0 as described above; 1 loads the pair indices of the next tile's gather
 * before the current tile's matrix products.
 * Results are bit-identical across shapes (every output element sums in the
 * same order).  Drops any captured graph of the plan. */
int mlff_plan_set_tiles(mlff_plan* plan, int fwd_pairs, int bwd_pairs, int msg_unroll_fwd,
                        int msg_unroll_bwd, int bwd_variant);

/* Number of streams one evaluation uses: 2 (default) runs the pair filter
 * kernels on a stream of the plan's own, forked from and joined back into the
 * caller's stream with events, so they overlap the atom-side kernels; 1 keeps
 * every kernel on the caller's stream.  Results are bit-identical either way.
 * The filter stream is created on the device current at mlff_plan_bind. */
int mlff_plan_set_streams(mlff_plan* plan, int n_streams);

/* ---------------------------------------------------------------------------
 * Molecular dynamics.  Velocity Verlet as ASE's VelocityVerlet.step
 * (called from nve_harness.py:214-235): p += dt/2 F; x += dt p/m; F = F(x);
 * p += dt/2 F.  State is float64 (as ASE keeps it); the model sees a float32
 * copy of x.  Every step appends (E_pot, E_kin) per structure to a float64
 * ledger [ledger_len, B, 2] at row step % ledger_len.
 * ------------------------------------------------------------------------- */
int mlff_md_bind(mlff_plan* plan, double* x_dev, double* p_dev, const double* inv_mass_dev,
                 float* pos32_dev, float* forces_dev, double* energy_dev,
                 double* ledger_dev, int32_t ledger_len, double dt);
/* Per-step frames, the positions, velocities and forces the reference's
 * NVEMDHarness._collect_data copies through the host every step
 * (nve_harness.py:263-287).  With n_frames > 0, every step also stores x and
 * p (float64) and F (float32), each [N, 3], into row step % n_frames of
 * x_frames / p_frames / f_frames ([n_frames, N, 3]); the row index is the
 * ledger's step counter, so graph replays record too.  n_frames = 0 unbinds.
 * Needs mlff_md_bind first; a later mlff_md_bind keeps the binding. */
int mlff_md_bind_frames(mlff_plan* plan, double* x_frames_dev, double* p_frames_dev,
                        float* f_frames_dev, int32_t n_frames);
/* Forces at the current positions (call once before the first step). */
int mlff_md_init_forces(mlff_plan* plan, void* stream);
/* Enqueue n_steps steps (eager launches). */
int mlff_md_run(mlff_plan* plan, void* stream, int n_steps);
/* Capture steps_per_graph steps into a HIP graph (once), then replay it
 * n_launches times. */
int mlff_md_graph_run(mlff_plan* plan, void* stream, int steps_per_graph, int n_launches);

/* ---------------------------------------------------------------------------
 * Multiple time step (no counterpart in the reference, whose MD is plain
 * Velocity Verlet, nve_harness.py:214).  r-RESPA with the inner plan's model
 * (a compact student) as the fast force at dt and the outer plan's model as
 * the slow correction F_outer - F_inner every n_inner steps.  The outer
 * model evaluates on the inner plan's neighbour list and pair geometry (same
 * positions, same cutoff; both plans need equal capacities).  Per structure,
 * flag != 0 makes the outer model the inner force (fallback: plain
 * Velocity Verlet with the outer model).  dis_max [B] receives the running
 * max of |F_outer - F_inner_model| at outer steps.  The ledger gets
 * (E_outer, E_kin) per structure every outer step.
 * ------------------------------------------------------------------------- */
int mlff_mts_bind(mlff_plan* outer, mlff_plan* inner, double* x_dev, double* p_dev,
                  const double* inv_mass_dev, float* pos32_dev, float* f_outer_dev,
                  float* f_inner_dev, float* f_mix_dev, double* e_outer_dev,
                  double* e_inner_dev, const uint8_t* flag_dev, float* dis_max_dev,
                  double* ledger_dev, int32_t ledger_len, double dt, int32_t n_inner);
/* Forces of both models at the current positions. */
int mlff_mts_init_forces(mlff_plan* outer, void* stream);
/* Enqueue n_outer outer steps; fallback != 0 honours the flags (and
 * evaluates the outer model at every inner step). */
int mlff_mts_run(mlff_plan* outer, void* stream, int n_outer, int fallback);
/* Capture outer_per_graph outer steps (no fallback) once, replay n_launches times. */
int mlff_mts_graph_run(mlff_plan* outer, void* stream, int outer_per_graph, int n_launches);

/* ---------------------------------------------------------------------------
 * Force-matching training pass.  The reference differentiates the loss
 * wE MSE(E) + wF MSE(F) by autograd double backward
 * (models/distillation_wrapper.py:86-91, training/trainer.py:325-328,
 * scripts/train_compact_models.py:229-372); the package's train.py forms the
 * same gradient explicitly (a reverse pass and its tangent along dL/dF) and
 * runs its edge- and pair-sized parts here.  All buffers are float32 device
 * memory (features [N][H] and [N][3][H], 16-byte aligned), graphs are a
 * symmetric receiver CSR (row_ptr [N+1], col [E] ascending within a row)
 * with edge_pair [E] the unordered pair (lo < hi) of each directed edge.
 * Every sum runs in a fixed order (bit-reproducible, no atomics).
 * hidden_dim is 32, 64 or 128.
 * ------------------------------------------------------------------------- */

/* Message block of PaiNNMessage.forward (student_model.py:357-387) on
 * given filters W [P][3H] = (Ws, Wv1, Wv2) and pair unit vectors u [P][3]
 * (from lo to hi):
 *   s_out[i] = s[i] + sum_e s[j] Ws[p],
 *   v_out[i] = v[i] + sum_e (v[j] Wv1[p] + sg u[p] (x) Wv2[p]),
 * sg = +1 when i is the pair's lower atom, -1 otherwise.  With ds_dev non-NULL
 * also the tangent along (ds, dv, dW, du) into ds_out, dv_out. */
int mlff_train_message(void* stream, int hidden_dim, int n_atoms, const int32_t* row_ptr_dev,
                       const int32_t* col_dev, const int32_t* edge_pair_dev, const float* s_dev,
                       const float* v_dev, const float* w_dev, const float* u_dev,
                       const float* ds_dev, const float* dv_dev, const float* dw_dev,
                       const float* du_dev, float* s_out_dev, float* v_out_dev,
                       float* ds_out_dev, float* dv_out_dev);

/* Adjoint of the message gathers (the backward of student_model.py:359-385,
 * index_add_ transposed): gs_out[i] = gs[i] + sum_e gs[j] Ws[p],
 * gv_out[i] = gv[i] + sum_e gv[j] Wv1[p]; with dgs_dev non-NULL also its
 * tangent along (dgs, dgv, dW). */
int mlff_train_message_adjoint(void* stream, int hidden_dim, int n_atoms,
                               const int32_t* row_ptr_dev, const int32_t* col_dev,
                               const int32_t* edge_pair_dev, const float* gs_dev,
                               const float* gv_dev, const float* w_dev, const float* dgs_dev,
                               const float* dgv_dev, const float* dw_dev, float* gs_out_dev,
                               float* gv_out_dev, float* dgs_out_dev, float* dgv_out_dev);

/* Gradient w.r.t. each pair's filter output W_p (the backward of
 * student_model.py:351-385 into rbf_to_scalar's output): for p = (a < b),
 *   gW[p] = (gs_a s_b + gs_b s_a,  sum_x gv_a,x v_b,x + gv_b,x v_a,x,
 *            sum_x u_x (gv_a,x - gv_b,x)),
 * and with dgw_dev non-NULL its tangent along (du, ds, dv, dgs, dgv). */
int mlff_train_pair_grad(void* stream, int hidden_dim, int n_pairs, const int32_t* pair_lo_dev,
                         const int32_t* pair_hi_dev, const float* u_dev, const float* du_dev,
                         const float* s_dev, const float* v_dev, const float* ds_dev,
                         const float* dv_dev, const float* gs_dev, const float* gv_dev,
                         const float* dgs_dev, const float* dgv_dev, float* gw_dev,
                         float* dgw_dev);

/* PaiNNMessage.forward on a caller-given directed edge list
 * (student_model.py:346-387, the interaction submodules' own forward):
 * each edge e (src -> dst) has its own filter row w[e] [3H] and edge vector
 * ev[e] [3].  Edges are grouped by receiver: row_ptr [N+1] over dst, col [E]
 * the sources, edge_id [E] each edge's index into w and ev; within a row they
 * are added in that order:
 *   s_out[i] = s[i] + sum_e s[src] Ws[e],
 *   v_out[i] = v[i] + sum_e (v[src] Wv1[e] + ev[e] (x) Wv2[e]). */
int mlff_edge_message(void* stream, int hidden_dim, int n_atoms, const int32_t* row_ptr_dev,
                      const int32_t* col_dev, const int32_t* edge_id_dev, const float* s_dev,
                      const float* v_dev, const float* w_dev, const float* ev_dev,
                      float* s_out_dev, float* v_out_dev);

/* Per-edge gradients of mlff_edge_message given the gradients (gs, gv) of
 * its outputs: gw[e] = (gs[dst] s[src], sum_x gv[dst]_x v[src]_x,
 * sum_x ev_x gv[dst]_x) [E][3H] and gev[e]_x = sum_h gv[dst]_x,h Wv2[e]_h
 * [E][3].  (The gradients w.r.t. s and v are mlff_train_message_adjoint on
 * the edges grouped by source.) */
int mlff_edge_grad(void* stream, int hidden_dim, int n_edges, const int32_t* src_dev,
                   const int32_t* dst_dev, const float* s_dev, const float* v_dev,
                   const float* w_dev, const float* ev_dev, const float* gs_dev,
                   const float* gv_dev, float* gw_dev, float* gev_dev);

/* Fixed-order segment sums, replacing index_add_ reductions (per-structure
 * energies, student_model.py:739-755; the embedding gradient):
 * out[k][c] = sum over r in [seg_ptr[k], seg_ptr[k+1]) of
 * src[perm ? perm[r] : r][c], into exactly one of out64_dev (float64) or
 * out32_dev (float32), [n_segments][width].  The rows are added in a fixed
 * partition and order (strided row groups, then the groups in order). */
int mlff_segment_sum(void* stream, int n_segments, int width, const int32_t* seg_ptr_dev,
                     const int32_t* perm_dev, const float* src_dev, double* out64_dev,
                     float* out32_dev);

/* Weight and bias gradients of a Linear layer over K rows (pairs or atoms),
 * the parameter-gradient products of the reference's autograd backward
 * through nn.Linear (student_model.py:318 and 350 rbf_to_scalar, 407 and 444
 * update_mlp, 599 and 736 energy_head; the 3x3 mixing_matrix, 414 and 460, as
 * the block-diagonal sums of one [3H][3H] product):
 *   out[m][n] = sum_k w_k (a0[k][m] b0[k][n] + a1[k][m] b1[k][n])   [M][N]
 *   bias[m]   = sum_k w_k a0[k][m]                                  [M]
 * a0, a1 [K][lda], b0, b1 [K][ldb] row-major float32; a1/b1 NULL for one
 * term, row_w_dev [K] NULL for w = 1, bias_out_dev NULL to skip the bias
 * (K = 0 writes zeros; a0 and b0 may then be NULL).
 * M and N multiples of 4, strides multiples of 4, buffers 16-byte aligned.
 * The K rows are split at a partition that depends on (K, M, N) only and the
 * partial sums are added in a fixed order (bit-reproducible).  ws_dev is
 * device scratch of at least mlff_weight_grad_workspace_bytes(K, M, N). */
size_t mlff_weight_grad_workspace_bytes(int K, int M, int N);
int mlff_weight_grad(void* stream, int K, int M, int N, const float* a0_dev, int lda0,
                     const float* b0_dev, int ldb0, const float* a1_dev, int lda1,
                     const float* b1_dev, int ldb1, const float* row_w_dev, void* ws_dev,
                     size_t ws_bytes, float* out_dev, float* bias_out_dev);

/* The same weight gradients in two steps, so that the partial sums of many
 * gradients (one training pass's) are added by one launch.  The products the
 * reference's backward computes are the same as for mlff_weight_grad.
 * mlff_weight_grad_partials takes mlff_weight_grad's inputs and workspace and
 * writes only the partial sums: S = mlff_weight_grad_partial_count(K, M, N)
 * weight partials [S][M][N] at the start of ws_dev, then, with with_bias,
 * S bias partials [S][M].  mlff_weight_grad_reduce_group then adds partials
 * for n_rows outputs (at most 256): row r of table_dev (int64, 16-byte
 * aligned, on the device) is {S, n4, partials address, output address}, n4
 * being the output's float count / 4; partials and outputs 16-byte aligned.
 * Any n_blocks >= 1 covers every row; the sum of ceil(n4 / 256) over the
 * rows runs each block once.  Every element is added in the partition and
 * order of mlff_weight_grad, so the outputs are bit-identical to it.  Each
 * row's S must be the partial count of the gradient it adds. */
int mlff_weight_grad_partial_count(int K, int M, int N);
int mlff_weight_grad_partials(void* stream, int K, int M, int N, const float* a0_dev, int lda0,
                              const float* b0_dev, int ldb0, const float* a1_dev, int lda1,
                              const float* b1_dev, int ldb1, const float* row_w_dev,
                              int with_bias, void* ws_dev, size_t ws_bytes);
int mlff_weight_grad_reduce_group(void* stream, int n_rows, const int64_t* table_dev,
                                  int n_blocks);

/* Pair geometry and radial basis of the training pass over n_pairs pairs
 * (lo_dev, hi_dev int32; x_dev [N][3] float32 positions; shift_dev [P][3] the
 * periodic image shift already times the cell, or NULL), as in
 * student_model.py:706-722 (edge vectors, distances, unit vectors) and
 * 249-253, 285-290 (Gaussian basis times cosine cutoff):
 *   vec = x[hi] - x[lo] + shift, d = |vec|, den = d + eps, u = vec / den,
 *   e [P][num_rbf] the basis and de = de/dd.
 * With lam_dev ([N][3], a position direction) it writes the tangent instead:
 *   dd = vec . (lam[hi] - lam[lo]) / d, du = d(u) along lam, de_t = de dd
 * (dd_dev [P], du_dev [P][3], de_t_dev [P][num_rbf]; the forward outputs may
 * then be NULL). */
int mlff_train_geometry(void* stream, int n_pairs, int num_rbf, const float* x_dev,
                        const float* lam_dev, const int32_t* lo_dev, const int32_t* hi_dev,
                        const float* shift_dev, const float* centers_dev,
                        const float* widths_dev, float cutoff, float eps, float* vec_dev,
                        float* d_dev, float* den_dev, float* u_dev, float* e_dev,
                        float* de_dev, float* dd_dev, float* du_dev, float* de_t_dev);

/* SiLU of the MLPs (student_model.py:318-322, 407-411, 599-605) and its
 * derivatives, elementwise over n float32 values (n a multiple of 4, buffers
 * 16-byte aligned), each output skipped when its pointer is NULL:
 *   h = silu(a),  dh = silu'(a) da,  ga = g silu'(a),
 *   dga = dg silu'(a) + g silu''(a) da   (dg NULL: zero). */
int mlff_silu_chain(void* stream, int64_t n, const float* a_dev, const float* da_dev,
                    const float* g_dev, const float* dg_dev, float* h_dev, float* dh_dev,
                    float* ga_dev, float* dga_dev);

/* Gradient of PaiNNUpdate's 3x3 mixing_matrix (student_model.py:414 and 460
 * under the reference's autograd) over n_atoms atoms:
 *   out[x][z] = sum_n w_n sum_h (a0[n][x][h] b0[n][z][h] + a1[n][x][h] b1[n][z][h])
 * a0, b0, a1, b1 [n_atoms][3][hidden] float32, 16-byte aligned (a1/b1 NULL for
 * one term, row_w_dev NULL for w = 1; all may be NULL when n_atoms is 0);
 * hidden a multiple of 4.  Writes
 * S = mlff_xyz_outer_partial_count(n_atoms) partial sums of 12 floats each
 * (the 9 sums row-major, then zeros) to part_dev [S][12]; the
 * mlff_weight_grad_reduce_group row {S, 3, part_dev, out} adds them (in a
 * fixed partition and order) into out[12]. */
int mlff_xyz_outer_partial_count(int n_atoms);
int mlff_xyz_outer_partials(void* stream, int n_atoms, int hidden, const float* a0_dev,
                            const float* b0_dev, const float* a1_dev, const float* b1_dev,
                            const float* row_w_dev, float* part_dev);

/* The filter MLP of PaiNNMessage (student_model.py:313-322, 350-355) over
 * n_pairs pairs in the training pass, fused in one pair-tile kernel:
 *   tangent = 0:  a = x W1^T + b1,  h = silu(a),         W = h W2^T + b2
 *   tangent = 1:  a = x W1^T,       h = silu'(a_dev) a,  W = h W2^T
 * x [P][num_rbf] (the basis e, or its tangent), a and h [P][hidden],
 * W [P][3 hidden], row-major float32.  w1_padded_dev is W1 [hidden][RP]
 * with RP = 16 (num_rbf <= 16) or 32 and zeros past num_rbf; w2_planes_dev
 * is W2 [3 hidden][hidden] split by mlff_split_bf16x3.  The second product
 * runs on the bf16 matrix cores with both operands in three bf16 parts
 * (float32 accuracy).  hidden 32, 64 or 128; num_rbf <= 32.  Replaces
 * rbf_to_scalar's forward and its autograd tangent in the reference's
 * double backward (distillation_wrapper.py:86-91). */
int mlff_train_filter(void* stream, int tangent, int hidden, int n_pairs, int num_rbf,
                      const float* x_dev, const float* w1_padded_dev, const float* b1_dev,
                      const uint16_t* w2_planes_dev, const float* b2_dev, const float* a_dev,
                      float* a_out_dev, float* h_out_dev, float* w_out_dev);

/* The reverse of the same MLP's first layer in the training pass (the
 * autograd backward of rbf_to_scalar, student_model.py:318-322, in the
 * reference's double backward, distillation_wrapper.py:86-91), over n_pairs
 * pairs, fused in one pair-tile kernel:
 *   ga  = silu'(a) (gw W2)
 *   dga = silu'(a) (dgw W2) + silu''(a) da (gw W2)     (tangent = 1 only)
 * gw, dgw [P][3 hidden] (the adjoint of the filter output W and its tangent),
 * a, da, ga, dga [P][hidden], row-major float32; w2t_dev is W2 transposed,
 * [hidden][3 hidden].  The products run on the FP32 matrix cores.  gw, dgw and
 * w2t 16-byte aligned; hidden 32, 64 or 128. */
int mlff_train_filter_adjoint(void* stream, int tangent, int hidden, int n_pairs,
                              const float* gw_dev, const float* dgw_dev, const float* w2t_dev,
                              const float* a_dev, const float* da_dev, float* ga_out_dev,
                              float* dga_out_dev);

/* Split n float32 values into three bfloat16 planes (planes_dev [3][n]):
 * x = hi + mid + lo, each part the round-to-nearest-even bfloat16 of what the
 * previous parts leave (24 significant bits together). */
int mlff_split_bf16x3(void* stream, int64_t n, const float* src_dev, uint16_t* planes_dev);

/* Elementwise parts of PaiNNUpdate.forward (student_model.py:430-470) in the
 * training pass, around the update MLP's two products.  s [N][H],
 * v [N][3][H], xin [N][2H], y = MLP(xin) = (ds, g1, g2) [N][3H], M the 3x3
 * mixing matrix (row-major, device); hidden a multiple of 4, buffers
 * 16-byte aligned float32.
 *   update_norm:          xin = (s, |v|_xyz)
 *   update_apply:         s' = s + ds, Mv = M v, v' = v g1 + Mv g2
 *   update_norm_tangent:  dxin = (ds, dnrm), dnrm = sum_x v dv / |v| (0 at |v| = 0)
 *   update_apply_tangent: ds' = ds + dds, dMv = M dv,
 *                         dv' = dv g1 + v dg1 + dMv g2 + Mv dg2   (dy = (dds, dg1, dg2))
 *   update_gate_adjoint:  gy = (gs, sum_x gv v, sum_x gv Mv), and with dgy_dev
 *                         non-NULL its tangent along (dgs, dgv, dv, dMv)
 *   update_input_adjoint: with gx = (gs_x, gn) = dL/dxin: gvg2 = gv g2,
 *                         gs_m = gs + gs_x, gv_m = gv g1 + M^T gvg2 + gn v/|v|,
 *                         and with dgs_out_dev non-NULL their tangents along
 *                         (dgx, dgs, dgv, dy, dv, dxin). */
int mlff_update_norm(void* stream, int n_atoms, int hidden, const float* s_dev,
                     const float* v_dev, float* xin_dev);
int mlff_update_apply(void* stream, int n_atoms, int hidden, const float* m_dev,
                      const float* y_dev, const float* s_dev, const float* v_dev,
                      float* s_out_dev, float* v_out_dev, float* mv_out_dev);
int mlff_update_norm_tangent(void* stream, int n_atoms, int hidden, const float* v_dev,
                             const float* xin_dev, const float* ds_dev, const float* dv_dev,
                             float* dxin_dev);
int mlff_update_apply_tangent(void* stream, int n_atoms, int hidden, const float* m_dev,
                              const float* y_dev, const float* dy_dev, const float* ds_dev,
                              const float* v_dev, const float* dv_dev, const float* mv_dev,
                              float* ds_out_dev, float* dv_out_dev, float* dmv_out_dev);
int mlff_update_gate_adjoint(void* stream, int n_atoms, int hidden, const float* gs_dev,
                             const float* gv_dev, const float* v_dev, const float* mv_dev,
                             const float* dgs_dev, const float* dgv_dev, const float* dv_dev,
                             const float* dmv_dev, float* gy_dev, float* dgy_dev);
int mlff_update_input_adjoint(void* stream, int n_atoms, int hidden, const float* m_dev,
                              const float* gx_dev, const float* gs_dev, const float* gv_dev,
                              const float* y_dev, const float* v_dev, const float* xin_dev,
                              const float* dgx_dev, const float* dgs_dev, const float* dgv_dev,
                              const float* dy_dev, const float* dv_dev, const float* dxin_dev,
                              float* gs_out_dev, float* gv_out_dev, float* gvg2_out_dev,
                              float* dgs_out_dev, float* dgv_out_dev, float* dgvg2_out_dev);

/* ---------------------------------------------------------------------------
 * Matrix-core throughput probe, for the measured peak next to the published
 * one in roofline figures (no counterpart in the reference, which has no GPU
 * roofline).  blocks x 256 threads; each wave issues iters x 8 independent
 * register-only MFMAs of one shape (kind 0: f32 16x16x4, kind 1: bf16
 * 16x16x32) and writes one float per thread to out_dev [blocks*256].
 * mlff_probe_mfma_flops returns the launch's FLOP count (-1 on bad input).
 * ------------------------------------------------------------------------- */
int64_t mlff_probe_mfma_flops(int kind, int blocks, int iters);
int mlff_probe_mfma(void* stream, int kind, int blocks, int iters, float* out_dev);

#ifdef __cplusplus
}
#endif

#endif /* MLFF_AMD_H */
