// extern "C" entry points (include/mlff_amd.h).  Host-side bookkeeping
// only: the plan records the caller's device buffers, the workspace carve-up
// and the weight layout; no device memory is allocated here.

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

#include "common.hpp"
#include "internal.hpp"

namespace mlff {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

static int hip_fail(hipError_t e, const char* what) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return MLFF_ERR_HIP;
}

static inline int64_t align_floats(int64_t x) { return (x + 63) & ~(int64_t)63; }

int build_layout(int H, int L, int R, int max_z, ModelW* mw, mlff_slot* slots, int max_slots,
                 int* n_slots) {
  if (H != 32 && H != 64 && H != 128) {
    set_error("hidden_dim %d not supported (kernels cover 32, 64, 128)", H);
    return MLFF_ERR_UNSUPPORTED;
  }
  if (L < 1 || L > 8) { set_error("num_interactions %d outside 1..8", L); return MLFF_ERR_UNSUPPORTED; }
  // the filter adjoint keeps d e_k/d d of a tile in LDS for at most 32 basis functions
  if (R < 1 || R > 32) { set_error("num_rbf %d outside 1..32", R); return MLFF_ERR_UNSUPPORTED; }
  if (max_z < 1) { set_error("max_z must be >= 1"); return MLFF_ERR_INVALID; }
  ModelW m;
  memset(&m, 0, sizeof(m));
  m.H = H; m.L = L; m.R = R; m.RP = pad16(R); m.HP2 = pad16(H / 2); m.HP4 = pad16(H / 4);
  m.max_z = max_z;
  int64_t off = 0;
  int ns = 0;
  auto add = [&](int tensor, int layer, int kind, int rows, int cols, int ld) -> int64_t {
    const int64_t o = off;
    if (slots != nullptr && ns < max_slots) {
      slots[ns].tensor = tensor; slots[ns].layer = layer; slots[ns].kind = kind;
      slots[ns].rows = rows; slots[ns].cols = cols; slots[ns].ld = ld; slots[ns].offset = o;
    }
    ++ns;
    off = align_floats(off + (int64_t)rows * ld);
    return o;
  };
  const int RP = m.RP, HP2 = m.HP2, HP4 = m.HP4;
  m.emb = add(MLFF_T_EMBEDDING, -1, 0, max_z + 1, H, H);
  m.centers = add(MLFF_T_RBF_CENTERS, -1, 0, 1, R, RP);
  m.widths = add(MLFF_T_RBF_WIDTHS, -1, 0, 1, R, RP);
  for (int l = 0; l < L; ++l) {
    LayerW& w = m.layer[l];
    w.f1w = add(MLFF_T_F1_W, l, 0, H, R, RP);
    w.f1b = add(MLFF_T_F1_B, l, 0, 1, H, H);
    w.f1t = add(MLFF_T_F1_W, l, 1, RP, H, H);
    w.f2w = add(MLFF_T_F2_W, l, 0, 3 * H, H, H);
    w.f2b = add(MLFF_T_F2_B, l, 0, 1, 3 * H, 3 * H);
    w.f2t = add(MLFF_T_F2_W, l, 1, H, 3 * H, 3 * H);
    w.u1w = add(MLFF_T_U1_W, l, 0, H, 2 * H, 2 * H);
    w.u1b = add(MLFF_T_U1_B, l, 0, 1, H, H);
    w.u1t = add(MLFF_T_U1_W, l, 1, 2 * H, H, H);
    w.u2w = add(MLFF_T_U2_W, l, 0, 3 * H, H, H);
    w.u2b = add(MLFF_T_U2_B, l, 0, 1, 3 * H, 3 * H);
    w.u2t = add(MLFF_T_U2_W, l, 1, H, 3 * H, 3 * H);
    w.mix = add(MLFF_T_MIX, l, 0, 3, 3, 3);
  }
  m.r1w = add(MLFF_T_R1_W, -1, 0, HP2, H, H);
  m.r1b = add(MLFF_T_R1_B, -1, 0, 1, H / 2, HP2);
  m.r1t = add(MLFF_T_R1_W, -1, 1, H, H / 2, HP2);
  m.r2w = add(MLFF_T_R2_W, -1, 0, HP4, H / 2, HP2);
  m.r2b = add(MLFF_T_R2_B, -1, 0, 1, H / 4, HP4);
  m.r2t = add(MLFF_T_R2_W, -1, 1, HP2, H / 4, HP4);
  m.r3w = add(MLFF_T_R3_W, -1, 0, 1, H / 4, HP4);
  m.r3b = add(MLFF_T_R3_B, -1, 0, 1, 1, 16);
  m.total = off + 64;   // zero tail
  if (mw) *mw = m;
  if (n_slots) *n_slots = ns;
  return MLFF_OK;
}

}  // namespace mlff

using namespace mlff;

struct mlff_plan {
  mlff_config cfg;
  ModelW mw;
  Buffers buf;
  NLArgs nl;
  int64_t ws_bytes;
  bool bound;
  // MD
  bool md_bound;
  double* x; double* p; const double* inv_mass; float* pos32; float* forces; double* energy;
  double* ledger; int ledger_len; double dt;
  double* frame_x = nullptr; double* frame_p = nullptr; float* frame_f = nullptr;
  int n_frames = 0;
  hipGraph_t graph; hipGraphExec_t graph_exec; int graph_steps; hipStream_t graph_stream;
  int4* grid_dev;
  // multiple time step (this plan is the outer one)
  bool mts_bound;
  mlff_plan* mts_inner;
  float* f_outer; float* f_inner; float* f_mix; double* e_outer; double* e_inner;
  const uint8_t* flag; float* dis_max; int n_inner;
  hipGraph_t mts_graph; hipGraphExec_t mts_exec; int mts_graph_steps;
  // filter stream (created at bind on the current device) and its fork/join events
  int n_streams;
  int stream_dev;
  hipStream_t aux;
  hipEvent_t ev[2 * 8 + 2];
  // workspace carve-up (byte offsets)
  struct { int64_t off, bytes; } seg[48];
  int nseg;
};

namespace {

enum Seg {
  S_COUNTS, S_DEG, S_UDEG, S_ROWPTR, S_PAIRPTR, S_COL, S_EPAIR, S_PLO, S_PHI, S_PSHIFT, S_GEOM,
  S_PACC, S_GVEC, S_WF, S_S, S_V, S_SMID, S_VMID, S_AUPD, S_G12, S_GS0, S_GV0, S_GS1, S_GV1,
  S_EPS, S_VIR, S_GRID, S_CELLCNT, S_CELLSTART, S_CELLATOMS, S_ATOMCELL, S_TMPSHIFT, S_WBF,
  S_COUNT_
};

void carve(mlff_plan* pl) {
  const mlff_config& c = pl->cfg;
  const int64_t N = c.n_atoms, H = c.hidden_dim, L = c.num_interactions;
  const int64_t E = c.edge_capacity, P = c.pair_capacity;
  static_assert(S_COUNT_ <= (int)(sizeof(pl->seg) / sizeof(pl->seg[0])), "segment table too small");
  int64_t sz[S_COUNT_] = {};
  sz[S_COUNTS] = 4 * MLFF_COUNT_SLOTS;
  sz[S_DEG] = 4 * N;
  sz[S_UDEG] = 4 * N;
  sz[S_ROWPTR] = 4 * (N + 1);
  sz[S_PAIRPTR] = 4 * (N + 1);
  sz[S_COL] = 4 * E;
  sz[S_EPAIR] = 4 * E;
  sz[S_PLO] = 4 * P;
  sz[S_PHI] = 4 * P;
  sz[S_PSHIFT] = c.any_periodic ? 12 * P : 0;
  // at least one pair row: edges whose pair is switched off (p = -1) read
  // row 0 of the pair geometry and of the filter outputs
  const int64_t P1 = P > 0 ? P : 1;
  sz[S_GEOM] = 16 * P1;
  sz[S_PACC] = 16 * P;
  sz[S_GVEC] = 16 * P;
  sz[S_WF] = 4 * L * P1 * 3 * H;
  sz[S_S] = 4 * (L + 1) * N * H;
  sz[S_V] = 4 * (L + 1) * N * 3 * H;
  sz[S_SMID] = 4 * L * N * H;
  sz[S_VMID] = 4 * L * N * 3 * H;
  sz[S_AUPD] = 4 * L * N * H;
  sz[S_G12] = 4 * L * N * 2 * H;
  sz[S_GS0] = 4 * N * H;
  sz[S_GV0] = 4 * N * 3 * H;
  sz[S_GS1] = 4 * L * N * H;
  sz[S_GV1] = 4 * L * N * 3 * H;
  sz[S_EPS] = 4 * N;
  sz[S_VIR] = 4 * 9 * N;
  const int64_t CC = c.cell_capacity > 0 ? c.cell_capacity : 0;
  sz[S_GRID] = CC ? 16 * (int64_t)c.n_structures : 0;
  sz[S_CELLCNT] = 4 * CC;
  sz[S_CELLSTART] = CC ? 4 * (CC + 1) : 0;
  sz[S_CELLATOMS] = CC ? 4 * N : 0;
  sz[S_ATOMCELL] = CC ? 4 * N : 0;
  sz[S_TMPSHIFT] = CC ? 12 * E : 0;
  sz[S_WBF] = 2 * wbf_elems((int)H, (int)L);
  int64_t off = 0;
  for (int k = 0; k < S_COUNT_; ++k) {
    pl->seg[k].off = off;
    pl->seg[k].bytes = sz[k];
    off += (sz[k] + 255) & ~(int64_t)255;
  }
  pl->nseg = S_COUNT_;
  pl->ws_bytes = off;
}

void release_streams(mlff_plan* pl) {
  for (hipEvent_t& e : pl->ev)
    if (e) { (void)hipEventDestroy(e); e = nullptr; }
  if (pl->aux) { (void)hipStreamDestroy(pl->aux); pl->aux = nullptr; }
  pl->stream_dev = -1;
}

// The filter stream lives on the device current at bind time (the device of
// the workspace).  Default priority, non-blocking: only the events order it.
hipError_t acquire_streams(mlff_plan* pl) {
  pl->buf.aux = nullptr;
  pl->buf.ev = pl->ev;
  if (pl->n_streams < 2) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (pl->aux && pl->stream_dev != dev) release_streams(pl);
  if (!pl->aux) {
    e = hipStreamCreateWithFlags(&pl->aux, hipStreamNonBlocking);
    if (e != hipSuccess) { pl->aux = nullptr; return e; }
    for (hipEvent_t& ev : pl->ev) {
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) { ev = nullptr; release_streams(pl); return e; }
    }
    pl->stream_dev = dev;
  }
  pl->buf.aux = pl->aux;
  return hipSuccess;
}

// Workgroups of the persistent pair kernels: the caller's choice, else 32 per
// CU, sixteen rounds of two resident blocks (cus < 0: ask the current device).
// Many short-lived blocks balance the tiles better than a few long-lived ones:
// one 256-molecule evaluation took 3.53 ms at 32 per CU against 3.62 ms at 6.
int persistent_grid(int blocks, int cus) {
  if (blocks <= 0) {
    if (cus <= 0) {
      int dev = 0;
      cus = 256;
      if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
          cus = prop.multiProcessorCount;
      }
    }
    blocks = 32 * cus;
  }
  return (blocks + 7) / 8 * 8;
}

void set_default_shapes(mlff_plan* pl) {
  const int H = pl->cfg.hidden_dim;
  pl->buf.tile_fwd = default_shape(H, 0);
  pl->buf.tile_bwd = default_shape(H, 1);
  pl->buf.unroll_fwd = default_shape(H, 2);
  pl->buf.unroll_bwd = default_shape(H, 3);
  pl->buf.bwd_variant = default_shape(H, 4);
}

template <typename T>
T* segp(mlff_plan* pl, void* ws, int k) {
  return pl->seg[k].bytes ? reinterpret_cast<T*>(static_cast<char*>(ws) + pl->seg[k].off) : nullptr;
}

}  // namespace

extern "C" {

int mlff_version(void) { return 100; }

const char* mlff_last_error(void) { return g_err; }

int mlff_weights_layout(int hidden_dim, int num_interactions, int num_rbf, int max_z,
                        mlff_slot* slots, int max_slots, int* n_slots, int64_t* total_floats) {
  ModelW mw;
  int rc = build_layout(hidden_dim, num_interactions, num_rbf, max_z, &mw, slots, max_slots,
                        n_slots);
  if (rc != MLFF_OK) return rc;
  if (total_floats) *total_floats = mw.total;
  return MLFF_OK;
}

int64_t mlff_neighbor_workspace_bytes(int n_atoms) {
  if (n_atoms < 0) return -1;
  // deg, udeg, pair_ptr
  return (int64_t)4 * (3 * (int64_t)n_atoms + 1) + 256;
}

int mlff_neighbor_list(void* stream, const float* pos_dev, int n_atoms, int n_structures,
                       const int32_t* mol_ptr_dev, const int32_t* atom_mol_dev,
                       const float* cells_dev, const uint8_t* pbc_dev, float cutoff,
                       int32_t* row_ptr_dev, int32_t* col_dev, int32_t* shift_dev,
                       int32_t edge_capacity, int32_t* counts_dev, void* workspace_dev) {
  if (n_atoms < 0 || n_structures < 0 || edge_capacity < 0 || !(cutoff > 0.f)) {
    set_error("mlff_neighbor_list: invalid sizes or cutoff");
    return MLFF_ERR_INVALID;
  }
  if (n_atoms > 0 && (!pos_dev || !mol_ptr_dev || !atom_mol_dev || !row_ptr_dev ||
                      !counts_dev || !workspace_dev || (edge_capacity > 0 && !col_dev))) {
    set_error("mlff_neighbor_list: null pointer");
    return MLFF_ERR_INVALID;
  }
  NLArgs a;
  memset(&a, 0, sizeof(a));
  int32_t* ws = static_cast<int32_t*>(workspace_dev);
  a.pos = pos_dev; a.N = n_atoms; a.B = n_structures;
  a.mol_ptr = mol_ptr_dev; a.atom_mol = atom_mol_dev;
  a.cells = cells_dev; a.pbc = pbc_dev; a.rc = cutoff;
  a.deg = ws; a.udeg = ws + n_atoms; a.pair_ptr = ws + 2 * n_atoms;
  a.row_ptr = row_ptr_dev; a.col = col_dev; a.edge_shift = shift_dev; a.ecap = edge_capacity;
  a.counts = counts_dev;
  if (n_atoms == 0) return MLFF_OK;
  hipError_t e = launch_neighbor_list(a, static_cast<hipStream_t>(stream), false);
  if (e != hipSuccess) return hip_fail(e, "mlff_neighbor_list");
  return MLFF_OK;
}

int mlff_neighbor_pairs(void* stream, const float* pos_dev, int n_atoms, int n_structures,
                        const int32_t* mol_ptr_dev, const int32_t* atom_mol_dev, float cutoff,
                        int32_t* row_ptr_dev, int32_t* col_dev, int32_t edge_capacity,
                        int32_t* edge_pair_dev, int32_t* pair_lo_dev, int32_t* pair_hi_dev,
                        float* geom_dev, int32_t pair_capacity, int32_t* counts_dev,
                        void* workspace_dev) {
  if (n_atoms < 0 || n_structures < 0 || edge_capacity < 0 || pair_capacity < 0 ||
      !(cutoff > 0.f)) {
    set_error("mlff_neighbor_pairs: invalid sizes or cutoff");
    return MLFF_ERR_INVALID;
  }
  if (n_atoms > 0 && (!pos_dev || !mol_ptr_dev || !atom_mol_dev || !row_ptr_dev ||
                      !counts_dev || !workspace_dev ||
                      (edge_capacity > 0 && (!col_dev || !edge_pair_dev)) ||
                      (pair_capacity > 0 && (!pair_lo_dev || !pair_hi_dev || !geom_dev)))) {
    set_error("mlff_neighbor_pairs: null pointer");
    return MLFF_ERR_INVALID;
  }
  if (reinterpret_cast<uintptr_t>(geom_dev) & 15) {
    set_error("mlff_neighbor_pairs: geom must be 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  NLArgs a;
  memset(&a, 0, sizeof(a));
  int32_t* ws = static_cast<int32_t*>(workspace_dev);
  a.pos = pos_dev; a.N = n_atoms; a.B = n_structures;
  a.mol_ptr = mol_ptr_dev; a.atom_mol = atom_mol_dev;
  a.rc = cutoff;
  a.deg = ws; a.udeg = ws + n_atoms; a.pair_ptr = ws + 2 * n_atoms;
  a.row_ptr = row_ptr_dev; a.col = col_dev; a.ecap = edge_capacity;
  a.edge_pair = edge_pair_dev; a.pair_lo = pair_lo_dev; a.pair_hi = pair_hi_dev;
  a.geom = reinterpret_cast<float4*>(geom_dev); a.pcap = pair_capacity;
  a.counts = counts_dev;
  if (n_atoms == 0) return MLFF_OK;
  hipError_t e = launch_neighbor_list(a, static_cast<hipStream_t>(stream), true);
  if (e != hipSuccess) return hip_fail(e, "mlff_neighbor_pairs");
  return MLFF_OK;
}

int mlff_plan_create(mlff_plan** out, const mlff_config* cfg) {
  if (!out || !cfg) { set_error("mlff_plan_create: null argument"); return MLFF_ERR_INVALID; }
  *out = nullptr;
  if (cfg->n_atoms < 1 || cfg->n_structures < 1 || cfg->edge_capacity < 0 ||
      cfg->pair_capacity < 0 || !(cfg->cutoff > 0.f)) {
    set_error("mlff_plan_create: invalid sizes");
    return MLFF_ERR_INVALID;
  }
  ModelW mw;
  int rc = build_layout(cfg->hidden_dim, cfg->num_interactions, cfg->num_rbf, cfg->max_z, &mw,
                        nullptr, 0, nullptr);
  if (rc != MLFF_OK) return rc;
  mlff_plan* pl = new (std::nothrow) mlff_plan;
  if (!pl) { set_error("out of host memory"); return MLFF_ERR_INVALID; }
  memset(pl, 0, sizeof(*pl));
  pl->cfg = *cfg;
  pl->mw = mw;
  carve(pl);
  pl->buf.blocks = persistent_grid(cfg->persistent_blocks, -1);
  set_default_shapes(pl);
  pl->n_streams = 2;
  pl->stream_dev = -1;
  *out = pl;
  return MLFF_OK;
}

int mlff_plan_destroy(mlff_plan* pl) {
  if (!pl) return MLFF_OK;
  if (pl->graph_exec) (void)hipGraphExecDestroy(pl->graph_exec);
  if (pl->graph) (void)hipGraphDestroy(pl->graph);
  if (pl->mts_exec) (void)hipGraphExecDestroy(pl->mts_exec);
  if (pl->mts_graph) (void)hipGraphDestroy(pl->mts_graph);
  release_streams(pl);
  delete pl;
  return MLFF_OK;
}

int64_t mlff_plan_workspace_bytes(const mlff_plan* pl) { return pl ? pl->ws_bytes : -1; }

}  // extern "C"

namespace {

// Point the plan's buffers and neighbour-list arguments at the workspace
// segments and the per-batch constants (no device work).
int bind_buffers(mlff_plan* pl, const char* who, void* ws, const float* weights_dev,
                 const int64_t* numbers_dev, const int32_t* mol_ptr_dev,
                 const int32_t* atom_mol_dev, const float* cells_dev, const uint8_t* pbc_dev) {
  if (!weights_dev || !numbers_dev || !mol_ptr_dev || !atom_mol_dev) {
    set_error("%s: null pointer", who);
    return MLFF_ERR_INVALID;
  }
  // the kernels read and write these with 16-byte vector accesses
  if ((reinterpret_cast<uintptr_t>(ws) | reinterpret_cast<uintptr_t>(weights_dev)) & 15) {
    set_error("%s: workspace and weights must be 16-byte aligned", who);
    return MLFF_ERR_INVALID;
  }
  if (pl->cfg.any_periodic && (!cells_dev || !pbc_dev)) {
    set_error("%s: periodic plan needs cells and pbc", who);
    return MLFF_ERR_INVALID;
  }
  Buffers& b = pl->buf;
  const mlff_config& c = pl->cfg;
  b.N = c.n_atoms; b.B = c.n_structures; b.H = c.hidden_dim; b.L = c.num_interactions;
  b.R = c.num_rbf; b.rc = c.cutoff; b.ecap = c.edge_capacity; b.pcap = c.pair_capacity;
  b.w = weights_dev; b.mw = &pl->mw; b.numbers = numbers_dev; b.mol_ptr = mol_ptr_dev;
  b.atom_mol = atom_mol_dev; b.cells = c.any_periodic ? cells_dev : nullptr;
  b.counts = segp<int32_t>(pl, ws, S_COUNTS);
  b.row_ptr = segp<int32_t>(pl, ws, S_ROWPTR);
  b.col = segp<int32_t>(pl, ws, S_COL);
  b.edge_pair = segp<int32_t>(pl, ws, S_EPAIR);
  b.pair_lo = segp<int32_t>(pl, ws, S_PLO);
  b.pair_hi = segp<int32_t>(pl, ws, S_PHI);
  b.pair_shift = segp<int32_t>(pl, ws, S_PSHIFT);
  b.geom = segp<float4>(pl, ws, S_GEOM);
  b.pacc = segp<float4>(pl, ws, S_PACC);
  b.gvec = segp<float4>(pl, ws, S_GVEC);
  b.Wf = segp<float>(pl, ws, S_WF);
  b.S = segp<float>(pl, ws, S_S);
  b.V = segp<float>(pl, ws, S_V);
  b.Smid = segp<float>(pl, ws, S_SMID);
  b.Vmid = segp<float>(pl, ws, S_VMID);
  b.Aupd = segp<float>(pl, ws, S_AUPD);
  b.G12 = segp<float>(pl, ws, S_G12);
  b.gS0 = segp<float>(pl, ws, S_GS0);
  b.gV0 = segp<float>(pl, ws, S_GV0);
  b.gS1 = segp<float>(pl, ws, S_GS1);
  b.gV1 = segp<float>(pl, ws, S_GV1);
  b.eps = segp<float>(pl, ws, S_EPS);
  b.vir_atom = segp<float>(pl, ws, S_VIR);
  b.wbf = segp<uint16_t>(pl, ws, S_WBF);
  NLArgs& a = pl->nl;
  memset(&a, 0, sizeof(a));
  a.N = b.N; a.B = b.B; a.mol_ptr = mol_ptr_dev; a.atom_mol = atom_mol_dev;
  a.cells = c.any_periodic ? cells_dev : nullptr; a.pbc = c.any_periodic ? pbc_dev : nullptr;
  a.rc = c.cutoff;
  a.deg = segp<int32_t>(pl, ws, S_DEG); a.udeg = segp<int32_t>(pl, ws, S_UDEG);
  a.row_ptr = segp<int32_t>(pl, ws, S_ROWPTR); a.pair_ptr = segp<int32_t>(pl, ws, S_PAIRPTR);
  a.col = segp<int32_t>(pl, ws, S_COL); a.edge_shift = nullptr;
  a.edge_pair = segp<int32_t>(pl, ws, S_EPAIR); a.ecap = b.ecap;
  a.pair_lo = segp<int32_t>(pl, ws, S_PLO); a.pair_hi = segp<int32_t>(pl, ws, S_PHI);
  a.pair_shift = segp<int32_t>(pl, ws, S_PSHIFT); a.geom = segp<float4>(pl, ws, S_GEOM);
  a.pcap = b.pcap; a.counts = b.counts;
  a.grid = nullptr;   // brute force until mlff_plan_set_cell_grid
  a.n_cells = 0;
  a.cell_count = segp<int32_t>(pl, ws, S_CELLCNT);
  a.cell_start = segp<int32_t>(pl, ws, S_CELLSTART);
  a.cell_atoms = segp<int32_t>(pl, ws, S_CELLATOMS);
  a.atom_cell = segp<int32_t>(pl, ws, S_ATOMCELL);
  a.edge_tmp_shift = segp<int32_t>(pl, ws, S_TMPSHIFT);
  pl->grid_dev = segp<int4>(pl, ws, S_GRID);
  return MLFF_OK;
}

}  // namespace

extern "C" {

int mlff_plan_bind(mlff_plan* pl, void* ws, int64_t ws_bytes, const float* weights_dev,
                   const int64_t* numbers_dev, const int32_t* mol_ptr_dev,
                   const int32_t* atom_mol_dev, const float* cells_dev, const uint8_t* pbc_dev) {
  if (!pl) { set_error("mlff_plan_bind: null plan"); return MLFF_ERR_INVALID; }
  if (!ws || ws_bytes < pl->ws_bytes) {
    set_error("mlff_plan_bind: workspace of %lld bytes, need %lld", (long long)ws_bytes,
              (long long)pl->ws_bytes);
    return MLFF_ERR_WORKSPACE;
  }
  int rc = bind_buffers(pl, "mlff_plan_bind", ws, weights_dev, numbers_dev, mol_ptr_dev,
                        atom_mol_dev, cells_dev, pbc_dev);
  if (rc != MLFF_OK) return rc;
  Buffers& b = pl->buf;
  hipError_t e = hipMemset(b.counts, 0, 4 * MLFF_COUNT_SLOTS);
  if (e != hipSuccess) return hip_fail(e, "mlff_plan_bind");
  e = acquire_streams(pl);
  if (e != hipSuccess) return hip_fail(e, "mlff_plan_bind: filter stream");
  e = split_filter_weights(b);
  if (e != hipSuccess) return hip_fail(e, "mlff_plan_bind: bf16 weight planes");
  pl->bound = true;
  pl->md_bound = false;
  pl->mts_bound = false;
  if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
  if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  return MLFF_OK;
}

int mlff_plan_set_weights(mlff_plan* pl, void* stream, const float* weights_dev) {
  if (!pl || !pl->bound) { set_error("mlff_plan_set_weights: plan not bound"); return MLFF_ERR_STATE; }
  if (!weights_dev || (reinterpret_cast<uintptr_t>(weights_dev) & 15)) {
    set_error("mlff_plan_set_weights: weights must be non-null and 16-byte aligned");
    return MLFF_ERR_INVALID;
  }
  pl->buf.w = weights_dev;
  // the bf16 planes follow the new weights in stream order on the caller's stream
  hipError_t e = split_filter_weights_async(pl->buf, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "mlff_plan_set_weights: bf16 weight planes");
  // captured graphs hold the old weights
  if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
  if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  return MLFF_OK;
}

int32_t* mlff_plan_counts(mlff_plan* pl) { return (pl && pl->bound) ? pl->buf.counts : nullptr; }

static int run_ef(mlff_plan* pl, hipStream_t s, const float* pos, double* energy, float* forces,
                  double* virial, float* atom_energy) {
  pl->nl.pos = pos;
  hipError_t e = launch_neighbor_list(pl->nl, s, true);
  if (e != hipSuccess) return hip_fail(e, "neighbour list");
  e = launch_forward(pl->buf, s);
  if (e != hipSuccess) return hip_fail(e, "forward");
  if (forces != nullptr || virial != nullptr) {
    e = launch_backward(pl->buf, s, forces, virial != nullptr);
    if (e != hipSuccess) return hip_fail(e, "adjoint");
  }
  e = launch_reduce(pl->buf, s, energy, virial, atom_energy);
  if (e != hipSuccess) return hip_fail(e, "reduce");
  return MLFF_OK;
}

int mlff_energy_forces(mlff_plan* pl, void* stream, const float* pos_dev, double* energy_dev,
                       float* forces_dev, double* virial_dev, float* atom_energy_dev) {
  if (!pl || !pl->bound) { set_error("mlff_energy_forces: plan not bound"); return MLFF_ERR_STATE; }
  if (!pos_dev) { set_error("mlff_energy_forces: null positions"); return MLFF_ERR_INVALID; }
  return run_ef(pl, static_cast<hipStream_t>(stream), pos_dev, energy_dev, forces_dev, virial_dev,
                atom_energy_dev);
}

int mlff_launch_stage(mlff_plan* pl, void* stream, int stage, int layer) {
  if (!pl || !pl->bound) { set_error("mlff_launch_stage: plan not bound"); return MLFF_ERR_STATE; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (stage == MLFF_STAGE_NEIGHBOR) {
    if (!pl->nl.pos) { set_error("mlff_launch_stage: no evaluation yet"); return MLFF_ERR_STATE; }
    e = launch_neighbor_list(pl->nl, s, true);
  } else {
    e = launch_stage(pl->buf, s, stage, layer);
  }
  if (e != hipSuccess) return hip_fail(e, "mlff_launch_stage");
  return MLFF_OK;
}

int mlff_plan_set_cell_grid(mlff_plan* pl, const int32_t* grid_host, int n_structures) {
  if (!pl || !pl->bound) { set_error("mlff_plan_set_cell_grid: plan not bound"); return MLFF_ERR_STATE; }
  if (!grid_host || n_structures != pl->cfg.n_structures) {
    set_error("mlff_plan_set_cell_grid: need one grid entry per structure");
    return MLFF_ERR_INVALID;
  }
  if (pl->cfg.cell_capacity <= 0 || !pl->cfg.any_periodic) {
    set_error("mlff_plan_set_cell_grid: plan has no cell capacity or no periodic structure");
    return MLFF_ERR_STATE;
  }
  int used = 0;
  for (int b = 0; b < n_structures; ++b) {
    const int32_t* g = grid_host + 4 * b;
    if (g[0] == 0) continue;
    if (g[0] < 3 || g[1] < 3 || g[2] < 3 || g[3] < 0) {
      set_error("mlff_plan_set_cell_grid: structure %d needs >= 3 cells per axis", b);
      return MLFF_ERR_INVALID;
    }
    const int64_t end = (int64_t)g[3] + (int64_t)g[0] * g[1] * g[2];
    if (end > pl->cfg.cell_capacity) {
      set_error("mlff_plan_set_cell_grid: %lld cells exceed the capacity %d", (long long)end,
                pl->cfg.cell_capacity);
      return MLFF_ERR_INVALID;
    }
    used = (int)(end > used ? end : used);
  }
  hipError_t e = hipMemcpy(pl->grid_dev, grid_host, 16 * (size_t)n_structures,
                           hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "mlff_plan_set_cell_grid");
  pl->nl.grid = used > 0 ? pl->grid_dev : nullptr;
  pl->nl.n_cells = used;
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  return MLFF_OK;
}

int mlff_plan_set_streams(mlff_plan* pl, int n_streams) {
  if (!pl || n_streams < 1 || n_streams > 2) {
    set_error("mlff_plan_set_streams: 1 or 2 streams");
    return MLFF_ERR_INVALID;
  }
  pl->n_streams = n_streams;
  if (pl->bound) {
    if (n_streams < 2) release_streams(pl);
    hipError_t e = acquire_streams(pl);
    if (e != hipSuccess) return hip_fail(e, "mlff_plan_set_streams");
  }
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
  if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
  return MLFF_OK;
}

int mlff_plan_set_blocks(mlff_plan* pl, int blocks) {
  if (!pl || blocks < 1) { set_error("mlff_plan_set_blocks: invalid"); return MLFF_ERR_INVALID; }
  pl->buf.blocks = blocks;
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
  if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
  return MLFF_OK;
}

int mlff_plan_set_tiles(mlff_plan* pl, int fwd_pairs, int bwd_pairs, int msg_unroll_fwd,
                        int msg_unroll_bwd, int bwd_variant) {
  if (!pl) { set_error("mlff_plan_set_tiles: null plan"); return MLFF_ERR_INVALID; }
  const int H = pl->cfg.hidden_dim;
  const int want[4] = {fwd_pairs, bwd_pairs, msg_unroll_fwd, msg_unroll_bwd};
  int got[4];
  for (int k = 0; k < 4; ++k) {
    const int d = default_shape(H, k), a = alternative_shape(H, k);
    got[k] = want[k] == 0 ? d : want[k];
    if (got[k] != d && got[k] != a) {
      set_error("mlff_plan_set_tiles: shape %d of kind %d not instantiated for hidden_dim %d "
                "(choices %d, %d)", want[k], k, H, d, a);
      return MLFF_ERR_UNSUPPORTED;
    }
  }
  if (bwd_variant < 0 || bwd_variant > alternative_shape(H, 4)) {
    set_error("mlff_plan_set_tiles: bwd_variant must be 0..%d", alternative_shape(H, 4));
    return MLFF_ERR_INVALID;
  }
  if (bwd_variant && got[1] != default_shape(H, 1)) {
    set_error("mlff_plan_set_tiles: filter-adjoint variants exist for the default adjoint tile "
              "(%d pairs) only", default_shape(H, 1));
    return MLFF_ERR_UNSUPPORTED;
  }
  pl->buf.tile_fwd = got[0];
  pl->buf.tile_bwd = got[1];
  pl->buf.unroll_fwd = got[2];
  pl->buf.unroll_bwd = got[3];
  pl->buf.bwd_variant = bwd_variant;
  // captured graphs hold the old launches
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
  if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
  return MLFF_OK;
}

int mlff_md_bind(mlff_plan* pl, double* x_dev, double* p_dev, const double* inv_mass_dev,
                 float* pos32_dev, float* forces_dev, double* energy_dev, double* ledger_dev,
                 int32_t ledger_len, double dt) {
  if (!pl || !pl->bound) { set_error("mlff_md_bind: plan not bound"); return MLFF_ERR_STATE; }
  if (!x_dev || !p_dev || !inv_mass_dev || !pos32_dev || !forces_dev || !energy_dev ||
      ledger_len < 0 || (ledger_len > 0 && !ledger_dev)) {
    set_error("mlff_md_bind: invalid argument");
    return MLFF_ERR_INVALID;
  }
  pl->x = x_dev; pl->p = p_dev; pl->inv_mass = inv_mass_dev; pl->pos32 = pos32_dev;
  pl->forces = forces_dev; pl->energy = energy_dev; pl->ledger = ledger_dev;
  pl->ledger_len = ledger_len; pl->dt = dt;
  pl->md_bound = true;
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  return MLFF_OK;
}

int mlff_md_bind_frames(mlff_plan* pl, double* x_frames_dev, double* p_frames_dev,
                        float* f_frames_dev, int32_t n_frames) {
  if (!pl || !pl->md_bound) { set_error("mlff_md_bind_frames: MD not bound"); return MLFF_ERR_STATE; }
  if (n_frames < 0 || (n_frames > 0 && (!x_frames_dev || !p_frames_dev || !f_frames_dev))) {
    set_error("mlff_md_bind_frames: invalid argument");
    return MLFF_ERR_INVALID;
  }
  pl->frame_x = n_frames > 0 ? x_frames_dev : nullptr;
  pl->frame_p = n_frames > 0 ? p_frames_dev : nullptr;
  pl->frame_f = n_frames > 0 ? f_frames_dev : nullptr;
  pl->n_frames = n_frames;
  if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
  if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
  return MLFF_OK;
}

int mlff_md_init_forces(mlff_plan* pl, void* stream) {
  if (!pl || !pl->md_bound) { set_error("mlff_md_init_forces: MD not bound"); return MLFF_ERR_STATE; }
  return run_ef(pl, static_cast<hipStream_t>(stream), pl->pos32, pl->energy, pl->forces, nullptr,
                nullptr);
}

static int md_step(mlff_plan* pl, hipStream_t s) {
  const Buffers& b = pl->buf;
  hipError_t e = launch_vv_first(b.N, pl->x, pl->p, pl->inv_mass, pl->forces, pl->pos32, pl->dt, s);
  if (e != hipSuccess) return hip_fail(e, "vv first half");
  int rc = run_ef(pl, s, pl->pos32, pl->energy, pl->forces, nullptr, nullptr);
  if (rc != MLFF_OK) return rc;
  e = launch_vv_second(b.N, b.B, b.mol_ptr, pl->p, pl->inv_mass, pl->forces, pl->dt, pl->energy,
                       pl->ledger, pl->ledger_len, b.counts, pl->x, pl->frame_x, pl->frame_p,
                       pl->frame_f, pl->n_frames, s);
  if (e != hipSuccess) return hip_fail(e, "vv second half");
  return MLFF_OK;
}

int mlff_md_run(mlff_plan* pl, void* stream, int n_steps) {
  if (!pl || !pl->md_bound) { set_error("mlff_md_run: MD not bound"); return MLFF_ERR_STATE; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int k = 0; k < n_steps; ++k) {
    int rc = md_step(pl, s);
    if (rc != MLFF_OK) return rc;
  }
  return MLFF_OK;
}

int mlff_md_graph_run(mlff_plan* pl, void* stream, int steps_per_graph, int n_launches) {
  if (!pl || !pl->md_bound) { set_error("mlff_md_graph_run: MD not bound"); return MLFF_ERR_STATE; }
  if (steps_per_graph < 1 || n_launches < 0) {
    set_error("mlff_md_graph_run: invalid counts");
    return MLFF_ERR_INVALID;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (s == nullptr) {
    set_error("mlff_md_graph_run: graph capture needs a non-default stream");
    return MLFF_ERR_INVALID;
  }
  if (!pl->graph_exec || pl->graph_steps != steps_per_graph) {
    if (pl->graph_exec) { (void)hipGraphExecDestroy(pl->graph_exec); pl->graph_exec = nullptr; }
    if (pl->graph) { (void)hipGraphDestroy(pl->graph); pl->graph = nullptr; }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int rc = MLFF_OK;
    for (int k = 0; k < steps_per_graph && rc == MLFF_OK; ++k) rc = md_step(pl, s);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc != MLFF_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) { (void)hipGraphDestroy(g); return hip_fail(e, "hipGraphInstantiate"); }
    pl->graph = g;
    pl->graph_exec = ge;
    pl->graph_steps = steps_per_graph;
  }
  for (int k = 0; k < n_launches; ++k) {
    hipError_t e = hipGraphLaunch(pl->graph_exec, s);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  }
  return MLFF_OK;
}

// Model of `pl` on the neighbour list and pair geometry of `src` (same
// positions, same cutoff, equal capacities).
static int run_ef_on(mlff_plan* pl, hipStream_t s, double* energy, float* forces,
                     const Buffers& src) {
  Buffers b = pl->buf;
  b.counts = src.counts;
  b.row_ptr = src.row_ptr; b.col = src.col; b.edge_pair = src.edge_pair;
  b.pair_lo = src.pair_lo; b.pair_hi = src.pair_hi; b.pair_shift = src.pair_shift;
  b.geom = src.geom;
  hipError_t e = launch_forward(b, s);
  if (e != hipSuccess) return hip_fail(e, "outer forward");
  e = launch_backward(b, s, forces, false);
  if (e != hipSuccess) return hip_fail(e, "outer adjoint");
  e = launch_reduce(b, s, energy, nullptr, nullptr);
  if (e != hipSuccess) return hip_fail(e, "outer reduce");
  return MLFF_OK;
}

int mlff_mts_bind(mlff_plan* pl, mlff_plan* in, double* x_dev, double* p_dev,
                  const double* inv_mass_dev, float* pos32_dev, float* f_outer_dev,
                  float* f_inner_dev, float* f_mix_dev, double* e_outer_dev,
                  double* e_inner_dev, const uint8_t* flag_dev, float* dis_max_dev,
                  double* ledger_dev, int32_t ledger_len, double dt, int32_t n_inner) {
  if (!pl || !in || !pl->bound || !in->bound) {
    set_error("mlff_mts_bind: both plans must be bound");
    return MLFF_ERR_STATE;
  }
  const mlff_config &a = pl->cfg, &b = in->cfg;
  if (a.n_atoms != b.n_atoms || a.n_structures != b.n_structures || a.cutoff != b.cutoff ||
      a.edge_capacity != b.edge_capacity || a.pair_capacity != b.pair_capacity ||
      a.any_periodic != b.any_periodic) {
    set_error("mlff_mts_bind: plans differ in batch, cutoff or capacity");
    return MLFF_ERR_INVALID;
  }
  if (!x_dev || !p_dev || !inv_mass_dev || !pos32_dev || !f_outer_dev || !f_inner_dev ||
      !f_mix_dev || !e_outer_dev || !e_inner_dev || !flag_dev || !dis_max_dev || n_inner < 1 ||
      ledger_len < 0 || (ledger_len > 0 && !ledger_dev)) {
    set_error("mlff_mts_bind: invalid argument");
    return MLFF_ERR_INVALID;
  }
  pl->x = x_dev; pl->p = p_dev; pl->inv_mass = inv_mass_dev; pl->pos32 = pos32_dev;
  pl->ledger = ledger_dev; pl->ledger_len = ledger_len; pl->dt = dt;
  pl->mts_inner = in;
  pl->f_outer = f_outer_dev; pl->f_inner = f_inner_dev; pl->f_mix = f_mix_dev;
  pl->e_outer = e_outer_dev; pl->e_inner = e_inner_dev; pl->flag = flag_dev;
  pl->dis_max = dis_max_dev; pl->n_inner = n_inner;
  pl->mts_bound = true;
  if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
  if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
  return MLFF_OK;
}

int mlff_mts_init_forces(mlff_plan* pl, void* stream) {
  if (!pl || !pl->mts_bound) { set_error("mlff_mts_init_forces: MTS not bound"); return MLFF_ERR_STATE; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  mlff_plan* in = pl->mts_inner;
  int rc = run_ef(in, s, pl->pos32, pl->e_inner, pl->f_inner, nullptr, nullptr);
  if (rc != MLFF_OK) return rc;
  return run_ef_on(pl, s, pl->e_outer, pl->f_outer, in->buf);
}

static int mts_outer_step(mlff_plan* pl, hipStream_t s, bool fallback) {
  mlff_plan* in = pl->mts_inner;
  const int N = pl->buf.N;
  const double dt = pl->dt, Dt = pl->n_inner * dt;
  const float* Fin = fallback ? pl->f_mix : pl->f_inner;
  hipError_t e = launch_mts_kick(N, pl->p, pl->f_outer, Fin, 0.5 * Dt, s);
  if (e != hipSuccess) return hip_fail(e, "mts outer kick");
  for (int k = 0; k < pl->n_inner; ++k) {
    e = launch_vv_first(N, pl->x, pl->p, pl->inv_mass, Fin, pl->pos32, dt, s);
    if (e != hipSuccess) return hip_fail(e, "mts drift");
    int rc = run_ef(in, s, pl->pos32, pl->e_inner, pl->f_inner, nullptr, nullptr);
    if (rc != MLFF_OK) return rc;
    if (fallback) {
      rc = run_ef_on(pl, s, pl->e_outer, pl->f_outer, in->buf);
      if (rc != MLFF_OK) return rc;
      e = launch_mts_select(N, pl->buf.atom_mol, pl->flag, pl->f_outer, pl->f_inner, pl->f_mix, s);
      if (e != hipSuccess) return hip_fail(e, "mts select");
    }
    e = launch_mts_kick(N, pl->p, Fin, nullptr, 0.5 * dt, s);
    if (e != hipSuccess) return hip_fail(e, "mts inner kick");
  }
  if (!fallback) {
    int rc = run_ef_on(pl, s, pl->e_outer, pl->f_outer, in->buf);
    if (rc != MLFF_OK) return rc;
  }
  e = launch_mts_kick(N, pl->p, pl->f_outer, Fin, 0.5 * Dt, s);
  if (e != hipSuccess) return hip_fail(e, "mts outer kick");
  e = launch_mts_disagree(N, pl->buf.atom_mol, pl->f_outer, pl->f_inner, pl->dis_max, s);
  if (e != hipSuccess) return hip_fail(e, "mts disagreement");
  e = launch_mts_ledger(pl->buf.B, pl->buf.mol_ptr, pl->p, pl->inv_mass, pl->e_outer, pl->ledger,
                        pl->ledger_len, pl->buf.counts, s);
  if (e != hipSuccess) return hip_fail(e, "mts ledger");
  return MLFF_OK;
}

int mlff_mts_run(mlff_plan* pl, void* stream, int n_outer, int fallback) {
  if (!pl || !pl->mts_bound) { set_error("mlff_mts_run: MTS not bound"); return MLFF_ERR_STATE; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fallback) {
    hipError_t e = launch_mts_select(pl->buf.N, pl->buf.atom_mol, pl->flag, pl->f_outer,
                                     pl->f_inner, pl->f_mix, s);
    if (e != hipSuccess) return hip_fail(e, "mts select");
  }
  for (int k = 0; k < n_outer; ++k) {
    int rc = mts_outer_step(pl, s, fallback != 0);
    if (rc != MLFF_OK) return rc;
  }
  return MLFF_OK;
}

int mlff_mts_graph_run(mlff_plan* pl, void* stream, int outer_per_graph, int n_launches) {
  if (!pl || !pl->mts_bound) { set_error("mlff_mts_graph_run: MTS not bound"); return MLFF_ERR_STATE; }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (s == nullptr || outer_per_graph < 1 || n_launches < 0) {
    set_error("mlff_mts_graph_run: needs a non-default stream and positive counts");
    return MLFF_ERR_INVALID;
  }
  if (!pl->mts_exec || pl->mts_graph_steps != outer_per_graph) {
    if (pl->mts_exec) { (void)hipGraphExecDestroy(pl->mts_exec); pl->mts_exec = nullptr; }
    if (pl->mts_graph) { (void)hipGraphDestroy(pl->mts_graph); pl->mts_graph = nullptr; }
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    int rc = MLFF_OK;
    for (int k = 0; k < outer_per_graph && rc == MLFF_OK; ++k) rc = mts_outer_step(pl, s, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    if (rc != MLFF_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) { (void)hipGraphDestroy(g); return hip_fail(e, "hipGraphInstantiate"); }
    pl->mts_graph = g;
    pl->mts_exec = ge;
    pl->mts_graph_steps = outer_per_graph;
  }
  for (int k = 0; k < n_launches; ++k) {
    hipError_t e = hipGraphLaunch(pl->mts_exec, s);
    if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  }
  return MLFF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Evaluation without device allocation: the workspace is a file-scope device
// array, one copy per GPU, held until the process ends.
// ---------------------------------------------------------------------------

namespace {

static __device__ __align__(256) unsigned char scratch[MLFF_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

// Per device: the device address of `scratch` (the host-side name is only a
// stand-in and must never reach a kernel) and the CU count.
struct StaticScratch {
  std::mutex mu;
  void* ptr[kMaxDevices] = {};
  int cus[kMaxDevices] = {};
};
StaticScratch g_static;

int static_scratch(int* dev_out, void** ptr_out, int* cus_out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "mlff_energy_forces_static: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("mlff_energy_forces_static: device %d beyond the %d this library tracks", dev,
              kMaxDevices);
    return MLFF_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_static.mu);
  if (!g_static.ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "mlff_energy_forces_static: hipGetSymbolAddress");
    if (!p) {
      set_error("mlff_energy_forces_static: no device address for the scratch array");
      return MLFF_ERR_HIP;
    }
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
      cus = prop.multiProcessorCount;
    g_static.cus[dev] = cus;
    g_static.ptr[dev] = p;
  }
  *dev_out = dev;
  *ptr_out = g_static.ptr[dev];
  *cus_out = g_static.cus[dev];
  return MLFF_OK;
}

__global__ void fill_i32_kernel(int32_t* __restrict__ p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

extern "C" int mlff_energy_forces_static(void* stream, const mlff_config* cfg,
                                         const float* weights_dev, const int64_t* numbers_dev,
                                         const int32_t* mol_ptr_dev, const int32_t* atom_mol_dev,
                                         const float* cells_dev, const uint8_t* pbc_dev,
                                         const float* pos_dev, double* energy_dev,
                                         float* forces_dev, double* virial_dev,
                                         float* atom_energy_dev, int32_t* counts_dev) {
  static const char* who = "mlff_energy_forces_static";
  if (!cfg || !pos_dev || !energy_dev || !counts_dev) {
    set_error("%s: null argument", who);
    return MLFF_ERR_INVALID;
  }
  if (cfg->n_atoms < 1 || cfg->n_structures < 1 || cfg->edge_capacity < 0 ||
      cfg->pair_capacity < 0 || !(cfg->cutoff > 0.f)) {
    set_error("%s: invalid sizes", who);
    return MLFF_ERR_INVALID;
  }
  if (reinterpret_cast<uintptr_t>(counts_dev) & 3) {
    set_error("%s: counts must be 4-byte aligned", who);
    return MLFF_ERR_INVALID;
  }
  // the plan lives on this call's stack: nothing to allocate, nothing kept
  mlff_plan pl;
  memset(&pl, 0, sizeof(pl));
  pl.cfg = *cfg;
  pl.cfg.cell_capacity = 0;       // periodic structures use the brute-force image search
  int rc = build_layout(cfg->hidden_dim, cfg->num_interactions, cfg->num_rbf, cfg->max_z,
                        &pl.mw, nullptr, 0, nullptr);
  if (rc != MLFF_OK) return rc;
  carve(&pl);
  if (pl.ws_bytes > (int64_t)MLFF_STATIC_SCRATCH_BYTES) {
    set_error("%s: this batch needs %lld bytes of scratch; the static scratch holds %lld "
              "(use mlff_energy_forces with a workspace)", who, (long long)pl.ws_bytes,
              (long long)MLFF_STATIC_SCRATCH_BYTES);
    return MLFF_ERR_WORKSPACE;
  }
  int dev = 0, cus = 0;
  void* ws = nullptr;
  rc = static_scratch(&dev, &ws, &cus);
  if (rc != MLFF_OK) return rc;
  pl.buf.blocks = persistent_grid(cfg->persistent_blocks, cus);
  set_default_shapes(&pl);
  pl.n_streams = 1;
  pl.stream_dev = -1;
  rc = bind_buffers(&pl, who, ws, weights_dev, numbers_dev, mol_ptr_dev, atom_mol_dev, cells_dev,
                    pbc_dev);
  if (rc != MLFF_OK) return rc;
  // the caller's counts replace the scratch's slots, cleared by a kernel (not
  // a memset node, so the call can be captured into a HIP graph)
  pl.buf.counts = counts_dev;
  pl.nl.counts = counts_dev;
  pl.buf.aux = nullptr;
  pl.buf.ev = pl.ev;
  hipStream_t s = static_cast<hipStream_t>(stream);
  fill_i32_kernel<<<1, 64, 0, s>>>(counts_dev, MLFF_COUNT_SLOTS, 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, who);
  e = split_filter_weights_async(pl.buf, s);
  if (e != hipSuccess) return hip_fail(e, "mlff_energy_forces_static: bf16 weight planes");
  return run_ef(&pl, s, pos_dev, energy_dev, forces_dev, virial_dev, atom_energy_dev);
}
