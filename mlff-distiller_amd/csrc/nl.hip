// Radius-graph neighbour list on gfx950.
//
// Replaces radius_graph_native (student_model.py:63-109): every ordered pair
// i != j inside one structure with ||p_j - p_i|| <= rc.  The distance is
// rounded exactly like torch.norm on the CPU (norm3_exact), so the edge set
// matches the reference bit for bit, including pairs a few ulp from rc.
//
// Passes (all deterministic, no atomics):
//   count : per atom, number of neighbours (deg) and of neighbours j > i (udeg)
//   scan  : exclusive prefix sums -> row_ptr (edges) and pair_ptr (pairs), per
//           1024-atom slice, then the offsets of the slices before
//   fill  : per atom, neighbours in ascending j -> col; for j > i the pair
//           (lo=i, hi=j) gets its index, lattice shift and geometry
//           {p_hi - p_lo (+ shift @ cell), d}
//   link  : per atom (one wave), edges with j < i find their pair in row j
//           by binary search among row j's sorted upper neighbours.
// Every write is guarded by the edge / pair capacity; a run that exceeds a
// capacity sets counts[MLFF_COUNT_OVERFLOW] and never writes out of bounds.
//
// Periodic structures use the minimum image: the fractional difference is
// rounded to the nearest lattice vector along periodic axes (orthorhombic
// cells need nothing more when rc is below half the box); for skewed cells
// the 27 neighbouring images of that vector are searched for the shortest.

#include "common.hpp"
#include "internal.hpp"

namespace mlff {

namespace {

struct CellInfo {
  float c[9];     // rows = lattice vectors
  float inv[9];   // inverse (row-vector convention: frac = d @ inv)
  int per[3];
  int ortho;
};

__device__ __forceinline__ void load_cell(const float* cells, const uint8_t* pbc, int b,
                                          CellInfo& ci, bool& periodic) {
  periodic = false;
  // an isolated structure's shift vectors (always zero shifts) must read a
  // defined cell: zeros
#pragma unroll
  for (int k = 0; k < 9; ++k) ci.c[k] = 0.f;
  if (cells == nullptr || pbc == nullptr) return;
  ci.per[0] = pbc[3 * b]; ci.per[1] = pbc[3 * b + 1]; ci.per[2] = pbc[3 * b + 2];
  periodic = ci.per[0] | ci.per[1] | ci.per[2];
  if (!periodic) return;
  const float* c = cells + 9 * b;
#pragma unroll
  for (int k = 0; k < 9; ++k) ci.c[k] = c[k];
  // inverse via adjugate, in double for a stable inverse
  double a = c[0], bb = c[1], cc = c[2], d = c[3], e = c[4], f = c[5], g = c[6], h = c[7], k = c[8];
  double det = a * (e * k - f * h) - bb * (d * k - f * g) + cc * (d * h - e * g);
  double id = 1.0 / det;
  ci.inv[0] = (float)((e * k - f * h) * id); ci.inv[1] = (float)((cc * h - bb * k) * id);
  ci.inv[2] = (float)((bb * f - cc * e) * id); ci.inv[3] = (float)((f * g - d * k) * id);
  ci.inv[4] = (float)((a * k - cc * g) * id); ci.inv[5] = (float)((cc * d - a * f) * id);
  ci.inv[6] = (float)((d * h - e * g) * id); ci.inv[7] = (float)((bb * g - a * h) * id);
  ci.inv[8] = (float)((a * e - bb * d) * id);
  ci.ortho = (c[1] == 0.f && c[2] == 0.f && c[3] == 0.f && c[5] == 0.f && c[6] == 0.f && c[7] == 0.f);
}

// shift vector t @ cell in a fixed operation order (the geometry of a pair
// and the neighbour test must round identically)
__device__ __forceinline__ void shift_vec(const CellInfo& ci, int t0, int t1, int t2, float& sx,
                                          float& sy, float& sz) {
#pragma clang fp contract(off)
  float f0 = (float)t0, f1 = (float)t1, f2 = (float)t2;
  sx = __builtin_fmaf(f2, ci.c[6], __builtin_fmaf(f1, ci.c[3], f0 * ci.c[0]));
  sy = __builtin_fmaf(f2, ci.c[7], __builtin_fmaf(f1, ci.c[4], f0 * ci.c[1]));
  sz = __builtin_fmaf(f2, ci.c[8], __builtin_fmaf(f1, ci.c[5], f0 * ci.c[2]));
}

// Edge test for (i -> j): vector v = p_j - p_i (+ image shift), distance d.
__device__ __forceinline__ bool within(float dx, float dy, float dz, float rc, bool periodic,
                                       const CellInfo& ci, float& vx, float& vy, float& vz,
                                       float& d, int& s0, int& s1, int& s2) {
#pragma clang fp contract(off)
  s0 = s1 = s2 = 0;
  if (!periodic) {
    vx = dx; vy = dy; vz = dz;
    d = norm3_exact(dx, dy, dz);
    return d <= rc;
  }
  float fr0 = __builtin_fmaf(dz, ci.inv[6], __builtin_fmaf(dy, ci.inv[3], dx * ci.inv[0]));
  float fr1 = __builtin_fmaf(dz, ci.inv[7], __builtin_fmaf(dy, ci.inv[4], dx * ci.inv[1]));
  float fr2 = __builtin_fmaf(dz, ci.inv[8], __builtin_fmaf(dy, ci.inv[5], dx * ci.inv[2]));
  int n0 = ci.per[0] ? -(int)__builtin_rintf(fr0) : 0;
  int n1 = ci.per[1] ? -(int)__builtin_rintf(fr1) : 0;
  int n2 = ci.per[2] ? -(int)__builtin_rintf(fr2) : 0;
  if (ci.ortho) {
    float sx, sy, sz;
    shift_vec(ci, n0, n1, n2, sx, sy, sz);
    vx = dx + sx; vy = dy + sy; vz = dz + sz;
    d = norm3_exact(vx, vy, vz);
    s0 = n0; s1 = n1; s2 = n2;
    return d <= rc;
  }
  float best = 3.4e38f;
  for (int a = -ci.per[0]; a <= ci.per[0]; ++a)
    for (int b = -ci.per[1]; b <= ci.per[1]; ++b)
      for (int c = -ci.per[2]; c <= ci.per[2]; ++c) {
        float sx, sy, sz;
        shift_vec(ci, n0 + a, n1 + b, n2 + c, sx, sy, sz);
        float wx = dx + sx, wy = dy + sy, wz = dz + sz;
        float dd = norm3_exact(wx, wy, wz);
        if (dd < best) {
          best = dd; vx = wx; vy = wy; vz = wz; s0 = n0 + a; s1 = n1 + b; s2 = n2 + c;
        }
      }
  d = best;
  return d <= rc;
}

// ---------------------------------------------------------------------------
// cell list (periodic structures with at least 3 cells of width >= rc per axis)
// ---------------------------------------------------------------------------

struct CellList {
  const int4* grid;          // per structure {nc0, nc1, nc2, first cell}; nc0 == 0: brute force
  const int32_t* cell_start; // [C+1]
  const int32_t* cell_atoms; // atoms grouped by cell, ascending within a cell
  const int32_t* atom_cell;  // [N], -1 for brute-force atoms
};

// Visit every atom of the 27 cells around atom i's cell (distinct because
// every axis has at least 3 cells).
template <class Fn>
__device__ __forceinline__ void for_each_cell_candidate(const CellList& cl, int i, int b, Fn fn) {
  const int4 g = cl.grid[b];
  const int lc = cl.atom_cell[i] - g.w;
  const int c2 = lc % g.z, c1 = (lc / g.z) % g.y, c0 = lc / (g.y * g.z);
  for (int da = -1; da <= 1; ++da) {
    const int n0 = (c0 + da + g.x) % g.x;
    for (int db = -1; db <= 1; ++db) {
      const int n1 = (c1 + db + g.y) % g.y;
      for (int dc = -1; dc <= 1; ++dc) {
        const int n2 = (c2 + dc + g.z) % g.z;
        const int cell = g.w + (n0 * g.y + n1) * g.z + n2;
        for (int k = cl.cell_start[cell]; k < cl.cell_start[cell + 1]; ++k) fn(cl.cell_atoms[k]);
      }
    }
  }
}

__device__ __forceinline__ int wave_inclusive_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// The same 27 cells, walked by a whole wave: lane q < 27 holds cell q's first
// slot in cell_atoms and the number of candidates before it; candidate f of
// the flattened list is found by a wave-uniform pass over the 27 cells.
struct CellWalk {
  int start, off, total;
};

__device__ __forceinline__ CellWalk cell_walk(const CellList& cl, int i, int b) {
  const int lane = threadIdx.x & 63;
  const int4 g = cl.grid[b];
  int start = 0, cnt = 0;
  if (lane < 27) {
    const int lc = cl.atom_cell[i] - g.w;
    const int c2 = lc % g.z, c1 = (lc / g.z) % g.y, c0 = lc / (g.y * g.z);
    const int n0 = (c0 + lane / 9 - 1 + g.x) % g.x;
    const int n1 = (c1 + (lane / 3) % 3 - 1 + g.y) % g.y;
    const int n2 = (c2 + lane % 3 - 1 + g.z) % g.z;
    const int cell = g.w + (n0 * g.y + n1) * g.z + n2;
    start = cl.cell_start[cell];
    cnt = cl.cell_start[cell + 1] - start;
  }
  const int incl = wave_inclusive_scan(cnt);
  return CellWalk{start, incl - cnt, __shfl(incl, 63, 64)};
}

// atom index of candidate f (0 <= f < total); every lane of the wave calls it
__device__ __forceinline__ int cell_candidate(const CellList& cl, const CellWalk& w, int f) {
  int slot = 0;
#pragma unroll
  for (int q = 0; q < 27; ++q) {
    const int oq = __shfl(w.off, q, 64), sq = __shfl(w.start, q, 64);
    if (f >= oq) slot = sq + (f - oq);
  }
  return f < w.total ? cl.cell_atoms[slot] : -1;
}

// cell-list rows of up to this many neighbours are sorted in LDS by a wave;
// longer rows take the one-lane path
constexpr int kCellRowMax = 256;
constexpr int kNlBlock = 256;

__global__ void cl_zero_kernel(int n, int32_t* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// cell of every cell-list atom from its wrapped fractional coordinates
__global__ void cl_bin_kernel(const float* __restrict__ pos, int N,
                              const int32_t* __restrict__ atom_mol,
                              const float* __restrict__ cells, const uint8_t* __restrict__ pbc,
                              const int4* __restrict__ grid, int32_t* __restrict__ atom_cell,
                              int32_t* __restrict__ cell_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int b = atom_mol[i];
  const int4 g = grid[b];
  if (g.x == 0) { atom_cell[i] = -1; return; }
  CellInfo ci;
  bool periodic;
  load_cell(cells, pbc, b, ci, periodic);
  const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
  float f[3];
  f[0] = __builtin_fmaf(z, ci.inv[6], __builtin_fmaf(y, ci.inv[3], x * ci.inv[0]));
  f[1] = __builtin_fmaf(z, ci.inv[7], __builtin_fmaf(y, ci.inv[4], x * ci.inv[1]));
  f[2] = __builtin_fmaf(z, ci.inv[8], __builtin_fmaf(y, ci.inv[5], x * ci.inv[2]));
  const int nc[3] = {g.x, g.y, g.z};
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float w = f[a] - floorf(f[a]);   // [0, 1]
    int k = (int)(w * (float)nc[a]);
    c[a] = k < 0 ? 0 : (k >= nc[a] ? nc[a] - 1 : k);
  }
  const int cell = g.w + (c[0] * g.y + c[1]) * g.z + c[2];
  atom_cell[i] = cell;
  atomicAdd(&cell_count[cell], 1);
}

// exclusive scan of the cell counts (one workgroup), then reset the counts
// so they serve as insertion cursors
__global__ __launch_bounds__(1024) void cl_scan_kernel(int32_t* __restrict__ cell_count, int C,
                                                       int32_t* __restrict__ cell_start) {
  __shared__ int s[1024];
  const int t = threadIdx.x;
  const int per = (C + 1023) / 1024;
  const int lo = min(C, t * per), hi = min(C, lo + per);
  int a = 0;
  for (int k = lo; k < hi; ++k) a += cell_count[k];
  s[t] = a;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = (t >= off) ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int r = s[t] - a;
  for (int k = lo; k < hi; ++k) {
    cell_start[k] = r;
    r += cell_count[k];
    cell_count[k] = 0;
  }
  if (t == 1023) cell_start[C] = s[1023];
}

__global__ void cl_place_kernel(int N, const int32_t* __restrict__ atom_cell,
                                const int32_t* __restrict__ cell_start,
                                int32_t* __restrict__ cursor, int32_t* __restrict__ cell_atoms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int cell = atom_cell[i];
  if (cell < 0) return;
  const int slot = atomicAdd(&cursor[cell], 1);
  cell_atoms[cell_start[cell] + slot] = i;
}

// ascending atom order inside every cell (placement order is not deterministic)
// One wave per cell.  A cell of at most 64 atoms is ranked in registers: lane
// q holds the cell's q-th atom and its rank is the number of smaller indices
// (indices are distinct), so each lane writes its atom to slot rank; every
// lane has loaded its atom before any lane stores.  A larger cell is sorted
// by one lane (insertion sort).
__global__ void cl_sort_cells_kernel(int C, const int32_t* __restrict__ cell_start,
                                     int32_t* __restrict__ cell_atoms) {
  const int c = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= C) return;
  const int lo = cell_start[c], n = cell_start[c + 1] - lo;
  if (n <= 64) {
    const int v = lane < n ? cell_atoms[lo + lane] : 0x7fffffff;
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += (__shfl(v, j, 64) < v) ? 1 : 0;
    if (lane < n) cell_atoms[lo + rank] = v;
  } else if (lane == 0) {
    for (int k = lo + 1; k < lo + n; ++k) {
      const int v = cell_atoms[k];
      int m = k - 1;
      while (m >= lo && cell_atoms[m] > v) { cell_atoms[m + 1] = cell_atoms[m]; --m; }
      cell_atoms[m + 1] = v;
    }
  }
}

__global__ void nl_count_kernel(const float* __restrict__ pos, int N,
                                const int32_t* __restrict__ mol_ptr,
                                const int32_t* __restrict__ atom_mol,
                                const float* __restrict__ cells, const uint8_t* __restrict__ pbc,
                                float rc, int32_t* __restrict__ deg, int32_t* __restrict__ udeg,
                                CellList cl) {
  // one wave per atom: the lanes test 64 candidates at a time, the atoms of
  // the structure (brute force) or of the 27 cells around the atom's cell
  const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  const int b = atom_mol[i];
  CellInfo ci;
  bool periodic;
  load_cell(cells, pbc, b, ci, periodic);
  const float xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  int n = 0, nu = 0;
  auto hit = [&](int j) {
    float vx, vy, vz, d;
    int s0, s1, s2;
    return within(pos[3 * j] - xi, pos[3 * j + 1] - yi, pos[3 * j + 2] - zi, rc, periodic, ci,
                  vx, vy, vz, d, s0, s1, s2);
  };
  if (cl.grid != nullptr && cl.atom_cell[i] >= 0) {
    const CellWalk w = cell_walk(cl, i, b);
    for (int fb = 0; fb < w.total; fb += 64) {
      const int j = cell_candidate(cl, w, fb + lane);
      const bool h = (j >= 0) && (j != i) && hit(j);
      n += __popcll(__ballot(h));
      nu += __popcll(__ballot(h && j > i));
    }
  } else {
    const int j0 = mol_ptr[b], j1 = mol_ptr[b + 1];
    for (int jb = j0; jb < j1; jb += 64) {
      const int j = jb + lane;
      const bool h = (j < j1) && (j != i) && hit(j);
      n += __popcll(__ballot(h));
      nu += __popcll(__ballot(h && j > i));
    }
  }
  if (lane == 0) {
    deg[i] = n;
    udeg[i] = nu;
  }
}

// Exclusive scan of deg and udeg in two passes over 1024-atom slices, one
// workgroup per slice (any N; a slice's work is a handful of loads, so the
// scan costs two short launches at any N instead of a chain of barriers in
// one workgroup):
//   nl_scan_slice_kernel  : slice-local exclusive sums -> row_ptr, pair_ptr;
//                           the slice totals are parked in deg[first atom of
//                           the slice] and udeg[...] (deg and udeg are not
//                           read after the scan)
//   nl_scan_offset_kernel : adds the sum of the totals of the slices before,
//                           and writes row_ptr[N], pair_ptr[N] and the counts.
constexpr int kScanThreads = 256;
constexpr int kScanSlice = 1024;
constexpr int kScanPer = kScanSlice / kScanThreads;   // 4 consecutive atoms per thread

__global__ __launch_bounds__(kScanThreads) void nl_scan_slice_kernel(
    int32_t* __restrict__ deg, int32_t* __restrict__ udeg, int N, int32_t* __restrict__ row_ptr,
    int32_t* __restrict__ pair_ptr) {
  __shared__ __align__(16) int b1[kScanSlice];
  __shared__ __align__(16) int b2[kScanSlice];
  __shared__ int w1[kScanThreads / 64], w2[kScanThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int base = blockIdx.x * kScanSlice;
  const int n = min(kScanSlice, N - base);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int i = k * kScanThreads + t;
    b1[i] = i < n ? deg[base + i] : 0;
    b2[i] = i < n ? udeg[base + i] : 0;
  }
  __syncthreads();
  const int4 a1 = reinterpret_cast<const int4*>(b1)[t];
  const int4 a2 = reinterpret_cast<const int4*>(b2)[t];
  const int s1 = a1.x + a1.y + a1.z + a1.w, s2 = a2.x + a2.y + a2.z + a2.w;
  const int i1 = wave_inclusive_scan(s1), i2 = wave_inclusive_scan(s2);
  if (lane == 63) { w1[wv] = i1; w2[wv] = i2; }
  __syncthreads();
  int o1 = i1 - s1, o2 = i2 - s2, tot1 = 0, tot2 = 0;
#pragma unroll
  for (int q = 0; q < kScanThreads / 64; ++q) {
    if (q < wv) { o1 += w1[q]; o2 += w2[q]; }
    tot1 += w1[q];
    tot2 += w2[q];
  }
  reinterpret_cast<int4*>(b1)[t] = make_int4(o1, o1 + a1.x, o1 + a1.x + a1.y,
                                             o1 + a1.x + a1.y + a1.z);
  reinterpret_cast<int4*>(b2)[t] = make_int4(o2, o2 + a2.x, o2 + a2.x + a2.y,
                                             o2 + a2.x + a2.y + a2.z);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int i = k * kScanThreads + t;
    if (i < n) {
      row_ptr[base + i] = b1[i];
      pair_ptr[base + i] = b2[i];
    }
  }
  // thread 0 read deg[base] itself before the first barrier; no other
  // workgroup touches this slice
  if (t == 0) {
    deg[base] = tot1;
    udeg[base] = tot2;
  }
}

__global__ __launch_bounds__(kScanThreads) void nl_scan_offset_kernel(
    const int32_t* __restrict__ deg, const int32_t* __restrict__ udeg, int N,
    int32_t* __restrict__ row_ptr, int32_t* __restrict__ pair_ptr, int32_t* __restrict__ counts,
    int32_t ecap, int32_t pcap) {
  __shared__ int w1[kScanThreads / 64], w2[kScanThreads / 64];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int g = blockIdx.x, G = gridDim.x;
  // totals of the slices before g (the last workgroup also adds its own:
  // that sum is the number of edges and pairs)
  const int upto = g == G - 1 ? G : g;
  int s1 = 0, s2 = 0;
  for (int q = t; q < upto; q += kScanThreads) {
    s1 += deg[q * kScanSlice];
    s2 += udeg[q * kScanSlice];
  }
  s1 = wave_inclusive_scan(s1);
  s2 = wave_inclusive_scan(s2);
  if (lane == 63) { w1[wv] = s1; w2[wv] = s2; }
  __syncthreads();
  int all1 = 0, all2 = 0;
#pragma unroll
  for (int q = 0; q < kScanThreads / 64; ++q) { all1 += w1[q]; all2 += w2[q]; }
  int o1 = all1, o2 = all2;
  if (g == G - 1) {   // own slice total out of the offset
    o1 -= deg[g * kScanSlice];
    o2 -= udeg[g * kScanSlice];
  }
  const int base = g * kScanSlice;
  const int n = min(kScanSlice, N - base);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const int i = k * kScanThreads + t;
    if (i < n) {
      row_ptr[base + i] += o1;
      pair_ptr[base + i] += o2;
    }
  }
  if (g == G - 1 && t == 0) {
    row_ptr[N] = all1;
    pair_ptr[N] = all2;
    counts[MLFF_COUNT_EDGES] = all1;
    counts[MLFF_COUNT_PAIRS] = all2;
    // the largest count of a run of lists (an MD chunk), for the regrowth
    if (all1 > counts[MLFF_COUNT_EDGES_MAX]) counts[MLFF_COUNT_EDGES_MAX] = all1;
    if (all1 > ecap || all2 > pcap) counts[MLFF_COUNT_OVERFLOW] = 1;
  }
}

__global__ void nl_fill_kernel(const float* __restrict__ pos, int N,
                               const int32_t* __restrict__ mol_ptr,
                               const int32_t* __restrict__ atom_mol,
                               const float* __restrict__ cells, const uint8_t* __restrict__ pbc,
                               float rc, const int32_t* __restrict__ row_ptr,
                               const int32_t* __restrict__ pair_ptr, int32_t* __restrict__ col,
                               int32_t* __restrict__ edge_shift, int32_t* __restrict__ edge_pair,
                               int32_t ecap, int32_t* __restrict__ pair_lo,
                               int32_t* __restrict__ pair_hi, int32_t* __restrict__ pair_shift,
                               float4* __restrict__ geom, int32_t pcap, CellList cl,
                               int32_t* __restrict__ tmp_shift) {
  // one wave per atom, as the count pass
  __shared__ int4 hits_lds[(kNlBlock / 64) * kCellRowMax];
  const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  const int b = atom_mol[i];
  const int j0 = mol_ptr[b], j1 = mol_ptr[b + 1];
  CellInfo ci;
  bool periodic;
  load_cell(cells, pbc, b, ci, periodic);
  const float xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  int e = row_ptr[i];
  int p = pair_ptr[i];
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int deg = row_ptr[i + 1] - e;
  if (cl.grid != nullptr && cl.atom_cell[i] >= 0 && deg <= kCellRowMax) {
    // cell-list row: the wave collects the hits {j, shift} in candidate order
    // into LDS, then every hit's place in the row is the number of hits with
    // a smaller neighbour index (the neighbours of a row are distinct)
    int4* hl = hits_lds + (threadIdx.x >> 6) * kCellRowMax;
    const CellWalk w = cell_walk(cl, i, b);
    int k = 0;
    for (int fb = 0; fb < w.total; fb += 64) {
      const int j = cell_candidate(cl, w, fb + lane);
      float vx, vy, vz, d;
      int s0 = 0, s1 = 0, s2 = 0;
      const bool h = (j >= 0) && (j != i) &&
                     within(pos[3 * j] - xi, pos[3 * j + 1] - yi, pos[3 * j + 2] - zi, rc,
                            periodic, ci, vx, vy, vz, d, s0, s1, s2);
      const uint64_t mh = __ballot(h);
      const int slot = k + __popcll(mh & below);
      if (h && slot < kCellRowMax) hl[slot] = make_int4(j, s0, s1, s2);
      k += __popcll(mh);
    }
    k = min(k, kCellRowMax);
    // the LDS writes above are visible to every lane of the wave below
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int n_lower = deg - (pair_ptr[i + 1] - p);   // neighbours j < i
    for (int q = lane; q < k; q += 64) {
      const int4 hq = hl[q];
      int rank = 0;
      for (int m = 0; m < k; ++m) rank += hl[m].x < hq.x;
      const int j = hq.x, s0 = hq.y, s1 = hq.z, s2 = hq.w;
      const int ek = e + rank;
      const bool e_ok = ek < ecap;
      if (e_ok) {
        col[ek] = j;
        if (edge_shift) {
          edge_shift[3 * ek] = s0; edge_shift[3 * ek + 1] = s1; edge_shift[3 * ek + 2] = s2;
        }
      }
      if (j > i) {
        const int pk = p + rank - n_lower;
        if (pair_lo != nullptr && pk < pcap) {
          // the pair vector rounds exactly as inside within()
          float sx, sy, sz;
          shift_vec(ci, s0, s1, s2, sx, sy, sz);
          const float vx = (pos[3 * j] - xi) + sx, vy = (pos[3 * j + 1] - yi) + sy,
                      vz = (pos[3 * j + 2] - zi) + sz;
          pair_lo[pk] = i;
          pair_hi[pk] = j;
          if (pair_shift) {
            pair_shift[3 * pk] = s0; pair_shift[3 * pk + 1] = s1; pair_shift[3 * pk + 2] = s2;
          }
          geom[pk] = make_float4(vx, vy, vz, norm3_exact(vx, vy, vz));
        }
        if (edge_pair != nullptr && e_ok) edge_pair[ek] = (pk < pcap) ? pk : -1;
      }
    }
    return;
  }
  if (cl.grid != nullptr && cl.atom_cell[i] >= 0) {
    if (lane != 0) return;
    // a row longer than the LDS list: lane 0 collects in cell order, sorts by
    // neighbour index, then emits
    const int e0 = e;
    for_each_cell_candidate(cl, i, b, [&](int j) {
      if (j == i) return;
      float vx, vy, vz, d;
      int s0, s1, s2;
      if (!within(pos[3 * j] - xi, pos[3 * j + 1] - yi, pos[3 * j + 2] - zi, rc, periodic, ci,
                  vx, vy, vz, d, s0, s1, s2))
        return;
      if (e < ecap) {
        col[e] = j;
        tmp_shift[3 * e] = s0; tmp_shift[3 * e + 1] = s1; tmp_shift[3 * e + 2] = s2;
      }
      ++e;
    });
    const int e1 = min(e, ecap);
    for (int k = e0 + 1; k < e1; ++k) {
      const int v = col[k];
      const int t0 = tmp_shift[3 * k], t1 = tmp_shift[3 * k + 1], t2 = tmp_shift[3 * k + 2];
      int m = k - 1;
      while (m >= e0 && col[m] > v) {
        col[m + 1] = col[m];
        tmp_shift[3 * (m + 1)] = tmp_shift[3 * m];
        tmp_shift[3 * (m + 1) + 1] = tmp_shift[3 * m + 1];
        tmp_shift[3 * (m + 1) + 2] = tmp_shift[3 * m + 2];
        --m;
      }
      col[m + 1] = v;
      tmp_shift[3 * (m + 1)] = t0; tmp_shift[3 * (m + 1) + 1] = t1; tmp_shift[3 * (m + 1) + 2] = t2;
    }
    for (int k = e0; k < e1; ++k) {
      const int j = col[k];
      const int s0 = tmp_shift[3 * k], s1 = tmp_shift[3 * k + 1], s2 = tmp_shift[3 * k + 2];
      if (edge_shift) {
        edge_shift[3 * k] = s0; edge_shift[3 * k + 1] = s1; edge_shift[3 * k + 2] = s2;
      }
      if (j > i) {
        if (pair_lo != nullptr && p < pcap) {
          // the pair vector rounds exactly as inside within()
          float sx, sy, sz;
          shift_vec(ci, s0, s1, s2, sx, sy, sz);
          const float vx = (pos[3 * j] - xi) + sx, vy = (pos[3 * j + 1] - yi) + sy,
                      vz = (pos[3 * j + 2] - zi) + sz;
          pair_lo[p] = i;
          pair_hi[p] = j;
          if (pair_shift) {
            pair_shift[3 * p] = s0; pair_shift[3 * p + 1] = s1; pair_shift[3 * p + 2] = s2;
          }
          geom[p] = make_float4(vx, vy, vz, norm3_exact(vx, vy, vz));
        }
        if (edge_pair != nullptr) edge_pair[k] = (p < pcap) ? p : -1;
        ++p;
      }
    }
    return;
  }
  // 64 candidates per step; a hit's slot is the number of hits in lower
  // lanes, so the row stays in ascending neighbour order
  for (int jb = j0; jb < j1; jb += 64) {
    const int j = jb + lane;
    float vx = 0.f, vy = 0.f, vz = 0.f, d = 0.f;
    int s0 = 0, s1 = 0, s2 = 0;
    const bool h = (j < j1) && (j != i) &&
                   within(pos[3 * j] - xi, pos[3 * j + 1] - yi, pos[3 * j + 2] - zi, rc, periodic,
                          ci, vx, vy, vz, d, s0, s1, s2);
    const uint64_t mh = __ballot(h);
    const uint64_t mu = __ballot(h && j > i);
    if (h) {
      const int ek = e + __popcll(mh & below);
      const bool e_ok = ek < ecap;
      if (e_ok) {
        col[ek] = j;
        if (edge_shift) {
          edge_shift[3 * ek] = s0; edge_shift[3 * ek + 1] = s1; edge_shift[3 * ek + 2] = s2;
        }
      }
      if (j > i) {
        const int pk = p + __popcll(mu & below);
        if (pair_lo != nullptr && pk < pcap) {
          pair_lo[pk] = i;
          pair_hi[pk] = j;
          if (pair_shift) {
            pair_shift[3 * pk] = s0; pair_shift[3 * pk + 1] = s1; pair_shift[3 * pk + 2] = s2;
          }
          geom[pk] = make_float4(vx, vy, vz, d);
        }
        if (edge_pair != nullptr && e_ok) edge_pair[ek] = (pk < pcap) ? pk : -1;
      }
    }
    e += __popcll(mh);
    p += __popcll(mu);
  }
}

// One wave per row i.  Row i's lower neighbours j < i are its first
// deg_i - udeg_i entries (rows are sorted ascending); the lanes take them in
// parallel and find the pair (j, i) by binary search for i among row j's
// upper neighbours, the last udeg_j entries of row j.  Edges past the edge
// capacity were not written and are skipped; pairs past the pair capacity
// get -1.
__global__ void nl_link_kernel(int N, const int32_t* __restrict__ row_ptr,
                               const int32_t* __restrict__ pair_ptr,
                               const int32_t* __restrict__ col, int32_t* __restrict__ edge_pair,
                               int32_t ecap, int32_t pcap) {
  const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= N) return;
  const int r0 = row_ptr[i];
  const int n_lower = (row_ptr[i + 1] - r0) - (pair_ptr[i + 1] - pair_ptr[i]);
  for (int q = lane; q < n_lower; q += 64) {
    const int e = r0 + q;
    if (e >= ecap) break;
    const int j = col[e];
    const int end_j = row_ptr[j + 1];
    const int up_j = end_j - (pair_ptr[j + 1] - pair_ptr[j]);   // first upper entry of row j
    int lo = up_j, hi = min(end_j, ecap);
    int p = -1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      const int v = col[mid];
      if (v < i) lo = mid + 1;
      else if (v > i) hi = mid;
      else { p = pair_ptr[j] + (mid - up_j); break; }
    }
    edge_pair[e] = (p < pcap) ? p : -1;
  }
}

}  // namespace

hipError_t launch_neighbor_list(const NLArgs& a, hipStream_t s, bool build_pairs) {
  const int tb = 128;
  const int g = (a.N + tb - 1) / tb;
  if (a.N == 0) return hipSuccess;
  CellList cl{nullptr, nullptr, nullptr, nullptr};
  if (a.grid != nullptr && a.n_cells > 0) {
    const int C = a.n_cells;
    cl_zero_kernel<<<(C + 255) / 256, 256, 0, s>>>(C, a.cell_count);
    cl_bin_kernel<<<g, tb, 0, s>>>(a.pos, a.N, a.atom_mol, a.cells, a.pbc, a.grid, a.atom_cell,
                                   a.cell_count);
    cl_scan_kernel<<<1, 1024, 0, s>>>(a.cell_count, C, a.cell_start);
    cl_place_kernel<<<g, tb, 0, s>>>(a.N, a.atom_cell, a.cell_start, a.cell_count, a.cell_atoms);
    cl_sort_cells_kernel<<<(unsigned)(((int64_t)C * 64 + 255) / 256), 256, 0, s>>>(
        C, a.cell_start, a.cell_atoms);
    cl = CellList{a.grid, a.cell_start, a.cell_atoms, a.atom_cell};
  }
  // count and fill: one wave per atom
  const int gw = (int)(((int64_t)a.N * 64 + kNlBlock - 1) / kNlBlock);
  nl_count_kernel<<<gw, kNlBlock, 0, s>>>(a.pos, a.N, a.mol_ptr, a.atom_mol, a.cells, a.pbc, a.rc,
                                   a.deg, a.udeg, cl);
  const int slices = (a.N + kScanSlice - 1) / kScanSlice;
  nl_scan_slice_kernel<<<slices, kScanThreads, 0, s>>>(a.deg, a.udeg, a.N, a.row_ptr, a.pair_ptr);
  nl_scan_offset_kernel<<<slices, kScanThreads, 0, s>>>(a.deg, a.udeg, a.N, a.row_ptr, a.pair_ptr,
                                                        a.counts, a.ecap,
                                                        build_pairs ? a.pcap : 0x7fffffff);
  nl_fill_kernel<<<gw, kNlBlock, 0, s>>>(a.pos, a.N, a.mol_ptr, a.atom_mol, a.cells, a.pbc, a.rc,
                                  a.row_ptr, a.pair_ptr, a.col, a.edge_shift,
                                  build_pairs ? a.edge_pair : nullptr, a.ecap,
                                  build_pairs ? a.pair_lo : nullptr, a.pair_hi, a.pair_shift,
                                  a.geom, a.pcap, cl, a.edge_tmp_shift);
  if (build_pairs) {
    nl_link_kernel<<<gw, kNlBlock, 0, s>>>(a.N, a.row_ptr, a.pair_ptr, a.col, a.edge_pair,
                                            a.ecap, a.pcap);
  }
  return hipGetLastError();
}

}  // namespace mlff
