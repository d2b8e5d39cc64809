"""Device-side engine: weight packing, plans and workspaces, launches.

Everything here drives libmlff_amd.so through ctypes; tensors are plain
device buffers (PyTorch's caching allocator), never arguments of the C ABI.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _capi as C

def device_of(device) -> torch.device:
    """A device with its index: torch.device('cuda') becomes cuda:<current>,
    so device comparisons and cache keys see one name per GPU."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        try:
            idx = torch.cuda.current_device()
        except (RuntimeError, AssertionError):     # no ROCm device visible
            idx = 0
        d = torch.device("cuda", idx)
    return d


# ---------------------------------------------------------------------------
# weights
# ---------------------------------------------------------------------------

_NAMES = {
    C.T_EMBEDDING: "embedding.weight",
    C.T_RBF_CENTERS: "rbf.centers",
    C.T_RBF_WIDTHS: "rbf.widths",
    C.T_F1_W: "interactions.{l}.message.rbf_to_scalar.0.weight",
    C.T_F1_B: "interactions.{l}.message.rbf_to_scalar.0.bias",
    C.T_F2_W: "interactions.{l}.message.rbf_to_scalar.2.weight",
    C.T_F2_B: "interactions.{l}.message.rbf_to_scalar.2.bias",
    C.T_U1_W: "interactions.{l}.update.update_mlp.0.weight",
    C.T_U1_B: "interactions.{l}.update.update_mlp.0.bias",
    C.T_U2_W: "interactions.{l}.update.update_mlp.2.weight",
    C.T_U2_B: "interactions.{l}.update.update_mlp.2.bias",
    C.T_MIX: "interactions.{l}.update.mixing_matrix",
    C.T_R1_W: "energy_head.0.weight",
    C.T_R1_B: "energy_head.0.bias",
    C.T_R2_W: "energy_head.2.weight",
    C.T_R2_B: "energy_head.2.bias",
    C.T_R3_W: "energy_head.4.weight",
    C.T_R3_B: "energy_head.4.bias",
}


def hyper_from_state(sd: Dict[str, torch.Tensor]) -> Tuple[int, int, int, int]:
    H = sd["embedding.weight"].shape[1]
    max_z = sd["embedding.weight"].shape[0] - 1
    L = len({int(k.split(".")[1]) for k in sd if k.startswith("interactions.")})
    R = sd["rbf.centers"].shape[0]
    return H, L, R, max_z


_LAYOUTS: Dict[tuple, tuple] = {}
# device slot tables of mlff_pack_weights, per (layout, tensor addresses)
_PACK_TABLES: Dict[tuple, torch.Tensor] = {}


def _layout(sd: Dict[str, torch.Tensor]):
    H, L, R, max_z = hyper_from_state(sd)
    key = (H, L, R, max_z)
    if key not in _LAYOUTS:
        _LAYOUTS[key] = C.weights_layout(H, L, R, max_z)
    return key, _LAYOUTS[key]


def pack_weights(sd: Dict[str, torch.Tensor], device) -> torch.Tensor:
    """Pack a StudentForceField state dict into the library's float32 blob.

    When every tensor is a contiguous float32 tensor on the device already
    (a model being trained), one kernel (mlff_pack_weights) writes the whole
    blob from the tensors in place, so a weight update costs one launch and no
    host round trip; otherwise the blob is assembled on the host and copied
    over.  Either way the blob is a fresh tensor."""
    key, (slots, total) = _layout(sd)
    device = device_of(device)
    srcs = []
    for s in slots:
        srcs.append(sd[_NAMES[s.tensor].format(l=s.layer)].detach())
    if device.type == "cuda" and all(t.device == device and t.dtype == torch.float32 and
                                     t.is_contiguous() for t in srcs):
        tkey = (key, device) + tuple(t.data_ptr() for t in srcs)
        table = _PACK_TABLES.get(tkey)
        if table is None:
            rows = []
            for s, t in sorted(zip(slots, srcs), key=lambda st: st[0].offset):
                tr, tc = (1, t.numel()) if t.dim() == 1 else tuple(t.shape)
                if s.kind == 1:
                    tr, tc = tc, tr
                if tr > s.rows or tc > s.ld:
                    raise ValueError(f"slot too small for {tuple(t.shape)}")
                rows.append([s.offset, s.rows, s.ld, tr, tc, s.kind, t.data_ptr(), 0])
            if len(_PACK_TABLES) > 32:
                _PACK_TABLES.clear()
            table = torch.tensor(rows, dtype=torch.int64).to(device)
            _PACK_TABLES[tkey] = table
        blob = torch.empty(total, dtype=torch.float32, device=device)
        C.check("mlff_pack_weights", C.lib().mlff_pack_weights(
            C.stream_handle(), table.shape[0], C.ptr(table), total, C.ptr(blob)))
        return blob
    host = torch.zeros(total, dtype=torch.float32)
    for s, t in zip(slots, srcs):
        t = t.to("cpu", torch.float32)
        if t.dim() == 1:
            t = t.unsqueeze(0)
        if s.kind == 1:
            t = t.t()
        rows, cols = t.shape
        if rows > s.rows or cols > s.ld:
            raise ValueError(f"slot too small ({rows}x{cols} into {s.rows}x{s.ld})")
        host[s.offset:s.offset + s.rows * s.ld].view(s.rows, s.ld)[:rows, :cols] = t
    return host if device.type == "cpu" else host.to(device)


# ---------------------------------------------------------------------------
# batch description
# ---------------------------------------------------------------------------

@dataclass
class BatchLayout:
    """Host description of a flattened batch (atoms of a structure contiguous)."""
    n_atoms: int
    sizes: np.ndarray           # [B]
    mol_ptr: torch.Tensor       # [B+1] int32, device
    atom_mol: torch.Tensor      # [N] int32, device
    cells: Optional[torch.Tensor]  # [B,3,3] float32 device or None
    pbc: Optional[torch.Tensor]    # [B,3] uint8 device or None
    periodic: bool
    cells_host: Optional[np.ndarray] = None
    pbc_host: Optional[np.ndarray] = None

    @property
    def n_structures(self) -> int:
        return len(self.sizes)


def make_layout(sizes: Sequence[int], device, cells=None, pbc=None) -> BatchLayout:
    sizes = np.asarray(sizes, dtype=np.int64)
    mol_ptr = np.zeros(len(sizes) + 1, dtype=np.int32)
    np.cumsum(sizes, out=mol_ptr[1:])
    atom_mol = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    periodic = False
    ct = pt = None
    cells_np = pbc_np = None
    if cells is not None and pbc is not None:
        pbc_np = np.asarray(pbc, dtype=bool).reshape(len(sizes), 3)
        if pbc_np.any():
            periodic = True
            cells_np = np.asarray(cells, dtype=np.float32).reshape(len(sizes), 3, 3)
            ct = torch.from_numpy(np.ascontiguousarray(cells_np)).to(device)
            pt = torch.from_numpy(pbc_np.astype(np.uint8)).to(device)
    return BatchLayout(int(mol_ptr[-1]), sizes, torch.from_numpy(mol_ptr).to(device),
                       torch.from_numpy(atom_mol).to(device), ct, pt, periodic,
                       cells_np if periodic else None, pbc_np if periodic else None)


# structures smaller than this keep the brute-force scan (cheaper than 27 cells)
CELL_LIST_MIN_ATOMS = 256


def cell_grid(layout: BatchLayout, cutoff: float, min_atoms: Optional[int] = None,
              max_cells: int = 1 << 22):
    """Cell-list grid per structure: {nc0, nc1, nc2, first cell} for fully
    periodic structures of at least ``min_atoms`` atoms whose cell holds
    three cells of perpendicular width >= cutoff along every axis; zeros
    (brute force) elsewhere.  The widths carry a 1e-5 relative margin so a
    position rounded onto a cell boundary still lands within one cell."""
    B = layout.n_structures
    if min_atoms is None:
        min_atoms = CELL_LIST_MIN_ATOMS
    grid = np.zeros((B, 4), dtype=np.int32)
    if not layout.periodic:
        return grid, 0
    base = 0
    for b in range(B):
        if layout.sizes[b] < min_atoms or not np.all(layout.pbc_host[b]):
            continue
        c = layout.cells_host[b].astype(np.float64)
        vol = abs(np.linalg.det(c))
        widths = np.array([vol / np.linalg.norm(np.cross(c[(a + 1) % 3], c[(a + 2) % 3]))
                           for a in range(3)])
        nc = np.floor(widths / (cutoff * (1.0 + 1e-5))).astype(np.int64)
        if np.any(nc < 3) or base + int(np.prod(nc)) > max_cells:
            continue
        grid[b] = (nc[0], nc[1], nc[2], base)
        base += int(np.prod(nc))
    return grid, base


def check_cells(cells: np.ndarray, pbc: np.ndarray, cutoff: float):
    """Minimum image needs the cutoff below half of every periodic width."""
    cells = np.asarray(cells, dtype=np.float64).reshape(-1, 3, 3)
    pbc = np.asarray(pbc, dtype=bool).reshape(-1, 3)
    for c, p in zip(cells, pbc):
        if not p.any():
            continue
        vol = abs(np.linalg.det(c))
        if vol < 1e-9:
            raise ValueError("periodic cell is degenerate")
        for a in range(3):
            if not p[a]:
                continue
            b, d = c[(a + 1) % 3], c[(a + 2) % 3]
            width = vol / np.linalg.norm(np.cross(b, d))
            if cutoff >= 0.5 * width:
                raise ValueError(
                    f"cutoff {cutoff} A is not below half the periodic cell width {width:.3f} A "
                    "(minimum image would be ambiguous)")


def edge_capacity_bound(sizes: Sequence[int], periodic: bool, max_neighbors: int = 256) -> int:
    """Upper bound on directed edges: complete graphs for small structures,
    ``max_neighbors`` per atom for large or periodic ones."""
    tot = 0
    for n in np.asarray(sizes).tolist():
        tot += n * min(n - 1, max_neighbors) if not periodic else n * max_neighbors
    return max(int(tot), 16)


# ---------------------------------------------------------------------------
# plans
# ---------------------------------------------------------------------------

# streams per evaluation for new plans (Plan.set_streams)
DEFAULT_STREAMS = 2


def launch_shapes(H: int):
    """The pre-instantiated launch shapes of hidden size H, default first:
    pairs per filter-forward tile, pairs per filter-adjoint tile, edges in
    flight per message-forward row, per message-adjoint row, and the filter
    adjoint's variant (0 plain, 1 next-tile index prefetch; default adjoint
    tile only) (mlff_plan_set_tiles)."""
    small = H <= 32
    return ((128, 64) if small else (64, 32), (64, 32) if small else (32, 64), (5, 3), (4, 2),
            (0, 1))


def default_tiles(H: int) -> Tuple[int, ...]:
    """The default of every launch-shape kind of hidden size H."""
    return tuple(values[0] for values in launch_shapes(H))


class Plan:
    """One bound plan: workspace + per-batch constants + weights."""

    def __init__(self, H, L, R, max_z, cutoff, layout: BatchLayout, numbers: torch.Tensor,
                 weights: torch.Tensor, edge_capacity: int, pair_capacity: Optional[int] = None,
                 persistent_blocks: int = 0, streams: Optional[int] = None,
                 tiles: Optional[Tuple[int, ...]] = None):
        self.lib = C.lib()
        self.layout = layout
        self.H, self.L, self.R, self.max_z, self.cutoff = H, L, R, max_z, float(cutoff)
        self.ecap = int(edge_capacity)
        self.pcap = int(pair_capacity if pair_capacity is not None else (self.ecap + 1) // 2)
        grid, n_cells = cell_grid(layout, cutoff)
        self.cell_grid = grid
        self.cell_capacity = n_cells
        cfg = C.Config(H, L, R, max_z, float(cutoff), layout.n_atoms, layout.n_structures,
                       self.ecap, self.pcap, int(layout.periodic), persistent_blocks, n_cells)
        handle = C.vp()
        C.check("mlff_plan_create", self.lib.mlff_plan_create(ctypes_byref(handle), ctypes_byref(cfg)))
        self.handle = handle
        nbytes = self.lib.mlff_plan_workspace_bytes(handle)
        dev = numbers.device
        self.workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self.numbers = numbers
        self.weights = weights
        C.check("mlff_plan_bind", self.lib.mlff_plan_bind(
            handle, C.ptr(self.workspace), int(nbytes), C.ptr(weights), C.ptr(numbers),
            C.ptr(layout.mol_ptr), C.ptr(layout.atom_mol), C.ptr(layout.cells),
            C.ptr(layout.pbc)))
        counts_addr = self.lib.mlff_plan_counts(handle)
        # view of the counts slots inside the workspace
        off = counts_addr - self.workspace.data_ptr()
        self.counts = self.workspace[off:off + 4 * C.MLFF_COUNT_SLOTS].view(torch.int32)
        self._apply_cell_grid()
        self.streams = int(streams if streams is not None else DEFAULT_STREAMS)
        if self.streams != 2:
            self.set_streams(self.streams)
        self.tiles = default_tiles(H)
        if tiles is not None:
            self.set_tiles(*tiles)

    def set_tiles(self, fwd_pairs: int = 0, bwd_pairs: int = 0, msg_unroll_fwd: int = 0,
                  msg_unroll_bwd: int = 0, bwd_variant: int = 0):
        """Launch shapes of the pair and message kernels (0: default); each
        must be one of launch_shapes(H).  Bit-identical results either way."""
        C.check("mlff_plan_set_tiles", self.lib.mlff_plan_set_tiles(
            self.handle, int(fwd_pairs), int(bwd_pairs), int(msg_unroll_fwd),
            int(msg_unroll_bwd), int(bwd_variant)))
        want = (fwd_pairs, bwd_pairs, msg_unroll_fwd, msg_unroll_bwd, bwd_variant)
        self.tiles = tuple(int(w) if w else d for w, d in zip(want, default_tiles(self.H)))

    def _apply_cell_grid(self):
        if int(self.cell_grid[:, 0].max(initial=0)) > 0:
            g = np.ascontiguousarray(self.cell_grid, dtype=np.int32)
            C.check("mlff_plan_set_cell_grid", self.lib.mlff_plan_set_cell_grid(
                self.handle, g.ctypes.data, self.layout.n_structures))

    def set_weights(self, weights: torch.Tensor, stream=None):
        """Rebind with a new weight blob (same layout), in stream order on
        ``stream`` (default: the current stream) without a host sync."""
        self.weights = weights
        C.check("mlff_plan_set_weights", self.lib.mlff_plan_set_weights(
            self.handle, C.stream_handle(stream), C.ptr(weights)))

    def energy_forces(self, positions: torch.Tensor, energy: torch.Tensor,
                      forces: Optional[torch.Tensor], virial: Optional[torch.Tensor] = None,
                      atom_energy: Optional[torch.Tensor] = None, stream=None):
        C.check("mlff_energy_forces", self.lib.mlff_energy_forces(
            self.handle, C.stream_handle(stream), C.ptr(positions), C.ptr(energy),
            C.ptr(forces), C.ptr(virial), C.ptr(atom_energy)))

    def set_streams(self, n_streams: int):
        """1: every kernel on the caller's stream; 2 (default): the pair filter
        kernels on a second stream of the plan's own, overlapping the
        atom-side kernels.  Bit-identical results either way."""
        C.check("mlff_plan_set_streams", self.lib.mlff_plan_set_streams(self.handle, int(n_streams)))
        self.streams = int(n_streams)

    def read_counts(self) -> np.ndarray:
        return self.counts.cpu().numpy()

    def __del__(self):
        h = getattr(self, "handle", None)
        self.handle = None
        if h is not None and h.value:
            try:
                self.lib.mlff_plan_destroy(h)
            except Exception:
                pass


def ctypes_byref(x):
    import ctypes
    return ctypes.byref(x)


def count_edges(positions: torch.Tensor, layout: BatchLayout, cutoff: float) -> int:
    """Number of directed edges of the current positions (the neighbour-list
    count and scan passes only: no edge is written; one device sync)."""
    lib = C.lib()
    N = layout.n_atoms
    dev = positions.device
    row_ptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
    counts = torch.zeros(C.MLFF_COUNT_SLOTS, dtype=torch.int32, device=dev)
    ws = torch.empty(int(lib.mlff_neighbor_workspace_bytes(N)), dtype=torch.uint8, device=dev)
    C.check("mlff_neighbor_list", lib.mlff_neighbor_list(
        C.stream_handle(), C.ptr(positions), N, layout.n_structures, C.ptr(layout.mol_ptr),
        C.ptr(layout.atom_mol), C.ptr(layout.cells), C.ptr(layout.pbc), float(cutoff),
        C.ptr(row_ptr), None, None, 0, C.ptr(counts), C.ptr(ws)))
    return int(counts[C.MLFF_COUNT_EDGES].item())


def neighbor_pairs(positions: torch.Tensor, layout: BatchLayout, cutoff: float):
    """Receiver-CSR neighbour list of isolated structures with its unordered
    pairs, on the device (mlff_neighbor_pairs).

    Returns (row_ptr [N+1], col [E], edge_pair [E], pair_lo [P], pair_hi [P]),
    int32, pairs numbered in CSR order of the upper edges; the only wait for
    the device is the count read that sizes the outputs.
    """
    if layout.periodic:
        raise ValueError("neighbor_pairs covers isolated structures only")
    lib = C.lib()
    N = layout.n_atoms
    dev = positions.device
    cap = edge_capacity_bound(layout.sizes, False)
    for _ in range(2):
        pcap = cap // 2 + 1
        row_ptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
        col = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        edge_pair = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        pair_lo = torch.empty(pcap, dtype=torch.int32, device=dev)
        pair_hi = torch.empty(pcap, dtype=torch.int32, device=dev)
        geom = torch.empty(pcap, 4, dtype=torch.float32, device=dev)
        counts = torch.zeros(C.MLFF_COUNT_SLOTS, dtype=torch.int32, device=dev)
        ws = torch.empty(int(lib.mlff_neighbor_workspace_bytes(N)), dtype=torch.uint8, device=dev)
        C.check("mlff_neighbor_pairs", lib.mlff_neighbor_pairs(
            C.stream_handle(), C.ptr(positions), N, layout.n_structures, C.ptr(layout.mol_ptr),
            C.ptr(layout.atom_mol), float(cutoff), C.ptr(row_ptr), C.ptr(col), cap,
            C.ptr(edge_pair), C.ptr(pair_lo), C.ptr(pair_hi), C.ptr(geom), pcap, C.ptr(counts),
            C.ptr(ws)))
        c = counts.cpu().numpy()
        E, P = int(c[C.MLFF_COUNT_EDGES]), int(c[C.MLFF_COUNT_PAIRS])
        if E <= cap and P <= pcap:
            return row_ptr, col[:E], edge_pair[:E], pair_lo[:P], pair_hi[:P]
        cap = max(E, 2 * P)
    raise RuntimeError("neighbour list capacity did not converge")


def neighbor_list(positions: torch.Tensor, layout: BatchLayout, cutoff: float,
                  edge_capacity: Optional[int] = None, with_shifts: bool = False):
    """Receiver-CSR neighbour list on the device.

    Returns (row_ptr [N+1] int32, col [E] int32, shift [E,3] int32 or None).
    """
    lib = C.lib()
    N = layout.n_atoms
    dev = positions.device
    cap = edge_capacity if edge_capacity is not None else edge_capacity_bound(layout.sizes,
                                                                             layout.periodic)
    for _ in range(2):
        row_ptr = torch.empty(N + 1, dtype=torch.int32, device=dev)
        col = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        shift = torch.empty((max(cap, 1), 3), dtype=torch.int32, device=dev) if with_shifts else None
        counts = torch.zeros(C.MLFF_COUNT_SLOTS, dtype=torch.int32, device=dev)
        ws = torch.empty(int(lib.mlff_neighbor_workspace_bytes(N)), dtype=torch.uint8, device=dev)
        C.check("mlff_neighbor_list", lib.mlff_neighbor_list(
            C.stream_handle(), C.ptr(positions), N, layout.n_structures, C.ptr(layout.mol_ptr),
            C.ptr(layout.atom_mol), C.ptr(layout.cells), C.ptr(layout.pbc), float(cutoff),
            C.ptr(row_ptr), C.ptr(col), C.ptr(shift), cap, C.ptr(counts), C.ptr(ws)))
        c = counts.cpu().numpy()
        E = int(c[C.MLFF_COUNT_EDGES])
        if E <= cap:
            col = col[:E]
            if shift is not None:
                shift = shift[:E]
            return row_ptr, col, shift
        cap = E
    raise RuntimeError("neighbour list capacity did not converge")
