"""radius_graph drop-in on the gfx950 neighbour-list kernels.

Same contract as radius_graph / radius_graph_native
(student_model.py:63-109, 165-191): ordered pairs i != j of one structure
with ||p_j - p_i|| <= r, as an int64 [2, E] (src, dst) tensor in torch.where
order (sorted by src, then dst).  The kernel's distance rounds exactly like
torch.norm on the CPU, so the edge set is identical.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from . import engine


def _sizes_from_batch(batch: Optional[torch.Tensor], n: int):
    if batch is None:
        return np.array([n], dtype=np.int64), None
    b = batch.detach().to("cpu", torch.int64).numpy()
    if b.size and np.any(b[1:] < b[:-1]):
        order = np.argsort(b, kind="stable")
    else:
        order = None
    nstruct = int(b.max()) + 1 if b.size else 1
    return np.bincount(b, minlength=nstruct).astype(np.int64), order


def radius_graph_csr(positions: torch.Tensor, r: float, batch: Optional[torch.Tensor] = None,
                     cell=None, pbc=None):
    """Receiver CSR (row_ptr [N+1], col [E], shift [E,3] or None) of a
    structure-sorted batch.  Periodic cells use the minimum image."""
    n = positions.shape[0]
    sizes, order = _sizes_from_batch(batch, n)
    if order is not None:
        raise ValueError("radius_graph_csr needs atoms sorted by structure")
    periodic = pbc is not None and bool(np.asarray(pbc).any())
    if periodic:
        engine.check_cells(cell, pbc, r)
    layout = engine.make_layout(sizes, positions.device, cell if periodic else None,
                                pbc if periodic else None)
    pos = positions.detach().to(torch.float32).contiguous()
    return engine.neighbor_list(pos, layout, r, with_shifts=periodic)


def radius_graph(positions: torch.Tensor, r: float, batch: Optional[torch.Tensor] = None,
                 loop: bool = False, use_torch_cluster: bool = True) -> torch.Tensor:
    """[2, E] (src, dst) edge index, reference order; ``loop`` adds i -> i."""
    if positions.device.type != "cuda":
        raise RuntimeError("radius_graph (MI355X build) needs positions on a ROCm device")
    n = positions.shape[0]
    sizes, order = _sizes_from_batch(batch, n)
    pos = positions
    if order is not None:
        perm = torch.as_tensor(order, device=positions.device)
        pos = positions[perm]
    layout = engine.make_layout(sizes, positions.device)
    row_ptr, col, _ = engine.neighbor_list(pos.detach().to(torch.float32).contiguous(), layout,
                                           float(np.float32(r)))
    deg = row_ptr[1:] - row_ptr[:-1]
    src = torch.repeat_interleave(torch.arange(n, device=positions.device), deg.long())
    dst = col.long()
    if order is not None:
        src, dst = perm[src], perm[dst]
    if loop:
        idx = torch.arange(n, device=positions.device)
        src = torch.cat([src, idx])
        dst = torch.cat([dst, idx])
    if order is not None or loop:
        key = src * n + dst
        key, srt = torch.sort(key)
        src, dst = src[srt], dst[srt]
    return torch.stack([src, dst], dim=0)


radius_graph_native = radius_graph
