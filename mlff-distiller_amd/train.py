"""Force-matching training on the GPU: hand-derived second-order pass.

The reference trains with torch autograd double backward
(models/distillation_wrapper.py:86-91, training/trainer.py:325-328, and the
compact recipe scripts/train_compact_models.py:229-372 with loss
wE * MSE(E) + wF * MSE(F), :340-342).  Here the gradient of the loss with
respect to the weights is formed explicitly:

  dL/dtheta = sum_b dL/dE_b dE_b/dtheta  -  d/dtheta < lambda, grad_x E >,
  lambda = dL/dF.

The first term is the reverse (adjoint) pass with parameter gradients; the
second is the tangent of that reverse pass along the position direction
lambda (forward-over-reverse).  Both come from one reverse pass seeded with
unit structure weights: its tangent gives the second term, and because
structures do not interact the first term is the same pass's parameter
terms with every atom or pair row scaled by dL/dE of its structure.

Where the work runs:
* edge- and pair-sized parts on HIP kernels (mlff_train_message,
  mlff_train_message_adjoint, mlff_train_pair_grad): the message gathers of
  the forward pass and its tangent as receiver rows in ascending neighbour
  order, the adjoint gathers the same way, and the filter-output gradients
  per unordered pair from both endpoints -- no scatter-adds, no atomics;
* per-structure energies and the embedding gradient by fixed-order segment
  sums (mlff_segment_sum);
* the dense products (filter and update MLPs, weight gradients) as plain
  library GEMMs; weight gradients over many rows are split into 2048-row
  chunks and summed, so no library split-K reduction is involved.
Every reduction runs in a fixed order: two runs give bit-identical gradients.

The forces the loss needs come from the fused HIP evaluation of the current
weights (the trainer calls model.evaluate).  The computation is written
over unordered pairs (lo < hi), like the HIP inference kernels: a pair's
filter is evaluated once and feeds both directed edges.  float32 on the
device; energies are summed in float64.
"""

from __future__ import annotations

import math
import weakref
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np
import torch

from . import _capi as C

Tensor = torch.Tensor
EPS = 1e-8   # edge_vector / (d + 1e-8), student_model.py:715


@dataclass
class PairGraph:
    """Receiver CSR of a structure-sorted batch and its unordered pairs."""
    row_ptr: Tensor          # [N+1] int32
    col: Tensor              # [E] int32, ascending within a row
    edge_pair: Tensor        # [E] int32, pair of each directed edge
    lo: Tensor               # [P] int64 (lo < hi)
    hi: Tensor               # [P] int64
    lo32: Tensor             # [P] int32
    hi32: Tensor             # [P] int32
    shift: Optional[Tensor]  # [P,3] float shift vector (already times the cell) or None
    batch: Tensor            # [N] int64 structure of each atom
    mol_ptr: Tensor          # [B+1] int32
    n_structures: int


def pair_graph_from_csr(row_ptr: Tensor, col: Tensor, batch: Tensor, n_structures: int,
                        shift_vec: Optional[Tensor] = None, n_pairs: Optional[int] = None,
                        mol_ptr: Optional[Tensor] = None) -> PairGraph:
    """Pairs (row < col) of a symmetric receiver CSR with rows in ascending
    column order, and the pair of every directed edge: an upper edge
    (row < col) numbers its pair in CSR order; a lower edge takes the
    number of its reverse edge, found by a binary search of the sorted
    (row, col) keys.  With the pair count (``n_pairs``, known on the host
    when no edge is a self-image) and the structure offsets (``mol_ptr``)
    given, nothing here waits for the device."""
    dev = col.device
    n = row_ptr.numel() - 1
    E = col.numel()
    deg = (row_ptr[1:] - row_ptr[:-1]).long()
    row = torch.repeat_interleave(torch.arange(n, device=dev), deg, output_size=E)
    colL = col.long()
    up = colL > row
    up_id = torch.cumsum(up.to(torch.int64), 0) - 1
    key = row * n + colL
    rev = torch.searchsorted(key, colL * n + row)
    edge_pair = torch.where(up, up_id, up_id[rev]).to(torch.int32)
    if n_pairs is None:
        lo, hi = row[up], colL[up]
        sv = None if shift_vec is None else shift_vec[up].contiguous()
    else:
        # every upper edge to its pair slot; the other edges to a spare slot
        dst = torch.where(up, up_id, torch.full_like(up_id, n_pairs))
        pick = lambda v: torch.empty(n_pairs + 1, dtype=v.dtype, device=dev).scatter_(
            0, dst, v)[:n_pairs]
        lo, hi = pick(row), pick(colL)
        sv = None
        if shift_vec is not None:
            sv = torch.empty(n_pairs + 1, 3, dtype=shift_vec.dtype, device=dev).index_copy_(
                0, dst, shift_vec)[:n_pairs].contiguous()
    if mol_ptr is None:
        sizes = torch.bincount(batch, minlength=n_structures)
        mol_ptr = torch.zeros(n_structures + 1, dtype=torch.int32, device=dev)
        mol_ptr[1:] = torch.cumsum(sizes, 0).to(torch.int32)
    return PairGraph(row_ptr.to(torch.int32).contiguous(), col.to(torch.int32).contiguous(),
                     edge_pair.contiguous(), lo, hi, lo.to(torch.int32).contiguous(),
                     hi.to(torch.int32).contiguous(), sv, batch.long(),
                     mol_ptr.to(torch.int32).contiguous(), n_structures)


# device layouts of non-periodic batches by (sizes, device): the training loop
# builds a pair graph per step, usually of the same batch shape
_GRAPH_LAYOUTS: "Dict[tuple, object]" = {}


def device_graph(Z: Tensor, x: Tensor, sizes: Sequence[int], rc: float, cells=None,
                 pbc=None) -> PairGraph:
    """Pair graph of a structure-sorted batch from the HIP neighbour list."""
    from . import engine
    periodic = pbc is not None and bool(np.asarray(pbc).any())
    if periodic:
        layout = engine.make_layout(sizes, x.device, cells, pbc)
    else:
        key = (tuple(int(v) for v in sizes), str(x.device))
        layout = _GRAPH_LAYOUTS.get(key)
        if layout is None:
            if len(_GRAPH_LAYOUTS) > 16:
                _GRAPH_LAYOUTS.clear()
            layout = engine.make_layout(sizes, x.device)
            _GRAPH_LAYOUTS[key] = layout
    pos32 = x.detach().to(torch.float32).contiguous()
    batch = layout.atom_mol.long()
    if not periodic:
        # isolated structures: the neighbour-list kernels number the pairs
        # (upper edges in CSR order) and link every edge to its pair
        row_ptr, col, edge_pair, lo32, hi32 = engine.neighbor_pairs(pos32, layout, rc)
        return PairGraph(row_ptr, col, edge_pair, lo32.long(), hi32.long(), lo32, hi32, None,
                         batch, layout.mol_ptr, len(sizes))
    row_ptr, col, shift = engine.neighbor_list(pos32, layout, rc, with_shifts=True)
    cell_t = torch.as_tensor(np.asarray(cells, dtype=np.float32).reshape(-1, 3, 3),
                             device=x.device)
    n = row_ptr.numel() - 1
    deg = (row_ptr[1:] - row_ptr[:-1]).long()
    row = torch.repeat_interleave(torch.arange(n, device=x.device), deg, output_size=col.numel())
    cells_e = cell_t[batch[row]]
    sh = shift.to(torch.float32)
    shift_vec = (sh[:, 0:1] * cells_e[:, 0] + sh[:, 1:2] * cells_e[:, 1] +
                 sh[:, 2:3] * cells_e[:, 2])
    # a periodic list may hold self-image edges (no upper partner): count on the device
    return pair_graph_from_csr(row_ptr, col, batch, len(sizes), shift_vec, mol_ptr=layout.mol_ptr)


# ---------------------------------------------------------------------------
# HIP launches
# ---------------------------------------------------------------------------

def _p(t: Optional[Tensor]) -> int:
    return 0 if t is None else t.data_ptr()


def _dense(t: Tensor) -> Tensor:
    """t as a contiguous, 16-byte aligned float32 device buffer."""
    if t.dtype != torch.float32 or not t.is_cuda:
        raise TypeError("the training pass runs on float32 device tensors (HIP kernels)")
    if not t.is_contiguous():
        t = t.contiguous()
    if t.data_ptr() % 16:
        t = t.clone()
    return t


def _message(g: PairGraph, H: int, s, v, W, u, ds=None, dv=None, dW=None, du=None):
    """Message outputs (s + sum s_j Ws, v + sum (v_j Wv1 + sg u Wv2)) and,
    with the tangent inputs, their tangents (mlff_train_message)."""
    s, v, W, u = _dense(s), _dense(v), _dense(W), _dense(u)
    so, vo = torch.empty_like(s), torch.empty_like(v)
    dso = dvo = None
    if ds is not None:
        ds, dv, dW, du = _dense(ds), _dense(dv), _dense(dW), _dense(du)
        dso, dvo = torch.empty_like(s), torch.empty_like(v)
    C.check("mlff_train_message", C.lib().mlff_train_message(
        C.stream_handle(), H, s.shape[0], _p(g.row_ptr), _p(g.col), _p(g.edge_pair), _p(s),
        _p(v), _p(W), _p(u), _p(ds), _p(dv), _p(dW), _p(du), _p(so), _p(vo), _p(dso), _p(dvo)))
    return so, vo, dso, dvo


def _message_adjoint(g: PairGraph, H: int, gs, gv, W, dgs=None, dgv=None, dW=None):
    """gs + sum gs_j Ws, gv + sum gv_j Wv1 and, with the tangent inputs,
    their tangents (mlff_train_message_adjoint)."""
    gs, gv, W = _dense(gs), _dense(gv), _dense(W)
    go, gvo = torch.empty_like(gs), torch.empty_like(gv)
    dgo = dgvo = None
    if dgs is not None:
        dgs, dgv, dW = _dense(dgs), _dense(dgv), _dense(dW)
        dgo, dgvo = torch.empty_like(gs), torch.empty_like(gv)
    C.check("mlff_train_message_adjoint", C.lib().mlff_train_message_adjoint(
        C.stream_handle(), H, gs.shape[0], _p(g.row_ptr), _p(g.col), _p(g.edge_pair), _p(gs),
        _p(gv), _p(W), _p(dgs), _p(dgv), _p(dW), _p(go), _p(gvo), _p(dgo), _p(dgvo)))
    return go, gvo, dgo, dgvo


def _pair_grad(g: PairGraph, H: int, u, s, v, gs, gv, du=None, ds=None, dv=None, dgs=None,
               dgv=None):
    """dL/dW per pair from both endpoints and, with the tangents, its tangent
    (mlff_train_pair_grad)."""
    P = g.lo.numel()
    u, s, v, gs, gv = _dense(u), _dense(s), _dense(v), _dense(gs), _dense(gv)
    gW = torch.empty(P, 3 * H, dtype=torch.float32, device=s.device)
    dgW = None
    if dgs is not None:
        du, ds, dv, dgs, dgv = _dense(du), _dense(ds), _dense(dv), _dense(dgs), _dense(dgv)
        dgW = torch.empty_like(gW)
    C.check("mlff_train_pair_grad", C.lib().mlff_train_pair_grad(
        C.stream_handle(), H, P, _p(g.lo32), _p(g.hi32), _p(u), _p(du), _p(s), _p(v), _p(ds),
        _p(dv), _p(gs), _p(gv), _p(dgs), _p(dgv), _p(gW), _p(dgW)))
    return gW, dgW


def _segment_sum(seg_ptr: Tensor, src: Tensor, perm: Optional[Tensor] = None,
                 f64: bool = False) -> Tensor:
    """out[k] = sum of the rows of segment k in row order (mlff_segment_sum)."""
    src = _dense(src)
    n_seg = seg_ptr.numel() - 1
    width = src[0].numel() if src.dim() > 1 else 1
    shape = (n_seg,) + tuple(src.shape[1:])
    out = torch.empty(shape, dtype=torch.float64 if f64 else torch.float32, device=src.device)
    C.check("mlff_segment_sum", C.lib().mlff_segment_sum(
        C.stream_handle(), n_seg, width, _p(seg_ptr), _p(perm), _p(src),
        _p(out) if f64 else 0, 0 if f64 else _p(out)))
    return out


# ---------------------------------------------------------------------------
# elementwise pieces and dense products
# ---------------------------------------------------------------------------

def _glue(name: str, *args):
    """Call one of the update-block kernels (mlff_update_*): leading ints,
    then tensors (made dense float32) or None."""
    conv = []
    for a in args:
        if isinstance(a, Tensor):
            a = _p(_dense(a))
        elif a is None:
            a = None
        conv.append(a)
    C.check(name, getattr(C.lib(), name)(C.stream_handle(), *conv))


def _silu_call(a, da=None, g=None, dg=None, h=False, dh=False, dga=False):
    a = _dense(a)
    opt = lambda t: None if t is None else _dense(t)
    da, g, dg = opt(da), opt(g), opt(dg)
    outs = [torch.empty_like(a) if want else None for want in (h, dh, dga)]
    C.check("mlff_silu_chain", C.lib().mlff_silu_chain(
        C.stream_handle(), a.numel(), C.ptr(a), C.ptr(da), C.ptr(g), C.ptr(dg),
        C.ptr(outs[0]), C.ptr(outs[1]), None, C.ptr(outs[2])))
    return [o for o in outs if o is not None][0]


def _silu_pair(x, g, dx, dg=None):
    """The adjoint g silu'(x) and its tangent dg silu'(x) + g silu''(x) dx
    (dg None is zero) from one pass, stacked in one [2, *x.shape] buffer, so
    the products that follow take both as one operand."""
    x, g, dx = _dense(x), _dense(g), _dense(dx)
    dg = None if dg is None else _dense(dg)
    out = torch.empty((2,) + tuple(x.shape), dtype=torch.float32, device=x.device)
    C.check("mlff_silu_chain", C.lib().mlff_silu_chain(
        C.stream_handle(), x.numel(), C.ptr(x), C.ptr(dx), C.ptr(g), C.ptr(dg), None, None,
        C.ptr(out[0]), C.ptr(out[1])))
    return out


def _silu(x):
    """silu(x) = x sigmoid(x) (mlff_silu_chain)."""
    return _silu_call(x, h=True)


def _silu_d(x, t):
    """silu'(x) t: a tangent along t, or the adjoint of a gradient t."""
    return _silu_call(x, da=t, dh=True)


def _silu_dd(x, g, dx, dg=None):
    """dg silu'(x) + g silu''(x) dx: the tangent of the adjoint g silu'(x)
    along (dx, dg); dg None is zero."""
    return _silu_call(x, da=dx, g=g, dg=dg, dga=True)


# prepared filter weights per (W1, W2) tensor pair: W1 padded to 16 or 32
# columns, W2 as three bf16 planes and W2 transposed.  An entry belongs to the two tensor
# objects themselves (weak references: a tensor allocated later at a freed
# one's address, or under its id, never matches) and is remade when the
# optimizer has written either tensor (its version counter moved).
_FILTER_W: Dict[tuple, tuple] = {}


def _filter_weights(p: Dict[str, Tensor], pre: str):
    W1o, W2o = p[pre + ".0.weight"], p[pre + ".2.weight"]
    key = (id(W1o), id(W2o))
    state = (W1o.data_ptr(), W2o.data_ptr(), W1o._version, W2o._version)
    ent = _FILTER_W.get(key)
    same = ent is not None and ent[0]() is W1o and ent[1]() is W2o
    if same and ent[2] == state:
        return ent[3], ent[4], ent[5]
    W1, W2 = _dense(W1o), _dense(W2o)
    H, R = W1.shape
    RP = 16 if R <= 16 else 32
    if len(_FILTER_W) > 64:
        for k in [k for k, e in _FILTER_W.items() if e[0]() is None or e[1]() is None]:
            del _FILTER_W[k]
    if (same and ent[3].shape == (H, RP) and ent[4].numel() == 3 * W2.numel()
            and ent[5].shape == W2.T.shape):
        # the same tensors, written by the optimizer: refill the buffers in place
        # (W1's padding columns stay zero)
        w1p, planes, w2t = ent[3], ent[4], ent[5]
    else:
        w1p = torch.zeros(H, RP, dtype=torch.float32, device=W1.device)
        planes = torch.empty(3 * W2.numel(), dtype=torch.int16, device=W2.device)
        w2t = torch.empty(W2.shape[1], W2.shape[0], dtype=torch.float32, device=W2.device)
    w1p[:, :R].copy_(W1)
    w2t.copy_(W2.T)
    C.check("mlff_split_bf16x3", C.lib().mlff_split_bf16x3(
        C.stream_handle(), W2.numel(), C.ptr(W2), C.ptr(planes)))
    _FILTER_W[key] = (weakref.ref(W1o), weakref.ref(W2o), state, w1p, planes, w2t)
    return w1p, planes, w2t


def _filter(p: Dict[str, Tensor], pre: str, x: Tensor, a: Optional[Tensor] = None):
    """The filter MLP on the fused pair kernel (mlff_train_filter).  Forward
    (a None): x = e gives (a1, h, W) with a1 = e W1^T + b1, h = silu(a1),
    W = h W2^T + b2.  Tangent: x = de and a = a1 give (da1, dh, dW) with
    da1 = de W1^T, dh = silu'(a1) da1, dW = dh W2^T."""
    w1p, planes, _ = _filter_weights(p, pre)
    x = _dense(x)
    P, R = x.shape
    H = p[pre + ".0.weight"].shape[0]
    if w1p.shape[0] != H or planes.numel() != 3 * p[pre + ".2.weight"].numel():
        raise RuntimeError("prepared filter weights do not match the layer")
    a_out = torch.empty(P, H, dtype=torch.float32, device=x.device)
    h_out = torch.empty_like(a_out)
    w_out = torch.empty(P, 3 * H, dtype=torch.float32, device=x.device)
    tan = a is not None
    b1 = None if tan else _dense(p[pre + ".0.bias"])
    b2 = None if tan else _dense(p[pre + ".2.bias"])
    C.check("mlff_train_filter", C.lib().mlff_train_filter(
        C.stream_handle(), int(tan), H, P, R, C.ptr(x), C.ptr(w1p), C.ptr(b1), C.ptr(planes),
        C.ptr(b2), C.ptr(_dense(a) if tan else None), C.ptr(a_out), C.ptr(h_out),
        C.ptr(w_out)))
    return a_out, h_out, w_out


def _filter_adjoint(p: Dict[str, Tensor], pre: str, gW: Tensor, a1: Tensor,
                    dgW: Optional[Tensor] = None, da1: Optional[Tensor] = None):
    """Reverse of the filter's first layer on the fused pair kernel
    (mlff_train_filter_adjoint): ga1 = silu'(a1) (gW W2) and, with the
    tangent (dgW, da1), dga1 = silu'(a1) (dgW W2) + silu''(a1) da1 (gW W2).
    Returns (ga1, dga1 or None)."""
    _, _, w2t = _filter_weights(p, pre)
    gW, a1 = _dense(gW), _dense(a1)
    P, H = a1.shape
    if w2t.shape != (H, 3 * H) or gW.shape != (P, 3 * H):
        raise RuntimeError("filter adjoint operands do not match the layer")
    tan = dgW is not None
    ga = torch.empty(P, H, dtype=torch.float32, device=a1.device)
    dga = torch.empty_like(ga) if tan else None
    C.check("mlff_train_filter_adjoint", C.lib().mlff_train_filter_adjoint(
        C.stream_handle(), int(tan), H, P, C.ptr(gW), C.ptr(_dense(dgW) if tan else None),
        C.ptr(w2t), C.ptr(a1), C.ptr(_dense(da1) if tan else None), C.ptr(ga), C.ptr(dga)))
    return ga, dga


def _edge_features(d, centers, widths, rc):
    """e_k(d) = exp(-(d-mu_k)^2 / w_k^2) * 0.5 (cos(pi d/rc) + 1) [d < rc]
    (student_model.py:249-253, 285-290, 720-722) and de/dd."""
    diff = d.unsqueeze(-1) - centers
    gamma = 1.0 / widths ** 2
    phi = torch.exp(-gamma * diff ** 2)
    inside = (d < rc).to(d.dtype).unsqueeze(-1)
    arg = math.pi * d.unsqueeze(-1) / rc
    fc = 0.5 * (torch.cos(arg) + 1.0) * inside
    dfc = -0.5 * (math.pi / rc) * torch.sin(arg) * inside
    return phi * fc, (-2.0 * gamma * diff * phi) * fc + phi * dfc


def _geometry(g: PairGraph, x: Tensor, centers: Tensor, widths: Tensor, rc: float,
              lam: Optional[Tensor] = None):
    """Pair geometry and basis on one HIP launch (mlff_train_geometry):
    (vec, d, den, u, e, de) with vec = x[hi] - x[lo] (+ shift), d = |vec|,
    den = d + EPS, u = vec / den, e = _edge_features(d) and de = de/dd; with
    a position direction ``lam``, the tangent (dd, du, de_t) instead."""
    P = g.lo32.numel()
    R = centers.numel()
    dev = x.device
    f = lambda *shape: torch.empty(*shape, dtype=torch.float32, device=dev)
    shift = None if g.shift is None else _dense(g.shift.to(torch.float32))
    if lam is None:
        outs = (f(P, 3), f(P), f(P), f(P, 3), f(P, R), f(P, R))
        ptrs = [C.ptr(t) for t in outs] + [None] * 3
    else:
        outs = (f(P), f(P, 3), f(P, R))
        ptrs = [None] * 6 + [C.ptr(t) for t in outs]
    C.check("mlff_train_geometry", C.lib().mlff_train_geometry(
        C.stream_handle(), P, R, C.ptr(_dense(x)), C.ptr(None if lam is None else _dense(lam)),
        C.ptr(g.lo32), C.ptr(g.hi32), C.ptr(shift), C.ptr(_dense(centers)),
        C.ptr(_dense(widths)), float(rc), EPS, *ptrs))
    return outs


def _edge_feature_param_derivs(d, centers, widths, rc):
    """de_k/dmu_k, de_k/dw_k and their d-derivatives, each [P, R]
    (learnable_rbf, student_model.py:226-236): with phi = exp(-(d-mu)^2/w^2),
    dphi/dmu = 2 gamma (d-mu) phi and dphi/dw = 2 (d-mu)^2 phi / w^3."""
    diff = d.unsqueeze(-1) - centers
    gamma = 1.0 / widths ** 2
    phi = torch.exp(-gamma * diff ** 2)
    inside = (d < rc).to(d.dtype).unsqueeze(-1)
    arg = math.pi * d.unsqueeze(-1) / rc
    fc = 0.5 * (torch.cos(arg) + 1.0) * inside
    dfc = -0.5 * (math.pi / rc) * torch.sin(arg) * inside
    w3 = widths ** 3
    A = 2.0 * gamma * diff * phi
    Bw = 2.0 * diff ** 2 * phi / w3
    dA = 2.0 * gamma * phi * (1.0 - 2.0 * gamma * diff ** 2)
    dBw = (2.0 * diff * phi / w3) * (2.0 - 2.0 * gamma * diff ** 2)
    return fc * A, fc * Bw, dfc * A + fc * dA, dfc * Bw + fc * dBw


def _wg_operand(t: Tensor) -> Tensor:
    """A [K, C] float32 operand of mlff_weight_grad: unit column stride, row
    stride and C multiples of 4 (zero columns appended), 16-byte aligned."""
    t = t.reshape(t.shape[0], math.prod(t.shape[1:]))
    if t.shape[1] % 4:
        t = torch.nn.functional.pad(t, (0, 4 - t.shape[1] % 4))
    if t.stride(1) != 1 or t.stride(0) % 4:
        t = t.contiguous()
    if t.data_ptr() % 16:
        t = t.clone()
    return t


_WG_WS: Dict[tuple, int] = {}   # workspace floats of mlff_weight_grad by (K, M, N)
_WG_S: Dict[tuple, int] = {}    # partial sums of mlff_weight_grad by (K, M, N)


class _WgGroup:
    """Weight gradients whose split-K partial sums are added later, all in one
    launch (mlff_weight_grad_reduce_group), instead of one launch each.  The
    outputs _wgrad returns with a group hold their values only after
    flush(); the partial buffers are kept alive until then."""

    def __init__(self):
        self.rows = []
        self.keep = []

    def add(self, S: int, n4: int, part: int, out: int, buf: Tensor):
        self.rows.append((S, n4, part, out))
        self.keep.append(buf)

    def flush(self):
        if not self.rows:
            return
        lib = C.lib()
        dev = self.keep[0].device
        for i in range(0, len(self.rows), 256):   # at most 256 rows per launch
            rows = self.rows[i:i + 256]
            table = torch.tensor(rows, dtype=torch.int64, pin_memory=True).to(
                dev, non_blocking=True)
            n_blocks = sum((n4 + 255) // 256 for _, n4, _, _ in rows)
            C.check("mlff_weight_grad_reduce_group", lib.mlff_weight_grad_reduce_group(
                C.stream_handle(), len(rows), table.data_ptr(), min(n_blocks, 1 << 16)))
        self.rows, self.keep = [], []


def _wgrad(a0: Tensor, b0: Tensor, a1: Optional[Tensor] = None, b1: Optional[Tensor] = None,
           w: Optional[Tensor] = None, bias: bool = False, group: Optional[_WgGroup] = None):
    """sum_k w_k (a0[k] (x) b0[k] + a1[k] (x) b1[k]) [M, N] and, with
    ``bias``, sum_k w_k a0[k] [M] over the K rows (pairs or atoms): a Linear
    layer's weight and bias gradients, on the MFMA kernel with a fixed split
    and order of the row sum (mlff_weight_grad).  Returns (weight, bias or
    None).  With ``group`` only the partial sums are launched here; the
    outputs are filled, with the same bits, by group.flush()."""
    M, N = a0.shape[1], b0.shape[1]
    K = a0.shape[0]
    A0, B0 = _wg_operand(a0), _wg_operand(b0)
    A1 = B1 = None
    if a1 is not None:
        A1, B1 = _wg_operand(a1), _wg_operand(b1)
    M4, N4 = A0.shape[1], B0.shape[1]
    lib = C.lib()
    key = (K, M4, N4)
    ws_floats = _WG_WS.get(key)
    if ws_floats is None:
        ws_floats = (int(lib.mlff_weight_grad_workspace_bytes(K, M4, N4)) + 15) // 16 * 4
        _WG_WS[key] = ws_floats
    # one allocation: workspace, then the weight gradient, then the bias
    buf = torch.empty(ws_floats + M4 * N4 + M4, dtype=torch.float32, device=a0.device)
    out = buf[ws_floats:ws_floats + M4 * N4].view(M4, N4)
    bout = buf[ws_floats + M4 * N4:] if bias else None
    wv = None
    if w is not None:
        wv = w if (w.dtype == torch.float32 and w.is_contiguous()) else w.to(torch.float32).contiguous()
    operands = (C.stream_handle(), K, M4, N4, A0.data_ptr(), A0.stride(0), B0.data_ptr(),
                B0.stride(0), C.ptr(A1), A1.stride(0) if A1 is not None else 0,
                C.ptr(B1), B1.stride(0) if B1 is not None else 0, C.ptr(wv))
    if group is None:
        C.check("mlff_weight_grad", lib.mlff_weight_grad(
            *operands, buf.data_ptr(), ws_floats * 4, out.data_ptr(), C.ptr(bout)))
    else:
        C.check("mlff_weight_grad_partials", lib.mlff_weight_grad_partials(
            *operands, int(bias), buf.data_ptr(), ws_floats * 4))
        S = _WG_S.get(key)
        if S is None:
            S = _WG_S[key] = int(lib.mlff_weight_grad_partial_count(K, M4, N4))
        group.add(S, M4 * N4 // 4, buf.data_ptr(), out.data_ptr(), buf)
        if bout is not None:
            group.add(S, M4 // 4, buf.data_ptr() + S * M4 * N4 * 4, bout.data_ptr(), buf)
    if M4 != M or N4 != N:
        out = out[:M, :N]
    if bout is not None and M4 != M:
        bout = bout[:M]
    return out, bout


def _lin(x, p, name):
    return torch.addmm(p[name + ".bias"], x, p[name + ".weight"].T)


def _xyz_outer(a, b, a1=None, b1=None, w=None, group: Optional[_WgGroup] = None):
    """sum_n w_n sum_h (a[n,x,h] b[n,z,h] + a1[n,x,h] b1[n,z,h]) -> [3,3]
    (gradient of a 3x3 mixing matrix), on one partial-sum kernel
    (mlff_xyz_outer_partials) whose partials the weight-gradient reduce adds:
    at ``group``'s flush when one is given, else at once."""
    n, _, H = a.shape
    lib = C.lib()
    S = int(lib.mlff_xyz_outer_partial_count(n))
    buf = torch.empty(12 * S + 12, dtype=torch.float32, device=a.device)
    opt = lambda t: None if t is None else _dense(t)
    wv = None if w is None else _dense(w.to(torch.float32))
    C.check("mlff_xyz_outer_partials", lib.mlff_xyz_outer_partials(
        C.stream_handle(), n, H, C.ptr(_dense(a)), C.ptr(_dense(b)), C.ptr(opt(a1)),
        C.ptr(opt(b1)), C.ptr(wv), buf.data_ptr()))
    out = buf[12 * S:]
    grp = group if group is not None else _WgGroup()
    grp.add(S, 3, buf.data_ptr(), out.data_ptr(), buf)
    if group is None:
        grp.flush()
    return out[:9].view(3, 3)


# ---------------------------------------------------------------------------
# forward (with saved activations) and its tangent
# ---------------------------------------------------------------------------

def forward(p: Dict[str, Tensor], Z: Tensor, x: Tensor, g: PairGraph, rc: float,
            lam: Optional[Tensor] = None):
    """Primal pass (student_model.py:634-757); returns (E [B] float64, cache).
    With ``lam`` [N,3] the tangent along lam is computed too."""
    H = p["embedding.weight"].shape[1]
    L = len({int(k.split(".")[1]) for k in p if k.startswith("interactions.")})
    n = Z.shape[0]
    x = _dense(x.to(torch.float32))
    vec, d, den, u, e, de = _geometry(g, x, p["rbf.centers"], p["rbf.widths"], rc)
    s = p["embedding.weight"][Z].contiguous()
    v = torch.zeros(n, 3, H, dtype=torch.float32, device=x.device)
    layers = []
    for k in range(L):
        pre = f"interactions.{k}.message.rbf_to_scalar"
        a1, h, W = _filter(p, pre, e)
        s_m, v_m, _, _ = _message(g, H, s, v, W, u)
        upre = f"interactions.{k}.update"
        # PaiNNUpdate.forward (student_model.py:430-470)
        xin = torch.empty(n, 2 * H, dtype=torch.float32, device=x.device)
        _glue("mlff_update_norm", n, H, s_m, v_m, xin)
        a2 = _lin(xin, p, upre + ".update_mlp.0")
        h2 = _silu(a2)
        y = _lin(h2, p, upre + ".update_mlp.2")
        M = _dense(p[upre + ".mixing_matrix"])
        s_new = torch.empty_like(s_m)
        v_new = torch.empty_like(v_m)
        Mv = torch.empty_like(v_m)
        _glue("mlff_update_apply", n, H, M, y, s_m, v_m, s_new, v_new, Mv)
        layers.append(dict(s=s, v=v, a1=a1, h=h, W=W, s_m=s_m, v_m=v_m, xin=xin, a2=a2, h2=h2,
                           y=y, Mv=Mv, M=M))
        s, v = s_new, v_new
    r1 = _lin(s, p, "energy_head.0")
    h3 = _silu(r1)
    r2 = _lin(h3, p, "energy_head.2")
    h4 = _silu(r2)
    eps = _lin(h4, p, "energy_head.4")[:, 0]
    E = _segment_sum(g.mol_ptr, eps, f64=True)
    cache = dict(H=H, L=L, n=n, x=x, vec=vec, d=d, den=den, u=u, e=e, de=de, layers=layers,
                 sL=s, r1=r1, h3=h3, r2=r2, h4=h4, Z=Z, rc=rc, lam=None)
    if lam is not None:
        tangent(p, cache, g, lam)
    return E, cache


def tangent(p: Dict[str, Tensor], c: dict, g: PairGraph, lam: Tensor):
    """Tangent of the forward pass along the position direction lam [N,3],
    stored in the cache (the message tangents on the HIP row kernel)."""
    H, n = c["H"], c["n"]
    lam32 = lam.to(torch.float32)
    dd, du, de_t = _geometry(g, c["x"], p["rbf.centers"], p["rbf.widths"], c["rc"], lam32)
    vec = c["vec"]
    ds_ = torch.zeros(n, H, dtype=torch.float32, device=vec.device)
    dv_ = torch.zeros(n, 3, H, dtype=torch.float32, device=vec.device)
    for k, lay in enumerate(c["layers"]):
        pre = f"interactions.{k}.message.rbf_to_scalar"
        lay["ds"], lay["dv"] = ds_, dv_
        da1, dh, dW = _filter(p, pre, de_t, lay["a1"])
        lay["da1"], lay["dh"], lay["dW"] = da1, dh, dW
        _, _, dsm, dvm = _message(g, H, lay["s"], lay["v"], lay["W"], c["u"], ds_, dv_, dW, du)
        v_m = lay["v_m"]
        dxin = torch.empty(n, 2 * H, dtype=torch.float32, device=vec.device)
        _glue("mlff_update_norm_tangent", n, H, v_m, lay["xin"], dsm, dvm, dxin)
        upre = f"interactions.{k}.update"
        da2 = dxin @ p[upre + ".update_mlp.0.weight"].T
        dh2 = _silu_d(lay["a2"], da2)
        dy = dh2 @ p[upre + ".update_mlp.2.weight"].T
        ds_ = torch.empty_like(dsm)
        dv_ = torch.empty_like(dvm)
        dMv = torch.empty_like(dvm)
        _glue("mlff_update_apply_tangent", n, H, lay["M"], lay["y"], dy, dsm, v_m, dvm,
              lay["Mv"], ds_, dv_, dMv)
        lay.update(dsm=dsm, dvm=dvm, dxin=dxin, da2=da2, dh2=dh2, dy=dy, dMv=dMv)
    dr1 = ds_ @ p["energy_head.0.weight"].T
    dh3 = _silu_d(c["r1"], dr1)
    dr2 = dh3 @ p["energy_head.2.weight"].T
    dh4 = _silu_d(c["r2"], dr2)
    c.update(lam=lam, dd=dd, du=du, de_t=de_t, dsL=ds_, dr1=dr1, dh3=dh3, dr2=dr2, dh4=dh4)


# ---------------------------------------------------------------------------
# reverse pass (parameter gradients), optionally with its tangent
# ---------------------------------------------------------------------------

# element segments of the last element tensor seen: a training loop over one
# batch passes the same tensor every step (weak reference: a tensor made later
# at the same address never matches; the version counter catches writes)
_Z_SEG: Dict[str, tuple] = {}


def _z_segments(Z: Tensor, n_rows: int):
    """Atoms ordered by element (stable sort) and the element segments, for
    the fixed-order embedding-gradient sum."""
    ent = _Z_SEG.get("last")
    if ent is not None and ent[0]() is Z and ent[1] == (Z._version, n_rows):
        return ent[2], ent[3]
    zs, order = torch.sort(Z, stable=True)
    # first position of each element in the sorted list (no host sync)
    ptr = torch.searchsorted(zs, torch.arange(n_rows + 1, dtype=zs.dtype, device=Z.device))
    ptr, order = ptr.to(torch.int32), order.to(torch.int32)
    _Z_SEG["last"] = (weakref.ref(Z), (Z._version, n_rows), ptr, order)
    return ptr, order


def reverse(p: Dict[str, Tensor], c: dict, g: PairGraph, gE: Tensor,
            lam: Optional[Tensor] = None, gw: Optional[Tensor] = None,
            rbf_grads: bool = False):
    """Parameter gradients of sum_b gw_b E_b (gw defaults to gE) and -- with
    the tangent along ``lam`` [N,3] (reused from the cache when forward()
    already computed it for the same lam) -- the parameter gradients of
    < lam, grad_x (sum_b gE_b E_b) >.  Returns (grads, tgrads or None).

    gw is applied as a per-row weight on the gE adjoint (structures do not
    interact), so gE must then be all ones.  ``rbf_grads`` adds the basis
    centres and widths (learnable_rbf)."""
    H, L, n = c["H"], c["L"], c["n"]
    T = lam is not None
    if T and c.get("lam") is not lam:
        tangent(p, c, g, lam)
    grads: Dict[str, Tensor] = {}
    tg: Dict[str, Tensor] = {}
    dev = c["d"].device

    # the weight gradients' partial sums are added by one launch at the end
    grp = _WgGroup()

    def acc(store, name, val):
        if name in store:
            grp.flush()   # both terms must hold their values
            store[name] = store[name] + val
        else:
            store[name] = val

    gE = gE.to(torch.float32)
    w_atom = w_pair = None
    if gw is not None:
        gw = gw.to(torch.float32)
        w_atom = gw[g.batch]
        w_pair = w_atom[g.lo]
        wa = lambda t: t * w_atom.view(-1, *([1] * (t.dim() - 1)))
        wp = lambda t: t * w_pair.view(-1, *([1] * (t.dim() - 1)))
    else:
        wa = wp = lambda t: t

    # ---- readout reverse ------------------------------------------------------
    geps = gE[g.batch]
    r1, h3, r2, h4, sL = c["r1"], c["h3"], c["r2"], c["h4"], c["sL"]
    Q3 = p["energy_head.4.weight"]
    wgeps = wa(geps)
    acc(grads, "energy_head.4.weight", (wgeps @ h4).reshape(1, -1))
    acc(grads, "energy_head.4.bias", wgeps.sum().reshape(1))
    gh4 = geps.unsqueeze(1) * Q3
    R2w, R1w = p["energy_head.2.weight"], p["energy_head.0.weight"]
    if T:
        # each adjoint is stacked with its tangent: one product for both
        dr1, dh3, dr2, dh4 = c["dr1"], c["dh3"], c["dr2"], c["dh4"]
        GR2 = _silu_pair(r2, gh4, dr2)
        gr2, dgr2 = GR2[0], GR2[1]
        GH3 = GR2.view(2 * n, -1) @ R2w
        gh3, dgh3 = GH3[:n], GH3[n:]
        GR1 = _silu_pair(r1, gh3, dr1, dgh3)
        gr1, dgr1 = GR1[0], GR1[1]
        GS = GR1.view(2 * n, -1) @ R1w
        gs, dgs = GS[:n], GS[n:]
    else:
        gr2 = _silu_d(r2, gh4)
        gr1 = _silu_d(r1, gr2 @ R2w)
        gs = (gr1 @ R1w).contiguous()
    gw_, gb_ = _wgrad(gr2, h3, w=w_atom, bias=True, group=grp)
    acc(grads, "energy_head.2.weight", gw_)
    acc(grads, "energy_head.2.bias", gb_)
    gw_, gb_ = _wgrad(gr1, sL, w=w_atom, bias=True, group=grp)
    acc(grads, "energy_head.0.weight", gw_)
    acc(grads, "energy_head.0.bias", gb_)
    gv = torch.zeros(n, 3, H, dtype=torch.float32, device=dev)
    if T:
        acc(tg, "energy_head.4.weight", (geps @ dh4).reshape(1, -1))
        acc(tg, "energy_head.4.bias", torch.zeros(1, dtype=torch.float32, device=dev))
        gw_, gb_ = _wgrad(dgr2, h3, gr2, dh3, bias=True, group=grp)
        acc(tg, "energy_head.2.weight", gw_)
        acc(tg, "energy_head.2.bias", gb_)
        gw_, gb_ = _wgrad(dgr1, sL, gr1, c["dsL"], bias=True, group=grp)
        acc(tg, "energy_head.0.weight", gw_)
        acc(tg, "energy_head.0.bias", gb_)
        dgv = torch.zeros_like(gv)

    u = c["u"]
    if rbf_grads:
        e_mu, e_w, de_mu, de_w = _edge_feature_param_derivs(c["d"], p["rbf.centers"],
                                                           p["rbf.widths"], c["rc"])
    for k in range(L - 1, -1, -1):
        lay = c["layers"][k]
        upre = f"interactions.{k}.update"
        mpre = f"interactions.{k}.message.rbf_to_scalar"
        U1 = p[upre + ".update_mlp.0.weight"]
        U2 = p[upre + ".update_mlp.2.weight"]
        M, v_m, Mv, y, xin = lay["M"], lay["v_m"], lay["Mv"], lay["y"], lay["xin"]
        # update reverse (student_model.py:434-470), gates and their tangent
        # the adjoint and its tangent stacked ([2, n, ...]): one product for both
        YY = torch.empty(2 if T else 1, n, 3 * H, dtype=torch.float32, device=dev)
        gy, dgy = YY[0], (YY[1] if T else None)
        dvm, dMv, dxin = (lay["dvm"], lay["dMv"], lay["dxin"]) if T else (None, None, None)
        _glue("mlff_update_gate_adjoint", n, H, gs, gv, v_m, Mv, dgs if T else None,
              dgv if T else None, dvm, dMv, gy, dgy)
        if T:
            GH2 = YY.view(2 * n, 3 * H) @ U2
            GA2 = _silu_pair(lay["a2"], GH2[:n], lay["da2"], GH2[n:])
            ga2, dga2 = GA2[0], GA2[1]
            GX = GA2.view(2 * n, -1) @ U1
            gx, dgx = GX[:n], GX[n:]
        else:
            ga2 = _silu_d(lay["a2"], gy @ U2)
            gx, dgx = ga2 @ U1, None
        gw_, gb_ = _wgrad(gy, lay["h2"], w=w_atom, bias=True, group=grp)
        acc(grads, upre + ".update_mlp.2.weight", gw_)
        acc(grads, upre + ".update_mlp.2.bias", gb_)
        gw_, gb_ = _wgrad(ga2, xin, w=w_atom, bias=True, group=grp)
        acc(grads, upre + ".update_mlp.0.weight", gw_)
        acc(grads, upre + ".update_mlp.0.bias", gb_)
        if T:
            gw_, gb_ = _wgrad(dgy, lay["h2"], gy, lay["dh2"], bias=True, group=grp)
            acc(tg, upre + ".update_mlp.2.weight", gw_)
            acc(tg, upre + ".update_mlp.2.bias", gb_)
            gw_, gb_ = _wgrad(dga2, xin, ga2, dxin, bias=True, group=grp)
            acc(tg, upre + ".update_mlp.0.weight", gw_)
            acc(tg, upre + ".update_mlp.0.bias", gb_)
        gs_m = torch.empty(n, H, dtype=torch.float32, device=dev)
        gv_m = torch.empty(n, 3, H, dtype=torch.float32, device=dev)
        gvg2 = torch.empty_like(gv_m)
        dgs_m = torch.empty_like(gs_m) if T else None
        dgv_m = torch.empty_like(gv_m) if T else None
        dgvg2 = torch.empty_like(gv_m) if T else None
        _glue("mlff_update_input_adjoint", n, H, M, gx, gs, gv, y, v_m, xin, dgx,
              dgs if T else None, dgv if T else None, lay["dy"] if T else None, dvm, dxin,
              gs_m, gv_m, gvg2, dgs_m, dgv_m, dgvg2)
        acc(grads, upre + ".mixing_matrix", _xyz_outer(gvg2, v_m, w=w_atom, group=grp))
        if T:
            acc(tg, upre + ".mixing_matrix", _xyz_outer(dgvg2, v_m, gvg2, dvm, group=grp))
        # message reverse (student_model.py:346-387): filter-output gradients per pair
        s, v, W = lay["s"], lay["v"], lay["W"]
        if T:
            gW, dgW = _pair_grad(g, H, u, s, v, gs_m, gv_m, c["du"], lay["ds"], lay["dv"],
                                 dgs_m, dgv_m)
        else:
            gW, _ = _pair_grad(g, H, u, s, v, gs_m, gv_m)
        W1 = p[mpre + ".0.weight"]
        gw_, gb_ = _wgrad(gW, lay["h"], w=w_pair, bias=True, group=grp)
        acc(grads, mpre + ".2.weight", gw_)
        acc(grads, mpre + ".2.bias", gb_)
        # gW W2 through the SiLU (and its tangent) on the fused pair kernel
        ga1, dga1 = _filter_adjoint(p, mpre, gW, lay["a1"], dgW if T else None,
                                    lay["da1"] if T else None)
        gw_, gb_ = _wgrad(ga1, c["e"], w=w_pair, bias=True, group=grp)
        acc(grads, mpre + ".0.weight", gw_)
        acc(grads, mpre + ".0.bias", gb_)
        if rbf_grads:
            ge = ga1 @ W1
            acc(grads, "rbf.centers", (wp(ge) * e_mu).sum(0))
            acc(grads, "rbf.widths", (wp(ge) * e_w).sum(0))
        if T:
            dh = lay["dh"]
            gw_, gb_ = _wgrad(dgW, lay["h"], gW, dh, bias=True, group=grp)
            acc(tg, mpre + ".2.weight", gw_)
            acc(tg, mpre + ".2.bias", gb_)
            gw_, gb_ = _wgrad(dga1, c["e"], ga1, c["de_t"], bias=True, group=grp)
            acc(tg, mpre + ".0.weight", gw_)
            acc(tg, mpre + ".0.bias", gb_)
            if rbf_grads:
                dge = dga1 @ W1
                dd = c["dd"].unsqueeze(-1)
                acc(tg, "rbf.centers", (dge * e_mu + ge * de_mu * dd).sum(0))
                acc(tg, "rbf.widths", (dge * e_w + ge * de_w * dd).sum(0))
            gs, gv, dgs, dgv = _message_adjoint(g, H, gs_m, gv_m, W, dgs_m, dgv_m, lay["dW"])
        else:
            gs, gv, _, _ = _message_adjoint(g, H, gs_m, gv_m, W)
    emb_rows = p["embedding.weight"].shape[0]
    zptr, zorder = _z_segments(c["Z"], emb_rows)
    acc(grads, "embedding.weight", _segment_sum(zptr, wa(gs), zorder))
    if T:
        acc(tg, "embedding.weight", _segment_sum(zptr, dgs, zorder))
    grp.flush()
    return grads, (tg if T else None)


# ---------------------------------------------------------------------------
# loss
# ---------------------------------------------------------------------------

def force_matching_grads(p: Dict[str, Tensor], Z: Tensor, x: Tensor, g: PairGraph, rc: float,
                         E_label: Tensor, F_label: Tensor, forces: Tensor,
                         w_energy: float = 1.0, w_force: float = 100.0,
                         rbf_grads: bool = False):
    """loss = wE MSE(E) + wF MSE(F) (train_compact_models.py:340-342) and its
    exact gradient with respect to every parameter in ``p``.  ``forces``
    are the model's forces at ``x`` (from the fused evaluation)."""
    F = forces.to(torch.float32)
    eF = F - F_label.to(torch.float32)
    dLdF = w_force * 2.0 * eF / eF.numel()
    E, cache = forward(p, Z, x, g, rc, lam=dLdF)
    B = E.shape[0]
    eE = E - E_label.to(torch.float64)
    loss = w_energy * (eE ** 2).mean() + w_force * (eF.double() ** 2).mean()
    dLdE = (w_energy * 2.0 * eE / B).to(torch.float32)
    # one reverse pass gives both parts:
    #   d/dtheta of wE MSE(E): the unit adjoint's terms weighted by dL/dE (gw)
    #   d/dtheta of wF MSE(F) = - d/dtheta < dL/dF, grad_x sum E >
    gE, tg = reverse(p, cache, g, torch.ones(B, dtype=torch.float32, device=x.device),
                     lam=dLdF, gw=dLdE, rbf_grads=rbf_grads)
    grads = {}
    both, A, B = [], [], []
    for name in p:
        a = gE.get(name)
        b = tg.get(name)
        if a is None and b is None:
            continue
        if a is not None and b is not None:
            both.append(name)
            A.append(a)
            B.append(b)
        else:
            grads[name] = a if b is None else torch.neg(b)
    # one multi-tensor launch for every parameter with both parts
    if both:
        for name, d in zip(both, torch._foreach_sub(A, B)):
            grads[name] = d
    return loss, E, F, grads


class ForceMatchingTrainer:
    """AdamW force matching on the hand-derived gradients.

    The loop mirrors the compact-student recipe
    (scripts/train_compact_models.py:302-372): loss wE MSE(E) + wF MSE(F),
    zero_grad, backward, gradient-norm clipping, optimizer step; the
    backward is force_matching_grads instead of autograd.
    """

    def __init__(self, model, lr: float = 1e-3, weight_decay: float = 0.0,
                 w_energy: float = 1.0, w_force: float = 100.0, grad_clip: float = 1.0):
        self.model = model
        self.rbf_grads = any(n.startswith("rbf.") for n, _ in model.named_parameters())
        self.w_energy, self.w_force, self.grad_clip = w_energy, w_force, grad_clip
        self.opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=weight_decay)
        self._p = None
        self._p_sig = None

    def _params(self) -> Dict[str, Tensor]:
        """Detached views of the model's state (the optimizer updates the
        tensors in place, so the views stay current); rebuilt when a tensor
        has been replaced (model.to(), load_state_dict(assign=True))."""
        sig = tuple(t.data_ptr() for t in self.model.parameters()) + tuple(
            t.data_ptr() for t in self.model.buffers())
        if self._p is None or sig != self._p_sig:
            self._p = {k: v.detach() for k, v in self.model.state_dict().items()}
            self._p_sig = sig
        return self._p

    def loss_and_grads(self, Z, x, sizes, E_label, F_label, cells=None, pbc=None):
        x = x.detach().to(torch.float32).contiguous()
        g = device_graph(Z, x, sizes, self.model.cutoff, cells, pbc)
        p = self._params()
        # forces of the current weights from the fused HIP evaluation.  The
        # gradient pass is enqueued behind it before its neighbour-capacity
        # check is read, so the host does not wait for the evaluation; if the
        # list overflowed, the pass runs again on the regrown evaluation.
        finish = self.model.evaluate_async(Z, x, np.asarray(sizes), cells, pbc)
        with torch.no_grad():
            res = force_matching_grads(p, Z, x, g, self.model.cutoff, E_label, F_label,
                                       finish.pending[1], self.w_energy, self.w_force,
                                       rbf_grads=self.rbf_grads)
            done = finish()
            if done is not finish.pending:
                res = force_matching_grads(p, Z, x, g, self.model.cutoff, E_label, F_label,
                                           done[1], self.w_energy, self.w_force,
                                           rbf_grads=self.rbf_grads)
        return res

    def step(self, Z, x, sizes, E_label, F_label, cells=None, pbc=None) -> float:
        loss, E, F, grads = self.loss_and_grads(Z, x, sizes, E_label, F_label, cells, pbc)
        for name, prm in self.model.named_parameters():
            gr = grads.get(name)
            prm.grad = torch.zeros_like(prm) if gr is None else gr.to(prm.dtype).reshape(prm.shape)
        if self.grad_clip and self.grad_clip > 0:
            torch.nn.utils.clip_grad_norm_(self.model.parameters(), self.grad_clip)
        self.opt.step()
        return float(loss)
