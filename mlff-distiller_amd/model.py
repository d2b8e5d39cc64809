"""Drop-in StudentForceField whose evaluation runs on the gfx950 kernels.

Interface kept from src/mlff_distiller/models/student_model.py:
  * constructor arguments and attributes (:566-583), submodule tree and
    state-dict keys (embedding, rbf.centers/widths, interactions.k.message.
    rbf_to_scalar.{0,2}, interactions.k.update.update_mlp.{0,2},
    interactions.k.update.mixing_matrix, energy_head.{0,2,4}), and the order
    in which __init__ consumes the random stream (:585-628), so
    ``torch.manual_seed(s); StudentForceField()`` yields the reference's
    weights;
  * forward(atomic_numbers, positions, cell, pbc, batch) -> energy (0-d, or
    [B] when batch has more than one structure) (:634-757), differentiable
    with respect to positions, so ``-autograd.grad(E, positions)`` gives
    forces exactly as the reference is used;
  * predict_energy_and_forces (:759-795), forward_with_analytical_forces
    (:797-932), num_parameters, save/load checkpoint format (:1075-1176).

Differences, by design: the whole evaluation (neighbour list, message
passing, readout, and forces from the hand-written adjoint) is one pipeline
of HIP kernels; periodic cells are honoured (the reference ignores
cell/pbc, :694-703); structure energies are accumulated in float64 and
returned in the positions' dtype.  Tensors must live on a ROCm device: there
is no CPU path.
"""

from __future__ import annotations

import logging
from collections import OrderedDict
from pathlib import Path
from typing import Optional, Tuple, Union

import numpy as np
import torch
import torch.nn as nn

from . import engine

logger = logging.getLogger(__name__)


class GaussianRBF(nn.Module):
    """Gaussian radial basis (student_model.py:196-255); centers/widths buffers
    (parameters when ``learnable``).  Its forward is a convenience for
    inspecting the basis; StudentForceField evaluates the basis inside its
    kernels."""

    def __init__(self, num_rbf: int = 20, cutoff: float = 5.0, learnable: bool = False):
        super().__init__()
        self.num_rbf = num_rbf
        self.cutoff = cutoff
        centers = torch.linspace(0, cutoff, num_rbf)
        widths = torch.ones(num_rbf) * (cutoff / num_rbf)
        if learnable:
            self.centers = nn.Parameter(centers)
            self.widths = nn.Parameter(widths)
        else:
            self.register_buffer("centers", centers)
            self.register_buffer("widths", widths)

    def forward(self, distances: torch.Tensor) -> torch.Tensor:
        diff = distances.unsqueeze(-1) - self.centers
        return torch.exp(-(1.0 / self.widths ** 2) * diff ** 2)


class CosineCutoff(nn.Module):
    """Cosine cutoff (student_model.py:258-292), strict ``d < cutoff``."""

    def __init__(self, cutoff: float = 5.0):
        super().__init__()
        self.cutoff = cutoff

    def forward(self, distances: torch.Tensor) -> torch.Tensor:
        f = 0.5 * (torch.cos(np.pi * distances / self.cutoff) + 1.0)
        return f * (distances < self.cutoff).to(distances.dtype)


def _edge_csr(key: torch.Tensor, other: torch.Tensor, n: int):
    """Edges grouped by ``key`` (stable, so each group keeps the caller's
    edge order): row_ptr [n+1], the other endpoint of each edge, its index."""
    order = torch.sort(key, stable=True).indices
    row_ptr = torch.zeros(n + 1, dtype=torch.int32, device=key.device)
    row_ptr[1:] = torch.cumsum(torch.bincount(key, minlength=n), 0).to(torch.int32)
    return row_ptr, other[order].to(torch.int32).contiguous(), order.to(torch.int32).contiguous()


def _device_f32(*ts):
    for t in ts:
        if not t.is_cuda or t.dtype != torch.float32:
            raise TypeError("the interaction blocks run on the HIP kernels: float32 tensors on "
                            "a ROCm device are required")


class _EdgeMessage(torch.autograd.Function):
    """The gathers of PaiNNMessage.forward (student_model.py:357-387) on a
    caller-given directed edge list: the receiver-row HIP kernel forward; the
    backward groups the edges by source for the feature gradients and takes
    the filter and edge-vector gradients per edge (all on HIP kernels, fixed
    order, no atomics)."""

    @staticmethod
    def forward(ctx, s, v, W, ev, src, dst):
        from . import _capi as C
        _device_f32(s, v, W, ev)
        N, H = s.shape
        s, v, W, ev = s.contiguous(), v.contiguous(), W.contiguous(), ev.contiguous()
        src, dst = src.long(), dst.long()
        if src.numel() and (int(src.min()) < 0 or int(src.max()) >= N or int(dst.min()) < 0
                            or int(dst.max()) >= N):
            raise ValueError("edge_index refers to atoms outside [0, num_atoms)")
        row_ptr, col, eid = _edge_csr(dst, src, N)
        so, vo = torch.empty_like(s), torch.empty_like(v)
        C.check("mlff_edge_message", C.lib().mlff_edge_message(
            C.stream_handle(), H, N, C.ptr(row_ptr), C.ptr(col), C.ptr(eid), C.ptr(s), C.ptr(v),
            C.ptr(W), C.ptr(ev), C.ptr(so), C.ptr(vo)))
        ctx.save_for_backward(s, v, W, ev, src, dst)
        return so, vo

    @staticmethod
    def backward(ctx, gso, gvo):
        from . import _capi as C
        s, v, W, ev, src, dst = ctx.saved_tensors
        N, H = s.shape
        E = src.numel()
        gso = torch.zeros_like(s) if gso is None else gso.contiguous().float()
        gvo = torch.zeros_like(v) if gvo is None else gvo.contiguous().float()
        row_ptr, col, eid = _edge_csr(src, dst, N)
        gs, gv = torch.empty_like(s), torch.empty_like(v)
        C.check("mlff_train_message_adjoint", C.lib().mlff_train_message_adjoint(
            C.stream_handle(), H, N, C.ptr(row_ptr), C.ptr(col), C.ptr(eid), C.ptr(gso),
            C.ptr(gvo), C.ptr(W), None, None, None, C.ptr(gs), C.ptr(gv), None, None))
        gW = torch.empty(E, 3 * H, dtype=torch.float32, device=s.device)
        gev = torch.empty(E, 3, dtype=torch.float32, device=s.device)
        s32, d32 = src.to(torch.int32).contiguous(), dst.to(torch.int32).contiguous()
        C.check("mlff_edge_grad", C.lib().mlff_edge_grad(
            C.stream_handle(), H, E, C.ptr(s32), C.ptr(d32), C.ptr(s), C.ptr(v), C.ptr(W),
            C.ptr(ev), C.ptr(gso), C.ptr(gvo), C.ptr(gW), C.ptr(gev)))
        return gs, gv, gW, gev, None, None


class PaiNNMessage(nn.Module):
    """Message block (student_model.py:297-387).  StudentForceField evaluates
    it inside its fused pipeline; its own forward takes a caller-given
    directed edge list, as the reference's, and runs on the HIP kernels."""

    def __init__(self, hidden_dim: int, num_rbf: int):
        super().__init__()
        self.hidden_dim = hidden_dim
        self.rbf_to_scalar = nn.Sequential(nn.Linear(num_rbf, hidden_dim), nn.SiLU(),
                                           nn.Linear(hidden_dim, hidden_dim * 3))

    def forward(self, scalar_features: torch.Tensor, vector_features: torch.Tensor,
                edge_index: torch.Tensor, edge_rbf: torch.Tensor,
                edge_vector: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """s_i + sum_{j->i} s_j Ws, v_i + sum_{j->i} (v_j Wv1 + e_ji (x) Wv2)
        with (Ws, Wv1, Wv2) = rbf_to_scalar(edge_rbf) per edge."""
        filt = self.rbf_to_scalar(edge_rbf)
        return _EdgeMessage.apply(scalar_features, vector_features, filt, edge_vector,
                                  edge_index[0], edge_index[1])


class PaiNNUpdate(nn.Module):
    """Update block (student_model.py:390-470): n = |v| over xyz,
    (ds, g1, g2) = update_mlp([s, n]), s + ds and v g1 + (M v) g2 (no
    residual on v; the 3x3 mixing acts on the spatial axis)."""

    def __init__(self, hidden_dim: int):
        super().__init__()
        self.hidden_dim = hidden_dim
        self.update_mlp = nn.Sequential(nn.Linear(hidden_dim * 2, hidden_dim), nn.SiLU(),
                                        nn.Linear(hidden_dim, hidden_dim * 3))
        self.mixing_matrix = nn.Parameter(torch.randn(3, 3) / np.sqrt(3))

    def forward(self, scalar_features: torch.Tensor,
                vector_features: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        _device_f32(scalar_features, vector_features)
        H = self.hidden_dim
        norms = torch.linalg.vector_norm(vector_features, dim=1)
        upd = self.update_mlp(torch.cat([scalar_features, norms], dim=-1))
        ds, g1, g2 = upd[:, :H], upd[:, H:2 * H], upd[:, 2 * H:]
        mixed = torch.matmul(self.mixing_matrix, vector_features)
        return (scalar_features + ds,
                vector_features * g1.unsqueeze(1) + mixed * g2.unsqueeze(1))


class PaiNNInteraction(nn.Module):
    """Message then update (student_model.py:473-527)."""

    def __init__(self, hidden_dim: int, num_rbf: int):
        super().__init__()
        self.message = PaiNNMessage(hidden_dim, num_rbf)
        self.update = PaiNNUpdate(hidden_dim)

    def forward(self, scalar_features: torch.Tensor, vector_features: torch.Tensor,
                edge_index: torch.Tensor, edge_rbf: torch.Tensor,
                edge_vector: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        s, v = self.message(scalar_features, vector_features, edge_index, edge_rbf, edge_vector)
        return self.update(s, v)


class _EnergyFunction(torch.autograd.Function):
    """E(positions) from the fused pipeline; dE/dx = -F from the adjoint
    computed in the same launch sequence."""

    @staticmethod
    def forward(ctx, positions, runner, batch_index):
        need = ctx.needs_input_grad[0]
        energy, forces = runner(positions.detach(), need)
        ctx.batch_index = batch_index
        if need:
            ctx.save_for_backward(forces)
        return energy

    @staticmethod
    def backward(ctx, grad_energy):
        (forces,) = ctx.saved_tensors
        if grad_energy.dim() == 0:
            g = grad_energy.to(forces.dtype)
            return -forces * g, None, None
        g = grad_energy.to(forces.dtype)[ctx.batch_index].unsqueeze(1)
        return -forces * g, None, None


class _Session:
    """One differentiable evaluation: energies/forces from the kernels, and the
    hand-derived parameter-gradient passes (train.py) evaluated lazily on the
    same positions and neighbour list."""

    def __init__(self, model, Z, pos, sizes, cells, pbc, batch_index, multi, perm):
        self.model, self.Z, self.pos, self.sizes = model, Z, pos, sizes
        self.cells, self.pbc, self.batch_index, self.multi, self.perm = cells, pbc, batch_index, multi, perm
        self.forces = None
        self._prep = None

    def run(self):
        e64, f, _ = self.model.evaluate(self.Z, self.pos, self.sizes, self.cells, self.pbc,
                                        need_forces=True)
        self.model.last_energy_f64 = e64
        if self.perm is not None:
            fo = torch.empty_like(f)
            fo[self.perm] = f
            f = fo
        self.forces = f
        e = e64.to(self.pos.dtype)
        return e if self.multi else e.reshape(())

    def _prepared(self):
        if self._prep is None:
            from . import train
            p = {k: v.detach() for k, v in self.model.state_dict().items()}
            g = train.device_graph(self.Z, self.pos, self.sizes, self.model.cutoff, self.cells,
                                   self.pbc)
            _, cache = train.forward(p, self.Z, self.pos.detach(), g, self.model.cutoff)
            self._prep = (p, g, cache)
        return self._prep

    def _to_sorted(self, t):
        return t[self.perm] if self.perm is not None else t

    def param_grads(self, gE):
        from . import train
        p, g, c = self._prepared()
        grads, _ = train.reverse(p, c, g, gE.reshape(-1), rbf_grads=self._rbf_grads())
        return grads

    def _rbf_grads(self) -> bool:
        return any(n.startswith("rbf.") for n, _ in self.model.named_parameters())

    def tangent_grads(self, gE, lam):
        from . import train
        p, g, c = self._prepared()
        _, tg = train.reverse(p, c, g, gE.reshape(-1), lam=self._to_sorted(lam),
                              rbf_grads=self._rbf_grads())
        return tg


def _grads_tuple(names, grads, params):
    out = []
    for n, prm in zip(names, params):
        gr = grads.get(n)
        out.append(None if gr is None else gr.to(prm.dtype).reshape(prm.shape))
    return tuple(out)


class _TrainableEnergy(torch.autograd.Function):
    """E(positions, parameters): first derivatives from the adjoint (positions)
    and the hand-derived reverse pass (parameters); when the position gradient
    is itself differentiated (create_graph), _EnergyGradient supplies the
    hand-derived second-order pass."""

    @staticmethod
    def forward(ctx, positions, session, names, *params):
        ctx.session, ctx.names = session, names
        return session.run()

    @staticmethod
    def backward(ctx, gE):
        s = ctx.session
        gE_vec = gE.reshape(-1)
        g_pos = None
        if ctx.needs_input_grad[0]:
            if torch.is_grad_enabled():
                g_pos = _EnergyGradient.apply(gE_vec, s, ctx.names, *[
                    p for _, p in s.model.named_parameters()])
            else:
                g = gE_vec.to(s.forces.dtype)
                g = g[s.batch_index] if s.batch_index is not None else g.expand(s.forces.shape[0])
                g_pos = -s.forces * g.unsqueeze(1)
        pg = (None,) * len(ctx.names)
        if any(ctx.needs_input_grad[3:]):
            params = [p for _, p in s.model.named_parameters()]
            pg = _grads_tuple(ctx.names, s.param_grads(gE_vec.detach()), params)
        return (g_pos, None, None) + pg


class _EnergyGradient(torch.autograd.Function):
    """grad_x (sum_b gE_b E_b) as a function of gE and the parameters."""

    @staticmethod
    def forward(ctx, gE, session, names, *params):
        ctx.session, ctx.names = session, names
        ctx.save_for_backward(gE)
        g = gE.to(session.forces.dtype)
        g = g[session.batch_index] if session.batch_index is not None else g.expand(
            session.forces.shape[0])
        return -session.forces * g.unsqueeze(1)

    @staticmethod
    def backward(ctx, gg):
        (gE,) = ctx.saved_tensors
        s = ctx.session
        # d/dgE_b of < gg, grad_x E_b >
        dot = -(gg * s.forces).sum(1)
        if s.batch_index is not None:
            g_gE = torch.zeros_like(gE).index_add_(0, s.batch_index, dot.to(gE.dtype))
        else:
            g_gE = dot.sum().reshape(gE.shape).to(gE.dtype)
        params = [p for _, p in s.model.named_parameters()]
        tg = _grads_tuple(ctx.names, s.tangent_grads(gE.detach(), gg.detach()), params)
        return (g_gE, None, None) + tg


class StudentForceField(nn.Module):
    """PaiNN student force field (student_model.py:532-1176) on MI355X."""

    def __init__(self, hidden_dim: int = 128, num_interactions: int = 3, num_rbf: int = 20,
                 cutoff: float = 5.0, max_z: int = 118, learnable_rbf: bool = False,
                 use_torch_cluster: bool = True):
        super().__init__()
        self.hidden_dim = hidden_dim
        self.num_interactions = num_interactions
        self.num_rbf = num_rbf
        self.cutoff = cutoff
        self.max_z = max_z
        self.use_torch_cluster = use_torch_cluster
        self.embedding = nn.Embedding(max_z + 1, hidden_dim)
        self.rbf = GaussianRBF(num_rbf, cutoff, learnable_rbf)
        self.cutoff_fn = CosineCutoff(cutoff)
        self.interactions = nn.ModuleList([PaiNNInteraction(hidden_dim, num_rbf)
                                           for _ in range(num_interactions)])
        self.energy_head = nn.Sequential(
            nn.Linear(hidden_dim, hidden_dim // 2), nn.SiLU(),
            nn.Linear(hidden_dim // 2, hidden_dim // 4), nn.SiLU(),
            nn.Linear(hidden_dim // 4, 1))
        self._initialize_parameters()
        self._packed = None
        self._packed_sig = None
        self._z_checked = None
        self._z_checked_version = -1
        self._plans: "OrderedDict[tuple, dict]" = OrderedDict()
        self.last_energy_f64: Optional[torch.Tensor] = None
        logger.info("Initialized StudentForceField: %s parameters, %dD hidden, %d interactions, "
                    "%sA cutoff", f"{self.num_parameters():,}", hidden_dim, num_interactions,
                    cutoff)

    def _initialize_parameters(self):
        """Same scheme and order as student_model.py:618-628."""
        nn.init.uniform_(self.embedding.weight, -np.sqrt(3), np.sqrt(3))
        for module in self.modules():
            if isinstance(module, nn.Linear):
                nn.init.xavier_uniform_(module.weight)
                if module.bias is not None:
                    nn.init.zeros_(module.bias)

    def num_parameters(self) -> int:
        return sum(p.numel() for p in self.parameters() if p.requires_grad)

    # ------------------------------------------------------------------
    # device plumbing
    # ------------------------------------------------------------------
    def packed_weights(self, device) -> torch.Tensor:
        device = engine.device_of(device)
        sd = self.state_dict()
        sig = tuple((t.data_ptr(), t._version) for t in sd.values())
        if self._packed is None or self._packed_sig != sig or self._packed.device != device:
            # a fresh blob: plans bound elsewhere (an MD run) keep the weights they
            # were bound with, consistent with their split bf16 copies
            self._packed = engine.pack_weights(sd, device)
            self._packed_sig = sig
            for entry in self._plans.values():
                entry["plan"].set_weights(self._packed)
        return self._packed

    def _plan_for(self, sizes: np.ndarray, device, periodic: bool, cells, pbc,
                  positions: Optional[torch.Tensor] = None) -> dict:
        key = (tuple(int(x) for x in sizes), periodic, str(device))
        entry = self._plans.get(key)
        if entry is not None:
            self._plans.move_to_end(key)
            if periodic:
                lay = entry["layout"]
                c32 = np.asarray(cells, dtype=np.float32).reshape(-1, 3, 3)
                pb = np.asarray(pbc, dtype=bool).reshape(-1, 3)
                if not (np.array_equal(c32, lay.cells_host) and np.array_equal(pb, lay.pbc_host)):
                    lay.cells.copy_(torch.as_tensor(c32))
                    lay.pbc.copy_(torch.as_tensor(pb.astype(np.uint8)))
                    lay.cells_host, lay.pbc_host = c32, pb
                    plan = entry["plan"]
                    grid, n_cells = engine.cell_grid(lay, self.cutoff)
                    cap = int(plan.cell_capacity)
                    if n_cells <= cap:
                        plan.cell_grid = grid
                        plan._apply_cell_grid()
                    else:
                        entry["plan"] = engine.Plan(
                            self.hidden_dim, self.num_interactions, self.num_rbf, self.max_z,
                            self.cutoff, lay, entry["numbers"], self.packed_weights(device),
                            plan.ecap)
            return entry
        layout = engine.make_layout(sizes, device, cells if periodic else None,
                                    pbc if periodic else None)
        weights = self.packed_weights(device)
        numbers = torch.zeros(layout.n_atoms, dtype=torch.int64, device=device)
        cap = engine.edge_capacity_bound(sizes, periodic)
        if positions is not None:
            # size the edge capacity from the first positions (25 % headroom)
            # instead of the worst case, which for large structures reserves
            # far more filter rows than exist; a later call that needs more
            # regrows the plan (evaluate)
            edges = engine.count_edges(positions, layout, self.cutoff)
            cap = min(cap, max(int(edges * 1.25) + 64, 16))
        plan = engine.Plan(self.hidden_dim, self.num_interactions, self.num_rbf, self.max_z,
                           self.cutoff, layout, numbers, weights, cap)
        entry = {"layout": layout, "plan": plan, "numbers": numbers}
        self._plans[key] = entry
        while len(self._plans) > 8:
            self._plans.popitem(last=False)
        return entry

    def _grow(self, key_entry: dict, sizes, device, periodic, cells, pbc, edges: int):
        key = (tuple(int(x) for x in sizes), periodic, str(device))
        layout = key_entry["layout"]
        plan = engine.Plan(self.hidden_dim, self.num_interactions, self.num_rbf, self.max_z,
                           self.cutoff, layout, key_entry["numbers"], self.packed_weights(device),
                           int(edges * 1.25) + 16)
        key_entry["plan"] = plan
        self._plans[key] = key_entry
        return key_entry

    def _check_numbers(self, atomic_numbers: torch.Tensor):
        """Element numbers must index the embedding table: one host read of
        their range, skipped for the tensor checked last while it has not
        been modified (same object, same version counter)."""
        if (atomic_numbers is self._z_checked and
                atomic_numbers._version == self._z_checked_version):
            return
        if atomic_numbers.numel():
            lo, hi = torch.aminmax(atomic_numbers)
            zmin, zmax = (int(v) for v in torch.stack([lo, hi]).tolist())
        else:
            zmin = zmax = 1
        if zmin < 0 or zmax > self.max_z:
            raise ValueError(f"atomic numbers must lie in [0, {self.max_z}]")
        self._z_checked = atomic_numbers
        self._z_checked_version = atomic_numbers._version

    def _empty_results(self, sizes, device, need_forces, need_virial):
        # no atoms: every sum is empty, as in the reference's forward
        # (student_model.py:734-757), so no plan is needed
        B = max(len(sizes), 1)
        return (torch.zeros(B, dtype=torch.float64, device=device),
                torch.zeros(0, 3, dtype=torch.float32, device=device) if need_forces else None,
                torch.zeros(B, 3, 3, dtype=torch.float64, device=device) if need_virial else None)

    def _prepare(self, atomic_numbers, positions, sizes, cells, pbc):
        device = positions.device
        if device.type != "cuda":
            raise RuntimeError("StudentForceField (MI355X build) needs tensors on a ROCm device; "
                               f"got {device}")
        periodic = pbc is not None and bool(np.asarray(pbc).any())
        if periodic:
            engine.check_cells(cells, pbc, self.cutoff)
        self._check_numbers(atomic_numbers)
        return device, periodic

    def _launch(self, entry, pos, B, need_forces, need_virial):
        device = pos.device
        plan = entry["plan"]
        N = pos.shape[0]
        energy = torch.empty(B, dtype=torch.float64, device=device)
        forces = torch.empty(N, 3, dtype=torch.float32, device=device) if need_forces else None
        virial = torch.empty(B, 9, dtype=torch.float64, device=device) if need_virial else None
        plan.counts[C_OVERFLOW].zero_()
        plan.energy_forces(pos, energy, forces, virial)
        return energy, forces, (virial.view(B, 3, 3) if virial is not None else None)

    def evaluate(self, atomic_numbers: torch.Tensor, positions: torch.Tensor,
                 sizes: np.ndarray, cells=None, pbc=None, need_forces: bool = True,
                 need_virial: bool = False):
        """Run the fused pipeline on a flattened, structure-contiguous batch.

        Returns (energy float64 [B], forces float32 [N,3] or None,
        virial float64 [B,3,3] or None).
        """
        device, periodic = self._prepare(atomic_numbers, positions, sizes, cells, pbc)
        if positions.shape[0] == 0:
            return self._empty_results(sizes, device, need_forces, need_virial)
        pos = positions.detach().to(torch.float32).contiguous()
        entry = self._plan_for(sizes, device, periodic, cells, pbc, pos)
        self.packed_weights(device)
        entry["numbers"].copy_(atomic_numbers.to(torch.int64))
        for _attempt in range(2):
            out = self._launch(entry, pos, len(sizes), need_forces, need_virial)
            counts = entry["plan"].read_counts()
            if counts[C_OVERFLOW] == 0:
                return out
            entry = self._grow(entry, sizes, device, periodic, cells, pbc,
                               int(counts[C_EDGES]))
        raise RuntimeError("neighbour-list capacity overflow persisted after regrowth")

    def evaluate_async(self, atomic_numbers: torch.Tensor, positions: torch.Tensor,
                       sizes: np.ndarray, cells=None, pbc=None, need_forces: bool = True,
                       need_virial: bool = False):
        """evaluate() enqueued on the current stream without waiting for it.

        Returns a function that completes the call: it waits for the
        evaluation's capacity check only (the counters travel to pinned host
        memory behind the kernels) and returns (energy, forces, virial) as
        evaluate() does.  If the neighbour list overflowed its capacity the
        enqueued results are discarded and evaluate() runs again with a
        regrown plan.  The results are ordered on the stream that was current
        here: other streams wait for it before reading them.  The enqueued
        results are also the function's ``pending`` attribute, for work that
        is enqueued behind them before the check (finish() returns that same
        tuple when the check passes, and new results when it does not).
        """
        device, periodic = self._prepare(atomic_numbers, positions, sizes, cells, pbc)
        if positions.shape[0] == 0:
            out = self._empty_results(sizes, device, need_forces, need_virial)

            def done():
                return out
            done.pending = out
            return done
        pos = positions.detach().to(torch.float32).contiguous()
        entry = self._plan_for(sizes, device, periodic, cells, pbc, pos)
        self.packed_weights(device)
        entry["numbers"].copy_(atomic_numbers.to(torch.int64))
        out = self._launch(entry, pos, len(sizes), need_forces, need_virial)
        # pinned (torch caches the pages): the copy stays asynchronous
        host = torch.empty(entry["plan"].counts.numel(), dtype=torch.int32, pin_memory=True)
        host.copy_(entry["plan"].counts, non_blocking=True)
        ready = torch.cuda.Event()
        ready.record()

        def finish():
            ready.synchronize()
            if int(host[C_OVERFLOW]) == 0:
                return out
            return self.evaluate(atomic_numbers, positions, sizes, cells, pbc, need_forces,
                                 need_virial)
        finish.pending = out
        return finish

    # ------------------------------------------------------------------
    # reference API
    # ------------------------------------------------------------------
    @staticmethod
    def _batch_sizes(batch: Optional[torch.Tensor], n: int):
        if batch is None:
            return np.array([n], dtype=np.int64), None, None, False
        b = batch.detach().to("cpu", torch.int64).numpy()
        nstruct = int(b.max()) + 1 if b.size else 1
        if b.size and np.any(b[1:] < b[:-1]):
            order = np.argsort(b, kind="stable")
        else:
            order = None
        sizes = np.bincount(b, minlength=nstruct).astype(np.int64)
        return sizes, order, b, nstruct > 1 or False

    def forward(self, atomic_numbers: torch.Tensor, positions: torch.Tensor,
                cell: Optional[torch.Tensor] = None, pbc: Optional[torch.Tensor] = None,
                batch: Optional[torch.Tensor] = None) -> torch.Tensor:
        n = atomic_numbers.shape[0]
        if n == 0:
            # no atoms: the reference's sums over atoms are empty and give 0
            # (student_model.py:734-757); differentiable in the positions
            return positions.sum() * 0.0
        sizes, order, bnp, multi = self._batch_sizes(batch, n)
        cells_np, pbc_np = self._cell_args(cell, pbc, len(sizes))
        Z = atomic_numbers
        perm = None
        if order is not None:
            perm = torch.as_tensor(order, device=positions.device)
            Z = atomic_numbers[perm]
        batch_index = None
        if multi:
            batch_index = torch.as_tensor(bnp, device=positions.device)

        params = [p for _, p in self.named_parameters()]
        if torch.is_grad_enabled() and any(p.requires_grad for p in params):
            # differentiable in the parameters (training): hand-derived passes
            pos_sorted = positions.detach()[perm] if perm is not None else positions.detach()
            session = _Session(self, Z, pos_sorted, sizes, cells_np, pbc_np, batch_index, multi,
                               perm)
            names = tuple(n_ for n_, _ in self.named_parameters())
            return _TrainableEnergy.apply(positions, session, names, *params)

        def runner(pos_detached, need):
            p = pos_detached[perm] if perm is not None else pos_detached
            e64, f, _ = self.evaluate(Z, p, sizes, cells_np, pbc_np, need_forces=need)
            self.last_energy_f64 = e64
            if f is not None and perm is not None:
                fo = torch.empty_like(f)
                fo[perm] = f
                f = fo
            e = e64.to(positions.dtype)
            if not multi:
                e = e.reshape(())
            return e, f

        return _EnergyFunction.apply(positions, runner, batch_index)

    @staticmethod
    def _cell_args(cell, pbc, nstruct):
        if cell is None or pbc is None:
            return None, None
        pbc_np = np.asarray(pbc.detach().cpu() if torch.is_tensor(pbc) else pbc, dtype=bool)
        if not pbc_np.any():
            return None, None
        cell_np = np.asarray(cell.detach().cpu() if torch.is_tensor(cell) else cell,
                             dtype=np.float64)
        cell_np = np.broadcast_to(cell_np.reshape(-1, 3, 3), (nstruct, 3, 3))
        pbc_np = np.broadcast_to(pbc_np.reshape(-1, 3), (nstruct, 3))
        return np.ascontiguousarray(cell_np), np.ascontiguousarray(pbc_np)

    def predict_energy_and_forces(self, atomic_numbers: torch.Tensor, positions: torch.Tensor,
                                  cell: Optional[torch.Tensor] = None,
                                  pbc: Optional[torch.Tensor] = None
                                  ) -> Tuple[torch.Tensor, torch.Tensor]:
        """Energy and forces (student_model.py:759-795) in one fused launch."""
        if self.training and torch.is_grad_enabled():
            positions = positions.requires_grad_(True)
            energy = self.forward(atomic_numbers, positions, cell, pbc)
            forces = -torch.autograd.grad(energy, positions, create_graph=True,
                                          retain_graph=True)[0]
            return energy, forces
        n = atomic_numbers.shape[0]
        cells_np, pbc_np = self._cell_args(cell, pbc, 1)
        e64, f, _ = self.evaluate(atomic_numbers, positions, np.array([n]), cells_np, pbc_np)
        self.last_energy_f64 = e64
        return e64.to(positions.dtype).reshape(()), f.to(positions.dtype)

    def forward_with_analytical_forces(self, atomic_numbers: torch.Tensor,
                                       positions: torch.Tensor,
                                       cell: Optional[torch.Tensor] = None,
                                       pbc: Optional[torch.Tensor] = None,
                                       batch: Optional[torch.Tensor] = None
                                       ) -> Tuple[torch.Tensor, torch.Tensor]:
        """Energy and forces from the hand-written adjoint (student_model.py:797-932
        names this path; here it really is analytical, with no autograd)."""
        n = atomic_numbers.shape[0]
        sizes, order, _, multi = self._batch_sizes(batch, n)
        cells_np, pbc_np = self._cell_args(cell, pbc, len(sizes))
        Z, p = atomic_numbers, positions
        if order is not None:
            perm = torch.as_tensor(order, device=positions.device)
            Z, p = atomic_numbers[perm], positions[perm]
        e64, f, _ = self.evaluate(Z, p, sizes, cells_np, pbc_np)
        if order is not None:
            fo = torch.empty_like(f)
            fo[perm] = f
            f = fo
        e = e64.to(positions.dtype)
        return (e if multi else e.reshape(())), f.to(positions.dtype)

    def energy_forces_virial(self, atomic_numbers: torch.Tensor, positions: torch.Tensor,
                             cell, pbc, sizes: Optional[np.ndarray] = None):
        """Energy (float64), forces and virial dE/d(strain) (float64 [B,3,3])."""
        n = atomic_numbers.shape[0]
        sizes = np.array([n]) if sizes is None else np.asarray(sizes)
        cells_np, pbc_np = self._cell_args(cell, pbc, len(sizes))
        return self.evaluate(atomic_numbers, positions, sizes, cells_np, pbc_np, True, True)

    # ------------------------------------------------------------------
    # checkpoints (student_model.py:1075-1176)
    # ------------------------------------------------------------------
    def save(self, path: Union[str, Path]):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        torch.save({
            "model_state_dict": self.state_dict(),
            "config": {"hidden_dim": self.hidden_dim, "num_interactions": self.num_interactions,
                       "num_rbf": self.num_rbf, "cutoff": self.cutoff, "max_z": self.max_z,
                       "use_torch_cluster": self.use_torch_cluster},
            "num_parameters": self.num_parameters(),
        }, path)

    @classmethod
    def load(cls, path: Union[str, Path], device: str = "cpu"):
        checkpoint = torch.load(path, map_location=device, weights_only=True)
        config = dict(checkpoint.get("config", {}))
        config.setdefault("use_torch_cluster", True)
        keep = {"hidden_dim", "num_interactions", "num_rbf", "cutoff", "max_z", "learnable_rbf",
                "use_torch_cluster"}
        cfg = {k: v for k, v in config.items() if k in keep}
        sd = checkpoint["model_state_dict"]
        if "hidden_dim" not in cfg and "embedding.weight" in sd:
            cfg["hidden_dim"] = sd["embedding.weight"].shape[1]
        if "num_rbf" not in cfg and "rbf.centers" in sd:
            cfg["num_rbf"] = sd["rbf.centers"].shape[0]
        if "num_interactions" not in cfg:
            ks = {int(k.split(".")[1]) for k in sd if k.startswith("interactions.")}
            if ks:
                cfg["num_interactions"] = len(ks)
        cfg.setdefault("cutoff", 5.0)
        if "max_z" not in cfg and "embedding.weight" in sd:
            cfg["max_z"] = sd["embedding.weight"].shape[0] - 1
        model = cls(**cfg)
        model.load_state_dict(sd)
        model.to(device)
        return model


C_EDGES = 0
C_OVERFLOW = 2

__all__ = ["StudentForceField", "PaiNNInteraction", "PaiNNMessage", "PaiNNUpdate",
           "GaussianRBF", "CosineCutoff"]
