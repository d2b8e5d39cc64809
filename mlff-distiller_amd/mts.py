"""Multiple-time-step MD with a learned inner force (no counterpart in the
reference; its MD is plain Velocity Verlet, nve_harness.py:214).

Impulse (r-RESPA) splitting F_full = F_fast + (F_full - F_fast):

  p += (Dt/2) (F_full - F_fast)(x)                          outer half kick
  n times:  p += (dt/2) F_fast ; x += dt p/m ; F_fast(x) ; p += (dt/2) F_fast
  p += (Dt/2) (F_full - F_fast)(x)                          outer half kick

with dt = 0.5 fs inner steps on the compact student and Dt = n dt = 2 fs
outer corrections from the full model.  The conserved quantity is
E_full + E_kin, which the FP64 ledger records every outer step.

Fallback: after every outer step the per-structure disagreement
max_i |F_full,i - F_fast,i| is recorded.  A chunk in which any structure
exceeds the threshold is replayed from its saved state with the full model
as that structure's inner force (so its slow force vanishes and it follows
plain full-model Velocity Verlet at dt); the flag is re-evaluated at the end
of every chunk.

Training data: frames and labels from the full model's own MD
(md_frames), and a compact student fitted with the hand-derived
force-matching gradients (train_student).
"""

from __future__ import annotations

import time
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import engine, synth
from . import _capi as C


# ---------------------------------------------------------------------------
# training data and the compact student
# ---------------------------------------------------------------------------

class FrameSet:
    """Molecule frames (positions, energy, forces) cut from batched MD."""

    def __init__(self, numbers: np.ndarray, sizes: np.ndarray, device):
        self.numbers = torch.as_tensor(numbers, dtype=torch.int64, device=device)
        self.sizes = np.asarray(sizes)
        self.mol_ptr = np.concatenate([[0], np.cumsum(self.sizes)])
        self.pos: List[torch.Tensor] = []
        self.forces: List[torch.Tensor] = []
        self.energy: List[torch.Tensor] = []

    def add(self, pos32, forces, energy):
        self.pos.append(pos32.detach().clone())
        self.forces.append(forces.detach().clone())
        self.energy.append(energy.detach().to(torch.float32).clone())

    @property
    def n_frames(self) -> int:
        return len(self.pos) * len(self.sizes)

    def batch(self, rng: np.random.Generator, n_mol: int):
        """A training batch of n_mol random (frame, molecule) pairs."""
        f = rng.integers(0, len(self.pos), size=n_mol)
        m = rng.integers(0, len(self.sizes), size=n_mol)
        idx = [torch.arange(self.mol_ptr[j], self.mol_ptr[j + 1], device=self.numbers.device)
               for j in m]
        Z = torch.cat([self.numbers[i] for i in idx])
        x = torch.cat([self.pos[fi][i] for fi, i in zip(f, idx)])
        F = torch.cat([self.forces[fi][i] for fi, i in zip(f, idx)])
        E = torch.stack([self.energy[fi][mj] for fi, mj in zip(f, m)])
        return Z, x, self.sizes[m], E, F


def md_frames(full_model, structures, n_frames: int, stride: int, seed: int = 0,
              temperature_K: float = 300.0, timestep_fs: float = 0.5) -> FrameSet:
    """Frames every ``stride`` steps of the full model's own NVE MD."""
    from .md import BatchedNVE
    md = BatchedNVE(full_model, structures, timestep_fs=timestep_fs, temperature_K=temperature_K,
                    seed=seed, max_steps=n_frames * stride + 1)
    fs = FrameSet(md.numbers.cpu().numpy(), md.sizes, md.device)
    fs.add(md.pos32, md.forces, md.energy)
    for _ in range(n_frames - 1):
        md.run(stride, graph=stride >= 5, steps_per_graph=stride if stride >= 5 else 1)
        fs.add(md.pos32, md.forces, md.energy)
    return fs


def train_student(student, frames: FrameSet, steps: int, batch_molecules: int = 64,
                  lr: float = 1e-3, w_energy: float = 1.0, w_force: float = 100.0,
                  seed: int = 0, log_every: int = 0):
    """Fit the compact student on the frames; returns the loss history."""
    from .train import ForceMatchingTrainer
    tr = ForceMatchingTrainer(student, lr=lr, w_energy=w_energy, w_force=w_force)
    rng = np.random.Generator(np.random.PCG64(seed))
    hist = []
    for k in range(steps):
        Z, x, sizes, E, F = frames.batch(rng, batch_molecules)
        hist.append(tr.step(Z, x, sizes, E, F))
        if log_every and (k % log_every == 0 or k == steps - 1):
            print(f"[train] step {k} loss {hist[-1]:.5f}", flush=True)
    return hist


# ---------------------------------------------------------------------------
# the integrator
# ---------------------------------------------------------------------------

class MTSNVE:
    """r-RESPA NVE: compact student inside, full model outside, with fallback."""

    def __init__(self, full_model, fast_model, structures: Sequence["synth.Structure"],
                 inner_fs: float = 0.5, n_inner: int = 4, temperature_K: float = 300.0,
                 seed: int = 0, threshold: float = 2.0, max_outer: int = 20000,
                 headroom: float = 1.3, device="cuda:0", momenta: Optional[np.ndarray] = None):
        dev = torch.device(device)
        self.device = dev
        self.full, self.fast = full_model, fast_model
        if abs(full_model.cutoff - fast_model.cutoff) > 0:
            raise ValueError("inner and outer models must share the cutoff (one neighbour list)")
        self.structures = list(structures)
        numbers, positions, mol_ptr = synth.flatten(self.structures)
        self.sizes = np.diff(mol_ptr)
        self.B, self.N = len(self.sizes), int(mol_ptr[-1])
        self.n_inner = n_inner
        self.dt = inner_fs * synth.FS
        self.inner_fs = inner_fs
        self.threshold = float(threshold)
        masses = synth.masses_of(numbers)
        if momenta is None:
            rng = np.random.Generator(np.random.PCG64(seed))
            momenta = np.concatenate([synth.maxwell_boltzmann(s.numbers, s.positions,
                                                              temperature_K, rng)
                                      for s in self.structures])
        self.layout = engine.make_layout(self.sizes, dev)
        self.batch = self.layout.atom_mol.long()
        self.numbers = torch.tensor(numbers, dtype=torch.int64, device=dev)
        self.x = torch.tensor(positions, dtype=torch.float64, device=dev)
        self.p = torch.tensor(momenta, dtype=torch.float64, device=dev)
        self.inv_m = torch.tensor(1.0 / masses, dtype=torch.float64, device=dev).unsqueeze(1)
        self.pos32 = self.x.to(torch.float32)
        f32 = dict(dtype=torch.float32, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.F_fast = torch.zeros(self.N, 3, **f32)
        self.F_full = torch.zeros(self.N, 3, **f32)
        self.F_in = torch.zeros(self.N, 3, **f32)
        self.E_fast = torch.zeros(self.B, **f64)
        self.E_full = torch.zeros(self.B, **f64)
        self.flag = torch.zeros(self.B, dtype=torch.bool, device=dev)
        self.dis_max = torch.zeros(self.B, **f32)
        self.max_outer = max_outer
        self.ledger = torch.zeros(max_outer + 1, self.B, 2, **f64)
        self.row = torch.ones(1, dtype=torch.int64, device=dev)
        self.headroom = headroom
        row_ptr, col, _ = engine.neighbor_list(self.pos32, self.layout, full_model.cutoff)
        self._make_plans(max(int(col.numel()), 16) * headroom)
        self.stream = torch.cuda.Stream(device=dev)
        self.graph = None
        self.graph_steps = 0
        self.outer_done = 0
        self.fallback_steps = 0
        self.replays = 0
        self._with_stream(self._init_forces)
        self.ledger[0, :, 0] = self.E_full
        self.ledger[0, :, 1] = self._ke()

    # ------------------------------------------------------------------
    def _make_plans(self, edges: float):
        bound = engine.edge_capacity_bound(self.sizes, False)
        cap = int(min(bound, int(edges) + 64))
        self.plan_full = engine.Plan(self.full.hidden_dim, self.full.num_interactions,
                                     self.full.num_rbf, self.full.max_z, self.full.cutoff,
                                     self.layout, self.numbers,
                                     self.full.packed_weights(self.device), cap)
        self.plan_fast = engine.Plan(self.fast.hidden_dim, self.fast.num_interactions,
                                     self.fast.num_rbf, self.fast.max_z, self.fast.cutoff,
                                     self.layout, self.numbers,
                                     self.fast.packed_weights(self.device), cap)
        self.graph = None

    def _with_stream(self, fn, *a):
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            out = fn(*a)
        self.stream.synchronize()
        return out

    def _init_forces(self):
        self.pos32.copy_(self.x)
        self.plan_fast.energy_forces(self.pos32, self.E_fast, self.F_fast)
        self.plan_full.energy_forces(self.pos32, self.E_full, self.F_full)
        self.F_in.copy_(self.F_fast)

    def _ke(self):
        ke = 0.5 * (self.p * self.p * self.inv_m).sum(1)
        return torch.zeros(self.B, dtype=torch.float64, device=self.device).index_add_(
            0, self.batch, ke)

    def _outer_step(self, fallback: bool):
        n, dt = self.n_inner, self.dt
        Dt = n * dt
        flag_atoms = self.flag[self.batch].unsqueeze(1) if fallback else None
        # outer half kick with the slow force F_full - F_in
        self.p.add_((self.F_full - self.F_in).to(torch.float64), alpha=0.5 * Dt)
        for _ in range(n):
            self.p.add_(self.F_in.to(torch.float64), alpha=0.5 * dt)
            self.x.add_(self.p * self.inv_m, alpha=dt)
            self.pos32.copy_(self.x)
            self.plan_fast.energy_forces(self.pos32, self.E_fast, self.F_fast)
            if fallback:
                self.plan_full.energy_forces(self.pos32, self.E_full, self.F_full)
                self.F_in.copy_(torch.where(flag_atoms, self.F_full, self.F_fast))
            else:
                self.F_in.copy_(self.F_fast)
            self.p.add_(self.F_in.to(torch.float64), alpha=0.5 * dt)
        if not fallback:
            self.plan_full.energy_forces(self.pos32, self.E_full, self.F_full)
        self.p.add_((self.F_full - self.F_in).to(torch.float64), alpha=0.5 * Dt)
        # disagreement of the two models at the new positions
        dis = (self.F_full - self.F_fast).norm(dim=1)
        dmax = torch.zeros(self.B, dtype=torch.float32, device=self.device).scatter_reduce(
            0, self.batch, dis, reduce="amax", include_self=True)
        torch.maximum(self.dis_max, dmax, out=self.dis_max)
        rec = torch.stack([self.E_full, self._ke()], dim=1).unsqueeze(0)
        self.ledger.index_copy_(0, self.row, rec)
        self.row.add_(1)

    def _chunk(self, k: int, fallback: bool):
        if fallback:
            for _ in range(k):
                self._outer_step(True)
            return
        if self.graph is None or self.graph_steps != k:
            # warm the allocator on the side stream, then capture k outer steps
            saved = self._save()
            self._with_stream(lambda: [self._outer_step(False) for _ in range(1)])
            self._restore(saved)
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(g, stream=self.stream):
                for _ in range(k):
                    self._outer_step(False)
            self._restore(saved)
            self.graph, self.graph_steps = g, k
        self.graph.replay()

    def _save(self):
        return [t.clone() for t in (self.x, self.p, self.pos32, self.F_fast, self.F_full,
                                    self.F_in, self.E_fast, self.E_full, self.row)]

    def _restore(self, s):
        for dst, src in zip((self.x, self.p, self.pos32, self.F_fast, self.F_full, self.F_in,
                             self.E_fast, self.E_full, self.row), s):
            dst.copy_(src)

    def run(self, n_outer: int, chunk: int = 10, graph: bool = True):
        """Advance n_outer outer steps (n_outer * n_inner inner steps)."""
        if self.outer_done + n_outer > self.max_outer:
            raise ValueError("ledger too short; raise max_outer")
        done = 0
        while done < n_outer:
            k = min(chunk, n_outer - done)
            fallback = bool(self.flag.any())
            # the inner force of this chunk's split: full model for flagged
            # structures (their slow force vanishes), the student elsewhere
            if fallback:
                self.F_in.copy_(torch.where(self.flag[self.batch].unsqueeze(1), self.F_full,
                                            self.F_fast))
            else:
                self.F_in.copy_(self.F_fast)
            saved = self._save()
            self.dis_max.zero_()
            if fallback or not graph:
                self._with_stream(lambda: [self._outer_step(fallback) for _ in range(k)])
            else:
                self._chunk(k, False)
                torch.cuda.synchronize()
            over = (self.plan_full.read_counts()[C.MLFF_COUNT_OVERFLOW] or
                    self.plan_fast.read_counts()[C.MLFF_COUNT_OVERFLOW])
            if over:
                E = int(max(self.plan_full.read_counts()[C.MLFF_COUNT_EDGES],
                            self.plan_fast.read_counts()[C.MLFF_COUNT_EDGES]))
                self._restore(saved)
                self._make_plans(E * max(self.headroom, 1.1))
                continue
            exceed = self.dis_max > self.threshold
            if not fallback and bool(exceed.any()):
                # replay this chunk with the full model inside for those structures
                self._restore(saved)
                self.flag.copy_(exceed)
                self.replays += 1
                continue
            if fallback:
                self.fallback_steps += k * int(self.flag.sum())
                self.flag.copy_(exceed & self.flag)
            done += k
            self.outer_done += k

    # ------------------------------------------------------------------
    def energies(self):
        led = self.ledger[:self.outer_done + 1].cpu().numpy()
        return led[..., 0], led[..., 1]

    def summary(self):
        epot, ekin = self.energies()
        etot = epot + ekin
        t_ps = np.arange(len(etot)) * self.n_inner * self.inner_fs / 1000.0
        rate = np.polyfit(t_ps, etot, 1)[0] / self.sizes if len(t_ps) > 1 else np.zeros(self.B)
        return {"outer_steps": self.outer_done, "inner_steps": self.outer_done * self.n_inner,
                "total_time_ps": float(t_ps[-1]),
                "energy_drift_per_atom": (etot[-1] - etot[0]) / self.sizes,
                "drift_rate_eV_per_atom_ps": rate,
                "max_abs_dev_per_atom": np.abs(etot - etot[0]).max(axis=0) / self.sizes,
                "fallback_structure_steps": self.fallback_steps, "replays": self.replays}
