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
    per_graph = max(d for d in range(1, 26) if stride % d == 0)
    for _ in range(n_frames - 1):
        md.run(stride, graph=per_graph >= 5, steps_per_graph=per_graph)
        fs.add(md.pos32, md.forces, md.energy)
    return fs


def train_student(student, frames: FrameSet, steps: int, batch_molecules: int = 64,
                  lr: float = 1e-3, w_energy: float = 1.0, w_force: float = 100.0,
                  seed: int = 0, log_every: int = 0, lr_final: float = 0.02):
    """Fit the compact student on the frames; returns the loss history.
    The learning rate follows a cosine from ``lr`` down to ``lr_final * lr``."""
    import math
    from .train import ForceMatchingTrainer
    tr = ForceMatchingTrainer(student, lr=lr, w_energy=w_energy, w_force=w_force)
    rng = np.random.Generator(np.random.PCG64(seed))
    hist = []
    for k in range(steps):
        c = 0.5 * (1.0 + math.cos(math.pi * k / max(steps - 1, 1)))
        for group in tr.opt.param_groups:
            group["lr"] = lr * (lr_final + (1.0 - lr_final) * c)
        Z, x, sizes, E, F = frames.batch(rng, batch_molecules)
        hist.append(tr.step(Z, x, sizes, E, F))
        if log_every and (k % log_every == 0 or k == steps - 1):
            print(f"[train] step {k} loss {hist[-1]:.5f}", flush=True)
    return hist


# ---------------------------------------------------------------------------
# the integrator
# ---------------------------------------------------------------------------

class MTSNVE:
    """r-RESPA NVE: compact student inside, full model outside, with fallback.

    The step loop runs in the C library (mlff_mts_run / mlff_mts_graph_run):
    fused kick/drift kernels, the student's neighbour list reused by the
    full model, chunks of outer steps captured in one HIP graph."""

    def __init__(self, full_model, fast_model, structures: Sequence["synth.Structure"],
                 inner_fs: float = 0.5, n_inner: int = 4, temperature_K: float = 300.0,
                 seed: int = 0, threshold: float = 2.0, max_outer: int = 20000,
                 headroom: float = 1.3, device="cuda:0", momenta: Optional[np.ndarray] = None):
        dev = engine.device_of(device)
        self.device = dev
        self.full, self.fast = full_model, fast_model
        if full_model.cutoff != fast_model.cutoff:
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
        self.inv_m = torch.tensor(1.0 / masses, dtype=torch.float64, device=dev)
        self.pos32 = self.x.to(torch.float32)
        f32 = dict(dtype=torch.float32, device=dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.F_full = torch.zeros(self.N, 3, **f32)
        self.F_fast = torch.zeros(self.N, 3, **f32)
        self.F_mix = torch.zeros(self.N, 3, **f32)
        self.E_full = torch.zeros(self.B, **f64)
        self.E_fast = torch.zeros(self.B, **f64)
        self.flag = torch.zeros(self.B, dtype=torch.uint8, device=dev)
        self.dis_max = torch.zeros(self.B, **f32)
        self.max_outer = max_outer
        self.ledger = torch.zeros(max_outer + 1, self.B, 2, **f64)
        self.headroom = headroom
        self.lib = C.lib()
        self.stream = torch.cuda.Stream(device=dev)
        self.outer_done = 0
        self.fallback_steps = 0
        self.replays = 0
        self.regrowths = 0
        row_ptr, col, _ = engine.neighbor_list(self.pos32, self.layout, full_model.cutoff)
        self._make_plans(max(int(col.numel()), 16) * headroom)
        self._on_stream(lambda: C.check("mlff_mts_init_forces", self.lib.mlff_mts_init_forces(
            self.plan_full.handle, self.stream.cuda_stream)))
        self._check_overflow()
        ke = 0.5 * (self.p * self.p).sum(1) * self.inv_m
        self.ledger[0, :, 0] = self.E_full
        self.ledger[0, :, 1] = torch.zeros(self.B, **f64).index_add_(0, self.batch, ke)

    # ------------------------------------------------------------------
    def _make_plans(self, edges: float, step: int = 1):
        bound = engine.edge_capacity_bound(self.sizes, False)
        cap = int(min(bound, int(edges) + 64))
        # one stream per plan: an outer step runs five evaluations, and their
        # fork/join events cost more than the overlap of the filter kernels
        # gains (measured 3.01 against 3.07 ms per outer step on 512 molecules)
        mk = lambda m: engine.Plan(m.hidden_dim, m.num_interactions, m.num_rbf, m.max_z,
                                   m.cutoff, self.layout, self.numbers,
                                   m.packed_weights(self.device), cap, streams=1)
        self.plan_full, self.plan_fast = mk(self.full), mk(self.fast)
        C.check("mlff_mts_bind", self.lib.mlff_mts_bind(
            self.plan_full.handle, self.plan_fast.handle, C.ptr(self.x), C.ptr(self.p),
            C.ptr(self.inv_m), C.ptr(self.pos32), C.ptr(self.F_full), C.ptr(self.F_fast),
            C.ptr(self.F_mix), C.ptr(self.E_full), C.ptr(self.E_fast), C.ptr(self.flag),
            C.ptr(self.dis_max), C.ptr(self.ledger), self.max_outer + 1, float(self.dt),
            self.n_inner))
        self.plan_full.counts[C.MLFF_COUNT_STEP] = step

    def _on_stream(self, fn):
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            fn()
        self.stream.synchronize()

    def _check_overflow(self):
        c = self.plan_fast.read_counts()
        if c[C.MLFF_COUNT_OVERFLOW]:
            # the largest count of the chunk (see BatchedNVE._check_overflow)
            raise OverflowError(int(max(c[C.MLFF_COUNT_EDGES], c[C.MLFF_COUNT_EDGES_MAX])))

    def _state(self):
        return [self.x, self.p, self.pos32, self.F_full, self.F_fast, self.F_mix, self.E_full,
                self.E_fast]

    def run(self, n_outer: int, chunk: int = 10, graph: bool = True):
        """Advance n_outer outer steps (n_outer * n_inner inner steps)."""
        if self.outer_done + n_outer > self.max_outer:
            raise ValueError("ledger too short; raise max_outer")
        done = 0
        while done < n_outer:
            k = min(chunk, n_outer - done)
            fallback = bool(self.flag.any())
            saved = [t.clone() for t in self._state()]
            step = int(self.plan_full.counts[C.MLFF_COUNT_STEP])
            self.dis_max.zero_()
            h, s = self.plan_full.handle, self.stream.cuda_stream
            if fallback or not graph:
                self._on_stream(lambda: C.check("mlff_mts_run", self.lib.mlff_mts_run(
                    h, s, k, int(fallback))))
            else:
                self._on_stream(lambda: C.check("mlff_mts_graph_run",
                                                self.lib.mlff_mts_graph_run(h, s, k, 1)))
            try:
                self._check_overflow()
            except OverflowError as exc:
                for dst, src in zip(self._state(), saved):
                    dst.copy_(src)
                need = int(exc.args[0])
                self.regrowths += 1
                if self.regrowths > 8 or need <= self.plan_fast.ecap:
                    raise RuntimeError(f"neighbour capacity did not converge: {need} edges "
                                       f"needed, capacity {self.plan_fast.ecap}") from exc
                self._make_plans(need * max(self.headroom, 1.1), step)
                continue
            exceed = self.dis_max > self.threshold
            if not fallback and bool(exceed.any()):
                # replay with the full model inside for the offending structures
                for dst, src in zip(self._state(), saved):
                    dst.copy_(src)
                self.plan_full.counts[C.MLFF_COUNT_STEP] = step
                self.flag.copy_(exceed.to(torch.uint8))
                self.replays += 1
                continue
            if fallback:
                self.fallback_steps += k * int(self.flag.sum())
                self.flag.copy_((exceed & self.flag.bool()).to(torch.uint8))
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
