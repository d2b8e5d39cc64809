"""Device-resident NVE molecular dynamics over a flattened batch.

Replaces the reference's per-step host loop (NVEMDHarness.run_simulation ->
ASE VelocityVerlet, src/mlff_distiller/testing/nve_harness.py:173-261,
which crosses host<->device twice per step) with a loop that stays on the
GPU: every step is neighbour list + energy/forces + Velocity Verlet in one
captured HIP graph, replayed; the per-step (E_pot, E_kin) of every
structure accumulates in a float64 ledger on the device.

Summary statistics follow NVEMDHarness._compute_summary
(nve_harness.py:311-360): drift % = 100 (E_T - E_0)/|E_0|, plus absolute
drift per atom and a fitted drift rate, which stay meaningful when E_0 is
near zero.
"""

from __future__ import annotations

import ctypes
import logging
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _capi as C
from . import engine, synth

logger = logging.getLogger(__name__)


class BatchedNVE:
    """Velocity Verlet NVE for many structures at once on one GPU."""

    def __init__(self, model, structures: Sequence["synth.Structure"], timestep_fs: float = 0.5,
                 temperature_K: float = 300.0, seed: int = 0, momenta: Optional[np.ndarray] = None,
                 max_steps: int = 100000, headroom: float = 1.3, device="cuda:0",
                 persistent_blocks: int = 0, streams: Optional[int] = None,
                 record_frames: bool = False, masses: Optional[np.ndarray] = None,
                 frame_window: Optional[int] = None):
        self.model = model
        self.device = engine.device_of(device)
        self.structures = list(structures)
        numbers, positions, mol_ptr = synth.flatten(self.structures)
        self.sizes = np.diff(mol_ptr)
        self.n_atoms = int(mol_ptr[-1])
        self.B = len(self.structures)
        self.dt = timestep_fs * synth.FS
        self.timestep_fs = timestep_fs
        masses = synth.masses_of(numbers) if masses is None else \
            np.asarray(masses, dtype=np.float64).reshape(-1)
        if len(masses) != len(numbers) or not np.all(masses > 0):
            raise ValueError("masses must be positive, one per atom")
        if momenta is None:
            rng = np.random.Generator(np.random.PCG64(seed))
            momenta = np.concatenate([
                synth.maxwell_boltzmann(s.numbers, s.positions, temperature_K, rng,
                                        masses=masses[a:b])
                for s, a, b in zip(self.structures, mol_ptr[:-1], mol_ptr[1:])])
        periodic = any(s.pbc is not None and np.any(s.pbc) for s in self.structures)
        cells = pbc = None
        if periodic:
            cells = np.stack([s.cell if s.cell is not None else np.eye(3) for s in self.structures])
            pbc = np.stack([s.pbc if s.pbc is not None else np.zeros(3, bool)
                            for s in self.structures])
            engine.check_cells(cells, pbc, model.cutoff)
        self.cells, self.pbc = cells, pbc
        dev = self.device
        self.numbers = torch.tensor(numbers, dtype=torch.int64, device=dev)
        self.x = torch.tensor(positions, dtype=torch.float64, device=dev)
        self.p = torch.tensor(momenta, dtype=torch.float64, device=dev)
        self.inv_mass = torch.tensor(1.0 / masses, dtype=torch.float64, device=dev)
        self.masses = masses
        self.pos32 = self.x.to(torch.float32)
        self.forces = torch.zeros(self.n_atoms, 3, dtype=torch.float32, device=dev)
        self.energy = torch.zeros(self.B, dtype=torch.float64, device=dev)
        self.max_steps = int(max_steps)
        self.ledger = torch.zeros(self.max_steps + 1, self.B, 2, dtype=torch.float64, device=dev)
        # per-step frames (x, p, F), written by the second half-kick on the device
        # into row (step mod rows): one row per step of the run, or a ring of
        # frame_window rows drained to the host after every chunk of steps
        self.frame_x = self.frame_p = self.frame_f = None
        self._frame_window = None
        self._host_frames = None
        if record_frames:
            rows = self.max_steps + 1
            if frame_window is not None:
                if int(frame_window) < 2:
                    raise ValueError("frame_window must hold at least 2 frames")
                rows = min(rows, int(frame_window))
                self._frame_window = rows
                self._host_frames = ([], [], [])
            shape = (rows, self.n_atoms, 3)
            self.frame_x = torch.zeros(shape, dtype=torch.float64, device=dev)
            self.frame_p = torch.zeros(shape, dtype=torch.float64, device=dev)
            self.frame_f = torch.zeros(shape, dtype=torch.float32, device=dev)
        self.layout = engine.make_layout(self.sizes, dev, cells, pbc)
        self.weights = model.packed_weights(dev)
        self._headroom = headroom
        self._persistent_blocks = persistent_blocks
        self._streams = streams
        self._tiles = None
        # size the neighbour capacity from the starting configuration
        self._make_plan(self._initial_edges() * headroom)
        self.steps_done = 0
        self.stream = torch.cuda.Stream(device=dev)
        for _attempt in range(3):
            self.stream.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(self.stream):
                C.check("mlff_md_init_forces", self.lib.mlff_md_init_forces(
                    self.plan.handle, self.stream.cuda_stream))
            self.stream.synchronize()
            try:
                self._check_overflow()
                break
            except OverflowError as exc:
                self._make_plan(int(exc.args[0]) * max(self._headroom, 1.0))
        else:
            raise RuntimeError("neighbour capacity did not converge")
        ke0 = 0.5 * (self.p.pow(2).sum(1) * self.inv_mass)
        ke0 = torch.zeros(self.B, dtype=torch.float64, device=dev).index_add_(
            0, self.layout.atom_mol.long(), ke0)
        self.ledger[0, :, 0] = self.energy
        self.ledger[0, :, 1] = ke0
        if self.frame_x is not None:
            self.frame_x[0] = self.x
            self.frame_p[0] = self.p
            self.frame_f[0] = self.forces
            if self._frame_window is not None:
                self._drain_frames(0, 1)
        self.tuning = None

    def autotune(self, **kw):
        """Choose the pair kernels' persistent grid on this workload
        (autotune.tune); later plans (after a capacity regrowth) keep it."""
        from . import autotune
        forces_saved = self.forces.clone()
        energy_saved = self.energy.clone()
        self.tuning = autotune.tune(self.plan, self.pos32, **kw)
        self._persistent_blocks = int(self.tuning["blocks"])
        self._tiles = tuple(self.tuning["tiles"])
        self.forces.copy_(forces_saved)
        self.energy.copy_(energy_saved)
        return self.tuning

    # ------------------------------------------------------------------
    def _initial_edges(self) -> int:
        row_ptr, col, _ = engine.neighbor_list(self.pos32, self.layout, self.model.cutoff)
        return max(int(col.numel()), 16)

    def _make_plan(self, edges: float):
        # the bound caps the first guess only: a measured count above it (a
        # dense box, more than 256 neighbours per atom) is honoured as is
        bound = engine.edge_capacity_bound(self.sizes, self.layout.periodic)
        cap = int(min(bound, int(edges) + 64)) if edges <= bound else int(edges) + 64
        self.plan = engine.Plan(self.model.hidden_dim, self.model.num_interactions,
                                self.model.num_rbf, self.model.max_z, self.model.cutoff,
                                self.layout, self.numbers, self.weights, cap,
                                persistent_blocks=self._persistent_blocks, streams=self._streams,
                                tiles=self._tiles)
        self.lib = C.lib()
        C.check("mlff_md_bind", self.lib.mlff_md_bind(
            self.plan.handle, C.ptr(self.x), C.ptr(self.p), C.ptr(self.inv_mass),
            C.ptr(self.pos32), C.ptr(self.forces), C.ptr(self.energy), C.ptr(self.ledger),
            self.max_steps + 1, float(self.dt)))
        if self.frame_x is not None:
            C.check("mlff_md_bind_frames", self.lib.mlff_md_bind_frames(
                self.plan.handle, C.ptr(self.frame_x), C.ptr(self.frame_p),
                C.ptr(self.frame_f), self.frame_x.shape[0]))
        # the ledger row index is the device step counter (+1: row 0 holds t = 0)
        self.plan.counts[C.MLFF_COUNT_STEP] = 1

    def _check_overflow(self):
        c = self.plan.read_counts()
        if c[C.MLFF_COUNT_OVERFLOW]:
            # the largest count of the chunk: the overflow may have happened at an
            # earlier step than the last, whose count alone is in MLFF_COUNT_EDGES
            raise OverflowError(int(max(c[C.MLFF_COUNT_EDGES], c[C.MLFF_COUNT_EDGES_MAX])))

    def run(self, n_steps: int, graph: bool = True, steps_per_graph: int = 10):
        """Advance n_steps; on a neighbour-capacity overflow the chunk is
        replayed from its saved state with a larger capacity."""
        if self.steps_done + n_steps > self.max_steps:
            raise ValueError("ledger too short; raise max_steps")
        done = 0
        chunk = max(steps_per_graph, 1) * 10
        if self._frame_window is not None:
            # a chunk's frames must not wrap onto each other before the drain
            chunk = min(chunk, self._frame_window)
        replays = 0
        while done < n_steps:
            k = min(chunk, n_steps - done)
            saved = (self.x.clone(), self.p.clone(), self.forces.clone(), self.energy.clone(),
                     self.plan.counts.clone())
            try:
                self._run_chunk(k, graph, steps_per_graph)
                self._check_overflow()
            except OverflowError as exc:
                self.x.copy_(saved[0]); self.p.copy_(saved[1]); self.forces.copy_(saved[2])
                self.energy.copy_(saved[3])
                step = int(saved[4][C.MLFF_COUNT_STEP])
                need = int(exc.args[0])
                replays += 1
                if replays > 8 or need <= self.plan.ecap:
                    raise RuntimeError(f"neighbour capacity did not converge: {need} edges "
                                       f"needed, capacity {self.plan.ecap} after "
                                       f"{replays} replays")
                self._make_plan(need * max(self._headroom, 1.1))
                self.plan.counts[C.MLFF_COUNT_STEP] = step
                self.pos32.copy_(self.x.to(torch.float32))
                continue
            if self._frame_window is not None:
                self._drain_frames(self.steps_done + 1, k)
            done += k
            self.steps_done += k

    def _drain_frames(self, first: int, count: int):
        """Copy the frames of steps first .. first+count-1 from the ring to
        the host."""
        rows = (torch.arange(first, first + count, device=self.device) % self._frame_window)
        for buf, dev_buf in zip(self._host_frames, (self.frame_x, self.frame_p, self.frame_f)):
            buf.append(dev_buf.index_select(0, rows).cpu())

    def _run_chunk(self, k: int, graph: bool, steps_per_graph: int):
        h = self.plan.handle
        s = self.stream
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            if graph and k >= steps_per_graph:
                n = k // steps_per_graph
                C.check("mlff_md_graph_run", self.lib.mlff_md_graph_run(h, s.cuda_stream,
                                                                        steps_per_graph, n))
                rest = k - n * steps_per_graph
            else:
                rest = k
            if rest:
                C.check("mlff_md_run", self.lib.mlff_md_run(h, s.cuda_stream, rest))
        s.synchronize()

    # ------------------------------------------------------------------
    def energies(self):
        """(E_pot, E_kin) [steps+1, B] float64 on the host; row 0 is t = 0."""
        led = self.ledger[:self.steps_done + 1].cpu().numpy()
        return led[..., 0], led[..., 1]

    def frames(self):
        """Positions, velocities (ASE units) and forces of every step so far,
        each [steps+1, N, 3] float64 on the host; row 0 is t = 0.  Needs
        record_frames=True."""
        if self.frame_x is None:
            raise RuntimeError("frames were not recorded (record_frames=False)")
        if self._frame_window is not None:
            fx, fp, ff = (torch.cat(b) for b in self._host_frames)
            v = fp * self.inv_mass.cpu()[None, :, None]
            return fx.numpy(), v.numpy(), ff.double().numpy()
        n = self.steps_done + 1
        v = self.frame_p[:n] * self.inv_mass[None, :, None]
        return (self.frame_x[:n].cpu().numpy(), v.cpu().numpy(),
                self.frame_f[:n].double().cpu().numpy())

    def summary(self):
        epot, ekin = self.energies()
        etot = epot + ekin
        t_ps = np.arange(len(etot)) * self.timestep_fs / 1000.0
        e0, ef = etot[0], etot[-1]
        with np.errstate(divide="ignore", invalid="ignore"):
            drift_pct = np.where(e0 != 0, 100.0 * (ef - e0) / np.abs(e0), 0.0)
        per_atom = (ef - e0) / self.sizes
        rate = np.polyfit(t_ps, etot, 1)[0] / self.sizes if len(t_ps) > 1 else np.zeros(self.B)
        kB = synth.KB_EV
        dof = 3 * self.sizes
        temps = 2 * ekin / (dof * kB)
        return {
            "n_steps": self.steps_done, "timestep_fs": self.timestep_fs,
            "total_time_ps": float(t_ps[-1]),
            "energy_drift_pct": drift_pct, "energy_drift_abs": ef - e0,
            "energy_drift_per_atom": per_atom, "drift_rate_eV_per_atom_ps": rate,
            "energy_std": etot.std(axis=0), "energy_range": np.ptp(etot, axis=0),
            "avg_temperature": temps.mean(axis=0), "std_temperature": temps.std(axis=0),
            "initial_energy": e0, "final_energy": ef,
        }

    def positions(self) -> np.ndarray:
        return self.x.cpu().numpy()

    def velocities(self) -> np.ndarray:
        return (self.p * self.inv_mass[:, None]).cpu().numpy()


# element symbols by atomic number (index 0 unused), for extended-XYZ output
_SYMBOLS = ("X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu "
            "Zn Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La "
            "Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At "
            "Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf Db Sg Bh Hs Mt Ds Rg Cn Nh "
            "Fl Mc Lv Ts Og").split()


def write_extxyz(path, numbers, positions, velocities, forces, info, cell=None, pbc=None):
    """Extended-XYZ frames (the format ase.io.write(format='extxyz') reads
    back): per frame the atom count, a comment line of key=value pairs
    (Lattice/pbc when periodic, then `info`'s scalars), then one line per atom
    with symbol, position, velocity and force.  `info` maps a key to one
    value per frame."""
    n_frames = len(positions)
    with open(path, "w") as fh:
        for t in range(n_frames):
            head = ['Properties=species:S:1:pos:R:3:vel:R:3:forces:R:3']
            if cell is not None and pbc is not None and np.any(pbc):
                head.insert(0, 'Lattice="%s"' % " ".join(f"{v:.10f}" for v in
                                                         np.asarray(cell).reshape(-1)))
                head.insert(2, 'pbc="%s"' % " ".join("T" if b else "F" for b in pbc))
            head += [f"{k}={float(v[t]):.12g}" for k, v in info.items()]
            fh.write(f"{len(numbers)}\n{' '.join(head)}\n")
            for z, x, v, f in zip(numbers, positions[t], velocities[t], forces[t]):
                fh.write(f"{_SYMBOLS[int(z)]:<2} " + " ".join(f"{c:.10f}" for c in x) + " "
                         + " ".join(f"{c:.10f}" for c in v) + " "
                         + " ".join(f"{c:.10f}" for c in f) + "\n")


# device frames kept by NVEMDHarness between host drains
FRAME_WINDOW = 1000


class NVEMDHarness:
    """The reference's NVE harness API (nve_harness.py:58-446) on the device
    integrator: same constructor, initialize_velocities, run_simulation,
    summary keys, reset, is_complete, n_atoms, get_trajectory_array and
    save_trajectory.  Every step's time, positions, velocities, forces and
    energies are collected, as the reference's _collect_data(interval=1)
    does (nve_harness.py:228-231, 263-287); here the second half-kick writes
    them into device frame buffers inside the replayed HIP graph and they
    cross to the host once per run.  Runs without reset accumulate, each
    starting its time at 0 (a new dynamics object per run in the
    reference)."""

    def __init__(self, atoms, calculator, temperature: float = 300.0, timestep: float = 0.5,
                 trajectory_file=None, log_interval: int = 10, remove_com_motion: bool = True,
                 seed: Optional[int] = None):
        if isinstance(atoms, (str,)) or hasattr(atoms, "__fspath__"):
            from ase.io import read  # only for file inputs
            atoms = read(str(atoms))
        self.atoms = atoms.copy()
        self.calculator = calculator
        if not hasattr(calculator, "model"):
            raise TypeError("NVEMDHarness needs a StudentForceFieldCalculator (its model runs "
                            "the device integrator)")
        self.temperature = temperature
        self.timestep_fs = timestep
        self.timestep = timestep * synth.FS
        self.trajectory_file = trajectory_file
        self.log_interval = log_interval
        self.remove_com_motion = remove_com_motion
        self._rng = np.random.Generator(np.random.PCG64(seed))
        self._momenta = None
        try:
            self.atoms.calc = calculator        # nve_harness.py:117
        except AttributeError:
            pass
        self.trajectory_data = {k: [] for k in ("time", "positions", "velocities", "forces",
                                                "potential_energy", "kinetic_energy",
                                                "total_energy", "temperature")}
        self._n_steps = 0
        self._wall_time = 0.0
        self._simulation_complete = False

    def _structure(self):
        a = self.atoms
        pbc = np.asarray(a.get_pbc(), dtype=bool)
        return synth.Structure(np.asarray(a.get_atomic_numbers(), dtype=np.int64),
                               np.asarray(a.get_positions(), dtype=np.float64),
                               np.asarray(a.get_cell(), dtype=np.float64) if pbc.any() else None,
                               pbc if pbc.any() else None)

    def _masses(self) -> np.ndarray:
        if hasattr(self.atoms, "get_masses"):
            return np.asarray(self.atoms.get_masses(), dtype=np.float64)
        return synth.masses_of(self.atoms.get_atomic_numbers())

    def _atoms_momenta(self) -> np.ndarray:
        """The atoms' own momenta (zeros when none were set, as ASE)."""
        if hasattr(self.atoms, "get_momenta"):
            p = self.atoms.get_momenta()
            if p is not None:
                return np.asarray(p, dtype=np.float64).reshape(-1, 3)
        if self._momenta is not None:
            return self._momenta
        return np.zeros((len(self.atoms), 3))

    def initialize_velocities(self, temperature: Optional[float] = None,
                              remove_com_translation: bool = True,
                              remove_com_rotation: bool = True):
        """nve_harness.py:141-171: Maxwell-Boltzmann momenta at the
        temperature, centre-of-mass translation and rotation removed, set on
        the atoms."""
        s = self._structure()
        T = self.temperature if temperature is None else temperature
        self._momenta = synth.maxwell_boltzmann(s.numbers, s.positions, T, self._rng,
                                                remove_com_translation, remove_com_rotation,
                                                masses=self._masses())
        if hasattr(self.atoms, "set_momenta"):
            self.atoms.set_momenta(self._momenta)

    @property
    def is_complete(self) -> bool:
        """nve_harness.py:379-382."""
        return self._simulation_complete

    @property
    def n_atoms(self) -> int:
        """nve_harness.py:384-387."""
        return len(self.atoms)

    def run_simulation(self, steps: int, initialize_velocities: bool = True,
                       save_interval: Optional[int] = None):
        """nve_harness.py:173-261.  save_interval (default log_interval)
        sets how often trajectory_file gets a frame, as the reference's
        attached Trajectory writer; the in-memory data holds every step."""
        import time as _time
        if steps <= 0:
            raise ValueError(f"steps must be positive, got {steps}")
        save_interval = max(int(save_interval or self.log_interval or 1), 1)
        if self.trajectory_file:
            self._writable(self.trajectory_file, None)     # fail before the run, not after it
        t0 = _time.perf_counter()
        try:
            if initialize_velocities:
                self.initialize_velocities()
            # the device keeps a ring of FRAME_WINDOW frames, drained to the host
            # after every chunk of steps
            md = BatchedNVE(self.calculator.model, [self._structure()],
                            timestep_fs=self.timestep_fs, momenta=self._atoms_momenta(),
                            max_steps=steps, device=self.calculator.device,
                            record_frames=True, masses=self._masses(),
                            frame_window=FRAME_WINDOW)
            md.run(steps, graph=steps >= 10, steps_per_graph=10)
            epot, ekin = md.energies()
            x, v, f = md.frames()
            n = md.n_atoms
            td = self.trajectory_data
            td["time"] += list(np.arange(steps + 1) * self.timestep_fs / 1000.0)
            td["positions"] += list(x)
            td["velocities"] += list(v)
            td["forces"] += list(f)
            td["potential_energy"] += list(epot[:, 0])
            td["kinetic_energy"] += list(ekin[:, 0])
            td["total_energy"] += list(epot[:, 0] + ekin[:, 0])
            td["temperature"] += list(2 * ekin[:, 0] / (3 * n * synth.KB_EV))
            if hasattr(self.atoms, "set_positions"):
                self.atoms.set_positions(x[-1])
            self._momenta = md.p.cpu().numpy()
            if hasattr(self.atoms, "set_momenta"):
                self.atoms.set_momenta(self._momenta)
            if self.trajectory_file:
                rows = np.arange(0, steps + 1, save_interval)
                k0 = len(td["time"]) - (steps + 1)
                self._write(self.trajectory_file, k0 + rows, None)
            self._n_steps = steps
            self._wall_time = _time.perf_counter() - t0
            self._simulation_complete = True
            k0 = len(td["time"]) - (steps + 1)
            for r in range(self.log_interval, steps + 1, max(self.log_interval, 1)):
                self._log_progress(k0 + r)
            results = self._compute_summary()
            logger.info(f"Simulation complete: {steps} steps in {self._wall_time:.2f}s "
                        f"({steps / max(self._wall_time, 1e-12):.1f} steps/s)")
            logger.info(f"Energy drift: {results['energy_drift_pct']:.4f}%, "
                        f"Avg T: {results['avg_temperature']:.2f}K")
            return results
        except Exception as e:
            raise RuntimeError(f"NVE MD simulation failed after "
                               f"{len(self.trajectory_data['time'])} steps: {e}") from e

    def _log_progress(self, k: int):
        """nve_harness.py:289-309, for collected row k (the reference logs as
        the run goes; here the rows come back from the device after it)."""
        td = self.trajectory_data
        etot, e0 = td["total_energy"][k], td["total_energy"][0]
        drift_pct = 100.0 * (etot - e0) / abs(e0) if e0 != 0 else 0.0
        logger.debug(f"Step {k:6d}: t={td['time'][k]:8.3f}ps, E={etot:12.6f}eV, "
                     f"T={td['temperature'][k]:7.2f}K, drift={drift_pct:+.4f}%")

    def _compute_summary(self):
        """nve_harness.py:311-360."""
        td = self.trajectory_data
        if not td["time"]:
            raise RuntimeError("No trajectory data collected")
        times = np.array(td["time"])
        etot = np.array(td["total_energy"])
        temps = np.array(td["temperature"])
        e0, ef = etot[0], etot[-1]
        return {
            "trajectory_data": td, "n_steps": self._n_steps,
            "total_time_ps": times[-1] - times[0], "timestep_fs": self.timestep_fs,
            "wall_time_s": self._wall_time,
            "energy_drift_pct": 100.0 * (ef - e0) / abs(e0) if e0 != 0 else 0.0,
            "energy_drift_abs": ef - e0, "energy_std": np.std(etot),
            "energy_range": np.ptp(etot), "avg_temperature": np.mean(temps),
            "std_temperature": np.std(temps), "initial_energy": e0, "final_energy": ef,
            "avg_potential_energy": np.mean(td["potential_energy"]),
            "avg_kinetic_energy": np.mean(td["kinetic_energy"]),
        }

    def reset(self):
        """nve_harness.py:362-377."""
        self.trajectory_data = {k: [] for k in self.trajectory_data}
        self._simulation_complete = False
        self._n_steps = 0
        self._wall_time = 0.0

    def get_trajectory_array(self, key: str) -> np.ndarray:
        """nve_harness.py:389-406."""
        if key not in self.trajectory_data:
            raise KeyError(f"Unknown trajectory key: {key}")
        data = self.trajectory_data[key]
        if not data:
            raise RuntimeError("No trajectory data collected")
        return np.array(data)

    def save_trajectory(self, filename, format: str = "traj"):
        """nve_harness.py:408-443: every collected frame with its time and
        energies.  Written through ase.io when ASE is installed; without it
        'xyz' and 'extxyz' are written as extended XYZ and other formats
        raise ValueError."""
        import warnings
        if not self._simulation_complete:
            warnings.warn("Simulation not complete, saving partial trajectory")
        if not self.trajectory_data["time"]:
            raise RuntimeError("No trajectory data to save")
        self._write(filename, np.arange(len(self.trajectory_data["time"])), format)

    @staticmethod
    def _writable(filename, format):
        """Without ASE only extended XYZ can be written."""
        try:
            import ase.io  # noqa: F401
            return
        except ImportError:
            pass
        fmt = format or str(filename).rsplit(".", 1)[-1]
        if fmt not in ("xyz", "extxyz"):
            raise ValueError(f"format {fmt!r} needs ASE (not installed); use 'xyz' or 'extxyz'")

    def _write(self, filename, rows, format):
        td = self.trajectory_data
        keys = ("time", "potential_energy", "kinetic_energy", "total_energy", "temperature")
        try:
            from ase.io import write as ase_write
        except ImportError:
            ase_write = None
        if ase_write is not None:
            frames = []
            for i in rows:
                a = self.atoms.copy()
                a.set_positions(td["positions"][i])
                a.set_velocities(td["velocities"][i])
                for k in keys:
                    a.info[k] = td[k][i]
                a.arrays["forces"] = td["forces"][i]
                frames.append(a)
            ase_write(str(filename), frames, format=format)
            return
        self._writable(filename, format)
        s = self._structure()
        write_extxyz(filename, s.numbers, [td["positions"][i] for i in rows],
                     [td["velocities"][i] for i in rows], [td["forces"][i] for i in rows],
                     {k: [td[k][i] for i in rows] for k in keys}, s.cell, s.pbc)
