"""Drop-in StudentForceFieldCalculator (ASE calculator API) on the gfx950 kernels.

Follows src/mlff_distiller/inference/ase_calculator.py: constructor
arguments (:105-119), implemented_properties (:139-141), calculate()
(:279-412) with its input validation (:414-459) and error contract
(ValueError re-raised, everything else wrapped in RuntimeError, :405-412),
calculate_batch() (:590-817), timing statistics (:819-883).

Differences, by design:
  * every evaluation is the fused HIP pipeline (energy + forces from the
    hand-written adjoint); use_compile / use_fp16 / use_torch_cluster /
    use_analytical_forces are accepted for compatibility and have nothing to
    switch;
  * stress is real: for periodic cells it is the symmetrised virial over the
    volume, (1/V) dE/d(strain) in ASE's Voigt order (xx, yy, zz, yz, xz,
    xy); the reference returns zeros (:548-588) and None in batch mode;
  * use_jit loads the parameters of the TorchScript export and evaluates
    them on the kernels.

When ASE is importable the class derives from ase's Calculator; otherwise
from a small stand-in with the same caching contract.
"""

from __future__ import annotations

import logging
import time
import warnings
from pathlib import Path
from typing import Any, Dict, List, Optional, Union

import numpy as np
import torch

from . import engine
from .model import StudentForceField

logger = logging.getLogger(__name__)

try:  # pragma: no cover - ASE is optional
    from ase.calculators.calculator import Calculator, all_changes
except ImportError:  # minimal stand-in for ase.calculators.calculator
    all_changes = ["positions", "numbers", "cell", "pbc", "initial_charges", "initial_magmoms"]

    class Calculator:
        """The part of ASE's Calculator the student calculator relies on:
        results caching keyed on the atoms' state."""

        implemented_properties: List[str] = []

        def __init__(self, **kwargs):
            self.parameters = dict(kwargs)
            self.results: Dict[str, Any] = {}
            self.atoms = None

        @staticmethod
        def _state(atoms):
            return (np.array(atoms.get_positions(), dtype=np.float64),
                    np.array(atoms.get_atomic_numbers()),
                    np.array(atoms.get_cell(), dtype=np.float64),
                    np.array(atoms.get_pbc(), dtype=bool))

        def check_state(self, atoms, tol=1e-15):
            if self.atoms is None:
                return list(all_changes)
            changes = []
            old, new = self._state(self.atoms), self._state(atoms)
            for name, a, b in zip(("positions", "numbers", "cell", "pbc"), old, new):
                if a.shape != b.shape or np.abs(a.astype(float) - b.astype(float)).max(initial=0) > tol:
                    changes.append(name)
            return changes

        def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
            if atoms is not None:
                self.atoms = atoms.copy()

        def get_property(self, name, atoms=None, allow_calculation=True):
            if atoms is None:
                atoms = self.atoms
            changes = self.check_state(atoms)
            if changes:
                self.results = {}
            if name not in self.results:
                if not allow_calculation:
                    return None
                self.calculate(atoms, [name], changes)
            return self.results[name]

        def get_potential_energy(self, atoms=None, force_consistent=False):
            return self.get_property("energy", atoms)

        def get_forces(self, atoms=None):
            return np.array(self.get_property("forces", atoms))

        def get_stress(self, atoms=None):
            return np.array(self.get_property("stress", atoms))

        def reset(self):
            self.atoms = None
            self.results = {}


class StudentForceFieldCalculator(Calculator):
    """ASE calculator for the PaiNN student on MI355X."""

    def __init__(self, checkpoint_path: Union[str, Path], device: str = "cuda",
                 dtype: torch.dtype = torch.float32, enable_stress: bool = False,
                 batch_size: Optional[int] = None, enable_timing: bool = False,
                 use_compile: bool = False, use_fp16: bool = False, use_jit: bool = False,
                 jit_path: Optional[Union[str, Path]] = None, use_torch_cluster: bool = True,
                 use_analytical_forces: bool = False, model: Optional[StudentForceField] = None,
                 **kwargs):
        super().__init__(**kwargs)
        self.checkpoint_path = Path(checkpoint_path) if checkpoint_path is not None else None
        self.device = engine.device_of(device)
        self.dtype = dtype
        self.enable_stress = enable_stress
        self.batch_size = batch_size
        self.enable_timing = enable_timing
        self.use_compile = use_compile
        self.use_fp16 = use_fp16
        self.use_jit = use_jit
        self.jit_path = Path(jit_path) if jit_path else None
        self.use_torch_cluster = use_torch_cluster
        self.use_analytical_forces = use_analytical_forces
        self.implemented_properties = ["energy", "forces"]
        if self.enable_stress:
            self.implemented_properties.append("stress")
        if dtype not in (torch.float32,):
            warnings.warn("the MI355X kernels compute in float32 (energies accumulated in "
                          f"float64); dtype={dtype} is ignored", stacklevel=2)
        if use_fp16:
            warnings.warn("use_fp16 has no effect: the kernels compute in float32", stacklevel=2)
        self.model = model if model is not None else self._load_model()
        self.model = self.model.to(self.device)
        self.model.eval()
        self._n_calls = 0
        self._total_time = 0.0
        self._call_times: List[float] = []

    # ------------------------------------------------------------------
    def _load_model(self) -> StudentForceField:
        if self.use_jit:
            if not self.jit_path:
                raise ValueError("use_jit=True but jit_path not provided")
            if not self.jit_path.exists():
                raise FileNotFoundError(f"TorchScript model not found: {self.jit_path}")
            try:
                scripted = torch.jit.load(str(self.jit_path), map_location="cpu")
                sd = {k.split("model.", 1)[-1] if k.startswith("model.") else k: v
                      for k, v in scripted.state_dict().items()}
                from .engine import hyper_from_state
                H, L, R, max_z = hyper_from_state(sd)
                m = StudentForceField(hidden_dim=H, num_interactions=L, num_rbf=R, max_z=max_z)
                m.load_state_dict(sd)
                return m
            except Exception as e:
                raise RuntimeError(f"Failed to load TorchScript model from {self.jit_path}: {e}") from e
        if self.checkpoint_path is None or not self.checkpoint_path.exists():
            raise FileNotFoundError(f"Checkpoint not found: {self.checkpoint_path}")
        try:
            m = StudentForceField.load(self.checkpoint_path, device="cpu")
            m.use_torch_cluster = self.use_torch_cluster
            return m
        except Exception as e:
            raise RuntimeError(f"Failed to load model from {self.checkpoint_path}: {e}") from e

    # ------------------------------------------------------------------
    @staticmethod
    def _validate_inputs(positions, numbers, cell, pbc):
        """ase_calculator.py:414-459."""
        if len(positions) == 0:
            raise ValueError("Cannot calculate properties for empty structure")
        if np.any(numbers < 1) or np.any(numbers > 118):
            raise ValueError(f"Invalid atomic numbers: must be 1-118, got {numbers}")
        if not np.isfinite(positions).all():
            raise ValueError("Positions contain NaN or Inf values")
        if np.any(pbc):
            if not np.isfinite(cell).all():
                raise ValueError("Cell contains NaN or Inf values")
            vol = abs(np.linalg.det(cell))
            if vol < 1e-6:
                warnings.warn(f"Cell volume very small ({vol:.2e} A^3), may indicate degenerate cell",
                              stacklevel=3)

    def _stress(self, virial: np.ndarray, cell: np.ndarray) -> np.ndarray:
        vol = abs(np.linalg.det(cell))
        s = 0.5 * (virial + virial.T) / vol
        return np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])

    def calculate(self, atoms=None, properties=("energy", "forces"),
                  system_changes=all_changes):
        Calculator.calculate(self, atoms, properties, system_changes)
        t0 = time.perf_counter() if self.enable_timing else 0.0
        try:
            positions = np.asarray(atoms.get_positions(), dtype=np.float64)
            numbers = np.asarray(atoms.get_atomic_numbers())
            cell = np.asarray(atoms.get_cell(), dtype=np.float64)
            pbc = np.asarray(atoms.get_pbc(), dtype=bool)
            self._validate_inputs(positions, numbers, cell, pbc)
            periodic = bool(pbc.any())
            want_stress = "stress" in properties and self.enable_stress
            Z = torch.as_tensor(numbers.astype(np.int64)).to(self.device)
            pos = torch.as_tensor(positions, dtype=torch.float32).to(self.device)
            e64, f, vir = self.model.evaluate(
                Z, pos, np.array([len(numbers)]), cell[None] if periodic else None,
                pbc[None] if periodic else None, need_forces=True,
                need_virial=want_stress and periodic)
            results = {"energy": float(e64[0].item()), "forces": f.cpu().numpy()}
            if want_stress:
                results["stress"] = (self._stress(vir[0].cpu().numpy(), cell) if periodic
                                     else np.zeros(6))
            self.results = results
            self._n_calls += 1
            if self.enable_timing:
                dt = time.perf_counter() - t0
                self._total_time += dt
                self._call_times.append(dt)
        except ValueError:
            raise
        except Exception as e:
            logger.error("Calculation failed: %s", e)
            raise RuntimeError(f"Failed to calculate properties for {len(atoms)} atoms: {e}") from e

    def calculate_batch(self, atoms_list: List[Any],
                        properties: List[str] = ("energy", "forces")) -> List[Dict[str, Any]]:
        """All structures in one fused evaluation (ase_calculator.py:590-645)."""
        if not atoms_list:
            return []
        if len(atoms_list) == 1:
            atoms = atoms_list[0]
            atoms.calc = self
            return [{
                "energy": atoms.get_potential_energy() if "energy" in properties else None,
                "forces": atoms.get_forces() if "forces" in properties else None,
                "stress": atoms.get_stress() if "stress" in properties and self.enable_stress
                else None,
            }]
        step = self.batch_size or len(atoms_list)
        out: List[Dict[str, Any]] = []
        for k in range(0, len(atoms_list), step):
            out.extend(self._batch(atoms_list[k:k + step], properties))
        return out

    def _batch(self, atoms_list, properties):
        pos_l, z_l, cells, pbcs = [], [], [], []
        for atoms in atoms_list:
            p = np.asarray(atoms.get_positions(), dtype=np.float64)
            z = np.asarray(atoms.get_atomic_numbers())
            c = np.asarray(atoms.get_cell(), dtype=np.float64)
            b = np.asarray(atoms.get_pbc(), dtype=bool)
            self._validate_inputs(p, z, c, b)
            pos_l.append(p); z_l.append(z); cells.append(c); pbcs.append(b)
        sizes = np.array([len(z) for z in z_l])
        periodic = any(b.any() for b in pbcs)
        want_stress = "stress" in properties and self.enable_stress
        Z = torch.as_tensor(np.concatenate(z_l).astype(np.int64)).to(self.device)
        pos = torch.as_tensor(np.concatenate(pos_l), dtype=torch.float32).to(self.device)
        e64, f, vir = self.model.evaluate(Z, pos, sizes, np.stack(cells) if periodic else None,
                                          np.stack(pbcs) if periodic else None, need_forces=True,
                                          need_virial=want_stress and periodic)
        energies = e64.cpu().numpy()
        forces = f.cpu().numpy()
        virs = vir.cpu().numpy() if vir is not None else None
        results, off = [], 0
        for i, n in enumerate(sizes):
            r: Dict[str, Any] = {}
            if "energy" in properties:
                r["energy"] = float(energies[i])
            if "forces" in properties:
                r["forces"] = forces[off:off + n]
            if want_stress:
                r["stress"] = (self._stress(virs[i], cells[i]) if pbcs[i].any() and virs is not None
                               else np.zeros(6))
            results.append(r)
            off += n
        self._n_calls += 1
        return results

    # ------------------------------------------------------------------
    def reset(self):
        Calculator.reset(self)

    @property
    def n_calls(self) -> int:
        return self._n_calls

    @property
    def avg_time(self) -> float:
        return self._total_time / self._n_calls if self._n_calls else 0.0

    def get_timing_stats(self) -> Dict[str, float]:
        if not self._call_times:
            return {"n_calls": 0, "total_time": 0.0, "avg_time": 0.0, "min_time": 0.0,
                    "max_time": 0.0, "median_time": 0.0}
        t = np.array(self._call_times)
        return {"n_calls": self._n_calls, "total_time": self._total_time,
                "avg_time": float(t.mean()), "min_time": float(t.min()),
                "max_time": float(t.max()), "median_time": float(np.median(t)),
                "std_time": float(t.std())}

    def __repr__(self) -> str:
        name = self.checkpoint_path.name if self.checkpoint_path else "<model>"
        return f"StudentForceFieldCalculator(checkpoint={name}, device={self.device}, calls={self._n_calls})"


__all__ = ["StudentForceFieldCalculator"]
