"""Training-frame datasets in the reference's ragged layout, and its
distillation batch interface.

Layout (the reference's HDF5 dataset writer, docs/hdf5_dataset_writer.md:36-68;
read back as in scripts/debug_pbc_bug.py:21-45):

  structures/atomic_numbers         int64   [N_total]
  structures/atomic_numbers_splits  int64   [S+1]   (cumulative atom offsets)
  structures/positions              float64 [N_total, 3]
  structures/cells                  float64 [S, 3, 3]
  structures/pbc                    bool    [S, 3]
  labels/energy                     float64 [S]
  labels/forces                     float32 [N_total, 3]
  labels/forces_splits              int64   [S+1]   (atom rows, as the atom splits)
  labels/stress                     float64 [S, 6]  (Voigt xx yy zz yz xz xy; 0 where absent)
  labels/stress_mask                bool    [S]
  labels/structure_indices          int64   [S]
  metadata                          JSON text (in HDF5: attributes of /metadata)

Two containers hold it: HDF5 through h5py when h5py is importable (the
reference's format), otherwise a NumPy ``.npz`` archive whose keys are the same
dataset paths.  Reading never unpickles (``allow_pickle=False``).

``distillation_collate`` builds the batch dictionary the reference's
DistillationWrapper.forward consumes (distillation_wrapper.py:52-60), and
``DistillationWrapper`` is that module on this package's StudentForceField:
energies from the model, forces from autograd (second-order when training).
"""

from __future__ import annotations

import json
import os
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np
import torch
from torch import nn

KEYS_STRUCT = ("atomic_numbers", "atomic_numbers_splits", "positions", "cells", "pbc")
KEYS_LABEL = ("energy", "forces", "forces_splits", "stress", "stress_mask", "structure_indices")


def _have_h5py() -> bool:
    try:
        import h5py  # noqa: F401
        return True
    except ImportError:
        return False


def _backend_for(path: str, backend: Optional[str]) -> str:
    if backend is None:
        if path.endswith(".npz"):
            return "npz"
        return "h5" if _have_h5py() else "npz"
    if backend not in ("h5", "npz"):
        raise ValueError(f"unknown backend {backend!r} (h5 or npz)")
    if backend == "h5" and not _have_h5py():
        raise ImportError("the h5 backend needs h5py, which is not installed")
    return backend


def _structure_arrays(atoms=None, numbers=None, positions=None, cell=None, pbc=None):
    """(numbers int64 [n], positions float64 [n,3], cell float64 [3,3], pbc bool [3])
    from an ASE-like Atoms object or explicit arrays."""
    if atoms is not None:
        numbers = atoms.get_atomic_numbers()
        positions = atoms.get_positions()
        cell = np.asarray(atoms.get_cell(), dtype=np.float64)
        pbc = atoms.get_pbc()
    numbers = np.asarray(numbers, dtype=np.int64).reshape(-1)
    positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    cell = np.zeros((3, 3)) if cell is None else np.asarray(cell, dtype=np.float64).reshape(3, 3)
    pbc = np.zeros(3, dtype=bool) if pbc is None else np.broadcast_to(
        np.asarray(pbc, dtype=bool), (3,)).copy()
    if len(numbers) != len(positions):
        raise ValueError(f"{len(numbers)} atomic numbers but {len(positions)} positions")
    return numbers, positions, cell, pbc


class FrameDatasetWriter:
    """Incremental writer (the reference's HDF5DatasetWriter interface:
    add_structure, add_batch, finalize, context manager, mode 'w' / 'a').

    Structures are kept in memory and written by ``finalize``.  With
    ``validate`` (default) non-finite energies, forces or stresses and shape
    mismatches raise ValueError.
    """

    def __init__(self, output_path: str, mode: str = "w", backend: Optional[str] = None,
                 validate: bool = True):
        if mode not in ("w", "a"):
            raise ValueError("mode must be 'w' or 'a'")
        self.path = str(output_path)
        self.backend = _backend_for(self.path, backend)
        self.validate = validate
        self._numbers: List[np.ndarray] = []
        self._positions: List[np.ndarray] = []
        self._cells: List[np.ndarray] = []
        self._pbc: List[np.ndarray] = []
        self._energy: List[float] = []
        self._forces: List[np.ndarray] = []
        self._stress: List[np.ndarray] = []
        self._stress_mask: List[bool] = []
        self._metadata: Dict = {}
        self._closed = False
        if mode == "a" and os.path.exists(self.path):
            old = FrameDataset(self.path, backend=self.backend)
            for i in range(len(old)):
                s = old[i]
                self._append(s["atomic_numbers"], s["positions"], s["cell"], s["pbc"],
                             s["energy"], s["forces"], s["stress"])
            self._metadata = dict(old.metadata)

    # -- adding ---------------------------------------------------------------
    def _append(self, numbers, positions, cell, pbc, energy, forces, stress):
        forces = np.asarray(forces, dtype=np.float32).reshape(-1, 3)
        energy = float(energy)
        if self.validate:
            if len(forces) != len(numbers):
                raise ValueError(f"forces for {len(forces)} atoms, structure has {len(numbers)}")
            if not np.isfinite(energy):
                raise ValueError("non-finite energy")
            if not np.all(np.isfinite(forces)):
                raise ValueError("non-finite forces")
            if not np.all(np.isfinite(positions)):
                raise ValueError("non-finite positions")
        has_stress = stress is not None
        st = np.zeros(6) if stress is None else np.asarray(stress, dtype=np.float64).reshape(-1)
        if has_stress and st.shape != (6,):
            raise ValueError("stress must be 6 Voigt components")
        if self.validate and has_stress and not np.all(np.isfinite(st)):
            raise ValueError("non-finite stress")
        self._numbers.append(numbers)
        self._positions.append(positions)
        self._cells.append(cell)
        self._pbc.append(pbc)
        self._energy.append(energy)
        self._forces.append(forces)
        self._stress.append(st)
        self._stress_mask.append(has_stress)

    def add_structure(self, atoms=None, energy=None, forces=None, stress=None, metadata=None, *,
                      numbers=None, positions=None, cell=None, pbc=None):
        if self._closed:
            raise RuntimeError("writer already finalized")
        if energy is None or forces is None:
            raise ValueError("energy and forces are required")
        n, p, c, b = _structure_arrays(atoms, numbers, positions, cell, pbc)
        self._append(n, p, c, b, energy, forces, stress)
        if metadata:
            self._metadata.setdefault("structure_metadata", {})[str(len(self._energy) - 1)] = metadata

    def add_batch(self, structures: Sequence, energies: Sequence, forces: Sequence,
                  stresses: Optional[Sequence] = None, show_progress: bool = False):
        if not (len(structures) == len(energies) == len(forces)):
            raise ValueError("structures, energies and forces differ in length")
        for k, atoms in enumerate(structures):
            self.add_structure(atoms, energies[k], forces[k],
                               None if stresses is None else stresses[k])

    def __len__(self) -> int:
        return len(self._energy)

    # -- writing --------------------------------------------------------------
    def arrays(self) -> Dict[str, np.ndarray]:
        S = len(self._energy)
        sizes = np.array([len(n) for n in self._numbers], dtype=np.int64)
        splits = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        cat = (lambda xs, shape, dt: np.concatenate(xs).astype(dt) if xs
               else np.zeros(shape, dtype=dt))
        return {
            "structures/atomic_numbers": cat(self._numbers, (0,), np.int64),
            "structures/atomic_numbers_splits": splits,
            "structures/positions": cat(self._positions, (0, 3), np.float64),
            "structures/cells": (np.stack(self._cells) if S else np.zeros((0, 3, 3))).astype(np.float64),
            "structures/pbc": (np.stack(self._pbc) if S else np.zeros((0, 3))).astype(bool),
            "labels/energy": np.asarray(self._energy, dtype=np.float64),
            "labels/forces": cat(self._forces, (0, 3), np.float32),
            "labels/forces_splits": splits.copy(),
            "labels/stress": (np.stack(self._stress) if S else np.zeros((0, 6))).astype(np.float64),
            "labels/stress_mask": np.asarray(self._stress_mask, dtype=bool),
            "labels/structure_indices": np.arange(S, dtype=np.int64),
        }

    def finalize(self, extra_metadata: Optional[Dict] = None):
        if self._closed:
            return
        meta = dict(self._metadata)
        if extra_metadata:
            meta.update(extra_metadata)
        arrays = self.arrays()
        meta.setdefault("n_structures", len(self._energy))
        meta.setdefault("n_atoms", int(len(arrays["structures/atomic_numbers"])))
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        if self.backend == "npz":
            tmp = self.path + ".tmp.npz"
            np.savez(tmp, metadata=np.array(json.dumps(meta, sort_keys=True)), **arrays)
            os.replace(tmp, self.path)
        else:
            import h5py
            with h5py.File(self.path, "w") as f:
                for key, val in arrays.items():
                    f.create_dataset(key, data=val)
                g = f.create_group("metadata")
                for k, v in meta.items():
                    g.attrs[k] = v if isinstance(v, (int, float, str)) else json.dumps(v)
        self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            self.finalize()


class FrameDataset:
    """Reader: ``len``, ``ds[i]`` -> dict(atomic_numbers, positions, cell, pbc,
    energy, forces, stress or None), ``metadata``."""

    def __init__(self, path: str, backend: Optional[str] = None):
        self.path = str(path)
        self.backend = _backend_for(self.path, backend)
        if self.backend == "npz":
            with np.load(self.path, allow_pickle=False) as z:
                self.arrays = {k: z[k] for k in z.files if k != "metadata"}
                self.metadata = json.loads(str(z["metadata"])) if "metadata" in z.files else {}
        else:
            import h5py
            with h5py.File(self.path, "r") as f:
                self.arrays = {f"{g}/{k}": f[g][k][()] for g, keys in
                               (("structures", KEYS_STRUCT), ("labels", KEYS_LABEL))
                               for k in keys if k in f[g]}
                self.metadata = dict(f["metadata"].attrs) if "metadata" in f else {}
        missing = [f"structures/{k}" for k in KEYS_STRUCT if f"structures/{k}" not in self.arrays]
        missing += [f"labels/{k}" for k in ("energy", "forces", "forces_splits")
                    if f"labels/{k}" not in self.arrays]
        if missing:
            raise ValueError(f"{self.path}: missing datasets {missing}")
        self.splits = self.arrays["structures/atomic_numbers_splits"]
        self.fsplits = self.arrays["labels/forces_splits"]

    def __len__(self) -> int:
        return len(self.arrays["labels/energy"])

    def __getitem__(self, i: int) -> Dict:
        if not 0 <= i < len(self):
            raise IndexError(i)
        a, b = int(self.splits[i]), int(self.splits[i + 1])
        fa, fb = int(self.fsplits[i]), int(self.fsplits[i + 1])
        A = self.arrays
        has_stress = bool(A["labels/stress_mask"][i]) if "labels/stress_mask" in A else False
        return {
            "atomic_numbers": A["structures/atomic_numbers"][a:b],
            "positions": A["structures/positions"][a:b],
            "cell": A["structures/cells"][i],
            "pbc": A["structures/pbc"][i].astype(bool),
            "energy": float(A["labels/energy"][i]),
            "forces": A["labels/forces"][fa:fb].reshape(-1, 3),
            "stress": A["labels/stress"][i] if has_stress else None,
        }


def distillation_collate(items: Sequence[Dict], device=None) -> Dict[str, torch.Tensor]:
    """Batch dictionary of the reference's distillation collate
    (distillation_wrapper.py:52-60): atomic_numbers int64 [N], positions
    float32 [N,3], cell float32 [B,3,3], pbc bool [B,3], batch int64 [N],
    n_atoms int64 [B], atom_splits int64 [B+1]; plus the labels energy
    float64 [B] and forces float32 [N,3]."""
    if not items:
        raise ValueError("empty batch")
    n = np.array([len(it["atomic_numbers"]) for it in items], dtype=np.int64)
    t = lambda x, dt: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=device)
    return {
        "atomic_numbers": t(np.concatenate([it["atomic_numbers"] for it in items]), torch.int64),
        "positions": t(np.concatenate([it["positions"] for it in items]), torch.float32),
        "cell": t(np.stack([it["cell"] for it in items]), torch.float32),
        "pbc": t(np.stack([it["pbc"] for it in items]), torch.bool),
        "batch": t(np.repeat(np.arange(len(items)), n), torch.int64),
        "n_atoms": t(n, torch.int64),
        "atom_splits": t(np.concatenate([[0], np.cumsum(n)]), torch.int64),
        "energy": t(np.array([it["energy"] for it in items]), torch.float64),
        "forces": t(np.concatenate([it["forces"] for it in items]), torch.float32),
    }


class DistillationWrapper(nn.Module):
    """The reference's DistillationWrapper (distillation_wrapper.py:48-96) on
    this package's StudentForceField: forward(batch) -> {'energy', 'forces'},
    forces = -dE/dx by autograd, with the graph kept while training so a
    force loss can be differentiated with respect to the weights."""

    def __init__(self, model: nn.Module):
        super().__init__()
        self.model = model

    def forward(self, batch: Dict[str, torch.Tensor]) -> Dict[str, torch.Tensor]:
        positions = batch["positions"].clone().detach().requires_grad_(True)
        with torch.set_grad_enabled(True):
            energy = self.model(atomic_numbers=batch["atomic_numbers"], positions=positions,
                                cell=batch.get("cell"), pbc=batch.get("pbc"),
                                batch=batch["batch"])
            forces = -torch.autograd.grad(energy.sum(), positions, create_graph=self.training,
                                          retain_graph=self.training)[0]
        return {"energy": energy, "forces": forces}

    def num_parameters(self) -> int:
        if hasattr(self.model, "num_parameters"):
            return self.model.num_parameters()
        return sum(p.numel() for p in self.parameters() if p.requires_grad)


def save_frames(frames, path: str, backend: Optional[str] = None,
                extra_metadata: Optional[Dict] = None) -> int:
    """Write an mts.FrameSet (molecule frames cut from batched MD) as one
    structure per (frame, molecule); returns the number of structures."""
    numbers = frames.numbers.cpu().numpy()
    ptr = frames.mol_ptr
    with FrameDatasetWriter(path, backend=backend) as w:
        for pos, forces, energy in zip(frames.pos, frames.forces, frames.energy):
            P = pos.cpu().numpy().astype(np.float64)
            F = forces.cpu().numpy()
            E = energy.cpu().numpy().astype(np.float64)
            for m in range(len(frames.sizes)):
                a, b = ptr[m], ptr[m + 1]
                w.add_structure(energy=E[m], forces=F[a:b], numbers=numbers[a:b], positions=P[a:b])
        w.finalize(extra_metadata=extra_metadata or {"source": "full-model MD frames"})
        return len(w)
