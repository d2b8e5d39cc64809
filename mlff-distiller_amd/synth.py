"""Seeded synthetic structures and initial velocities.

The reference ships no structures for the PaiNN student's E/F + MD benchmarks
(its trained checkpoint and datasets are not in the repository), so the
workloads are generated here from a fixed recipe:

* organic-like molecules: element fractions H 50 % / C 30 % / N 8 % / O 10 % /
  S 2 %, positions grown as a random bonded walk (each new atom is bonded to a
  randomly chosen atom already placed, bond length U[1.1, 1.5] Angstrom, random
  direction), with a 0.9 Angstrom minimum distance between any two atoms;
* a periodic box of rigid water-like triatomics (O-H 0.96 A, H-O-H 104.5 deg)
  on a jittered cubic lattice at 1.0 g/cm^3;
* Maxwell-Boltzmann velocities with centre-of-mass translation and rotation
  removed, following the reference's NVE harness set-up
  (src/mlff_distiller/testing/nve_harness.py:141-171, which calls ASE's
  MaxwellBoltzmannDistribution, Stationary and ZeroRotation).

Units follow ASE (eV, Angstrom, amu; time unit Angstrom*sqrt(amu/eV)) with
CODATA-2014 constants, which is what ``ase.units`` uses by default.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

# ---- physical constants (CODATA 2014, as in ase.units) ----------------------
_E = 1.6021766208e-19          # C
_AMU = 1.660539040e-27         # kg
_KB_SI = 1.38064852e-23        # J/K
KB_EV = _KB_SI / _E            # eV/K
FS = 1e-15 / (1e-10 * math.sqrt(_AMU / _E))   # 1 fs in ASE time units
AVOGADRO = 6.022140857e23

# atomic masses (amu), the values ase.data.atomic_masses holds for these elements
ATOMIC_MASSES = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 16: 32.06}

ELEMENTS = (1, 6, 7, 8, 16)
FRACTIONS = (0.50, 0.30, 0.08, 0.10, 0.02)


def masses_of(numbers: np.ndarray) -> np.ndarray:
    """Atomic masses (amu) for an array of atomic numbers."""
    out = np.empty(len(numbers), dtype=np.float64)
    for i, z in enumerate(np.asarray(numbers).tolist()):
        if z not in ATOMIC_MASSES:
            raise ValueError(f"no mass tabulated for Z={z}")
        out[i] = ATOMIC_MASSES[z]
    return out


@dataclass
class Structure:
    """One structure: atomic numbers, positions (A), optional cell/pbc."""
    numbers: np.ndarray            # [N] int64
    positions: np.ndarray          # [N, 3] float64
    cell: Optional[np.ndarray] = None   # [3, 3] rows are lattice vectors
    pbc: Optional[np.ndarray] = None    # [3] bool

    def __len__(self) -> int:
        return len(self.numbers)


def element_counts(n_atoms: int) -> List[int]:
    """Largest-remainder rounding of the H/C/N/O/S fractions to n_atoms."""
    raw = [f * n_atoms for f in FRACTIONS]
    counts = [int(math.floor(x)) for x in raw]
    rem = n_atoms - sum(counts)
    order = sorted(range(len(raw)), key=lambda k: (-(raw[k] - counts[k]), k))
    for k in order[:rem]:
        counts[k] += 1
    return counts


def random_molecule(n_atoms: int, rng: np.random.Generator,
                    step_min: float = 1.1, step_max: float = 1.5,
                    min_dist: float = 0.9, max_tries: int = 10000) -> Structure:
    """Grow one molecule as a random bonded walk.

    Atom k (k >= 1) is bonded to a uniformly chosen earlier atom at distance
    U[step_min, step_max] in a uniformly random direction; a placement closer
    than ``min_dist`` to any placed atom is rejected and redrawn.
    """
    if n_atoms < 1:
        raise ValueError("n_atoms must be >= 1")
    counts = element_counts(n_atoms)
    numbers = np.concatenate([np.full(c, z, dtype=np.int64)
                              for z, c in zip(ELEMENTS, counts)])
    rng.shuffle(numbers)
    pos = np.zeros((n_atoms, 3), dtype=np.float64)
    md2 = min_dist * min_dist
    for k in range(1, n_atoms):
        for _ in range(max_tries):
            anchor = int(rng.integers(0, k))
            direction = rng.normal(size=3)
            nrm = np.linalg.norm(direction)
            if nrm < 1e-12:
                continue
            step = rng.uniform(step_min, step_max)
            cand = pos[anchor] + direction / nrm * step
            d2 = np.sum((pos[:k] - cand) ** 2, axis=1)
            if np.all(d2 >= md2):
                pos[k] = cand
                break
        else:
            raise RuntimeError("could not place atom; increase max_tries")
    pos -= pos.mean(axis=0)
    return Structure(numbers=numbers, positions=pos)


def molecule_set(n_molecules: int, n_min: int, n_max: int, seed: int) -> List[Structure]:
    """``n_molecules`` molecules with atom counts uniform in [n_min, n_max]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.integers(n_min, n_max + 1, size=n_molecules)
    return [random_molecule(int(n), rng) for n in sizes]


def single_molecule(n_atoms: int, seed: int) -> Structure:
    rng = np.random.Generator(np.random.PCG64(seed))
    return random_molecule(n_atoms, rng)


def _rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a*a + b*b - c*c - d*d, 2*(b*c - a*d), 2*(b*d + a*c)],
        [2*(b*c + a*d), a*a - b*b + c*c - d*d, 2*(c*d - a*b)],
        [2*(b*d - a*c), 2*(c*d + a*b), a*a - b*b - c*c + d*d]])


def water_box(n_side: int = 10, density_g_cm3: float = 1.0, seed: int = 3,
              jitter: float = 0.15) -> Structure:
    """n_side^3 rigid waters on a jittered cubic lattice in a periodic cube."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n_mol = n_side ** 3
    m_water = 2 * ATOMIC_MASSES[1] + ATOMIC_MASSES[8]
    volume_A3 = n_mol * m_water / AVOGADRO / density_g_cm3 * 1e24
    L = volume_A3 ** (1.0 / 3.0)
    a = L / n_side
    r_oh = 0.96
    half = math.radians(104.5) / 2.0
    local = np.array([[0.0, 0.0, 0.0],
                      [r_oh * math.sin(half), 0.0, r_oh * math.cos(half)],
                      [-r_oh * math.sin(half), 0.0, r_oh * math.cos(half)]])
    local -= (local * np.array([[ATOMIC_MASSES[8]], [ATOMIC_MASSES[1]], [ATOMIC_MASSES[1]]])).sum(0) / m_water
    pos = np.empty((3 * n_mol, 3))
    numbers = np.tile(np.array([8, 1, 1], dtype=np.int64), n_mol)
    k = 0
    for ix in range(n_side):
        for iy in range(n_side):
            for iz in range(n_side):
                centre = (np.array([ix, iy, iz]) + 0.5) * a + rng.uniform(-jitter, jitter, 3)
                R = _rotation(rng)
                pos[3 * k:3 * k + 3] = centre + local @ R.T
                k += 1
    pos %= L
    return Structure(numbers=numbers, positions=pos, cell=np.eye(3) * L,
                     pbc=np.array([True, True, True]))


def maxwell_boltzmann(numbers: np.ndarray, positions: np.ndarray, temperature_K: float,
                      rng: np.random.Generator, zero_com: bool = True,
                      zero_rotation: bool = True,
                      masses: Optional[np.ndarray] = None) -> np.ndarray:
    """Momenta (amu*A/time-unit) drawn at ``temperature_K``.

    Mirrors ASE's MaxwellBoltzmannDistribution (p = N(0,1) * sqrt(m kT)),
    Stationary (subtract the centre-of-mass momentum share m_i P / M) and
    ZeroRotation (subtract omega x (r - r_com) with omega = I^-1 L), the
    calls made at src/mlff_distiller/testing/nve_harness.py:158-165.
    ``masses`` (amu) default to the tabulated masses of ``numbers``.
    """
    m = masses_of(numbers) if masses is None else np.asarray(masses, dtype=np.float64)
    kT = KB_EV * temperature_K
    p = rng.standard_normal((len(m), 3)) * np.sqrt(m * kT)[:, None]
    if zero_com:
        P = p.sum(axis=0)
        p -= m[:, None] * (P / m.sum())
    if zero_rotation and len(m) > 2:
        com = (m[:, None] * positions).sum(0) / m.sum()
        r = positions - com
        v = p / m[:, None]
        L = np.cross(r, p).sum(0)
        Imat = np.zeros((3, 3))
        for mi, ri in zip(m, r):
            Imat += mi * (np.dot(ri, ri) * np.eye(3) - np.outer(ri, ri))
        try:
            omega = np.linalg.solve(Imat, L)
        except np.linalg.LinAlgError:
            omega = np.zeros(3)
        v = v - np.cross(omega, r)
        p = v * m[:, None]
    return p


def flatten(structures: Sequence[Structure]):
    """Concatenate structures: numbers [N], positions [N,3], mol_ptr [B+1]."""
    numbers = np.concatenate([s.numbers for s in structures]).astype(np.int64)
    positions = np.concatenate([s.positions for s in structures]).astype(np.float64)
    sizes = np.array([len(s) for s in structures], dtype=np.int64)
    mol_ptr = np.zeros(len(structures) + 1, dtype=np.int64)
    np.cumsum(sizes, out=mol_ptr[1:])
    return numbers, positions, mol_ptr
