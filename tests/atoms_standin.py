"""A minimal stand-in for ase.Atoms (ASE is not installed here): the methods
the calculator, harness and tests call."""

import numpy as np


# standard atomic weights (amu) of the elements the tests use, as ase.data
_MASSES = {1: 1.008, 6: 12.011, 7: 14.007, 8: 15.999, 16: 32.06}
# Boltzmann constant in eV/K (CODATA 2014, as ase.units)
_KB = 1.38064852e-23 / 1.6021766208e-19


class Atoms:
    def __init__(self, numbers, positions, cell=None, pbc=False, masses=None, momenta=None):
        self.numbers = np.asarray(numbers, dtype=np.int64)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.zeros((3, 3)) if cell is None else np.asarray(cell, dtype=np.float64)
        self.pbc = np.broadcast_to(np.asarray(pbc, dtype=bool), (3,)).copy()
        self.masses = (np.array([_MASSES.get(int(z), 1.0) for z in self.numbers]) if masses is None
                       else np.asarray(masses, dtype=np.float64).copy())
        self.momenta = (np.zeros((len(self.numbers), 3)) if momenta is None
                        else np.asarray(momenta, dtype=np.float64).reshape(-1, 3).copy())
        self.calc = None

    def __len__(self):
        return len(self.numbers)

    def copy(self):
        a = Atoms(self.numbers.copy(), self.positions.copy(), self.cell.copy(), self.pbc.copy(),
                  self.masses, self.momenta)
        return a

    def get_masses(self):
        return self.masses.copy()

    def set_masses(self, m):
        self.masses = np.asarray(m, dtype=np.float64).copy()

    def get_momenta(self):
        return self.momenta.copy()

    def set_momenta(self, p):
        self.momenta = np.asarray(p, dtype=np.float64).reshape(-1, 3).copy()

    def get_velocities(self):
        return self.momenta / self.masses[:, None]

    def set_velocities(self, v):
        self.momenta = np.asarray(v, dtype=np.float64).reshape(-1, 3) * self.masses[:, None]

    def get_kinetic_energy(self):
        return 0.5 * float((self.momenta ** 2 / self.masses[:, None]).sum())

    def get_temperature(self):
        return 2 * self.get_kinetic_energy() / (3 * len(self) * _KB)

    def get_positions(self):
        return self.positions.copy()

    def set_positions(self, p):
        self.positions = np.asarray(p, dtype=np.float64).copy()

    def get_atomic_numbers(self):
        return self.numbers.copy()

    def get_cell(self):
        return self.cell.copy()

    def get_pbc(self):
        return self.pbc.copy()

    def get_potential_energy(self):
        return self.calc.get_potential_energy(self)

    def get_forces(self):
        return self.calc.get_forces(self)

    def get_stress(self):
        return self.calc.get_stress(self)
