"""A minimal stand-in for ase.Atoms (ASE is not installed here): the methods
the calculator, harness and tests call."""

import numpy as np


class Atoms:
    def __init__(self, numbers, positions, cell=None, pbc=False):
        self.numbers = np.asarray(numbers, dtype=np.int64)
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.zeros((3, 3)) if cell is None else np.asarray(cell, dtype=np.float64)
        self.pbc = np.broadcast_to(np.asarray(pbc, dtype=bool), (3,)).copy()
        self.calc = None

    def __len__(self):
        return len(self.numbers)

    def copy(self):
        a = Atoms(self.numbers.copy(), self.positions.copy(), self.cell.copy(), self.pbc.copy())
        return a

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
