"""Minimal stand-in for the parts of ASE the GAP calculator touches (ASE is
not installed here).  Test scaffolding only."""
import numpy as np


class Atoms:
    def __init__(self, numbers, positions, cell, pbc=True):
        self.numbers = np.asarray(numbers, dtype=np.int64)
        self.positions = np.asarray(positions, dtype=np.float64)
        self.cell = np.asarray(cell, dtype=np.float64)
        self.pbc = pbc
        self.calc = None

    def get_atomic_numbers(self):
        return self.numbers.copy()

    def copy(self):
        return Atoms(self.numbers, self.positions.copy(), self.cell.copy(), self.pbc)

    def get_potential_energy(self):
        self.calc.calculate(self, ["energy"])
        return self.calc.results["energy"]

    def get_forces(self):
        self.calc.calculate(self, ["forces"])
        return self.calc.results["forces"]

    def get_stress(self):
        self.calc.calculate(self, ["stress"])
        return self.calc.results["stress"]
