"""ASE calculator for GAP (the interface of gappy/ASE/gap_calc.py:19-52).

Install like the reference: copy this file into ase/calculators/ and use

    from ase.calculators.gap_calc import GAP
    atoms.calc = GAP(rcut=6.0)

Results: energy and free_energy (eV), forces (eV/A), stress (GPa, libgap's
order XX YY ZZ XY YZ ZX and VASP's sign, stored raw as the reference does),
variance (always 0).  The model is read from ./gap_parameters; the libgap
calculator keeps it resident on the GPU between steps.
"""
from __future__ import division

import numpy as np

from ase.calculators.calculator import Calculator, all_changes
import libgap.GAP as my_gap


class GAP(Calculator):
    implemented_properties = ['energy', 'forces', 'stress']

    nolabel = True

    def __init__(self, rcut=6.0, precision=None):
        Calculator.__init__(self)
        self.rcut = rcut
        self.precision = precision

    def calculate(self, atoms=None, properties=['energy'], system_changes=all_changes):
        Calculator.calculate(self, atoms, properties, system_changes)
        gap_calc = my_gap.Calculator(rcut=self.rcut, precision=self.precision)
        species = self.atoms.get_atomic_numbers()
        lat = np.asarray(self.atoms.cell)
        pos = self.atoms.positions
        lgrad = True
        energy, forces, stress, variance = gap_calc.gap_calc(species, lat, pos, lgrad)
        self.results['energy'] = energy
        self.results['free_energy'] = energy
        self.results['forces'] = forces
        self.results['stress'] = stress
        self.results['variance'] = variance
