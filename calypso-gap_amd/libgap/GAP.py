"""GAP calculator (the API of gappy/libgap/GAP.py:14-61).

    from libgap.GAP import Calculator
    gap = Calculator(rcut=6.0)
    ene, force, stress, variance = gap.gap_calc(species, lat, pos, lgrad)

``species`` may be chemical symbols or atomic numbers; ``lat`` rows are the
lattice vectors (A); ``pos`` Cartesian (A).  Energy in eV, forces in eV/A,
stress in GPa in the order XX YY ZZ XY YZ ZX with VASP's sign (gappy/README.md).
The model is read from ./gap_parameters, as in the reference.

Differences from the reference, all on purpose:
  * the model file is parsed once per (path, size, mtime) and the model stays
    resident on the GPU, instead of being re-read on every call;
  * errors raise ``RuntimeError`` instead of stopping the Python process;
  * ``precision`` selects the FP64 path (default, parity with the reference)
    or the mixed-precision path ("mixed", "mixed-bf16").
"""
import os

import numpy as np

from libgap.libgap import fgap_calc_status, fgap_last_error, fgap_read_dims, fgap_read2

_CHEMICAL_SYMBOLS = [
    'X', 'H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne', 'Na', 'Mg', 'Al', 'Si', 'P', 'S', 'Cl', 'Ar', 'K',
    'Ca', 'Sc', 'Ti', 'V', 'Cr', 'Mn', 'Fe', 'Co', 'Ni', 'Cu', 'Zn', 'Ga', 'Ge', 'As', 'Se', 'Br', 'Kr', 'Rb',
    'Sr', 'Y', 'Zr', 'Nb', 'Mo', 'Tc', 'Ru', 'Rh', 'Pd', 'Ag', 'Cd', 'In', 'Sn', 'Sb', 'Te', 'I', 'Xe', 'Cs', 'Ba',
    'La', 'Ce', 'Pr', 'Nd', 'Pm', 'Sm', 'Eu', 'Gd', 'Tb', 'Dy', 'Ho', 'Er', 'Tm', 'Yb', 'Lu', 'Hf', 'Ta', 'W',
    'Re', 'Os', 'Ir', 'Pt', 'Au', 'Hg', 'Tl', 'Pb', 'Bi', 'Po', 'At', 'Rn', 'Fr', 'Ra', 'Ac', 'Th', 'Pa', 'U',
    'Np', 'Pu', 'Am', 'Cm', 'Bk', 'Cf', 'Es', 'Fm', 'Md', 'No', 'Lr']
_ATOMIC_NUMBER = {s: z for z, s in enumerate(_CHEMICAL_SYMBOLS)}
_MODES = {None: -1, "fp64": 0, "mixed": 1, "mixed-bf16": 2}
_MODEL_CACHE = {}


def _read_model():
    path = os.path.abspath("gap_parameters")
    try:
        st = os.stat(path)
    except OSError:
        raise RuntimeError("gap_parameters file does not exist!")
    key = (path, st.st_size, st.st_mtime_ns)
    if key not in _MODEL_CACHE:
        nsparse, des_len, status = fgap_read_dims()
        if status != 0:
            raise RuntimeError("gap_parameters: cannot read the model")
        theta, mm, coeff, status = fgap_read2(nsparse, des_len)
        if status != 0:
            raise RuntimeError("gap_parameters: cannot read the model")
        _MODEL_CACHE.clear()
        _MODEL_CACHE[key] = (nsparse, des_len, theta, mm, coeff)
    return _MODEL_CACHE[key]


class Calculator(object):
    def __init__(self, rcut=6.0, precision=None):
        if precision not in _MODES:
            raise ValueError("precision must be one of %s" % sorted(k for k in _MODES if k))
        self.gap_read()
        self.rcut = rcut
        self.precision = precision

    def get_elenum(self, x):
        return [_ATOMIC_NUMBER[element] for element in x]

    def gap_read(self):
        self.nsparseX, self.des_len, self.theta, self.mm, self.coeff = _read_model()

    @property
    def invcmm(self):
        """Inverse sparse covariance; the reference never reads it and returns
        zeros (gap_calc.f90:358-361)."""
        return np.zeros((self.nsparseX, self.nsparseX))

    def gap_calc(self, species, lat, pos, lgrad):
        try:
            Species = self.get_elenum(species)
        except Exception:
            Species = species
        ene, force, stress, variance, status = fgap_calc_status(
            np.asarray(Species, dtype=np.int32), np.asarray(lat, dtype=np.float64),
            np.asarray(pos, dtype=np.float64), self.theta[0:self.des_len],
            self.mm[0:self.nsparseX, 0:self.des_len], self.coeff[0:self.nsparseX], self.rcut, lgrad,
            _MODES[self.precision])
        if status != 0:
            raise RuntimeError("gap_calc failed (%d): %s" % (status, fgap_last_error().decode(errors="replace").strip()))
        return ene, force, stress, variance
