"""Minimum interatomic distance (the API of gappy/libgap/BOND.py:14-52)."""
import numpy as np

from libgap.libgap import fget_bond
from libgap.GAP import _ATOMIC_NUMBER


class Bond(object):

    def __init__(self, nf=None, rcut=6.0, lgrad=True):
        self.nf = nf
        self.rcut = rcut
        self.lgrad = lgrad

    def get_elenum(self, x):
        return [_ATOMIC_NUMBER[element] for element in x]

    def get_min_bond(self, lat, elements, pos):
        try:
            Elements = self.get_elenum(elements)
        except Exception:
            Elements = elements
        return fget_bond(np.asarray(lat, dtype=np.float64), np.asarray(Elements, dtype=np.int32),
                         np.asarray(pos, dtype=np.float64), self.rcut)
