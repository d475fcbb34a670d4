"""Python access to the CPU FP64 oracle (test infrastructure only).

This module is the oracle the HIP evaluator is tested against.  Only the tests, the smoke
check and the benchmark's CPU-baseline leg import it; the product path never
does.  It wraps ``oracle/liboracle.so`` (the C restatement of libgap's
energy/force/stress path, gap_oracle.c) and, where it has been built,
``oracle/_ref/libgapref.so`` (the reference Fortran compiled from its own
sources, see oracle/Makefile).

Model files follow the reference's ``gap_parameters`` format
(gap_calc.f90:303-364 and wacsf.f90:40-56), read here with Fortran
list-directed semantics: a READ with k items consumes whitespace-separated
tokens across records until k values are read and then discards the rest of
the last record; a READ with no items skips one record.
"""
import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libgapref.so")

GPA = 6.24219e-3          # gap_calc.f90:9
PI_TRUNC = 3.141592654    # wacsf.f90:31


# --------------------------------------------------------------------------
# model
# --------------------------------------------------------------------------

@dataclass
class Model:
    z: np.ndarray                 # [nspecies] int
    w: np.ndarray                 # [nspecies] float
    sf_type: np.ndarray           # [nsf] int
    sf_alpha: np.ndarray          # [nsf] eta or Rs
    sf_cut: np.ndarray            # [nsf]
    theta: np.ndarray             # [D]
    mm: np.ndarray                # [M, D]
    coeff: np.ndarray             # [M]
    sf_zeta: np.ndarray = None    # [nsf] or None (all 1)
    extra: dict = field(default_factory=dict)

    @property
    def nsf(self):
        return len(self.sf_type)

    @property
    def des_len(self):
        return self.mm.shape[1]

    @property
    def nsparse(self):
        return self.mm.shape[0]

    def weight_of(self, species):
        table = {}
        for z, w in zip(self.z, self.w):
            table[int(z)] = float(w)
        return np.array([table[int(s)] for s in species])


class _Records:
    """Fortran list-directed reader over a text file."""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.pos = 0

    def skip(self):
        self.pos += 1

    def read(self, n, conv=float):
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise EOFError("end of model file while reading %d values" % n)
            toks = self.lines[self.pos].replace(",", " ").split()
            self.pos += 1
            vals.extend(toks)
        return [conv(float(v.replace("d", "e").replace("D", "e"))) for v in vals[:n]]


def parse_model(path):
    """Read ``gap_parameters`` (Appendix layout: species table, SF table,
    ``nsparseX des_len``, three skipped records, theta, MM rows, coeff)."""
    r = _Records(open(path).read())
    nspecies = r.read(1, int)[0]
    z, w = [], []
    for _ in range(nspecies):
        a, b = r.read(2)
        z.append(int(float(a)))
        w.append(float(b))
    nsf = r.read(1, int)[0]
    types, alphas, cuts, zetas = [], [], [], []
    for _ in range(nsf):
        line_start = r.pos
        t, a, c = r.read(3)
        types.append(int(float(t)))
        alphas.append(float(a))
        cuts.append(float(c))
        # optional 4th column (angular exponent zeta); the reference's read
        # ignores trailing values on the record, so the format stays compatible
        toks = r.lines[line_start].split() if r.pos == line_start + 1 else []
        zetas.append(float(toks[3]) if len(toks) >= 4 else 1.0)
    nsparse, des_len = (int(v) for v in r.read(2, int))
    r.skip()
    r.skip()
    r.skip()
    theta = np.array(r.read(des_len), dtype=np.float64)
    mm = np.empty((nsparse, des_len))
    for k in range(nsparse):
        mm[k] = r.read(des_len)
    coeff = np.array(r.read(nsparse), dtype=np.float64)
    zeta = np.array(zetas)
    return Model(z=np.array(z, dtype=np.int32), w=np.array(w), sf_type=np.array(types, dtype=np.int32),
                 sf_alpha=np.array(alphas), sf_cut=np.array(cuts), theta=theta, mm=mm, coeff=coeff,
                 sf_zeta=None if np.all(zeta == 1.0) else zeta)


def write_model(path, m):
    """Write a model in the ``gap_parameters`` layout (values in %.17g)."""
    out = ["%d" % len(m.z)]
    for z, w in zip(m.z, m.w):
        out.append("%d %.17g" % (z, w))
    out.append("%d" % m.nsf)
    for s in range(m.nsf):
        line = "%d %.17g %.17g" % (m.sf_type[s], m.sf_alpha[s], m.sf_cut[s])
        if m.sf_zeta is not None and m.sf_zeta[s] != 1.0:
            line += " %.17g" % m.sf_zeta[s]
        out.append(line)
    out.append("%d %d" % (m.nsparse, m.des_len))
    out += ["", "", ""]
    out.append(" ".join("%.17g" % v for v in m.theta))
    for k in range(m.nsparse):
        out.append(" ".join("%.17g" % v for v in m.mm[k]))
    out.append(" ".join("%.17g" % v for v in m.coeff))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# --------------------------------------------------------------------------
# ctypes binding of liboracle.so
# --------------------------------------------------------------------------

class _CModel(ctypes.Structure):
    _fields_ = [("nspecies", ctypes.c_int), ("z", ctypes.c_void_p), ("w", ctypes.c_void_p),
                ("nsf", ctypes.c_int), ("type", ctypes.c_void_p), ("alpha", ctypes.c_void_p),
                ("cut", ctypes.c_void_p), ("zeta", ctypes.c_void_p), ("des_len", ctypes.c_int),
                ("nsparse", ctypes.c_int), ("theta", ctypes.c_void_p), ("mm", ctypes.c_void_p),
                ("coeff", ctypes.c_void_p)]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        _lib.orc_gap_calc_faithful.argtypes = [vp, ci, vp, vp, vp, cd, ci, vp, vp, vp, vp, vp, vp, vp]
        _lib.orc_gap_calc_sparse.argtypes = [vp, ci, vp, vp, vp, cd, ci, ci, vp, vp, vp, vp, vp, vp]
        _lib.orc_gap_calc_sparse_range.argtypes = [vp, ci, vp, vp, vp, cd, ci, ci, ci, ci, vp, vp, vp]
        _lib.orc_descriptors.argtypes = [vp, ci, vp, vp, vp, cd, ci, vp]
        _lib.orc_neighbor_set.argtypes = [ci, vp, vp, cd, ctypes.c_long, vp, vp, vp]
        _lib.orc_neighbor_set.restype = ctypes.c_long
        _lib.orc_species_weights.argtypes = [vp, ci, vp, vp]
        _lib.orc_neighbors_ref.argtypes = [ci, vp, vp, vp, cd, ci, vp, vp]
        _lib.orc_car2acsf_dense.argtypes = [vp, ci, ci, vp, vp, vp, vp, vp, vp, ci]
        _lib.orc_car2acsf_dense.restype = None
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


class _Pinned:
    """Keeps the contiguous arrays behind a _CModel alive."""

    def __init__(self, m):
        self.z = np.ascontiguousarray(m.z, dtype=np.int32)
        self.w = np.ascontiguousarray(m.w, dtype=np.float64)
        self.t = np.ascontiguousarray(m.sf_type, dtype=np.int32)
        self.a = np.ascontiguousarray(m.sf_alpha, dtype=np.float64)
        self.c = np.ascontiguousarray(m.sf_cut, dtype=np.float64)
        self.zeta = None if m.sf_zeta is None else np.ascontiguousarray(m.sf_zeta, dtype=np.float64)
        self.theta = np.ascontiguousarray(m.theta, dtype=np.float64)
        self.mm = np.ascontiguousarray(m.mm, dtype=np.float64)
        self.coeff = np.ascontiguousarray(m.coeff, dtype=np.float64)
        self.cm = _CModel(len(self.z), _ptr(self.z), _ptr(self.w), len(self.t), _ptr(self.t),
                          _ptr(self.a), _ptr(self.c), _ptr(self.zeta), self.mm.shape[1],
                          self.mm.shape[0], _ptr(self.theta), _ptr(self.mm), _ptr(self.coeff))


_ERR = {-1: "more than 1000 neighbours", -2: "species not listed in the model",
        -3: "inconsistent model", -4: "bad atom range"}


def gap_calc(model, species, lat, pos, rcut, lgrad=True, mode="faithful", nthreads=1,
             want_details=False):
    """E (eV), F (N,3) eV/A, S (6,) GPa in libgap's order and sign."""
    pm = _Pinned(model)
    species = np.ascontiguousarray(species, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    na = len(species)
    D = model.des_len
    ene = np.zeros(1)
    force = np.zeros((na, 3))
    stress = np.zeros(6)
    var = np.zeros(1)
    xx = np.zeros((na, D)) if want_details else None
    dedg = np.zeros((na, D)) if want_details else None
    eatom = np.zeros(na) if want_details else None
    L = lib()
    if mode == "faithful":
        rc = L.orc_gap_calc_faithful(ctypes.byref(pm.cm), na, _ptr(species), _ptr(lat), _ptr(pos),
                                     rcut, int(lgrad), _ptr(ene), _ptr(force), _ptr(stress), _ptr(var),
                                     _ptr(xx), _ptr(dedg), _ptr(eatom))
    elif mode == "sparse":
        rc = L.orc_gap_calc_sparse(ctypes.byref(pm.cm), na, _ptr(species), _ptr(lat), _ptr(pos), rcut,
                                   int(lgrad), int(nthreads), _ptr(ene), _ptr(force), _ptr(stress),
                                   _ptr(xx), _ptr(dedg), _ptr(eatom))
    else:
        raise ValueError(mode)
    if rc != 0:
        raise RuntimeError("oracle: " + _ERR.get(rc, str(rc)))
    out = dict(energy=float(ene[0]), forces=force, stress=stress, variance=0.0)
    if want_details:
        out.update(x=xx, dedg=dedg, eatom=eatom)
    return out


def gap_calc_range(model, species, lat, pos, rcut, i0, i1, nthreads=1):
    """Contributions of centre atoms [i0, i1) (bounded CPU-baseline sample)."""
    pm = _Pinned(model)
    species = np.ascontiguousarray(species, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    na = len(species)
    ene = np.zeros(1)
    force = np.zeros((na, 3))
    stress = np.zeros(6)
    rc = lib().orc_gap_calc_sparse_range(ctypes.byref(pm.cm), na, _ptr(species), _ptr(lat), _ptr(pos),
                                         rcut, 1, int(nthreads), int(i0), int(i1), _ptr(ene),
                                         _ptr(force), _ptr(stress))
    if rc != 0:
        raise RuntimeError("oracle: " + _ERR.get(rc, str(rc)))
    return dict(energy=float(ene[0]), forces=force, stress=stress)


def descriptors(model, species, lat, pos, rcut, nthreads=0):
    pm = _Pinned(model)
    species = np.ascontiguousarray(species, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    xx = np.zeros((len(species), model.des_len))
    rc = lib().orc_descriptors(ctypes.byref(pm.cm), len(species), _ptr(species), _ptr(lat), _ptr(pos),
                               rcut, int(nthreads), _ptr(xx))
    if rc != 0:
        raise RuntimeError("oracle: " + _ERR.get(rc, str(rc)))
    return xx


def neighbor_table(model, species, lat, pos, rcut, max_nb=1000):
    """FGAP_CALC's neighbour table (gap_calc.f90:89-120) in the layout CAR2ACSF
    takes (wacsf.f90:5-6): Fortran NEIGHBOR(NA, max_nb, 6) = image x y z,
    distance, species weight, 1-based atom index; and the counts (int32)."""
    pm = _Pinned(model)
    species = np.ascontiguousarray(species, dtype=np.int32)
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    na = len(species)
    w = np.zeros(na)
    L = lib()
    if L.orc_species_weights(ctypes.byref(pm.cm), na, _ptr(species), _ptr(w)):
        raise RuntimeError("oracle: species not listed in the model")
    nb = np.zeros((na, max_nb, 6))
    cnt = np.zeros(na, dtype=np.int32)
    if L.orc_neighbors_ref(na, _ptr(lat), _ptr(pos), _ptr(w), rcut, max_nb, _ptr(nb), _ptr(cnt)):
        raise RuntimeError("oracle: more than %d neighbours" % max_nb)
    nb[:, :, 5] += 1.0  # 1-based, as FGAP_CALC stores real(j) (gap_calc.f90:115)
    for i in range(na):
        nb[i, cnt[i]:, :] = 0.0
    return np.asfortranarray(nb), cnt


def car2acsf(model, pos, neighbor, count, lgrad=True):
    """CAR2ACSF restated (wacsf.f90:2-796) on a neighbour table in the
    reference's layout (neighbor_table): xx(nf, na), dxdy(nf, na, na, 3),
    strs(3, 3, nf, na), Fortran-ordered arrays."""
    pm = _Pinned(model)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    na, max_nb = neighbor.shape[0], neighbor.shape[1]
    nb = np.ascontiguousarray(neighbor, dtype=np.float64).copy()
    nb[:, :, 5] -= 1.0  # the restatement indexes atoms from 0
    cnt = np.ascontiguousarray(count, dtype=np.int32)
    D = model.des_len
    xx = np.zeros(na * D)
    dxdy = np.zeros(na * na * 3 * D) if lgrad else None
    strs = np.zeros(na * 9 * D) if lgrad else None
    lib().orc_car2acsf_dense(ctypes.byref(pm.cm), na, max_nb, _ptr(pos), _ptr(nb), _ptr(cnt), _ptr(xx),
                             _ptr(dxdy), _ptr(strs), int(lgrad))
    out_x = np.asfortranarray(xx.reshape(na, D).T)
    if not lgrad:
        return out_x, np.zeros((D, na, na, 3), order="F"), np.zeros((3, 3, D, na), order="F")
    out_d = np.asfortranarray(dxdy.reshape(na, na, 3, D).transpose(3, 0, 1, 2))
    out_s = np.asfortranarray(strs.reshape(na, 3, 3, D).transpose(1, 2, 3, 0))
    return out_x, out_d, out_s


def min_bond(lat, pos, rcut, chunk=256):
    """FGET_BOND restated (gappy/libgap/get_bond.f90:31-78): min(10, the
    shortest |pos_i - (pos_j + n1 L1 + n2 L2 + n3 L3)| <= rcut over every
    pair and every image n in [-n_k, n_k], (i, i, 0) skipped), with
    n_k = ceil(rcut |b_k| / 2 pi) from the reciprocal vectors of the COLUMNS
    of LAT (get_bond.f90:40-43, 89-98) and the distance in the reference's
    order of operations (each product rounded, sums left to right)."""
    lat = np.asarray(lat, dtype=np.float64)
    pos = np.asarray(pos, dtype=np.float64)
    col = lat.T  # col[c] = LAT(:, c + 1)
    rc = np.stack([np.cross(col[1], col[2]), np.cross(col[2], col[0]), np.cross(col[0], col[1])])
    vol = abs(np.sum(col[0] * rc[0]))
    nabc = []
    for r in range(3):
        v = np.array([rc[0][r], rc[1][r], rc[2][r]]) / vol * 3.141592654 * 2.0
        nabc.append(int(math.ceil(rcut * math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) / 3.141592654 / 2.0)))
    best = 10.0
    na = len(pos)
    for n1 in range(-nabc[0], nabc[0] + 1):
        for n2 in range(-nabc[1], nabc[1] + 1):
            for n3 in range(-nabc[2], nabc[2] + 1):
                img = ((pos + float(n1) * lat[0]) + float(n2) * lat[1]) + float(n3) * lat[2]  # xyz of every j
                for i0 in range(0, na, chunk):
                    dr = pos[i0:i0 + chunk, None, :] - img[None, :, :]
                    dis = np.sqrt((dr[..., 0] * dr[..., 0] + dr[..., 1] * dr[..., 1]) + dr[..., 2] * dr[..., 2])
                    if n1 == 0 and n2 == 0 and n3 == 0:
                        ii = np.arange(i0, min(i0 + chunk, na))
                        dis[ii - i0, ii] = np.inf
                    dis = dis[dis <= rcut]
                    if dis.size:
                        best = min(best, float(dis.min()))
    return best


def neighbor_set(lat, pos, rcut):
    """Sorted (i, j, n1, n2, n3) rows and distances of every pair within rcut."""
    lat = np.ascontiguousarray(lat, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    na = len(pos)
    counts = np.zeros(na, dtype=np.int32)
    L = lib()
    total = L.orc_neighbor_set(na, _ptr(lat), _ptr(pos), rcut, 0, ctypes.c_void_p(0), ctypes.c_void_p(0),
                               _ptr(counts))
    out4 = np.zeros((max(total, 1), 4), dtype=np.int32)
    dist = np.zeros(max(total, 1))
    L.orc_neighbor_set(na, _ptr(lat), _ptr(pos), rcut, total, _ptr(out4), _ptr(dist), _ptr(counts))
    rows = np.concatenate([np.repeat(np.arange(na, dtype=np.int32), counts)[:, None], out4[:total]], axis=1)
    return rows, dist[:total]


def covariance(model, x, block=64):
    """k_ik = delta exp(-1/2 sum_d ((x_id - MM_kd) / theta_d)^2), delta = 1:
    FGAP_CALC's covariance (gap_calc.f90:270-287, delta from gap_calc.f90:8),
    from direct differences, a block of rows at a time."""
    x = np.asarray(x, dtype=np.float64)
    theta = np.asarray(model.theta, dtype=np.float64)
    mt = np.asarray(model.mm, dtype=np.float64) / theta
    out = np.empty((len(x), len(mt)))
    for r0 in range(0, len(x), block):
        xt = x[r0:r0 + block] / theta
        d2 = ((xt[:, None, :] - mt[None, :, :]) ** 2).sum(axis=2)
        out[r0:r0 + block] = np.exp(-0.5 * d2)
    return out


def gpr_variance(model, x, qmm, natoms=None):
    """Predictive variance: sigma2_i = delta - k_i Q k_i^T per atom, and the
    mean over each structure's atoms (natoms per structure; one structure if
    None).  This is the formula FGAP_CALC keeps commented out
    (gap_calc.f90:205-210: covf(i) = delta - dot_product(ckm(i,:),
    matmul(qmm, ckm(i,:))), VARIANCE = sum covf / na); the reference itself
    returns VARIANCE = 0.  qmm None means Q = 0 (FGAP_READ's INVCMM,
    gap_calc.f90:355-358)."""
    x = np.asarray(x, dtype=np.float64)
    if qmm is None:
        s2 = np.ones(len(x))
    else:
        K = covariance(model, x)
        s2 = 1.0 - np.einsum("ik,ik->i", K @ np.asarray(qmm, dtype=np.float64).T, K)
    if natoms is None:
        natoms = [len(x)]
    offs = np.concatenate([[0], np.cumsum(natoms)]).astype(int)
    var = np.array([s2[offs[k]:offs[k + 1]].mean() for k in range(len(natoms))])
    return var, s2


# --------------------------------------------------------------------------
# the reference itself (oracle/_ref), for writing golden vectors
# --------------------------------------------------------------------------

def ref_available():
    return os.path.exists(REF_PATH)


def ref_gap_calc(model_path, species, lat, pos, theta, mm, coeff, rcut, lgrad=True):
    """Call the reference FGAP_CALC (compiled Fortran) in the model's directory.

    FGAP_CALC reads species weights and the SF table from ./gap_parameters
    (gap_calc.f90:75, wacsf.f90:45), so the call runs with the working
    directory set to the folder holding the model file.
    """
    ref = ctypes.CDLL(REF_PATH)
    na = len(species)
    M, D = mm.shape
    args = dict(
        NA=ctypes.c_int(na), SPECIES=np.asfortranarray(species, dtype=np.int32),
        LAT=np.asfortranarray(lat, dtype=np.float64), POS=np.asfortranarray(pos, dtype=np.float64),
        ENE=ctypes.c_double(0.0), FORCE=np.zeros((na, 3), order="F"), STRESS=np.zeros(6),
        VAR=ctypes.c_double(0.0), NSP=ctypes.c_int(M), DL=ctypes.c_int(D),
        THETA=np.asfortranarray(theta, dtype=np.float64), MM=np.asfortranarray(mm, dtype=np.float64),
        QMM=np.zeros((M, M), order="F"), COEFF=np.asfortranarray(coeff, dtype=np.float64),
        RCUT=ctypes.c_double(rcut), LGRAD=ctypes.c_int(1 if lgrad else 0))
    cwd = os.getcwd()
    os.chdir(os.path.dirname(os.path.abspath(model_path)))
    try:
        def p(a):
            return a.ctypes.data_as(ctypes.c_void_p) if isinstance(a, np.ndarray) else ctypes.byref(a)
        ref.fgap_calc_(*(p(args[k]) for k in ("NA", "SPECIES", "LAT", "POS", "ENE", "FORCE", "STRESS",
                                              "VAR", "NSP", "DL", "THETA", "MM", "QMM", "COEFF", "RCUT",
                                              "LGRAD")))
    finally:
        os.chdir(cwd)
    return dict(energy=args["ENE"].value, forces=np.array(args["FORCE"]), stress=args["STRESS"].copy(),
                variance=args["VAR"].value)


def ref_car2acsf(model_path, pos, neighbor, count, nf, lgrad=True):
    """The reference CAR2ACSF (compiled Fortran, wacsf.f90:2-12), run in the
    model file's directory (it reads the SF table from ./gap_parameters)."""
    ref = ctypes.CDLL(REF_PATH)
    na = len(count)
    max_nb = neighbor.shape[1]
    P = np.asfortranarray(pos, dtype=np.float64)
    NB = np.asfortranarray(neighbor, dtype=np.float64)
    C = np.asfortranarray(count, dtype=np.int32)
    xx = np.zeros((nf, na), order="F")
    dxdy = np.zeros((nf, na, na, 3), order="F")
    strs = np.zeros((3, 3, nf, na), order="F")
    args = [ctypes.c_int(na), ctypes.c_int(max_nb), ctypes.c_int(nf)]
    lg = ctypes.c_int(1 if lgrad else 0)
    cwd = os.getcwd()
    os.chdir(os.path.dirname(os.path.abspath(model_path)))
    try:
        ref.car2acsf_(*(ctypes.byref(a) for a in args), P.ctypes.data_as(ctypes.c_void_p),
                      NB.ctypes.data_as(ctypes.c_void_p), C.ctypes.data_as(ctypes.c_void_p),
                      xx.ctypes.data_as(ctypes.c_void_p), dxdy.ctypes.data_as(ctypes.c_void_p),
                      strs.ctypes.data_as(ctypes.c_void_p), ctypes.byref(lg))
    finally:
        os.chdir(cwd)
    return xx, dxdy, strs


def ref_write_array_2dim(a, path):
    """The reference write_array_2dim (compiled Fortran, wacsf.f90:798-810)."""
    ref = ctypes.CDLL(REF_PATH)
    A = np.asfortranarray(a, dtype=np.float64)
    name = os.fsencode(path)
    ref.write_array_2dim_(ctypes.byref(ctypes.c_int(A.shape[0])), ctypes.byref(ctypes.c_int(A.shape[1])),
                          A.ctypes.data_as(ctypes.c_void_p), ctypes.c_char_p(name), ctypes.c_size_t(len(name)))


def ref_get_bond(lat, species, pos, rcut):
    """The reference FGET_BOND (compiled Fortran, get_bond.f90:4-117)."""
    ref = ctypes.CDLL(REF_PATH)
    na = ctypes.c_int(len(species))
    L = np.asfortranarray(lat, dtype=np.float64)
    P = np.asfortranarray(pos, dtype=np.float64)
    E = np.asfortranarray(species, dtype=np.int32)
    rc = ctypes.c_double(rcut)
    out = ctypes.c_double(0.0)
    ref.fget_bond_(ctypes.byref(na), L.ctypes.data_as(ctypes.c_void_p), E.ctypes.data_as(ctypes.c_void_p),
                   P.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rc), ctypes.byref(out))
    return out.value
