"""Python binding of the gapx C ABI (include/gapx.h) over ctypes.

``gapx`` loads ``calypso-gap_amd/lib/libgapx.so`` (built in-tree by
``calypso-gap_amd/build.py``) and raises ``GapxError`` when it is missing:
there is no CPU fallback.

    model = gapx.Model.read("gap_parameters")           # or Model.from_arrays(...)
    e, f, s, var = model.calc(species, lat, pos, rcut=6.0)   # FGAP_CALC semantics

    ctx = gapx.Context(model)                            # device-resident batch
    ctx.load([species...], [lat...], [pos...], rcut)
    ctx.run(gapx.GRAD | gapx.MIXED); ctx.results()
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libgapx.so")

OK = 0
GRAD = 1
MIXED = 2
LAYOUT_F = 4
REF_NEIGHBOR_CAP = 8
BF16 = 16
VARIANCE = 64   # gapx_calc: compute VARIANCE instead of the reference's 0
FS = 0.09822694788464063   # 1 fs in ASE time units (CODATA 2014, ASE 3.18.1)
KB = 8.6173303e-5          # eV/K, the same units

ERR_NAMES = {1: "GAPX_ERR_ARG", 2: "GAPX_ERR_FILE", 3: "GAPX_ERR_MODEL", 4: "GAPX_ERR_SPECIES",
             5: "GAPX_ERR_NEIGHBORS", 6: "GAPX_ERR_HIP", 7: "GAPX_ERR_NOMEM", 8: "GAPX_ERR_CAP"}

# every symbol include/gapx.h declares
EXPORTS = [
    "gapx_version", "gapx_last_error", "gapx_device_count", "gapx_model_create", "gapx_model_read",
    "gapx_model_dims", "gapx_model_arrays", "gapx_model_header", "gapx_model_free", "gapx_file_stamp",
    "gapx_calc", "gapx_calc_batch", "gapx_descriptors", "gapx_ctx_create", "gapx_ctx_load",
    "gapx_ctx_update", "gapx_ctx_run", "gapx_ctx_sync", "gapx_ctx_results", "gapx_ctx_features",
    "gapx_ctx_timing", "gapx_ctx_stage_count", "gapx_ctx_stage_info", "gapx_ctx_use_graph",
    "gapx_ctx_device_bytes", "gapx_ctx_free", "gapx_ctx_counts", "gapx_min_distance", "gapx_min_distance_noalloc",
    "gapx_ctx_autotune", "gapx_ctx_get_tuning", "gapx_ctx_set_tuning",
    "gapx_mlp_param_count", "gapx_mlp_create", "gapx_mlp_free", "gapx_mlp_get_params", "gapx_mlp_set_params",
    "gapx_mlp_forward", "gapx_mlp_loss_grad", "gapx_mlp_data_add", "gapx_mlp_data_clear", "gapx_mlp_data_count",
    "gapx_mlp_train", "gapx_ctx_export_labels", "gapx_ctx_run_surrogate", "gapx_ctx_relax",
    "gapx_ctx_positions", "gapx_model_set_qmm", "gapx_ctx_variance",
    "gapx_ctx_md", "gapx_ctx_momenta", "gapx_model_write", "gapx_model_save", "gapx_model_load",
    "gapx_model_read_cached", "gapx_probe_fma_rates", "gapx_model_precision", "gapx_car2acsf",
    "gapx_ctx_relax_lbfgs",
]


class GapxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERR_NAMES.get(code, str(code)), msg))
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GapxError(6, "HIP library not built: %s (run calypso-gap_amd/build.py)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.gapx_last_error.restype = ctypes.c_char_p
        L.gapx_model_create.argtypes = [ci, vp, vp, ci, vp, vp, vp, vp, ci, ci, vp, vp, ci, ci, vp, vp]
        L.gapx_model_read.argtypes = [ctypes.c_char_p, vp]
        L.gapx_model_dims.argtypes = [vp, vp, vp, vp, vp]
        L.gapx_model_arrays.argtypes = [vp, vp, vp, ci, ci, vp]
        L.gapx_model_header.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.gapx_model_free.argtypes = [vp]
        L.gapx_model_precision.argtypes = [vp, vp, vp, vp, vp]
        L.gapx_model_free.restype = None
        L.gapx_file_stamp.argtypes = [ctypes.c_char_p, vp, vp]
        L.gapx_calc.argtypes = [vp, ci, vp, vp, vp, cd, ci, vp, vp, vp, vp]
        L.gapx_calc_batch.argtypes = [vp, ci, vp, vp, vp, vp, cd, ci, vp, vp, vp]
        L.gapx_descriptors.argtypes = [vp, ci, vp, vp, vp, cd, ci, vp]
        L.gapx_ctx_create.argtypes = [vp, vp]
        L.gapx_ctx_load.argtypes = [vp, ci, vp, vp, vp, vp, cd]
        L.gapx_ctx_update.argtypes = [vp, vp, vp]
        L.gapx_ctx_run.argtypes = [vp, ci]
        L.gapx_ctx_sync.argtypes = [vp]
        L.gapx_ctx_results.argtypes = [vp, vp, vp, vp, vp]
        L.gapx_ctx_features.argtypes = [vp, vp, vp]
        L.gapx_ctx_timing.argtypes = [vp, ci]
        L.gapx_ctx_autotune.argtypes = [vp, ci, ci]
        cu = ctypes.c_ulonglong
        cf = ctypes.c_float
        L.gapx_mlp_param_count.argtypes = [vp]
        L.gapx_mlp_create.argtypes = [cu, vp]
        L.gapx_mlp_free.argtypes = [vp]
        L.gapx_mlp_free.restype = None
        L.gapx_mlp_get_params.argtypes = [vp, vp]
        L.gapx_mlp_set_params.argtypes = [vp, vp]
        L.gapx_mlp_forward.argtypes = [vp, ci, vp, ci, vp, vp]
        L.gapx_mlp_loss_grad.argtypes = [vp, ci, vp, vp, vp, cf, cf, ci, vp, vp]
        L.gapx_mlp_data_add.argtypes = [vp, ci, vp, vp, vp]
        L.gapx_mlp_data_clear.argtypes = [vp]
        L.gapx_mlp_data_count.argtypes = [vp, vp]
        L.gapx_mlp_train.argtypes = [vp, ci, ci, cf, cf, cf, cu, ci, vp]
        L.gapx_ctx_export_labels.argtypes = [vp, vp]
        L.gapx_ctx_run_surrogate.argtypes = [vp, vp, ci]
        L.gapx_ctx_relax.argtypes = [vp, ci, ci, cd, cd, vp, vp]
        L.gapx_ctx_relax_lbfgs.argtypes = [vp, ci, ci, cd, cd, ci, cd, vp, vp]
        L.gapx_ctx_positions.argtypes = [vp, vp]
        L.gapx_model_set_qmm.argtypes = [vp, vp, ci, ci]
        L.gapx_ctx_variance.argtypes = [vp, vp, vp]
        L.gapx_probe_fma_rates.argtypes = [vp]
        L.gapx_model_write.argtypes = [vp, ctypes.c_char_p]
        L.gapx_model_save.argtypes = [vp, ctypes.c_char_p]
        L.gapx_model_load.argtypes = [ctypes.c_char_p, vp]
        L.gapx_model_read_cached.argtypes = [ctypes.c_char_p, ctypes.c_char_p, vp]
        L.gapx_ctx_md.argtypes = [vp, ci, ci, cd, cd, cd, ci, vp, vp, vp]
        L.gapx_ctx_momenta.argtypes = [vp, vp]
        L.gapx_ctx_get_tuning.argtypes = [vp, ci, vp, vp, vp, vp, vp]
        L.gapx_ctx_set_tuning.argtypes = [vp, ci, ci, ci, ci, ci, ci]
        L.gapx_ctx_stage_count.argtypes = [vp, vp]
        L.gapx_ctx_stage_info.argtypes = [vp, ci, vp, vp, vp]
        L.gapx_ctx_use_graph.argtypes = [vp, ci]
        L.gapx_ctx_device_bytes.argtypes = [vp, vp]
        L.gapx_ctx_counts.argtypes = [vp, vp, vp]
        L.gapx_min_distance.argtypes = [ci, vp, vp, cd, ci, vp]
        L.gapx_min_distance_noalloc.argtypes = [ci, vp, vp, cd, ci, vp]
        L.gapx_car2acsf.argtypes = [vp, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci]
        L.gapx_ctx_free.argtypes = [vp]
        L.gapx_ctx_free.restype = None
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise GapxError(rc, lib().gapx_last_error().decode(errors="replace"))


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def device_count():
    n = ctypes.c_int(0)
    _check(lib().gapx_device_count(ctypes.byref(n)))
    return n.value


def probe_fma_rates():
    """Measured TFLOP/s of FP32 vector FMA (no packed instructions) and FP64 vector FMA."""
    out = np.zeros(2)
    _check(lib().gapx_probe_fma_rates(_p(out)))
    return dict(fma_f32=float(out[0]), fma_f64=float(out[1]))


def min_distance(lat, pos, rcut=6.0):
    """FGET_BOND: min(10, shortest interatomic distance within rcut over images)."""
    lat = np.ascontiguousarray(lat, np.float64)
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.zeros(1)
    _check(lib().gapx_min_distance(len(pos), _p(lat), _p(pos), float(rcut), 0, _p(out)))
    return float(out[0])


class Model:
    """Device-resident GAP model (species weights, SF table, theta, MM, coeff)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        ns, nsf, D, M = (ctypes.c_int() for _ in range(4))
        _check(lib().gapx_model_dims(self._h, ctypes.byref(ns), ctypes.byref(nsf), ctypes.byref(D),
                                     ctypes.byref(M)))
        self.nspecies, self.nsf, self.des_len, self.nsparse = ns.value, nsf.value, D.value, M.value

    @classmethod
    def read(cls, path, cache=False):
        """Parse a gap_parameters file; cache=True (or a path) goes through a
        binary copy keyed on the file's size and mtime (gapx_model_read_cached)."""
        h = ctypes.c_void_p()
        if cache:
            cp = None if cache is True else os.fsencode(cache)
            _check(lib().gapx_model_read_cached(os.fsencode(path), cp, ctypes.byref(h)))
        else:
            _check(lib().gapx_model_read(os.fsencode(path), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path):
        """A binary model file written by save()."""
        h = ctypes.c_void_p()
        _check(lib().gapx_model_load(os.fsencode(path), ctypes.byref(h)))
        return cls(h.value)

    def save(self, path):
        _check(lib().gapx_model_save(self._h, os.fsencode(path)))

    def write(self, path):
        """The model as a gap_parameters text file."""
        _check(lib().gapx_model_write(self._h, os.fsencode(path)))

    @classmethod
    def from_arrays(cls, z, w, sf_type, sf_alpha, sf_cut, theta, mm, coeff, sf_zeta=None):
        z = np.ascontiguousarray(z, np.int32)
        w = np.ascontiguousarray(w, np.float64)
        t = np.ascontiguousarray(sf_type, np.int32)
        a = np.ascontiguousarray(sf_alpha, np.float64)
        c = np.ascontiguousarray(sf_cut, np.float64)
        zt = None if sf_zeta is None else np.ascontiguousarray(sf_zeta, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        mm = np.ascontiguousarray(mm, np.float64)
        cf = np.ascontiguousarray(coeff, np.float64)
        h = ctypes.c_void_p()
        _check(lib().gapx_model_create(len(z), _p(z), _p(w), len(t), _p(t), _p(a), _p(c), _p(zt), mm.shape[1],
                                       mm.shape[0], _p(th), _p(mm), 0, 0, _p(cf), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_model(cls, m):
        """From any object with the oracle-style fields (z, w, sf_type, ...)."""
        zeta = getattr(m, "sf_zeta", None)
        if isinstance(m, dict):
            return cls.from_arrays(m["z"], m["w"], m["sf_type"], m["sf_alpha"], m["sf_cut"], m["theta"], m["mm"],
                                   m["coeff"], m.get("sf_zeta"))
        return cls.from_arrays(m.z, m.w, m.sf_type, m.sf_alpha, m.sf_cut, m.theta, m.mm, m.coeff, zeta)

    def arrays(self):
        th = np.zeros(self.des_len)
        mm = np.zeros((self.nsparse, self.des_len))
        cf = np.zeros(self.nsparse)
        _check(lib().gapx_model_arrays(self._h, _p(th), _p(mm), 0, 0, _p(cf)))
        return th, mm, cf

    def header(self):
        z = np.zeros(self.nspecies, np.int32)
        w = np.zeros(self.nspecies)
        t = np.zeros(self.nsf, np.int32)
        a, c, zt = np.zeros(self.nsf), np.zeros(self.nsf), np.zeros(self.nsf)
        _check(lib().gapx_model_header(self._h, _p(z), _p(w), _p(t), _p(a), _p(c), _p(zt)))
        return dict(z=z, w=w, sf_type=t, sf_alpha=a, sf_cut=c, sf_zeta=zt)

    def calc(self, species, lat, pos, rcut=6.0, grad=True, mixed=False, flags=0):
        """FGAP_CALC: (energy eV, forces (N,3) eV/A, stress (6,) GPa, variance)."""
        species = np.ascontiguousarray(species, np.int32)
        lat = np.ascontiguousarray(lat, np.float64)
        pos = np.ascontiguousarray(pos, np.float64)
        na = len(species)
        ene = np.zeros(1)
        force = np.zeros((na, 3))
        stress = np.zeros(6)
        var = np.zeros(1)
        fl = flags | (GRAD if grad else 0) | (MIXED if mixed else 0)
        _check(lib().gapx_calc(self._h, na, _p(species), _p(lat), _p(pos), float(rcut), fl, _p(ene), _p(force),
                               _p(stress), _p(var)))
        return float(ene[0]), force, stress, float(var[0])

    def calc_batch(self, structures, rcut=6.0, grad=True, mixed=False, flags=0):
        natoms = np.array([len(s["species"]) for s in structures], np.int32)
        species = np.ascontiguousarray(np.concatenate([s["species"] for s in structures]), np.int32)
        lat = np.ascontiguousarray(np.stack([s["lat"] for s in structures]), np.float64)
        pos = np.ascontiguousarray(np.concatenate([s["pos"] for s in structures]), np.float64)
        ene = np.zeros(len(structures))
        force = np.zeros((int(natoms.sum()), 3))
        stress = np.zeros((len(structures), 6))
        fl = flags | (GRAD if grad else 0) | (MIXED if mixed else 0)
        _check(lib().gapx_calc_batch(self._h, len(structures), _p(natoms), _p(species), _p(lat), _p(pos),
                                     float(rcut), fl, _p(ene), _p(force), _p(stress)))
        offs = np.concatenate([[0], np.cumsum(natoms)])
        return ene, [force[offs[k]:offs[k + 1]] for k in range(len(structures))], stress

    def precision(self):
        """How GAPX_MIXED runs treat the model: the cancellation probe A and
        whether they contract in FP64; the feature sensitivity B and whether
        they run wholly in FP64."""
        a, b = ctypes.c_double(), ctypes.c_double()
        f, g = ctypes.c_int(), ctypes.c_int()
        _check(lib().gapx_model_precision(self._h, ctypes.byref(a), ctypes.byref(f), ctypes.byref(b),
                                          ctypes.byref(g)))
        return dict(cancellation=a.value, mixed_contraction_fp64=bool(f.value), feature_sensitivity=b.value,
                    mixed_runs_fp64=bool(g.value))

    def set_qmm(self, qmm):
        """Q of the predictive variance (nsparse x nsparse), or None for Q = 0."""
        if qmm is None:
            _check(lib().gapx_model_set_qmm(self._h, None, 0, 0))
            return
        q = np.ascontiguousarray(qmm, np.float64)
        if q.shape != (self.nsparse, self.nsparse):
            raise GapxError(1, "qmm must be %d x %d" % (self.nsparse, self.nsparse))
        _check(lib().gapx_model_set_qmm(self._h, _p(q), 0, self.nsparse))

    def descriptors(self, species, lat, pos, rcut=6.0):
        species = np.ascontiguousarray(species, np.int32)
        x = np.zeros((len(species), self.des_len))
        _check(lib().gapx_descriptors(self._h, len(species), _p(species), _p(np.ascontiguousarray(lat, np.float64)),
                                      _p(np.ascontiguousarray(pos, np.float64)), float(rcut), 0, _p(x)))
        return x

    def car2acsf(self, pos, neighbor, neighbor_count, lgrad=True):
        """CAR2ACSF: (xx(nf, na), dxdy(nf, na, na, 3), strs(3, 3, nf, na)),
        Fortran-ordered, from a neighbour table NEIGHBOR(na, max_neighbor, 6)
        (image x y z, distance, weight, 1-based index) and its counts."""
        pos = np.asfortranarray(pos, np.float64)
        nb = np.asfortranarray(neighbor, np.float64)
        cnt = np.ascontiguousarray(neighbor_count, np.int32)
        na, max_nb = nb.shape[0], nb.shape[1]
        D = self.des_len
        xx = np.zeros((D, na), order="F")
        dxdy = np.zeros((D, na, na, 3), order="F")
        strs = np.zeros((3, 3, D, na), order="F")
        _check(lib().gapx_car2acsf(self._h, na, max_nb, D, _p(pos), _p(nb), _p(cnt), _p(xx), _p(dxdy), _p(strs),
                                   int(bool(lgrad))))
        return xx, dxdy, strs

    def close(self):
        if self._h:
            lib().gapx_model_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """Device-resident batch of structures and its workspace."""

    def __init__(self, model):
        self.model = model
        h = ctypes.c_void_p()
        _check(lib().gapx_ctx_create(model._h, ctypes.byref(h)))
        self._h = h
        self.nstruct = 0
        self.natoms = None

    def load(self, structures, rcut=6.0):
        self.natoms = np.array([len(s["species"]) for s in structures], np.int32)
        species = np.ascontiguousarray(np.concatenate([s["species"] for s in structures]), np.int32)
        lat = np.ascontiguousarray(np.stack([s["lat"] for s in structures]), np.float64)
        pos = np.ascontiguousarray(np.concatenate([s["pos"] for s in structures]), np.float64)
        self.nstruct = len(structures)
        _check(lib().gapx_ctx_load(self._h, self.nstruct, _p(self.natoms), _p(species), _p(lat), _p(pos),
                                   float(rcut)))

    def update(self, pos, lat=None):
        pos = np.ascontiguousarray(pos, np.float64)
        lat = None if lat is None else np.ascontiguousarray(lat, np.float64)
        _check(lib().gapx_ctx_update(self._h, _p(lat), _p(pos)))

    def run(self, flags=GRAD):
        _check(lib().gapx_ctx_run(self._h, int(flags)))

    def sync(self):
        _check(lib().gapx_ctx_sync(self._h))

    def results(self):
        n = int(self.natoms.sum())
        ene = np.zeros(self.nstruct)
        force = np.zeros((n, 3))
        stress = np.zeros((self.nstruct, 6))
        eatom = np.zeros(n)
        _check(lib().gapx_ctx_results(self._h, _p(ene), _p(force), _p(stress), _p(eatom)))
        return dict(energy=ene, forces=force, stress=stress, eatom=eatom)

    def features(self, dedx=True):
        n = int(self.natoms.sum())
        x = np.zeros((n, self.model.des_len))
        g = np.zeros((n, self.model.des_len)) if dedx else None
        _check(lib().gapx_ctx_features(self._h, _p(x), _p(g)))
        return x, g

    def md(self, masses, steps, dt_fs=1.0, temperature=300.0, taut_fs=500.0, fixcm=True, momenta=None,
           flags=GRAD, log=True):
        """Batched NVT MD (ASE NVTBerendsen steps) of the loaded batch on the device.
        masses per atom in amu; returns the log array [steps+1, nstruct, (E_pot, E_kin, T)] or None."""
        n = int(self.natoms.sum())
        m = np.ascontiguousarray(masses, np.float64)
        if m.shape != (n,):
            raise GapxError(1, "masses must have one entry per atom")
        p = None if momenta is None else np.ascontiguousarray(momenta, np.float64).reshape(n, 3)
        out = np.zeros((steps + 1, self.nstruct, 3)) if log else None
        taut = taut_fs * FS if taut_fs and taut_fs > 0 else 0.0
        _check(lib().gapx_ctx_md(self._h, int(flags), int(steps), float(dt_fs * FS), float(temperature), float(taut),
                                 int(bool(fixcm)), _p(m), _p(p), _p(out)))
        return out

    def momenta(self):
        p = np.zeros((int(self.natoms.sum()), 3))
        _check(lib().gapx_ctx_momenta(self._h, _p(p)))
        return p

    def variance(self):
        """Predictive variance of the last run: (per-structure mean, per-atom sigma^2)."""
        var = np.zeros(self.nstruct)
        s2 = np.zeros(int(self.natoms.sum()))
        _check(lib().gapx_ctx_variance(self._h, _p(var), _p(s2)))
        return var, s2

    def timing(self, enable=True):
        _check(lib().gapx_ctx_timing(self._h, int(enable)))

    def stages(self):
        n = ctypes.c_int()
        _check(lib().gapx_ctx_stage_count(self._h, ctypes.byref(n)))
        out = {}
        for k in range(n.value):
            name = ctypes.c_char_p()
            ms = ctypes.c_double()
            cnt = ctypes.c_int()
            _check(lib().gapx_ctx_stage_info(self._h, k, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def use_graph(self, on=True):
        _check(lib().gapx_ctx_use_graph(self._h, int(on)))

    def autotune(self, flags=GRAD, force=False):
        """Time launch shapes on the loaded batch and keep the fastest."""
        _check(lib().gapx_ctx_autotune(self._h, int(flags), int(force)))
        return self.tuning(flags)

    def tuning(self, flags=GRAD):
        v = [ctypes.c_int() for _ in range(5)]
        _check(lib().gapx_ctx_get_tuning(self._h, int(flags), *[ctypes.byref(x) for x in v]))
        return dict(zip(("waves_fwd", "waves_bwd", "gpr_rows", "gpr_blocks", "halves"), (x.value for x in v)))

    def relax(self, flags=GRAD, max_steps=500, fmax=0.05, maxstep=0.2, optimizer="fire", memory=100, alpha=70.0):
        """Relaxation of every loaded structure at fixed cell, on the device:
        optimizer "fire" (ASE's FIRE parameters) or "lbfgs" (ASE's LBFGS,
        the reference's optimiser; memory and alpha as there)."""
        steps = ctypes.c_int()
        conv = ctypes.c_int()
        if optimizer == "lbfgs":
            _check(lib().gapx_ctx_relax_lbfgs(self._h, int(flags), int(max_steps), float(fmax), float(maxstep),
                                              int(memory), float(alpha), ctypes.byref(steps), ctypes.byref(conv)))
        elif optimizer == "fire":
            _check(lib().gapx_ctx_relax(self._h, int(flags), int(max_steps), float(fmax), float(maxstep),
                                        ctypes.byref(steps), ctypes.byref(conv)))
        else:
            raise ValueError("optimizer must be 'fire' or 'lbfgs'")
        return dict(steps=steps.value, converged=conv.value)

    def positions(self):
        pos = np.zeros((int(self.natoms.sum()), 3))
        _check(lib().gapx_ctx_positions(self._h, _p(pos)))
        return pos

    def export_labels(self, mlp):
        """Append the last GAP run's (xt, per-atom E, dE/dxt) to mlp's training set."""
        _check(lib().gapx_ctx_export_labels(self._h, mlp._h))

    def run_surrogate(self, mlp, flags=GRAD):
        """Evaluate the loaded batch with the MLP in place of the GAP contraction."""
        _check(lib().gapx_ctx_run_surrogate(self._h, mlp._h, int(flags)))

    def set_tuning(self, flags=GRAD, waves_fwd=1, waves_bwd=1, gpr_rows=64, gpr_blocks=512, halves=1):
        _check(lib().gapx_ctx_set_tuning(self._h, int(flags), int(waves_fwd), int(waves_bwd), int(gpr_rows),
                                         int(gpr_blocks), int(halves)))

    def counts(self):
        """(pairs, triplets) of the last run."""
        p = ctypes.c_longlong()
        t = ctypes.c_longlong()
        _check(lib().gapx_ctx_counts(self._h, ctypes.byref(p), ctypes.byref(t)))
        return p.value, t.value

    def device_bytes(self):
        b = ctypes.c_longlong()
        _check(lib().gapx_ctx_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    def close(self):
        if self._h:
            lib().gapx_ctx_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MLP_SPLIT = 1
MLP_EXACT = 32   # run flag: hi + lo products in the surrogate path


class MLP:
    """Per-atom surrogate 128 -> 128 -> 128 -> 1 (SiLU) on the GPU."""

    def __init__(self, seed=0):
        h = ctypes.c_void_p()
        _check(lib().gapx_mlp_create(ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int()
        _check(lib().gapx_mlp_param_count(ctypes.byref(n)))
        self.n_params = n.value

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib().gapx_mlp_get_params(self._h, _p(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(lib().gapx_mlp_set_params(self._h, _p(p)))

    def forward(self, xt, grad=True, split=False):
        xt = np.ascontiguousarray(xt, np.float32)
        n = len(xt)
        e = np.zeros(n, np.float32)
        G = np.zeros((n, 128), np.float32) if grad else None
        _check(lib().gapx_mlp_forward(self._h, n, _p(xt), MLP_SPLIT if split else 0, _p(e), _p(G)))
        return (e, G) if grad else e

    def loss_grad(self, xt, et, Gt, cE=1.0, cG=1.0, split=False):
        xt = np.ascontiguousarray(xt, np.float32)
        et = np.ascontiguousarray(et, np.float32)
        Gt = np.ascontiguousarray(Gt, np.float32)
        sums = np.zeros(2)
        g = np.zeros(self.n_params, np.float32)
        _check(lib().gapx_mlp_loss_grad(self._h, len(xt), _p(xt), _p(et), _p(Gt), float(cE), float(cG),
                                        MLP_SPLIT if split else 0, _p(sums), _p(g)))
        n = len(xt)
        return 0.5 * cE / n * sums[0] + 0.5 * cG / n * sums[1], g

    def add_data(self, xt, et, Gt):
        xt = np.ascontiguousarray(xt, np.float32)
        _check(lib().gapx_mlp_data_add(self._h, len(xt), _p(xt), _p(np.ascontiguousarray(et, np.float32)),
                                       _p(np.ascontiguousarray(Gt, np.float32))))

    def clear_data(self):
        _check(lib().gapx_mlp_data_clear(self._h))

    def data_count(self):
        n = ctypes.c_longlong()
        _check(lib().gapx_mlp_data_count(self._h, ctypes.byref(n)))
        return n.value

    def train(self, steps, batch=4096, lr=1e-3, cE=1.0, cG=1.0, seed=0, split=False):
        """Adam steps on the device-resident training set; returns the last
        batch's (mean squared energy error, mean squared |G| error)."""
        sums = np.zeros(2)
        _check(lib().gapx_mlp_train(self._h, int(steps), int(batch), float(lr), float(cE), float(cG),
                                    ctypes.c_ulonglong(seed), MLP_SPLIT if split else 0, _p(sums)))
        return sums[0] / batch, sums[1] / batch

    def close(self):
        if self._h:
            lib().gapx_mlp_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
