"""Write golden E/F/S vectors by running the reference FGAP_CALC itself.

The reference Fortran (gappy/libgap/gap_calc.f90 + wacsf.f90) is compiled from
its own sources by ``make -C oracle ref`` into ``oracle/_ref/libgapref.so``;
this script calls its ``fgap_calc_`` through ctypes with the working directory
set to a temporary folder holding the model as ``gap_parameters`` (the
reference reads species weights and the symmetry-function table from there,
gap_calc.f90:75 and wacsf.f90:45).  It runs only in the build container, where
the reference is mounted.

Cases cover what the shipped ase.traj does not: negative and multi-species
weights (w_j w_k in angular terms), triclinic cells, cells smaller than the
cutoff (more than one image per axis, self-images), lgrad = .false., SF
type 3, per-line cutoffs below Rcut, the 64-atom / D = 96 / M = 500
single-species workload itself, and two dense cells with about 920 neighbours
per atom (the reference stops above 1000).

FGET_BOND (get_bond.f90:4-117) is run on the same structures at two cutoffs.

CAR2ACSF (wacsf.f90:2-12) is run on FGAP_CALC's neighbour tables of three of
the cases, and write_array_2dim (wacsf.f90:798-810) on a small matrix.

Output: tests/golden/ref_vectors.npz (inputs, model arrays, reference outputs)
and tests/golden/write_array_2dim.txt (the reference's text).
Usage: python tests/golden/make_ref_vectors.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "calypso-gap_amd"))

import oracle  # noqa: E402
from gapx import synth  # noqa: E402

REF_XYZ = "/root/reference/gappy/example/BC/sps_all.xyz"


def read_xyz_b8c56(path):
    """extended-XYZ reader for the one BC structure (species, Lattice, pos)."""
    lines = open(path).read().splitlines()
    n = int(lines[0])
    head = lines[1]
    lat = np.array([float(v) for v in head.split('Lattice="')[1].split('"')[0].split()]).reshape(3, 3)
    sym = {"B": 5, "C": 6}
    species, pos = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        species.append(sym[t[0]])
        pos.append([float(v) for v in t[1:4]])
    return np.array(species, np.int32), lat, np.array(pos)


def model_dict(m):
    return dict(z=m.z, w=m.w, sf_type=m.sf_type, sf_alpha=m.sf_alpha, sf_cut=m.sf_cut,
                sf_zeta=np.ones(m.nsf) if m.sf_zeta is None else m.sf_zeta, theta=m.theta, mm=m.mm,
                coeff=m.coeff)


def as_model(d):
    return oracle.Model(z=d["z"], w=d["w"], sf_type=d["sf_type"], sf_alpha=d["sf_alpha"],
                        sf_cut=d["sf_cut"], theta=d["theta"], mm=d["mm"], coeff=d["coeff"],
                        sf_zeta=None if np.all(d["sf_zeta"] == 1.0) else d["sf_zeta"])


def ternary_model(rng):
    """Ternary (Z 1, 8, 22; weights = Z) model with types 1-4, two cutoffs."""
    types = [1] * 6 + [3] * 4 + [2, 4] * 3 + [2, 4] * 2
    alphas = [0.01, 0.05, 0.2, 0.6, 1.5, 4.0] + [1.0, 2.0, 3.0, 4.0] + [0.005, 0.005, 0.05, 0.05, 0.5, 0.5] + \
        [0.02, 0.02, 0.2, 0.2]
    cuts = [6.0] * 6 + [5.0, 6.0, 5.0, 6.0] + [6.0] * 6 + [4.5] * 4
    header = dict(z=np.array([1, 8, 22], np.int32), w=np.array([1.0, 8.0, 22.0]),
                  sf=(np.array(types, np.int32), np.array(alphas), np.array(cuts), np.ones(len(types))))
    xs = []
    for p in range(4):
        s = synth.structure(24, 24 * 12.0, (1, 8, 22), np.random.default_rng(500 + p))
        xs.append(_desc(header, s["species"], s["lat"], s["pos"], 6.0))
    d = synth.make_model(header["z"], header["w"], header["sf"], np.concatenate(xs), 60, 0.05, rng)
    # widen the length scales so that most kernel values are far from 0 and
    # every feature column carries force
    d["theta"] = d["theta"] * 4.0
    return as_model(d)


def _desc(header, species, lat, pos, rc):
    t, a, c, z = header["sf"]
    m = oracle.Model(z=header["z"], w=header["w"], sf_type=t, sf_alpha=a, sf_cut=c, theta=np.ones(2 * len(t)),
                     mm=np.zeros((1, 2 * len(t))), coeff=np.zeros(1), sf_zeta=None if np.all(z == 1) else z)
    return oracle.descriptors(m, species, lat, pos, rc)


def c64_single_species_model():
    w = synth.WORKLOADS["c64_zeta1"]
    d = synth.workload_model("c64_zeta1", _desc)
    d["sf_zeta"] = np.ones(len(d["sf_type"]))
    assert d["mm"].shape == (w["M"], 96)
    return as_model(d)


def dense_model(base, species_pool, seed):
    """A base model's header (species weights and SF table) with theta, MM and
    coeff fitted on the descriptors of dense cells (1 A^3 per atom, about 920
    neighbours within 6 A), so that such cells have O(1) kernel values."""
    header = dict(z=base.z, w=base.w, sf=(base.sf_type, base.sf_alpha, base.sf_cut,
                                           np.ones(base.nsf) if base.sf_zeta is None else base.sf_zeta))
    xs = []
    for p in range(4):
        s = synth.structure(8, 8 * 1.0, species_pool, np.random.default_rng(seed + p), cell="cubic", min_dist=0.9)
        xs.append(_desc(header, s["species"], s["lat"], s["pos"], 6.0))
    d = synth.make_model(base.z, base.w, header["sf"], np.concatenate(xs), 24, 0.1, np.random.default_rng(seed))
    d["theta"] = d["theta"] * 2.0
    return as_model(d)


def run_ref(model, species, lat, pos, rcut, lgrad):
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "gap_parameters")
        oracle.write_model(path, model)
        return oracle.ref_gap_calc(path, species, lat, pos, model.theta, model.mm, model.coeff, rcut, lgrad)


def main():
    if not oracle.ref_available():
        raise SystemExit("build the reference first: make -C oracle ref")
    rng = np.random.default_rng(20240)
    models = {
        "example": oracle.parse_model(os.path.join(HERE, "c64_diamond", "gap_parameters")),
        "ternary": ternary_model(rng),
        "c64_single_species": c64_single_species_model(),
    }
    models["example_dense"] = dense_model(models["example"], (5, 6), 900)
    models["ternary_dense"] = dense_model(models["ternary"], (1, 8, 22), 910)
    cases = []
    sp, lat, pos = read_xyz_b8c56(REF_XYZ)
    cases.append(("example", "b8c56", sp, lat, pos, 6.0, True))
    r = np.random.default_rng(11)
    s = synth.structure(24, 24 * 9.0, (5, 6), r, min_dist=1.3)
    cases.append(("example", "bc24_triclinic", s["species"], s["lat"], s["pos"], 6.0, True))
    s = synth.structure(6, 6 * 10.0, (5, 6), np.random.default_rng(12), min_dist=1.3)
    cases.append(("example", "bc6_small_cell", s["species"], s["lat"], s["pos"], 6.0, True))
    cases.append(("example", "bc6_small_cell_nograd", s["species"], s["lat"], s["pos"], 6.0, False))
    s = synth.structure(16, 16 * 12.0, (1, 8, 22), np.random.default_rng(13))
    cases.append(("ternary", "t16", s["species"], s["lat"], s["pos"], 6.0, True))
    s = synth.structure(40, 40 * 11.0, (1, 8, 22), np.random.default_rng(14))
    cases.append(("ternary", "t40", s["species"], s["lat"], s["pos"], 6.0, True))
    cases.append(("ternary", "t40_nograd", s["species"], s["lat"], s["pos"], 6.0, False))
    s = synth.structure(8, 8 * 14.0, (1, 8, 22), np.random.default_rng(15))
    cases.append(("ternary", "t8_small_cell", s["species"], s["lat"], s["pos"], 6.0, True))
    s = synth.workload_structure("c64_zeta1")
    cases.append(("c64_single_species", "c64_single_species", s["species"], s["lat"], s["pos"], 6.0, True))
    # about 920 neighbours per atom within 6 A (8 atoms in a 2 A cube): near
    # the reference's limit of 1000 (gap_calc.f90:68, 107-111)
    s = synth.structure(8, 8 * 1.0, (5, 6), np.random.default_rng(16), cell="cubic", min_dist=0.9)
    cases.append(("example_dense", "bc8_dense_920", s["species"], s["lat"], s["pos"], 6.0, True))
    s = synth.structure(8, 8 * 1.0, (1, 8, 22), np.random.default_rng(17), cell="cubic", min_dist=0.9)
    cases.append(("ternary_dense", "t8_dense_920", s["species"], s["lat"], s["pos"], 6.0, True))

    out = {}
    for name, m in models.items():
        for k, v in model_dict(m).items():
            out["model_%s_%s" % (name, k)] = np.asarray(v)
    out["case_names"] = np.array([c[1] for c in cases])
    out["case_models"] = np.array([c[0] for c in cases])
    for mname, cname, sp, lat, pos, rcut, lgrad in cases:
        res = run_ref(models[mname], sp, lat, pos, rcut, lgrad)
        pre = "case_%s_" % cname
        out[pre + "species"] = np.asarray(sp, np.int32)
        out[pre + "lat"] = np.asarray(lat)
        out[pre + "pos"] = np.asarray(pos)
        out[pre + "rcut"] = np.array(rcut)
        out[pre + "lgrad"] = np.array(lgrad)
        out[pre + "energy"] = np.array(res["energy"])
        out[pre + "forces"] = res["forces"]
        out[pre + "stress"] = res["stress"]
        # FGET_BOND at the evaluation cutoff and at a cutoff below every distance
        out[pre + "min_bond"] = np.array([oracle.ref_get_bond(lat, sp, pos, rcut),
                                          oracle.ref_get_bond(lat, sp, pos, 0.9)])
        print("%-24s N=%3d E=%.10f |F|max=%.6f S0=%.6f" % (cname, len(sp), res["energy"],
                                                          np.abs(res["forces"]).max(), res["stress"][0]))
    # CAR2ACSF (wacsf.f90:2-12) on FGAP_CALC's own neighbour tables of three
    # of the cases: xx, dense dxdy and strs as the reference returns them
    car = ["bc6_small_cell", "t8_small_cell", "bc8_dense_920"]
    out["car_names"] = np.array(car)
    for mname, cname, sp, lat, pos, rcut, lgrad in cases:
        if cname not in car:
            continue
        m = models[mname]
        nb, cnt = oracle.neighbor_table(m, sp, lat, pos, rcut)
        nb = np.asfortranarray(nb[:, :max(int(cnt.max()), 1), :])  # max_neighbor = the largest count
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "gap_parameters")
            oracle.write_model(path, m)
            xx, dxdy, strs = oracle.ref_car2acsf(path, pos, nb, cnt, m.des_len, True)
        pre = "car_%s_" % cname
        out[pre + "model"] = np.array(mname)
        out[pre + "pos"] = np.asarray(pos)
        out[pre + "neighbor"] = nb
        out[pre + "count"] = cnt
        out[pre + "xx"] = xx
        out[pre + "dxdy"] = dxdy
        out[pre + "strs"] = strs
        print("car2acsf %-16s N=%d max_nb=%d |xx|max=%.4g |dxdy|max=%.4g" % (cname, len(sp), nb.shape[1],
                                                                           np.abs(xx).max(), np.abs(dxdy).max()))
    # write_array_2dim (wacsf.f90:798-810): the text the reference writes for a
    # small matrix with mixed signs and magnitudes
    wa = np.array([[1.5, -2.25, 1e-11, 123456.789012345], [-0.0, 3.0e-4, -9876.54321, 2.0 / 3.0],
                   [np.pi, -np.e, 1e5, -1e-10]])
    out["write_array_input"] = wa
    with tempfile.TemporaryDirectory() as tmp:
        oracle.ref_write_array_2dim(wa, os.path.join(tmp, "wa.txt"))
        text = open(os.path.join(tmp, "wa.txt"), "rb").read()
    with open(os.path.join(HERE, "write_array_2dim.txt"), "wb") as f:
        f.write(text)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
