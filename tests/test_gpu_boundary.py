"""The drop-in boundary on the GPU: libgap.libgap.fgap_calc (reference
signature), libgap.GAP.Calculator, the ASE adapter, FGET_BOND and CAR2ACSF
against the reference's own outputs; the batched and status entry points."""
import os
import shutil
import sys

import numpy as np
import pytest

import gapx
import oracle
from conftest import CAR_CASE_NAMES, GOLDEN, GOLDEN_CASE_NAMES, ROOT, car_case, golden_cases, golden_model

pytestmark = pytest.mark.gpu
PKG = os.path.join(ROOT, "calypso-gap_amd")
MODEL = os.path.join(GOLDEN, "c64_diamond", "gap_parameters")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")


@pytest.fixture
def in_model_dir(tmp_path, monkeypatch):
    shutil.copy(MODEL, tmp_path / "gap_parameters")
    monkeypatch.chdir(tmp_path)
    return tmp_path


def test_fgap_calc_reference_signature_on_traj_frames(in_model_dir, frames):
    import libgap.libgap as L
    n, d, theta, mm, invcmm, coeff = L.fgap_read()
    for k in (0, 5, 10):
        ene, force, stress, var = L.fgap_calc(frames["numbers"], frames["cells"][k], frames["positions"][k],
                                              theta[:d], mm[:n, :d], invcmm[:n, :n], coeff[:n], 6.0, True)
        assert abs(ene - frames["energies"][k]) / 64 < 1e-9
        assert np.abs(force - frames["forces"][k]).max() < 1e-7
        assert np.abs(stress - frames["stresses"][k]).max() < 1e-6
        assert var == 0.0


def test_gap_calculator_api(in_model_dir, frames):
    from libgap.GAP import Calculator
    gap = Calculator(rcut=6.0)
    symbols = ["C"] * 64
    ene, force, stress, _ = gap.gap_calc(symbols, frames["cells"][3], frames["positions"][3], True)
    assert abs(ene - frames["energies"][3]) / 64 < 1e-9
    assert np.abs(force - frames["forces"][3]).max() < 1e-7
    ene0, force0, stress0, _ = gap.gap_calc(symbols, frames["cells"][3], frames["positions"][3], False)
    assert ene0 == pytest.approx(ene, abs=1e-10) and not force0.any() and not stress0.any()
    mixed = Calculator(rcut=6.0, precision="mixed")
    em, fm, sm, _ = mixed.gap_calc(symbols, frames["cells"][3], frames["positions"][3], True)
    assert abs(em - ene) / 64 < 5e-3
    with pytest.raises(RuntimeError):
        gap.gap_calc([79] * 64, frames["cells"][3], frames["positions"][3], True)


def test_ase_adapter_on_traj_frames(in_model_dir, frames):
    sys.path.insert(0, os.path.join(ROOT, "tests", "stub_ase"))
    try:
        import importlib.util
        from ase import Atoms
        spec = importlib.util.spec_from_file_location("gap_calc", os.path.join(PKG, "ASE", "gap_calc.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        atoms = Atoms(frames["numbers"], frames["positions"][7], frames["cells"][7])
        atoms.calc = mod.GAP(rcut=6.0)
        assert abs(atoms.get_potential_energy() - frames["energies"][7]) / 64 < 1e-9
        assert np.abs(atoms.get_forces() - frames["forces"][7]).max() < 1e-7
        assert np.abs(atoms.get_stress() - frames["stresses"][7]).max() < 1e-6
        assert atoms.calc.results["free_energy"] == atoms.calc.results["energy"]
        assert atoms.calc.results["variance"] == 0.0
    finally:
        sys.path.remove(os.path.join(ROOT, "tests", "stub_ase"))


@pytest.mark.parametrize("case", GOLDEN_CASE_NAMES)
def test_fget_bond_matches_reference(ref_vectors, case):
    import libgap.libgap as L
    c = [x for x in golden_cases(ref_vectors) if x["name"] == case][0]
    want = ref_vectors["case_%s_min_bond" % case]
    got = [L.fget_bond(c["lat"], c["species"], c["pos"], c["rcut"]), L.fget_bond(c["lat"], c["species"], c["pos"], 0.9)]
    assert got == list(want)  # bit for bit


def bond_cases():
    from gapx import synth
    out = []
    s = synth.structure(1500, 1500 * 11.0, (1, 8, 22), np.random.default_rng(601))
    out.append(("1500 atoms, three-launch binning", s["lat"], s["pos"], 6.0))
    shift = np.random.default_rng(602).integers(-2, 3, size=(1500, 3)) @ s["lat"]
    out.append(("1500 atoms unwrapped by up to two cells", s["lat"], s["pos"] + shift, 6.0))
    s = synth.structure(3, 3 * 5.0, (1, 8), np.random.default_rng(603), min_dist=1.0)
    out.append(("3 atoms, cell far smaller than rcut", s["lat"], s["pos"], 6.0))
    lat = np.array([[9.0, 0.0, 0.0], [7.5, 4.0, 0.0], [-3.0, 2.5, 6.0]])  # strongly sheared
    pos = np.random.default_rng(604).random((40, 3)) @ lat
    out.append(("40 atoms, sheared cell", lat, pos, 6.0))
    out.append(("40 atoms, rcut below every distance", lat, pos, 0.05))
    return out


@pytest.mark.parametrize("k", range(5))
def test_min_distance_cell_list_equals_restatement(k):
    """FGET_BOND on the cell list against the restated reference (oracle,
    itself equal to the compiled reference on the golden cases), bit for bit:
    many atoms (the three-launch binning), positions outside the cell (the
    reference's image range counts from the positions as given), a cell far
    below rcut and a strongly sheared one (the image counts from the columns
    of LAT, as the reference takes them)."""
    name, lat, pos, rc = bond_cases()[k]
    got = gapx.min_distance(lat, pos, rc)
    want = oracle.min_bond(lat, pos, rc)
    print(name, got, want)
    assert got == want


def test_fgap_calc_batch_and_status(in_model_dir, ref_vectors):
    import libgap.libgap as L
    cases = [x for x in golden_cases(ref_vectors) if x["model"] == "example" and x["lgrad"]]
    n, d, theta, mm, invcmm, coeff = L.fgap_read()
    natoms = np.array([len(c["species"]) for c in cases], np.int32)
    species = np.concatenate([c["species"] for c in cases]).astype(np.int32)
    lat = np.stack([c["lat"] for c in cases], axis=2)        # LAT(3,3,nstruct)
    pos = np.concatenate([c["pos"] for c in cases]).T.copy()  # pos(3,ntot)
    ene, force, stress, status = L.fgap_calc_batch(natoms, species, lat, pos, theta[:d], mm[:n, :d], coeff[:n],
                                                   6.0, True, 0)
    assert status == 0
    off = np.concatenate([[0], np.cumsum(natoms)])
    for k, c in enumerate(cases):
        assert abs(ene[k] - c["energy"]) / natoms[k] < 1e-9
        assert np.abs(force[:, off[k]:off[k + 1]].T - c["forces"]).max() < 1e-7
        assert np.abs(stress[:, k] - c["stress"]).max() < 1e-6
    bad = cases[0]["species"].copy()
    bad[0] = 3
    out = L.fgap_calc_status(bad, cases[0]["lat"], cases[0]["pos"], theta[:d], mm[:n, :d], coeff[:n], 6.0, True, 0)
    assert out[-1] == 4
    assert b"atomic number 3" in L.fgap_last_error()


def test_fgap_descriptors_match_oracle(in_model_dir, frames):
    import libgap.libgap as L
    n, d, theta, mm, invcmm, coeff = L.fgap_read()
    xx, status = L.fgap_descriptors(frames["numbers"], frames["cells"][2], frames["positions"][2], theta[:d],
                                    mm[:n, :d], coeff[:n], 6.0)
    assert status == 0
    om = oracle.parse_model("gap_parameters")
    xo = oracle.descriptors(om, frames["numbers"], frames["cells"][2], frames["positions"][2], 6.0)
    np.testing.assert_allclose(xx.T, xo, rtol=1e-11, atol=1e-9)


def _close(got, ref, rel):
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=0, atol=rel * max(1.0, float(np.abs(ref).max())))


@pytest.mark.parametrize("case", CAR_CASE_NAMES)
def test_car2acsf_matches_reference(tmp_path, monkeypatch, ref_vectors, case):
    """CAR2ACSF through the f2py module with the reference's signature (the SF
    table read from ./gap_parameters) and through the C ABI, against the
    compiled reference's xx, dense dxdy and strs on FGAP_CALC's own neighbour
    tables (6 and 8 atoms; 8 atoms with 920 neighbours each under a 7-cutoff
    table, whose FP64 backward keeps its cutoff tables in global memory).
    FP64 bar: 1e-11 of each array's largest entry (the GPU sums in another
    order); lgrad = .false. gives xx and zeros."""
    c = car_case(ref_vectors, case)
    om = golden_model(ref_vectors, c["model"])
    oracle.write_model(str(tmp_path / "gap_parameters"), om)
    monkeypatch.chdir(tmp_path)
    import libgap.libgap as L
    nf = om.des_len
    xx, dxdy, strs = L.car2acsf(nf, np.asfortranarray(c["pos"]), c["neighbor"], c["count"], True)
    for got, ref in ((xx, c["xx"]), (dxdy, c["dxdy"]), (strs, c["strs"])):
        _close(got, ref, 1e-11)
    xx0, d0, s0 = L.car2acsf(nf, np.asfortranarray(c["pos"]), c["neighbor"], c["count"], False)
    _close(xx0, c["xx"], 1e-11)
    assert not np.any(d0) and not np.any(s0)
    m = gapx.Model.from_model(om)
    xx2, dxdy2, strs2 = m.car2acsf(c["pos"], c["neighbor"], c["count"])
    for got, ref in ((xx2, c["xx"]), (dxdy2, c["dxdy"]), (strs2, c["strs"])):
        _close(got, ref, 1e-11)
    m.close()
