"""The drop-in boundary on the CPU: the f2py module libgap.libgap exports
the reference's entry points with the reference's Python signatures, the
Fortran model reader agrees with the oracle's parser and keeps the
reference's caps and messages, and the Python/ASE layers import."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT

PKG = os.path.join(ROOT, "calypso-gap_amd")
MODEL = os.path.join(GOLDEN, "c64_diamond", "gap_parameters")


@pytest.fixture
def in_model_dir(tmp_path, monkeypatch):
    shutil.copy(MODEL, tmp_path / "gap_parameters")
    monkeypatch.chdir(tmp_path)
    return tmp_path


def test_f2py_signatures_match_the_reference():
    import libgap.libgap as L
    doc = L.__doc__
    # the signatures f2py generates from gappy/libgap/*.f90 (setup.py:36-45):
    # same names, same order
    assert "ene,force,stress,variance = fgap_calc(species,lat,pos,theta,mm,qmm,coeff,rcut,lgrad" in doc
    assert "nsparsex,des_len,theta,mm,invcmm,coeff = fgap_read()" in doc
    assert "min_bond = fget_bond(lat,elements,pos,rcut" in doc


def test_fortran_reader_matches_oracle_parser(in_model_dir):
    import libgap.libgap as L
    n, d, theta, mm, invcmm, coeff = L.fgap_read()
    om = oracle.parse_model("gap_parameters")
    assert (n, d) == (om.nsparse, om.des_len)
    assert theta.shape == (100,) and mm.shape == (4000, 100) and invcmm.shape == (4000, 4000)
    np.testing.assert_array_equal(theta[:d], om.theta)
    np.testing.assert_array_equal(mm[:n, :d], om.mm)
    np.testing.assert_array_equal(coeff[:n], om.coeff)
    assert not np.any(invcmm)
    n2, d2, st = L.fgap_read_dims()
    t2, m2, c2, st2 = L.fgap_read2(n2, d2)
    assert st == 0 and st2 == 0
    np.testing.assert_array_equal(m2, om.mm)


def _run(code, cwd):
    env = dict(os.environ, PYTHONPATH=PKG)
    return subprocess.run([sys.executable, "-c", code], cwd=cwd, env=env, capture_output=True, text=True, timeout=120)


def test_fgap_read_missing_file_stops_like_the_reference(tmp_path):
    r = _run("import libgap.libgap as L; L.fgap_read(); print('not reached')", tmp_path)
    assert "gap_parameters file does not exist!" in r.stdout
    assert "not reached" not in r.stdout


def test_fgap_read_caps_stop_like_the_reference(tmp_path):
    m = oracle.parse_model(MODEL)
    # 51 SF lines -> des_len 102 > nsf_max = 100 (gap_calc.f90:346-350)
    k = 51 - m.nsf
    m.sf_type = np.concatenate([m.sf_type, np.ones(k, np.int32)])
    m.sf_alpha = np.concatenate([m.sf_alpha, np.full(k, 0.3)])
    m.sf_cut = np.concatenate([m.sf_cut, np.full(k, 6.0)])
    m.theta = np.ones(102)
    m.mm = np.zeros((m.nsparse, 102))
    oracle.write_model(str(tmp_path / "gap_parameters"), m)
    r = _run("import libgap.libgap as L; L.fgap_read(); print('not reached')", tmp_path)
    assert "The length of descriptors large than nsf_max=" in r.stdout
    assert "not reached" not in r.stdout
    # the uncapped reader takes it
    r = _run("import libgap.libgap as L; print(L.fgap_read_dims())", tmp_path)
    assert "102" in r.stdout


def test_gap_calculator_reads_once_and_maps_symbols(in_model_dir):
    from libgap.GAP import Calculator
    c = Calculator(rcut=6.0)
    assert (c.nsparseX, c.des_len) == (129, 66)
    assert c.get_elenum(["B", "C", "Ti"]) == [5, 6, 22]
    assert c.invcmm.shape == (129, 129) and not c.invcmm.any()
    with pytest.raises(ValueError):
        Calculator(precision="fp8")


def test_ase_adapter_imports_against_ase_interface():
    sys.path.insert(0, os.path.join(ROOT, "tests", "stub_ase"))
    try:
        import importlib.util
        spec = importlib.util.spec_from_file_location("gap_calc", os.path.join(PKG, "ASE", "gap_calc.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        assert mod.GAP.implemented_properties == ["energy", "forces", "stress"]
    finally:
        sys.path.remove(os.path.join(ROOT, "tests", "stub_ase"))
