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
    assert "xx,dxdy,strs = car2acsf(nf,pos,neighbor,neighbor_count,lgrad" in doc
    assert "write_array_2dim(a,name" in doc


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


@pytest.mark.parametrize("bad", ["species", "nsf", "sf_line"])
def test_failed_header_read_is_reported_again_not_aborted(tmp_path, bad):
    """A gap_parameters header that fails part-way (a bad species line, SF
    count or SF line) is a status error on every call, not a Fortran abort on
    the second call: the reader frees what it allocated and forgets the
    file's stamp.  After the file is fixed, the next read goes past the header
    (here to the model upload, which needs a GPU: status 6, or 0 on one)."""
    lines = open(MODEL).read().splitlines()
    ns = int(lines[0].split()[0])
    k = {"species": 1, "nsf": 1 + ns, "sf_line": 2 + ns}[bad]
    broken = list(lines)
    broken[k] = "x y z"
    (tmp_path / "gap_parameters").write_text("\n".join(broken) + "\n")
    code = """
import numpy as np, os, time
import libgap.libgap as L
m = np.zeros((2, 66), order="F"); th = np.ones(66); c = np.zeros(2)
lat = np.eye(3) * 8.0; pos = np.zeros((1, 3), order="F"); sp = np.array([6], np.int32)
out = [L.fgap_calc_status(sp, lat, pos, th, m, c, 6.0, True, 0)[-1] for _ in range(3)]
os.replace("gap_parameters.good", "gap_parameters")
t = time.time() + 2.0
os.utime("gap_parameters", (t, t))
out.append(L.fgap_calc_status(sp, lat, pos, th, m, c, 6.0, True, 0)[-1])
print("statuses", out)
"""
    shutil.copy(MODEL, tmp_path / "gap_parameters.good")
    r = _run(code, tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("statuses")][0]
    st = eval(line.split(" ", 1)[1])
    assert st[:3] == [3, 3, 3]
    assert st[3] in (0, 6, 7)


def test_write_array_2dim_writes_the_reference_text(tmp_path):
    """write_array_2dim (wacsf.f90:798-810) writes, byte for byte, the text the
    compiled reference wrote for the same matrix (tests/golden)."""
    import libgap.libgap as L
    ref = np.load(os.path.join(GOLDEN, "ref_vectors.npz"))
    a = np.asfortranarray(ref["write_array_input"])
    L.write_array_2dim(a, "  " + str(tmp_path / "out.txt") + "  ")  # blanks around the name are dropped
    got = open(tmp_path / "out.txt", "rb").read()
    want = open(os.path.join(GOLDEN, "write_array_2dim.txt"), "rb").read()
    assert got == want
