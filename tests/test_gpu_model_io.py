"""Model I/O (gapx_model_write / save / load / read_cached).

* The text writer's output parses, with the oracle's list-directed reader
  (pinned on the reference's own gap_parameters), into exactly the same
  doubles, for the shipped example model and for a model with zeta columns;
  gapx_model_read of it gives the same arrays back.
* A binary save/load round trip keeps every array and gives bit-identical
  evaluations.
* read_cached writes its binary copy once, reuses it while the text file is
  unchanged, rebuilds it when the text changes or the copy is damaged.
"""
import os

import numpy as np
import pytest

import gapx
import oracle
from helpers import small_model, ternary_structure

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "c64_diamond", "gap_parameters")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")


def same_model(a, b):
    for k in ("z", "w", "sf_type", "sf_alpha", "sf_cut", "theta", "mm", "coeff"):
        assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k))), k
    za = np.ones(a.nsf) if a.sf_zeta is None else np.asarray(a.sf_zeta)
    zb = np.ones(b.nsf) if b.sf_zeta is None else np.asarray(b.sf_zeta)
    assert np.array_equal(za, zb)


def zeta_model():
    rng = np.random.default_rng(5)
    m, xs = small_model(rng, zeta=(1, 2, 4, 16), M=9)
    m.mm = xs * rng.uniform(0.5, 3.0, xs.shape[1])
    m.theta = rng.uniform(0.1, 1e6, m.des_len)
    m.coeff = rng.normal(0, 1e-3, m.nsparse)
    return m


@pytest.mark.parametrize("which", ["example", "zeta"])
def test_write_parses_back_to_the_same_doubles(tmp_path, which):
    om = oracle.parse_model(GOLDEN) if which == "example" else zeta_model()
    g = gapx.Model.from_model(om)
    path = str(tmp_path / "gap_parameters")
    g.write(path)
    same_model(oracle.parse_model(path), om)
    g2 = gapx.Model.read(path)
    th, mm, cf = g2.arrays()
    assert np.array_equal(th, om.theta) and np.array_equal(mm, om.mm) and np.array_equal(cf, om.coeff)
    h = g2.header()
    assert np.array_equal(h["sf_zeta"], np.ones(om.nsf) if om.sf_zeta is None else om.sf_zeta)
    g.close()
    g2.close()


def test_save_load_roundtrip_is_exact(tmp_path):
    om = zeta_model()
    om.theta = np.ones(om.des_len)
    g = gapx.Model.from_model(om)
    path = str(tmp_path / "model.gapxbin")
    g.save(path)
    g2 = gapx.Model.load(path)
    assert np.array_equal(g2.arrays()[1], om.mm) and np.array_equal(g2.header()["sf_zeta"], g.header()["sf_zeta"])
    s = ternary_structure(20, 12.0, 3)
    r1 = g.calc(s["species"], s["lat"], s["pos"], 6.0)
    r2 = g2.calc(s["species"], s["lat"], s["pos"], 6.0)
    assert r1[0] == r2[0] and np.array_equal(r1[1], r2[1]) and np.array_equal(r1[2], r2[2])
    with open(path, "r+b") as f:  # a flipped byte fails the checksum
        f.seek(40)
        b = f.read(1)
        f.seek(40)
        f.write(bytes([b[0] ^ 0xFF]))
    with pytest.raises(gapx.GapxError) as e:
        gapx.Model.load(path)
    assert e.value.code == 2
    with pytest.raises(gapx.GapxError):
        gapx.Model.load(GOLDEN)  # text is not a binary model
    g.close()
    g2.close()


def test_read_cached_reuses_and_refreshes(tmp_path):
    om = oracle.parse_model(GOLDEN)
    path = str(tmp_path / "gap_parameters")
    cache = path + ".gapxbin"
    gapx.Model.from_model(om).write(path)
    m1 = gapx.Model.read(path, cache=True)
    assert os.path.exists(cache)
    st1 = os.stat(cache)
    m2 = gapx.Model.read(path, cache=True)
    assert os.stat(cache).st_mtime_ns == st1.st_mtime_ns  # reused, not rewritten
    assert np.array_equal(m2.arrays()[2], om.coeff)
    # new text (different coefficients, new mtime): the copy is rebuilt
    om.coeff = om.coeff * 2.0
    gapx.Model.from_model(om).write(path)
    os.utime(path, ns=(st1.st_mtime_ns + 10**9, st1.st_mtime_ns + 10**9))
    m3 = gapx.Model.read(path, cache=True)
    assert np.array_equal(m3.arrays()[2], om.coeff)
    # a damaged copy is rebuilt as well
    with open(cache, "r+b") as f:
        f.seek(100)
        b = f.read(1)
        f.seek(100)
        f.write(bytes([b[0] ^ 0x55]))
    m4 = gapx.Model.read(path, cache=True)
    assert np.array_equal(m4.arrays()[1], om.mm)
    gapx.Model.load(cache).close()  # valid again
    for m in (m1, m2, m3, m4):
        m.close()
