"""Precision form at full size: 32768 ternary atoms at 0.07 atoms/A^3, rcut
6.5 A, D = 256 (64 radial + 64 angular lines), M = 8000.

* FP64 against the CPU sparse oracle on the whole structure: 1e-9 eV/atom,
  1e-7 eV/A, 1e-6 GPa.
* FP16 and BF16 split contractions against FP64 (errors printed): 1 meV/atom
  and 20 meV/A; sum of forces 1e-6 eV/A per atom-count scale.
"""
import numpy as np
import pytest

import gapx
import oracle
from gapx import synth
from helpers import workload_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg3():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    om = workload_model("ternary32k")
    s = synth.workload_structure("ternary32k")
    m = gapx.Model.from_model(om)
    r64 = m.calc(s["species"], s["lat"], s["pos"], 6.5)
    yield om, s, m, r64
    m.close()


def test_fp64_matches_oracle_full_size(cfg3):
    om, s, m, r = cfg3
    n = len(s["species"])
    assert m.des_len == 256 and m.nsparse == 8000 and n == 32768
    ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.5, mode="sparse", nthreads=16)
    de = abs(r[0] - ref["energy"]) / n
    df = float(np.abs(r[1] - ref["forces"]).max())
    ds = float(np.abs(r[2] - ref["stress"]).max())
    print("config 3 FP64 vs oracle: dE/atom %.2e  dF %.2e  dS %.2e" % (de, df, ds))
    assert de <= 1e-9 and df <= 1e-7 and ds <= 1e-6
    assert np.abs(r[1].sum(axis=0)).max() < 1e-7


@pytest.mark.parametrize("bf16", [False, True], ids=["fp16x3", "bf16x3"])
def test_mixed_against_fp64(cfg3, bf16):
    om, s, m, r64 = cfg3
    n = len(s["species"])
    rmx = m.calc(s["species"], s["lat"], s["pos"], 6.5, mixed=True, flags=gapx.BF16 if bf16 else 0)
    de = abs(rmx[0] - r64[0]) / n
    df = float(np.abs(rmx[1] - r64[1]).max())
    ds = float(np.abs(rmx[2] - r64[2]).max())
    print("config 3 mixed (%s) vs FP64: dE/atom %.3e eV  dF %.3e eV/A  dS %.3e GPa"
          % ("bf16" if bf16 else "fp16", de, df, ds))
    assert de <= 1e-3 and df <= 2e-2 and ds <= 5e-2
    assert np.abs(rmx[1].sum(axis=0)).max() < 1e-4
