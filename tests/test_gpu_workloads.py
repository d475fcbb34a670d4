"""The batched population and the small single-species cell at their full
sizes, against the CPU oracle.

* Population: 512 periodic structures of 8-256 atoms with 1-4 of the species
  1, 6, 8, 22 (10-20 A^3 per atom, triclinic cells), the D = 128, M = 4000
  model; the whole batch runs E+F+S from one context with graph replay, and a
  seeded sample of 32 structures (the smallest and the largest included)
  is compared with the oracle's sparse mode structure by structure: FP64 at
  1e-9 eV/atom, 1e-7 eV/A, 1e-6 GPa; mixed (FP16 and BF16 split
  contractions) at 1 meV/atom, 20 meV/A, 0.05 GPa.
* 64 carbon atoms at 0.06 atoms/A^3 with 24 radial and 24 angular lines,
  zeta in {1, 2, 4, 16} (D = 96, M = 500): FP64 and mixed against the oracle
  at the same bars.
"""
import numpy as np
import pytest

import gapx
import oracle
from gapx import synth
from helpers import workload_model

pytestmark = pytest.mark.gpu

FP64_BAR = (1e-9, 1e-7, 1e-6)
MIXED_BAR = (1e-3, 2e-2, 5e-2)


@pytest.fixture(scope="module")
def population():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    om = workload_model("population")
    structs = synth.population(synth.POPULATION["n_struct"])
    m = gapx.Model.from_model(om)
    n = np.array([len(s["species"]) for s in structs])
    rng = np.random.default_rng(2024)
    pick = set(rng.choice(len(structs), size=30, replace=False).tolist())
    pick |= {int(n.argmin()), int(n.argmax())}
    pick = sorted(pick)
    refs = {k: oracle.gap_calc(om, structs[k]["species"], structs[k]["lat"], structs[k]["pos"], 6.0,
                               mode="sparse", nthreads=16) for k in pick}
    yield om, structs, m, pick, refs
    m.close()


def run_batch(m, structs, flags):
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.use_graph(True)
    ctx.run(flags)
    ctx.results()
    ctx.run(flags)  # a replay of the captured graph
    r = ctx.results()
    ctx.close()
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    return r, offs


@pytest.mark.parametrize("mode", ["fp64", "fp16x3", "bf16x3"])
def test_population_sample_matches_oracle(population, mode):
    om, structs, m, pick, refs = population
    flags = gapx.GRAD | {"fp64": 0, "fp16x3": gapx.MIXED, "bf16x3": gapx.MIXED | gapx.BF16}[mode]
    bar = FP64_BAR if mode == "fp64" else MIXED_BAR
    r, offs = run_batch(m, structs, flags)
    assert len(r["energy"]) == 512 and np.all(np.isfinite(r["energy"])) and np.all(np.isfinite(r["forces"]))
    species_counts = set()
    worst = np.zeros(3)
    for k in pick:
        s, ref = structs[k], refs[k]
        n = len(s["species"])
        species_counts.add(len(set(s["species"].tolist())))
        f = r["forces"][offs[k]:offs[k + 1]]
        e = (abs(r["energy"][k] - ref["energy"]) / n, float(np.abs(f - ref["forces"]).max()),
             float(np.abs(r["stress"][k] - ref["stress"]).max()))
        worst = np.maximum(worst, e)
    n = np.array([len(structs[k]["species"]) for k in pick])
    print("%s: %d structures of %d-%d atoms, %s species; worst dE/atom %.2e dF %.2e dS %.2e"
          % (mode, len(pick), n.min(), n.max(), sorted(species_counts), *worst))
    assert n.min() <= 16 and n.max() >= 240 and species_counts == {1, 2, 3, 4}
    assert np.all(worst <= bar), worst


@pytest.mark.parametrize("mode", ["fp64", "fp16x3"])
def test_c64_zeta_workload_matches_oracle(mode):
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    om = workload_model("c64")
    s = synth.workload_structure("c64")
    assert om.des_len == 96 and om.nsparse == 500 and len(s["species"]) == 64
    assert sorted(set(np.asarray(om.sf_zeta).tolist())) == [1.0, 2.0, 4.0, 16.0]
    m = gapx.Model.from_model(om)
    r = m.calc(s["species"], s["lat"], s["pos"], 6.0, mixed=(mode != "fp64"))
    m.close()
    ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.0)
    e = (abs(r[0] - ref["energy"]) / 64, float(np.abs(r[1] - ref["forces"]).max()),
         float(np.abs(r[2] - ref["stress"]).max()))
    print("c64 zeta %s: dE/atom %.2e dF %.2e dS %.2e (E %.6f, |F|max %.3f)" % (mode, *e, ref["energy"],
                                                                                np.abs(ref["forces"]).max()))
    assert np.all(np.array(e) <= (FP64_BAR if mode == "fp64" else MIXED_BAR))
