"""FP64 GPU path against the reference's golden outputs and the CPU oracle.

Bar (FP64 build): 1e-9 eV/atom energy, 1e-7 eV/A forces, 1e-6 GPa stress
(or 1e-8 of the largest stress component where that is larger: the FP64
rounding level of a virial summed over some 10^4 terms in the dense golden
cells, whose stresses reach 4e4 GPa).
"""
import numpy as np
import pytest

import gapx
import oracle
from conftest import GOLDEN_CASE_NAMES, golden_cases, golden_model
from helpers import fitted_small_model, ternary_structure, workload_model

pytestmark = pytest.mark.gpu

TOL_E, TOL_F, TOL_S = 1e-9, 1e-7, 1e-6


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")


def check(res, ref, n, te=TOL_E, tf=TOL_F, ts=TOL_S):
    e, f, s = res[0], res[1], res[2]
    de = abs(e - ref["energy"]) / n
    df = float(np.abs(f - ref["forces"]).max())
    ds = float(np.abs(s - ref["stress"]).max())
    ts = max(ts, 1e-8 * float(np.abs(ref["stress"]).max()))
    assert de < te and df < tf and ds < ts, (de, df, ds)


def test_descriptors_match_oracle():
    rng = np.random.default_rng(21)
    s = ternary_structure(150, 11.0, 211)
    om = fitted_small_model(rng, s, zeta=(1, 2, 4, 16))
    m = gapx.Model.from_model(om)
    x = m.descriptors(s["species"], s["lat"], s["pos"], 6.0)
    xo = oracle.descriptors(om, s["species"], s["lat"], s["pos"], 6.0)
    np.testing.assert_allclose(x, xo, rtol=1e-11, atol=1e-13)


def test_ase_traj_frames(frames):
    import os
    from conftest import GOLDEN
    m = gapx.Model.read(os.path.join(GOLDEN, "c64_diamond", "gap_parameters"))
    n = len(frames["numbers"])
    for k in range(11):
        print("frame", k, flush=True)
        res = m.calc(frames["numbers"], frames["cells"][k], frames["positions"][k], 6.0)
        ref = dict(energy=frames["energies"][k], forces=frames["forces"][k], stress=frames["stresses"][k])
        check(res, ref, n)


@pytest.mark.parametrize("case", GOLDEN_CASE_NAMES)
def test_reference_vectors(ref_vectors, case):
    c = [x for x in golden_cases(ref_vectors) if x["name"] == case][0]
    m = gapx.Model.from_model(golden_model(ref_vectors, c["model"]))
    res = m.calc(c["species"], c["lat"], c["pos"], c["rcut"], grad=c["lgrad"])
    check(res, c, len(c["species"]))
    if not c["lgrad"]:
        assert not np.any(res[1]) and not np.any(res[2])


@pytest.mark.parametrize("n,vpa,seed", [(40, 10.0, 1), (300, 12.0, 2), (7, 13.0, 3)])
def test_small_models_match_oracle(n, vpa, seed):
    rng = np.random.default_rng(seed)
    s = ternary_structure(n, vpa, 300 + seed)
    om = fitted_small_model(rng, s, zeta=(1, 2, 4, 16))
    m = gapx.Model.from_model(om)
    res = m.calc(s["species"], s["lat"], s["pos"], 6.0)
    ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.0, mode="sparse", nthreads=8)
    check(res, ref, n)
    assert np.abs(res[1].sum(axis=0)).max() < 1e-9


def test_ternary4k_full_size_matches_oracle():
    """4096 atoms, D = 128, M = 4000 (the benchmark workload) against the sparse oracle."""
    from gapx import synth
    om = workload_model("ternary4k")
    s = synth.workload_structure("ternary4k")
    m = gapx.Model.from_model(om)
    res = m.calc(s["species"], s["lat"], s["pos"], 6.0)
    ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.0, mode="sparse", nthreads=16)
    check(res, ref, len(s["species"]))
    assert np.abs(res[1].sum(axis=0)).max() < 1e-8


def test_batch_equals_single_structures():
    from gapx import synth
    rng = np.random.default_rng(5)
    pop = synth.population(n_struct=12, nmin=8, nmax=64, seed=77)
    om = fitted_small_model(rng, pop[0], zeta=(1, 2), z=(1, 6, 8, 22), w=(1.0, 6.0, 8.0, 22.0))
    m = gapx.Model.from_model(om)
    ene, forces, stress = m.calc_batch(pop, 6.0)
    for k, s in enumerate(pop):
        ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.0, mode="sparse", nthreads=8)
        check((ene[k], forces[k], stress[k]), ref, len(s["species"]))


def test_energy_only_and_context_reuse():
    rng = np.random.default_rng(6)
    s = ternary_structure(64, 12.0, 66)
    om = fitted_small_model(rng, s)
    m = gapx.Model.from_model(om)
    ctx = gapx.Context(m)
    ctx.load([s], 6.0)
    ctx.run(0)
    r0 = ctx.results()
    ctx.run(gapx.GRAD)
    r1 = ctx.results()
    assert r0["energy"][0] == r1["energy"][0]
    assert not np.any(r0["forces"])
    assert abs(r1["eatom"].sum() - r1["energy"][0]) < 1e-12
    # repeated runs are bitwise reproducible
    ctx.run(gapx.GRAD)
    r2 = ctx.results()
    assert np.array_equal(r1["forces"], r2["forces"]) and np.array_equal(r1["stress"], r2["stress"])


def test_errors_are_reported_not_faults():
    rng = np.random.default_rng(7)
    s = ternary_structure(20, 12.0, 77)
    om = fitted_small_model(rng, s)
    m = gapx.Model.from_model(om)
    bad = s["species"].copy()
    bad[3] = 79
    with pytest.raises(gapx.GapxError) as e:
        m.calc(bad, s["lat"], s["pos"], 6.0)
    assert e.value.code == 4
    # more than 1000 neighbours with the reference's cap
    lat = np.eye(3) * 4.0
    pos = np.random.default_rng(1).random((40, 3)) * 4.0
    with pytest.raises(gapx.GapxError) as e:
        m.calc(np.ones(40, np.int32), lat, pos, 9.0, flags=gapx.REF_NEIGHBOR_CAP)
    assert e.value.code == 5


def test_stage_timing():
    rng = np.random.default_rng(8)
    s = ternary_structure(64, 12.0, 88)
    m = gapx.Model.from_model(fitted_small_model(rng, s))
    ctx = gapx.Context(m)
    ctx.load([s], 6.0)
    ctx.timing(True)
    for _ in range(3):
        ctx.run(gapx.GRAD)
    ctx.sync()
    st = ctx.stages()
    assert st["gpr_contraction"][1] == 3 and st["gpr_contraction"][0] > 0


def test_fma_probe_is_plausible():
    """The measurement aid returns finite rates below the published vector
    peaks (FP32 157.3 TFLOP/s with packed instructions, FP64 78.6)."""
    c = gapx.probe_fma_rates()
    print("measured TFLOP/s:", c)
    for k, peak in (("fma_f32", 157.3), ("fma_f64", 78.6)):
        assert np.isfinite(c[k]) and 0.02 * peak < c[k] < 1.05 * peak, k


@pytest.mark.parametrize("zeta", [(1, 2, 4, 8), (1, 2, 4, 16), (1, 4, 16), (2, 8), (1, 2, 4, 8, 16)],
                         ids=["1-2-4-8", "1-2-4-16", "1-4-16", "2-8", "1-2-4-8-16"])
def test_grid_zeta_lists_match_oracle(zeta):
    """The grid-form descriptor kernels for several zeta lists: the {1, 2, 4, 16}
    list runs a compile-time power chain, the others the run-time chain; both
    match the oracle in FP64, and the mixed path stays within its bar."""
    rng = np.random.default_rng(41)
    s = ternary_structure(60, 11.0, 411)
    om = fitted_small_model(rng, s, zeta=zeta)
    m = gapx.Model.from_model(om)
    res = m.calc(s["species"], s["lat"], s["pos"], 6.0)
    ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.0)
    check(res, ref, len(s["species"]))
    mix = m.calc(s["species"], s["lat"], s["pos"], 6.0, mixed=True)
    assert abs(mix[0] - res[0]) / len(s["species"]) <= 1e-3
    assert np.abs(mix[1] - res[1]).max() <= 2e-2
    m.close()
