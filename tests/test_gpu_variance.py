"""Predictive variance on the GPU (gapx_ctx_variance, gapx_calc with
GAPX_VARIANCE) against the oracle (oracle.gpr_variance).

Bars: FP64 runs 1e-10 absolute in sigma^2 (sigma^2 lies in [0, 1] for the
Q used here); mixed runs (FP32 features, widened) 1e-5 absolute.  At the
4096-atom workload size (M = 4000): with Q = (K_MM + 1e-6 I)^-1, every
sigma^2 lies in [-1e-9, 1 + 1e-12] (a Schur complement of a positive
definite kernel matrix), and 96 atoms spread over the batch match the oracle
evaluated on the run's own features, to 1e-9 (FP64 run; the norm expansion
costs about 1e-14 there).  A batch larger than one block of K (population
workload, mixed run) checks the block loop the same way, to 1e-6: the run
rounds the standardised features to FP32 after the raw ones it reports,
which moves sigma^2 by about 4e-8.
"""
import numpy as np
import pytest

import gapx
import oracle
from gapx import synth
from helpers import fitted_small_model, ternary_structure, workload_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")


def small_case():
    rng = np.random.default_rng(11)
    structs = [ternary_structure(n, 12.0, 20 + n) for n in (17, 40, 9)]
    m = fitted_small_model(rng, structs[1])
    Kmm = oracle.covariance(m, m.mm)
    Q = np.linalg.inv(Kmm + 0.05 * np.eye(m.nsparse))
    return m, structs, Q


@pytest.mark.parametrize("flags,tol", [(gapx.GRAD, 1e-10), (gapx.GRAD | gapx.MIXED, 1e-5), (0, 1e-10)],
                         ids=["fp64", "mixed", "fp64-energy-only"])
def test_variance_matches_oracle(flags, tol):
    om, structs, Q = small_case()
    m = gapx.Model.from_model(om)
    m.set_qmm(Q)
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.run(flags)
    var, s2 = ctx.variance()
    x = np.concatenate([oracle.descriptors(om, s["species"], s["lat"], s["pos"], 6.0) for s in structs])
    var_o, s2_o = oracle.gpr_variance(om, x, Q, natoms=[len(s["species"]) for s in structs])
    print("sigma2 range [%.3e, %.3e], max err %.2e" % (s2_o.min(), s2_o.max(), np.abs(s2 - s2_o).max()))
    assert s2_o.min() > 1e-3 and s2_o.max() < 1.0  # an informative Q
    assert np.abs(s2 - s2_o).max() <= tol
    assert np.abs(var - var_o).max() <= tol
    ctx.close()
    m.close()


def test_zero_q_and_calc_flag():
    om, structs, Q = small_case()
    m = gapx.Model.from_model(om)
    s = structs[0]
    # no Q: the reference's INVCMM = 0 gives sigma^2 = 1
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.run(gapx.GRAD)
    var, s2 = ctx.variance()
    assert np.array_equal(s2, np.ones(len(s2))) and np.array_equal(var, np.ones(len(structs)))
    # gapx_calc: 0 as FGAP_CALC unless GAPX_VARIANCE is given
    assert m.calc(s["species"], s["lat"], s["pos"], 6.0)[3] == 0.0
    m.set_qmm(Q)
    assert m.calc(s["species"], s["lat"], s["pos"], 6.0)[3] == 0.0
    e0, f0, _, v = m.calc(s["species"], s["lat"], s["pos"], 6.0, flags=gapx.VARIANCE)
    x = oracle.descriptors(om, s["species"], s["lat"], s["pos"], 6.0)
    assert abs(v - oracle.gpr_variance(om, x, Q)[0][0]) <= 1e-10
    e1, f1, _, _ = m.calc(s["species"], s["lat"], s["pos"], 6.0)
    assert e0 == e1 and np.array_equal(f0, f1)
    # clearing Q goes back to sigma^2 = 1
    m.set_qmm(None)
    ctx.run(gapx.GRAD)
    assert np.array_equal(ctx.variance()[1], np.ones(len(s2)))
    with pytest.raises(gapx.GapxError):
        m.set_qmm(np.zeros((3, 3)))
    ctx.close()
    m.close()


def check_large(om, structs, flags, n_check, seed, tol):
    M = om.nsparse
    Kmm = oracle.covariance(om, om.mm)
    Q = np.linalg.inv(Kmm + 1e-6 * np.eye(M))
    m = gapx.Model.from_model(om)
    m.set_qmm(Q)
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.run(flags)
    var, s2 = ctx.variance()
    x, _ = ctx.features(dedx=False)
    natoms = [len(s["species"]) for s in structs]
    print("N %d, M %d: sigma2 in [%.3e, %.3e], mean %.3e" % (len(s2), M, s2.min(), s2.max(), s2.mean()))
    assert s2.min() >= -1e-9 and s2.max() <= 1.0 + 1e-12
    offs = np.concatenate([[0], np.cumsum(natoms)])
    assert np.allclose(var, [s2[offs[k]:offs[k + 1]].mean() for k in range(len(structs))], rtol=1e-13, atol=1e-15)
    idx = np.sort(np.random.default_rng(seed).choice(len(s2), size=n_check, replace=False))
    _, s2_o = oracle.gpr_variance(om, x[idx], Q)
    err = np.abs(s2[idx] - s2_o).max()
    print("checked %d atoms against the oracle: max err %.2e" % (n_check, err))
    assert err <= tol
    ctx.close()
    m.close()


def test_variance_at_workload_size():
    om = workload_model("ternary4k")
    check_large(om, [synth.workload_structure("ternary4k")], gapx.GRAD, 96, 1, 1e-9)


def test_variance_block_loop_on_population():
    """The population workload (512 structures, 67,910 atoms): five blocks of
    K at M = 4000, mixed-precision features."""
    from test_gpu_batch import gpu_descriptors
    d = synth.population_model(gpu_descriptors)
    om = oracle.Model(z=d["z"], w=d["w"], sf_type=d["sf_type"], sf_alpha=d["sf_alpha"], sf_cut=d["sf_cut"],
                      theta=d["theta"], mm=d["mm"], coeff=d["coeff"], sf_zeta=d["sf_zeta"])
    check_large(om, synth.population(), gapx.MIXED, 128, 2, 1e-6)
