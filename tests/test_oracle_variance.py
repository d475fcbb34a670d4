"""The oracle's predictive variance (oracle.gpr_variance, oracle.covariance)
pinned by what fixes it independently:

* the covariance against the C oracle's energies, E = sum_i sum_k k_ik c_k
  (gap_calc.f90:153-154), which are pinned to the reference's golden vectors;
* closed forms of the covariance (a hand-computed value, k = 1 on a sparse
  point, invariance under a common rescaling of x, MM and theta);
* Q = 0 (the reference's INVCMM, gap_calc.f90:355-358) gives sigma^2 = delta = 1;
* Q = K_MM^-1: sigma^2 = 0 on every sparse point (interpolation), and in
  general the Schur complement det([[K_MM, k^T], [k, 1]]) / det(K_MM),
  computed through log-determinants instead of the quadratic form;
* the per-structure value is the mean over that structure's atoms
  (gap_calc.f90:209: VARIANCE += covf(i) / na).

The reference computes none of this (VARIANCE = 0, gap_calc.f90:206), so no
reference output exists for it beyond that 0.
"""
import numpy as np
import pytest

import oracle
from helpers import fitted_small_model, ternary_structure


def test_covariance_reproduces_oracle_energy():
    rng = np.random.default_rng(7)
    s = ternary_structure(24, 12.0, 8)
    m = fitted_small_model(rng, s)
    x = oracle.descriptors(m, s["species"], s["lat"], s["pos"], 6.0)
    K = oracle.covariance(m, x)
    e = oracle.gap_calc(m, s["species"], s["lat"], s["pos"], 6.0, lgrad=False)["energy"]
    e_k = float((K @ m.coeff).sum())
    assert np.abs(K).max() > 1e-3  # kernel values of order one, so the energy tests K
    assert abs(e_k - e) <= 1e-10 * max(1.0, abs(e))


def test_covariance_closed_forms():
    m = oracle.Model(z=np.array([1]), w=np.zeros(1), sf_type=np.array([1]), sf_alpha=np.ones(1),
                     sf_cut=np.array([6.0]), theta=np.array([2.0, 0.5]), mm=np.array([[3.0, 0.0], [1.0, 1.0]]),
                     coeff=np.zeros(2))
    K = oracle.covariance(m, np.array([[1.0, 0.0]]))
    assert K[0, 0] == pytest.approx(np.exp(-0.5 * ((1 - 3) / 2.0) ** 2), rel=1e-15)
    assert K[0, 1] == pytest.approx(np.exp(-0.5 * ((1 - 1) / 2.0) ** 2 - 0.5 * ((0 - 1) / 0.5) ** 2), rel=1e-15)
    rng = np.random.default_rng(3)
    mm = rng.normal(size=(7, 5))
    big = oracle.Model(z=np.array([1]), w=np.zeros(1), sf_type=np.array([1]), sf_alpha=np.ones(1),
                       sf_cut=np.array([6.0]), theta=rng.uniform(0.5, 2, 5), mm=mm, coeff=np.zeros(7))
    assert np.allclose(np.diag(oracle.covariance(big, mm)), 1.0, rtol=0, atol=1e-15)
    x = rng.normal(size=(4, 5))
    scaled = oracle.Model(z=big.z, w=big.w, sf_type=big.sf_type, sf_alpha=big.sf_alpha, sf_cut=big.sf_cut,
                          theta=3.0 * big.theta, mm=3.0 * mm, coeff=big.coeff)
    assert np.allclose(oracle.covariance(scaled, 3.0 * x), oracle.covariance(big, x), rtol=1e-13, atol=0)


def sparse_model(rng, M=30, D=6):
    mm = rng.normal(size=(M, D)) * 1.5
    return oracle.Model(z=np.array([1]), w=np.zeros(1), sf_type=np.array([1]), sf_alpha=np.ones(1),
                        sf_cut=np.array([6.0]), theta=rng.uniform(0.8, 1.5, D), mm=mm, coeff=np.zeros(M))


def test_zero_q_gives_delta():
    rng = np.random.default_rng(4)
    m = sparse_model(rng)
    x = rng.normal(size=(11, 6))
    for q in (None, np.zeros((30, 30))):
        var, s2 = oracle.gpr_variance(m, x, q, natoms=[5, 6])
        assert np.array_equal(s2, np.ones(11)) and np.array_equal(var, np.ones(2))


def test_exact_inverse_interpolates_and_matches_schur_complement():
    rng = np.random.default_rng(5)
    m = sparse_model(rng)
    Kmm = oracle.covariance(m, m.mm)
    assert np.linalg.cond(Kmm) < 1e6
    Q = np.linalg.inv(Kmm)
    _, s2 = oracle.gpr_variance(m, m.mm, Q)
    assert np.abs(s2).max() < 1e-9
    x = rng.normal(size=(9, 6))
    _, s2 = oracle.gpr_variance(m, x, Q)
    k = oracle.covariance(m, x)
    _, ld0 = np.linalg.slogdet(Kmm)
    for i in range(len(x)):
        A = np.block([[Kmm, k[i][:, None]], [k[i][None, :], np.ones((1, 1))]])
        sign, ld = np.linalg.slogdet(A)
        assert sign > 0
        assert s2[i] == pytest.approx(np.exp(ld - ld0), rel=1e-7, abs=1e-12)
    assert np.all(s2 >= 0) and np.all(s2 <= 1)


def test_structure_mean_and_asymmetric_q():
    rng = np.random.default_rng(6)
    m = sparse_model(rng)
    x = rng.normal(size=(10, 6))
    A = rng.normal(size=(30, 30)) * 0.01
    var, s2 = oracle.gpr_variance(m, x, A, natoms=[3, 7])
    assert var[0] == pytest.approx(s2[:3].mean(), rel=1e-15) and var[1] == pytest.approx(s2[3:].mean(), rel=1e-15)
    # the quadratic form sees only the symmetric part of Q
    _, s2s = oracle.gpr_variance(m, x, 0.5 * (A + A.T))
    assert np.allclose(s2, s2s, rtol=0, atol=1e-14)
