"""The host restatement of ASE's LBFGS step (tests/helpers.py), which the GPU
relaxation tests use as their checker, pinned against textbook BFGS.

* The two-loop recursion equals the explicit inverse-Hessian update
  H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T applied pair by pair
  from H0 = I / alpha (Nocedal and Wright, Numerical Optimization, eq. 6.17
  and algorithm 7.4), for random q.
* It satisfies the secant condition of its newest pair: H y = s.
* With no history the step is f / alpha; with a memory of m only the last m
  pairs count; a step longer than maxstep is scaled to it.
"""
import numpy as np

from helpers import lbfgs_step_host, lbfgs_two_loop


def pairs(rng, n, m):
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)  # positive definite: y . s > 0
    S = [rng.normal(size=n) for _ in range(m)]
    return S, [A @ s for s in S]


def explicit_h(S, Y, alpha, n):
    H = np.eye(n) / alpha
    for s, y in zip(S, Y):
        rho = 1.0 / np.dot(y, s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H


def test_two_loop_equals_explicit_bfgs_update():
    rng = np.random.default_rng(3)
    n = 12
    S, Y = pairs(rng, n, 5)
    H = explicit_h(S, Y, 70.0, n)
    for _ in range(4):
        q = rng.normal(size=n)
        np.testing.assert_allclose(lbfgs_two_loop(S, Y, q, 70.0), H @ q, rtol=1e-12, atol=1e-14)


def test_secant_condition_of_the_newest_pair():
    rng = np.random.default_rng(4)
    S, Y = pairs(rng, 9, 4)
    np.testing.assert_allclose(lbfgs_two_loop(S, Y, Y[-1], 70.0), S[-1], rtol=1e-11, atol=1e-14)


def test_first_step_memory_and_maxstep():
    rng = np.random.default_rng(5)
    x0 = rng.normal(size=(4, 3))
    f0 = rng.normal(size=(4, 3)) * 0.5
    # no history: x0 + f0 / alpha (0.5 / 70 is far below maxstep)
    np.testing.assert_allclose(lbfgs_step_host([x0], [f0]), x0 + f0 / 70.0, rtol=1e-15)
    # a step longer than maxstep is scaled so that its longest atom moves maxstep
    big = lbfgs_step_host([x0], [f0 * 1e3])
    assert abs(np.sqrt(((big - x0) ** 2).sum(1)).max() - 0.2) < 1e-12
    # memory m: the pairs before the last m do not count
    xs = [x0 + 0.01 * rng.normal(size=(4, 3)) for _ in range(6)]
    fs = [f0 + 0.01 * rng.normal(size=(4, 3)) for _ in range(6)]
    a = lbfgs_step_host(xs, fs, memory=2)
    b = lbfgs_step_host(xs[3:], fs[3:], memory=2)
    np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-15)
    assert np.abs(lbfgs_step_host(xs, fs, memory=3) - a).max() > 1e-9
