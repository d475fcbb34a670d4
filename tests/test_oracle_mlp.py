"""Pins the MLP surrogate oracle (oracle/mlp_oracle.py) without a GPU.

* The hand-derived reverse pass (including the second-order terms of the
  input-gradient loss) equals central finite differences of the loss for
  parameters drawn from every block.
* The same check rejects a planted mistake: the reverse pass with the SiLU
  second derivative dropped.
* The input gradient G = de/dxt equals finite differences of the energy.
* SiLU derivatives equal finite differences of SiLU.
"""
import numpy as np
import pytest

import mlp_oracle as mo

BLOCKS = {
    "W1": (0, mo.H * mo.H),
    "b1": (mo.H * mo.H, mo.H),
    "W2": (mo.H * mo.H + mo.H, mo.H * mo.H),
    "b2": (2 * mo.H * mo.H + mo.H, mo.H),
    "w3": (2 * mo.H * mo.H + 2 * mo.H, mo.H),
    "b3": (2 * mo.H * mo.H + 3 * mo.H, 1),
}


def problem(seed, n=7):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed)
    X = rng.normal(0, 1.0, (n, mo.H))
    et = rng.normal(0, 1.0, n)
    Gt = rng.normal(0, 0.05, (n, mo.H))
    return p, X, et, Gt


def fd_grad(p, idx, X, et, Gt, cE, cG, h=1e-5):
    out = []
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        out.append((mo.loss(pp, X, et, Gt, cE, cG) - mo.loss(pm, X, et, Gt, cE, cG)) / (2 * h))
    return np.array(out)


def sample(rng, block, k=12):
    o, n = BLOCKS[block]
    return o + rng.choice(n, size=min(k, n), replace=False)


def test_parameter_count():
    assert mo.N_PARAMS == 33153
    assert sum(n for _, n in BLOCKS.values()) == mo.N_PARAMS


@pytest.mark.parametrize("block", list(BLOCKS))
@pytest.mark.parametrize("cE,cG", [(1.0, 0.0), (0.0, 1.0), (1.0, 3.0)])
def test_loss_grad_matches_finite_differences(block, cE, cG):
    p, X, et, Gt = problem(11)
    L, g = mo.loss_grad(p, X, et, Gt, cE, cG)
    assert L == pytest.approx(mo.loss(p, X, et, Gt, cE, cG), rel=1e-12)
    idx = sample(np.random.default_rng(5), block)
    fd = fd_grad(p, idx, X, et, Gt, cE, cG)
    scale = max(np.abs(fd).max(), 1e-3)
    assert np.abs(g[idx] - fd).max() <= 1e-6 * scale + 1e-9


def test_planted_mistake_is_caught():
    """Dropping the s'' terms changes the gradient of the G loss by far more
    than the finite-difference tolerance used above."""
    p, X, et, Gt = problem(12)
    _, g_bad = mo.loss_grad(p, X, et, Gt, 0.0, 1.0, drop_second_order=True)
    rng = np.random.default_rng(6)
    idx = np.concatenate([sample(rng, "W1", 24), sample(rng, "b1", 24), sample(rng, "b2", 24)])
    fd = fd_grad(p, idx, X, et, Gt, 0.0, 1.0)
    scale = max(np.abs(fd).max(), 1e-3)
    assert np.abs(g_bad[idx] - fd).max() > 100 * (1e-6 * scale + 1e-9)


def test_input_gradient_matches_finite_differences():
    p, X, _, _ = problem(13, n=3)
    e, G = mo.input_grad(p, X)
    assert np.allclose(e, mo.forward(p, X), rtol=0, atol=1e-13)
    h = 1e-5
    rng = np.random.default_rng(7)
    for i in range(3):
        for d in rng.choice(mo.H, size=10, replace=False):
            Xp, Xm = X.copy(), X.copy()
            Xp[i, d] += h
            Xm[i, d] -= h
            fd = (mo.forward(p, Xp)[i] - mo.forward(p, Xm)[i]) / (2 * h)
            assert G[i, d] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_silu_derivatives():
    z = np.linspace(-8, 8, 41)
    h = 1e-5
    assert np.allclose(mo.dsilu(z), (mo.silu(z + h) - mo.silu(z - h)) / (2 * h), atol=1e-9)
    assert np.allclose(mo.d2silu(z), (mo.dsilu(z + h) - mo.dsilu(z - h)) / (2 * h), atol=1e-8)
    assert mo.silu(0.0) == 0.0 and mo.dsilu(0.0) == 0.5 and mo.d2silu(0.0) == 0.5
