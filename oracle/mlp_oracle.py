"""CPU FP64 oracle of the per-atom MLP surrogate (test infrastructure only).

The surrogate has no counterpart in the reference; this module restates its
mathematics in numpy so that the GPU kernels (calypso-gap_amd/csrc/mlp.hip)
can be checked:

    z1 = xt W1 + b1,  a1 = s(z1),  z2 = a1 W2 + b2,  a2 = s(z2),  e = a2 w3 + b3
    G  = de/dxt = ((s'(z2) o w3) W2^T o s'(z1)) W1^T
    L  = cE/(2n) sum_i (e_i - et_i)^2 + cG/(2n) sum_i |G_i - Gt_i|^2

with SiLU s(z) = z sigma(z).  `loss_grad` is the hand-derived reverse pass,
including the second-order terms that the G loss needs.  It is pinned in
tests/test_oracle_mlp.py against central finite differences of `loss` (every
parameter block, FP64) and against a planted mistake; `input_grad` is pinned
against finite differences of `forward`.

Parameter layout (flat, as the GPU library): W1[128][128] (row = input
feature), b1[128], W2[128][128], b2[128], w3[128], b3.
"""
import numpy as np

H = 128
N_PARAMS = H * H + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + H * H].reshape(H, H); o += H * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]])


def sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sig(z)


def dsilu(z):
    s = sig(z)
    return s * (1.0 + z * (1.0 - s))


def d2silu(z):
    s = sig(z)
    return s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


def init_params(seed):
    """A random parameter vector of the same scale as the library's init."""
    rng = np.random.default_rng(seed)
    s1 = np.sqrt(2.0 / H)
    return pack(rng.normal(0, s1, (H, H)), rng.normal(0, 0.1, H), rng.normal(0, s1, (H, H)),
                rng.normal(0, 0.1, H), rng.normal(0, 0.1, H), 0.3)


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = X @ W1 + b1
    a1 = silu(z1)
    z2 = a1 @ W2 + b2
    a2 = silu(z2)
    return a2 @ w3 + b3


def input_grad(p, X):
    """(e, G = de/dX)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    z1 = X @ W1 + b1
    a1 = silu(z1)
    z2 = a1 @ W2 + b2
    a2 = silu(z2)
    e = a2 @ w3 + b3
    d2 = dsilu(z2) * w3
    t1 = d2 @ W2.T
    d1 = dsilu(z1) * t1
    return e, d1 @ W1.T


def loss(p, X, et, Gt, cE, cG):
    n = len(X)
    e, G = input_grad(p, X)
    return 0.5 * cE / n * np.sum((e - et) ** 2) + 0.5 * cG / n * np.sum((G - Gt) ** 2)


def loss_grad(p, X, et, Gt, cE, cG, drop_second_order=False):
    """(L, dL/dp).  drop_second_order removes the s'' terms (a planted
    mistake used by the tests to show that the checks catch it)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = len(X)
    z1 = X @ W1 + b1
    a1 = silu(z1)
    z2 = a1 @ W2 + b2
    a2 = silu(z2)
    e = a2 @ w3 + b3
    d2 = dsilu(z2) * w3
    t1 = d2 @ W2.T
    d1 = dsilu(z1) * t1
    G = d1 @ W1.T
    L = 0.5 * cE / n * np.sum((e - et) ** 2) + 0.5 * cG / n * np.sum((G - Gt) ** 2)
    de = cE * (e - et) / n
    R = cG * (G - Gt) / n
    s2z1 = 0.0 if drop_second_order else d2silu(z1)
    s2z2 = 0.0 if drop_second_order else d2silu(z2)
    # through G = d1 W1^T
    gW1 = R.T @ d1
    U1 = R @ W1                      # dL/dd1
    dT1 = U1 * dsilu(z1)             # dL/dt1
    # through t1 = d2 W2^T
    gW2 = dT1.T @ d2
    U2 = dT1 @ W2                    # dL/dd2
    gw3 = np.sum(U2 * dsilu(z2), axis=0) + a2.T @ de
    gb3 = np.sum(de)
    q2 = U2 * w3 * s2z2 + np.outer(de, w3) * dsilu(z2)   # dL/dz2
    gW2 = gW2 + a1.T @ q2
    gb2 = np.sum(q2, axis=0)
    q1 = (q2 @ W2.T) * dsilu(z1) + U1 * t1 * s2z1         # dL/dz1
    gW1 = gW1 + X.T @ q1
    gb1 = np.sum(q1, axis=0)
    return L, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def standardise(x, theta, mu):
    """xt = x / theta - mu, the surrogate's input (the GAP model's scaling)."""
    return np.asarray(x) / np.asarray(theta) - np.asarray(mu)


def column_mu(mm, theta):
    """mu = column mean of MM / theta (the GAP path's shift)."""
    return (np.asarray(mm) / np.asarray(theta)).mean(axis=0)
