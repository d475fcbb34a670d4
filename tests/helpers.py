"""Small models and structures shared by the tests."""
import numpy as np

import oracle


def small_model(rng, zeta=(1, 2, 4), M=12, nrad=4, theta_scale=1.0, cuts=(6.0, 4.5),
                z=(1, 8, 22), w=(1.0, -0.5, 2.0)):
    types, alphas, cutl, zl = [], [], [], []
    for k in range(nrad):
        types.append(1)
        alphas.append(0.02 * 8 ** k)
        cutl.append(cuts[k % len(cuts)])
        zl.append(1.0)
    types += [3, 3]
    alphas += [1.5, 3.0]
    cutl += [6.0, 5.0]
    zl += [1.0, 1.0]
    for eta in (0.01, 0.1):
        for t in (2, 4):
            for zz in zeta:
                types.append(t)
                alphas.append(eta)
                cutl.append(cuts[0])
                zl.append(float(zz))
    nsf = len(types)
    D = 2 * nsf
    xs = rng.normal(size=(M, D))
    m = oracle.Model(z=np.array(z, np.int32), w=np.array(w, float),
                     sf_type=np.array(types, np.int32), sf_alpha=np.array(alphas), sf_cut=np.array(cutl),
                     theta=np.ones(D), mm=np.zeros((M, D)), coeff=rng.normal(0, 0.5, M),
                     sf_zeta=np.array(zl) if any(v != 1 for v in zl) else None)
    return m, xs


def fitted_small_model(rng, s, **kw):
    """small_model with theta/MM taken from the structure's own descriptors so
    that kernel values are O(1)."""
    m, _ = small_model(rng, **kw)
    x = oracle.descriptors(m, s["species"], s["lat"], s["pos"], 6.0)
    m.theta = x.std(axis=0) * 1.5 + 1e-3
    idx = rng.choice(len(x), size=m.nsparse, replace=len(x) < m.nsparse)
    m.mm = x[idx] + rng.normal(0, 0.2, size=(m.nsparse, m.des_len)) * m.theta
    return m


def ternary_structure(n, vpa, seed):
    from gapx import synth
    return synth.structure(n, n * vpa, (1, 8, 22), np.random.default_rng(seed))
