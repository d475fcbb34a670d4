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


_WL_CACHE = {}


def workload_model(name):
    """The benchmark workload's model ("population" for the 512-structure
    batch's), with descriptors from the CPU oracle (cached)."""
    if name in _WL_CACHE:
        return _WL_CACHE[name]
    from gapx import synth

    def desc(header, species, lat, pos, rc):
        t, a, c, z = header["sf"]
        hm = oracle.Model(z=header["z"], w=header["w"], sf_type=t, sf_alpha=a, sf_cut=c,
                          theta=np.ones(2 * len(t)), mm=np.zeros((1, 2 * len(t))), coeff=np.zeros(1),
                          sf_zeta=None if np.all(z == 1) else z)
        return oracle.descriptors(hm, species, lat, pos, rc, nthreads=16)
    d = synth.population_model(desc) if name == "population" else synth.workload_model(name, desc)
    _WL_CACHE[name] = oracle.Model(z=d["z"], w=d["w"], sf_type=d["sf_type"], sf_alpha=d["sf_alpha"], sf_cut=d["sf_cut"],
                        theta=d["theta"], mm=d["mm"], coeff=d["coeff"], sf_zeta=d["sf_zeta"])
    return _WL_CACHE[name]


def lbfgs_two_loop(S, Y, q, alpha):
    """H q for the L-BFGS inverse-Hessian approximation of the pairs (S, Y),
    oldest first, with H0 = 1 / alpha: the two-loop recursion as ASE 3.18.1
    runs it (ase/optimize/lbfgs.py, LBFGS.step; Nocedal and Wright,
    Numerical Optimization, algorithm 7.4)."""
    R = [1.0 / np.dot(y, s) for s, y in zip(S, Y)]
    L = len(S)
    a = np.empty(L)
    q = q.copy()
    for i in range(L - 1, -1, -1):
        a[i] = R[i] * np.dot(S[i], q)
        q -= a[i] * Y[i]
    z = q / alpha
    for i in range(L):
        b = R[i] * np.dot(Y[i], z)
        z += S[i] * (a[i] - b)
    return z


def lbfgs_step_host(xs, fs, memory=100, alpha=70.0, maxstep=0.2):
    """The position ASE 3.18.1's LBFGS (ase/optimize/lbfgs.py: LBFGS.step,
    update, determine_step) moves one structure to after the positions
    xs[0..k] and forces fs[0..k] of its first k + 1 steps: history pairs
    s = x_j - x_{j-1}, y = f_{j-1} - f_j, the last `memory` of them; the step
    -H (-f_k), scaled to maxstep by its longest per-atom displacement."""
    k = len(xs) - 1
    S = [(xs[j] - xs[j - 1]).ravel() for j in range(1, k + 1)][-memory:]
    Y = [(fs[j - 1] - fs[j]).ravel() for j in range(1, k + 1)][-memory:]
    dr = -lbfgs_two_loop(S, Y, -fs[k].ravel(), alpha).reshape(-1, 3)
    longest = np.sqrt((dr ** 2).sum(1)).max()
    if longest >= maxstep:
        dr *= maxstep / longest
    return xs[k] + dr
