"""MLP surrogate on the GPU against the CPU FP64 oracle (oracle/mlp_oracle.py),
its training on GPU GAP labels, and the pre-screen it serves.

Bars (the MLP runs in FP32 with BF16 MFMA products, or hi + lo BF16 products
with `split`):
* forward e and input gradient G: split 2e-5 of the largest value; BF16 3e-2;
* loss gradient: relative L2 error per parameter block, split 1e-3, BF16 6e-2;
* surrogate forces through the descriptor backward pass (FP64 descriptors,
  split MLP) against central differences of the oracle surrogate energy
  (FP64 descriptors from the C oracle, FP64 MLP): 1e-4 eV/A + 1e-4 relative;
* training on GAP labels of 600 structures: held-out per-atom energy error
  below half the label variance, and a pre-screen whose top 5% holds at least
  40% of the GAP's top 5% (a random pick holds 5%), Spearman >= 0.6.
"""
import numpy as np
import pytest

import gapx
import mlp_oracle as mo
import oracle
from gapx import synth
from helpers import workload_model

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", 0, mo.H * mo.H), ("b1", mo.H * mo.H, mo.H), ("W2", mo.H * mo.H + mo.H, mo.H * mo.H),
          ("b2", 2 * mo.H * mo.H + mo.H, mo.H), ("w3", 2 * mo.H * mo.H + 2 * mo.H, mo.H),
          ("b3", 2 * mo.H * mo.H + 3 * mo.H, 1)]


@pytest.fixture(scope="module", autouse=True)
def gpu():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")


def problem(seed, n):
    rng = np.random.default_rng(seed)
    p = mo.init_params(seed).astype(np.float32).astype(np.float64)
    X = rng.normal(0, 1.0, (n, mo.H)).astype(np.float32).astype(np.float64)
    et = rng.normal(0, 1.0, n).astype(np.float32).astype(np.float64)
    Gt = rng.normal(0, 0.05, (n, mo.H)).astype(np.float32).astype(np.float64)
    return p, X, et, Gt


@pytest.mark.parametrize("split,tol", [(True, 2e-5), (False, 3e-2)], ids=["split", "bf16"])
def test_forward_and_input_gradient_match_oracle(split, tol):
    p, X, _, _ = problem(21, 200)  # 200 rows: padded to 256 inside
    net = gapx.MLP(seed=1)
    net.set_params(p)
    e, G = net.forward(X, grad=True, split=split)
    eo, Go = mo.input_grad(p, X)
    de = np.abs(e - eo).max() / np.abs(eo).max()
    dg = np.abs(G - Go).max() / np.abs(Go).max()
    print("forward rel err %.2e, input gradient rel err %.2e" % (de, dg))
    assert de <= tol and dg <= tol
    assert np.array_equal(net.forward(X, grad=False, split=split), e)
    net.close()


@pytest.mark.parametrize("split,tol", [(True, 1e-3), (False, 6e-2)], ids=["split", "bf16"])
@pytest.mark.parametrize("cE,cG", [(1.0, 0.0), (0.0, 1.0), (1.0, 2.0)])
def test_loss_gradient_matches_oracle(split, tol, cE, cG):
    p, X, et, Gt = problem(22, 300)
    net = gapx.MLP(seed=2)
    net.set_params(p)
    L, g = net.loss_grad(X, et, Gt, cE, cG, split=split)
    Lo, go = mo.loss_grad(p, X, et, Gt, cE, cG)
    assert L == pytest.approx(Lo, rel=max(tol, 1e-5))
    for name, o, n in BLOCKS:
        ref = go[o:o + n]
        if np.abs(ref).max() == 0:
            continue
        err = np.linalg.norm(g[o:o + n] - ref) / max(np.linalg.norm(ref), 1e-12)
        print("%s rel L2 err %.2e" % (name, err))
        assert err <= tol, name
    net.close()


def test_surrogate_forces_match_oracle_finite_differences():
    om = workload_model("ternary4k")
    m = gapx.Model.from_model(om)
    theta = np.asarray(om.theta)
    mu = mo.column_mu(om.mm, theta)
    p = mo.init_params(23).astype(np.float32).astype(np.float64)
    net = gapx.MLP(seed=3)
    net.set_params(p)
    s = synth.jittered_structures(1, 24, 24, 0.07, (1, 8, 22), seed=31)[0]
    ctx = gapx.Context(m)
    ctx.load([s], 6.0)
    ctx.run_surrogate(net, gapx.GRAD | gapx.MLP_EXACT)
    r = ctx.results()

    def energy(pos):
        x = oracle.descriptors(om, s["species"], s["lat"], pos, 6.0)
        return float(mo.forward(p, mo.standardise(x, theta, mu)).sum())

    e0 = energy(s["pos"])
    assert abs(r["energy"][0] - e0) <= 1e-4 * max(1.0, abs(e0))
    h = 1e-4
    fd = np.zeros_like(s["pos"])
    for i in range(len(s["pos"])):
        for c in range(3):
            pp, pm = s["pos"].copy(), s["pos"].copy()
            pp[i, c] += h
            pm[i, c] -= h
            fd[i, c] = -(energy(pp) - energy(pm)) / (2 * h)
    err = np.abs(r["forces"] - fd).max()
    print("surrogate forces vs oracle FD: max err %.2e (max |F| %.2e)" % (err, np.abs(fd).max()))
    assert err <= 1e-4 + 1e-4 * np.abs(fd).max()
    assert np.abs(r["forces"].sum(axis=0)).max() <= 1e-5
    ctx.close()
    net.close()
    m.close()


def label(m, structs, net, theta, flags=gapx.GRAD | gapx.MIXED, chunk=200):
    """GAP labels into net's training set; returns (per-atom E, G = dE/dxt) on the host too."""
    ctx = gapx.Context(m)
    es, gs = [], []
    for k in range(0, len(structs), chunk):
        ctx.load(structs[k:k + chunk], 6.0)
        ctx.run(flags)
        ctx.export_labels(net)
        es.append(ctx.results()["eatom"])
        gs.append(ctx.features(dedx=True)[1] * theta)
    ctx.close()
    return np.concatenate(es), np.concatenate(gs)


def test_training_on_gap_labels_and_prescreen():
    """Train on GPU GAP labels (per-atom energies and dE/dxt) of structures
    drawn like the model's descriptor pool, then screen candidates: rank by
    surrogate energy per atom, judge the top 5% with the GAP.  The gain over
    screening without the surrogate (a random pick: 5% recall) is the recall
    of the GAP's own top 5%."""
    om = workload_model("ternary4k")
    m = gapx.Model.from_model(om)
    theta = np.asarray(om.theta)
    train = synth.random_ternary(600, seed=41, workers=8)
    net = gapx.MLP(seed=4)
    e_lab, g_lab = label(m, train, net, theta)
    assert net.data_count() == len(e_lab) == sum(len(s["species"]) for s in train)
    cE = 1.0 / max(np.var(e_lab), 1e-12)
    cG = 1.0 / max(np.mean(np.sum(g_lab ** 2, axis=1)), 1e-12)
    before = net.train(1, batch=4096, lr=0.0, cE=cE, cG=cG)
    after = None
    for k in range(8):
        after = net.train(250, batch=4096, lr=1e-3, cE=cE, cG=cG, seed=k)
    print("labels: %d atoms, var(E) %.3e, mean|G|^2 %.3e" % (len(e_lab), np.var(e_lab), 1.0 / cG))
    print("training: mse(E) %.3e -> %.3e, mse(G) %.3e -> %.3e" % (before[0], after[0], before[1], after[1]))
    # held-out atoms: the fraction of the per-atom energy variance left
    test = synth.random_ternary(100, seed=43, workers=8)
    probe = gapx.MLP(seed=5)
    e_t, _ = label(m, test, probe, theta)
    probe.close()
    ct = gapx.Context(m)
    ct.load(test, 6.0)
    ct.run_surrogate(net, gapx.MIXED)
    e_hat = ct.results()["eatom"]
    ct.close()
    resid = float(np.mean((e_hat - e_t) ** 2) / np.var(e_t))
    print("held-out per-atom mse(E) / var(E) = %.3f" % resid)
    assert after[0] < 0.5 * np.var(e_lab) and resid < 0.5

    cands = synth.random_ternary(400, seed=42, workers=8)
    n = np.array([len(s["species"]) for s in cands])
    cg = gapx.Context(m)
    cg.load(cands, 6.0)
    cg.run(gapx.MIXED)
    e_gap = cg.results()["energy"] / n
    cg.run_surrogate(net, gapx.MIXED)
    e_sur = cg.results()["energy"] / n
    cg.close()
    k = max(1, int(0.05 * len(cands)))
    top_gap = set(np.argsort(e_gap)[:k])
    top_sur = set(np.argsort(e_sur)[:k])
    recall = len(top_gap & top_sur) / k
    rank = lambda v: np.argsort(np.argsort(v))
    spearman = float(np.corrcoef(rank(e_gap), rank(e_sur))[0, 1])
    err = np.abs(e_sur - e_gap)
    print("pre-screen: recall of GAP top-5%% %.2f (random 0.05), Spearman %.2f, |dE|/atom mean %.2e max %.2e eV,"
          " spread of GAP E/atom %.2e" % (recall, spearman, err.mean(), err.max(), e_gap.std()))
    assert recall >= 0.4 and spearman >= 0.6
    net.close()
    m.close()
