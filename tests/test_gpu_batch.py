"""Batched population path: uneven periodic structures packed into one launch
per stage, graph capture and replay, and the autotuner.

Bars: the batched FP64 results equal the single-structure FP64 results to
1e-10 eV/atom, 1e-9 eV/A and 1e-8 GPa (only the contraction's split of the
sparse points, and so its summation order, depends on the batch); graph
replays equal the eager run bit for bit (same kernels, same order); the
mixed batched path is within 1 meV/atom and 20 meV/A of FP64.
"""
import numpy as np
import pytest

import gapx
from gapx import synth
from helpers import ternary_structure

pytestmark = pytest.mark.gpu


def gpu_descriptors(header, species, lat, pos, rc):
    t, a, c, z = header["sf"]
    D = 2 * len(t)
    m = gapx.Model.from_arrays(header["z"], header["w"], t, a, c, np.ones(D), np.zeros((1, D)), np.zeros(1), z)
    x = m.descriptors(species, lat, pos, rc)
    m.close()
    return x


@pytest.fixture(scope="module")
def pop():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    model = synth.population_model(gpu_descriptors)
    structs = synth.population(40, seed=3004)
    m = gapx.Model.from_model(model)
    yield m, structs
    m.close()


def evaluate(m, structs, flags, graph=False, reps=1, tune=None):
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    if tune is not None:
        ctx.set_tuning(flags, **tune)
    ctx.use_graph(graph)
    out = []
    for _ in range(reps):
        ctx.run(flags)
        out.append(ctx.results())
    ctx.close()
    return out


def split_forces(r, structs):
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    return [r["forces"][offs[k]:offs[k + 1]] for k in range(len(structs))]


def test_population_batch_equals_single_structures(pop):
    m, structs = pop
    r = evaluate(m, structs, gapx.GRAD)[0]
    fs = split_forces(r, structs)
    worst = [0.0, 0.0, 0.0]
    for k, s in enumerate(structs):
        e, f, st, _ = m.calc(s["species"], s["lat"], s["pos"], 6.0)
        n = len(s["species"])
        worst[0] = max(worst[0], abs(r["energy"][k] - e) / n)
        worst[1] = max(worst[1], float(np.abs(fs[k] - f).max()))
        worst[2] = max(worst[2], float(np.abs(r["stress"][k] - st).max()))
    print("batch vs single: dE/atom %.2e  dF %.2e  dS %.2e" % tuple(worst))
    assert worst[0] <= 1e-10 and worst[1] <= 1e-9 and worst[2] <= 1e-8


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_graph_replays_match_eager(pop, mixed):
    m, structs = pop
    flags = gapx.GRAD | (gapx.MIXED if mixed else 0)
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.run(flags)
    eager = ctx.results()
    ctx.use_graph(True)
    rng = np.random.default_rng(5)
    for rep in range(4):
        ctx.run(flags)
        r = ctx.results()
        for key in ("energy", "forces", "stress"):
            assert np.array_equal(r[key], eager[key]), (rep, key)
        # eager work on the same stream between replays
        x, _ = ctx.features(dedx=False)
        assert np.isfinite(x).all()
    # new positions through the same graph
    pos = np.concatenate([s["pos"] for s in structs])
    pos = pos + rng.normal(0, 0.02, size=pos.shape)
    ctx.update(pos)
    ctx.run(flags)
    rg = ctx.results()
    ctx.use_graph(False)
    ctx.run(flags)
    re = ctx.results()
    for key in ("energy", "forces", "stress"):
        assert np.array_equal(rg[key], re[key]), key
    assert not np.array_equal(rg["energy"], eager["energy"])
    ctx.close()


def test_autotune_keeps_results(pop):
    m, structs = pop
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.run(gapx.GRAD)
    ref = ctx.results()
    for flags in (gapx.GRAD, gapx.GRAD | gapx.MIXED):
        t = ctx.autotune(flags, force=True)
        assert t["waves_fwd"] in (1, 2, 4) and t["waves_bwd"] in (1, 2, 4)
        assert t["gpr_rows"] in (64, 128) and t["gpr_blocks"] >= 1 and t["halves"] in (0, 1)
        print("tuned", "mixed" if flags & gapx.MIXED else "fp64", t)
    ctx.run(gapx.GRAD)
    r = ctx.results()
    n = np.array([len(s["species"]) for s in structs])
    assert np.abs((r["energy"] - ref["energy"]) / n).max() <= 1e-10
    assert np.abs(r["forces"] - ref["forces"]).max() <= 1e-9
    with pytest.raises(gapx.GapxError):
        ctx.set_tuning(gapx.GRAD, waves_fwd=3)
    ctx.close()


@pytest.mark.parametrize("bf16", [False, True], ids=["fp16x3", "bf16x3"])
def test_population_mixed_vs_fp64(pop, bf16):
    m, structs = pop
    r64 = evaluate(m, structs, gapx.GRAD)[0]
    flags = gapx.GRAD | gapx.MIXED | (gapx.BF16 if bf16 else 0)
    rmx = evaluate(m, structs, flags, graph=True, reps=2)[-1]
    n = np.array([len(s["species"]) for s in structs])
    de = float(np.abs((rmx["energy"] - r64["energy"]) / n).max())
    df = float(np.abs(rmx["forces"] - r64["forces"]).max())
    ds = float(np.abs(rmx["stress"] - r64["stress"]).max())
    print("population mixed(%s) vs fp64: dE/atom %.3e eV  dF %.3e eV/A  dS %.3e GPa"
          % ("bf16" if bf16 else "fp16", de, df, ds))
    assert de <= 1e-3 and df <= 2e-2 and ds <= 5e-2


def test_reloads_equal_fresh_contexts(pop):
    """One context reloaded with batches of other sizes (more and fewer bins,
    the one-launch and the multi-launch binning) gives bit for bit what a
    fresh context gives: the binning leaves its counters cleared for the next
    run."""
    m, structs = pop
    big = [ternary_structure(1100, 11.5, 71), ternary_structure(1500, 14.0, 72)]  # above the one-launch limit
    batches = [big[:1], structs[:10], big, big[1:], structs[10:30], big[:1]]
    ctx = gapx.Context(m)
    for k, batch in enumerate(batches):
        ctx.load(batch, 6.0)
        for _ in range(2):
            ctx.run(gapx.GRAD)
        r = ctx.results()
        f = evaluate(m, batch, gapx.GRAD)[0]
        assert np.array_equal(r["energy"], f["energy"]) and np.array_equal(r["forces"], f["forces"]), k
        assert np.array_equal(r["stress"], f["stress"]), k
    ctx.close()


@pytest.mark.parametrize("flags", [gapx.GRAD, gapx.GRAD | gapx.MIXED], ids=["fp64", "mixed"])
def test_two_half_schedule_equals_one_piece(pop, flags):
    """Batches of at least 16 384 atoms run as two halves on two streams; the
    results equal the one-piece schedule (used while per-stage timing is on)
    bit for bit, with graph replay as well."""
    m, _ = pop
    structs = synth.population(160, seed=3005)
    assert sum(len(s["species"]) for s in structs) >= 16384
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.timing(True)
    ctx.run(flags)
    one = ctx.results()
    ctx.timing(False)
    ctx.run(flags)
    two = ctx.results()
    ctx.use_graph(True)
    for _ in range(2):
        ctx.run(flags)
    three = ctx.results()
    t = ctx.tuning(flags)
    t["halves"] = 0  # the autotuner's other choice: one piece without per-stage timing
    ctx.set_tuning(flags, **t)
    ctx.run(flags)
    four = ctx.results()
    for r in (two, three, four):
        for k in ("energy", "forces", "stress", "eatom"):
            assert np.array_equal(one[k], r[k]), k
    ctx.close()


def test_one_launch_binning_equals_three_launches(pop, monkeypatch):
    """Batches of structures of up to 1024 atoms are binned by one workgroup
    per structure in one launch; the three-launch path (wrap + count + scan,
    scatter, rank) gives the same neighbour order, hence bitwise-equal
    results."""
    m, _ = pop
    structs = [ternary_structure(1024, 1.0 / 0.07, 3101), ternary_structure(700, 1.0 / 0.07, 3102)]
    structs += synth.population(6, seed=3103)
    out = []
    for off in (False, True):
        if off:
            monkeypatch.setenv("GAPX_NO_FUSED_BINS", "1")
        ctx = gapx.Context(m)
        ctx.load(structs, 6.0)
        ctx.run(gapx.GRAD)
        out.append(ctx.results())
        pairs, _ = ctx.counts()
        ctx.close()
    monkeypatch.delenv("GAPX_NO_FUSED_BINS")
    assert pairs > 30 * sum(len(s["species"]) for s in structs)
    for k in ("energy", "forces", "stress", "eatom"):
        assert np.array_equal(out[0][k], out[1][k]), k


def test_lattice_update_with_same_bin_total_recaptures(pop):
    """A captured graph fixes the binning's launch parameters (which kernel,
    its LDS sizes, the bin count).  Two structures whose bin grids change from
    4x4x2 + 2x4x5 (32 + 40 bins) to 4x4x4 + 2x2x2 (64 + 8) keep the total at
    72 while the largest structure's bin count rises from 40 to 64; with graph
    replay on, a lattice update must not replay the old launch.  The updated
    context equals an eager run loaded with the new cells bit for bit."""
    m, _ = pop
    a2 = ternary_structure(40, 1728.0 / 40, 3201)   # 12 x 12 x 12 cell after the update
    b2 = ternary_structure(20, 216.0 / 20, 3202)    # 6 x 6 x 6
    a2["lat"], b2["lat"] = np.eye(3) * 12.0, np.eye(3) * 6.0
    a2["pos"] = synth.random_positions(a2["lat"], 40, np.random.default_rng(3203))
    b2["pos"] = synth.random_positions(b2["lat"], 20, np.random.default_rng(3204))
    lat1 = [np.diag([12.0, 12.0, 6.0]), np.diag([6.0, 12.0, 15.0])]   # 4x4x2 and 2x4x5 bins
    first = []
    for s, l1 in zip((a2, b2), lat1):
        frac = s["pos"] @ np.linalg.inv(s["lat"])
        first.append(dict(species=s["species"], lat=l1, pos=frac @ l1))
    flags = gapx.GRAD
    ctx = gapx.Context(m)
    ctx.load(first, 6.0)
    ctx.use_graph(True)
    for _ in range(2):
        ctx.run(flags)   # eager run, capture, replay
    ctx.results()
    ctx.update(np.concatenate([a2["pos"], b2["pos"]]), lat=np.stack([a2["lat"], b2["lat"]]))
    for _ in range(2):
        ctx.run(flags)
    rg = ctx.results()
    ctx.close()
    re = evaluate(m, [a2, b2], flags)[0]
    for k in ("energy", "forces", "stress"):
        assert np.array_equal(rg[k], re[k]), k
