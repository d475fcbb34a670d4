"""Batched relaxation on the device at fixed cell: FIRE, and LBFGS as ASE runs
it (the reference's optimiser).

* The FIRE loop equals a host restatement of the same FIRE steps driven by
  the same GAP forces (positions after 25 steps within 1e-9 A).
* Every LBFGS step equals ASE's step from the same positions and forces
  (within 1e-9 A), with the default memory and with a memory of 4.
* Relaxing a population with FIRE lowers every structure's energy, and every
  structure reported converged has max |F_i| <= fmax at its final positions.
"""
import numpy as np
import pytest

import gapx
from gapx import synth
from helpers import lbfgs_step_host, workload_model

pytestmark = pytest.mark.gpu

DT0, DTMAX, NMIN, FINC, FDEC, ASTART, FA = 0.1, 1.0, 5, 1.1, 0.5, 0.1, 0.99


@pytest.fixture(scope="module")
def setup():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    m = gapx.Model.from_model(workload_model("ternary4k"))
    structs = synth.random_ternary(12, nmin=16, nmax=40, seed=77)
    yield m, structs
    m.close()


def fire_host(m, structs, flags, steps, fmax, maxstep):
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    pos = np.concatenate([s["pos"] for s in structs]).copy()
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    v = np.zeros_like(pos)
    st = [dict(dt=DT0, a=ASTART, n=0, started=False, conv=False) for _ in structs]
    for _ in range(steps):
        ctx.update(pos)
        ctx.run(flags)
        F = ctx.results()["forces"]
        for k, S in enumerate(st):
            sl = slice(offs[k], offs[k + 1])
            f, vk = F[sl], v[sl]
            if S["conv"] or np.sqrt((f * f).sum(1)).max() <= fmax:
                S["conv"] = True
                continue
            vf, ff, vv = (f * vk).sum(), (f * f).sum(), (vk * vk).sum()
            cv, cf = 1.0, 0.0
            if S["started"]:
                if vf > 0:
                    cv, cf = 1.0 - S["a"], S["a"] * np.sqrt(vv / ff)
                    if S["n"] > NMIN:
                        S["dt"] = min(S["dt"] * FINC, DTMAX)
                        S["a"] *= FA
                    S["n"] += 1
                else:
                    cv, cf = 0.0, 0.0
                    S["a"], S["dt"], S["n"] = ASTART, S["dt"] * FDEC, 0
            S["started"] = True
            cf += S["dt"]
            vn = cv * vk + cf * f
            dr = S["dt"] * np.sqrt((vn * vn).sum())
            step = S["dt"] * maxstep / dr if dr > maxstep else S["dt"]
            v[sl] = vn
            pos[sl] += step * vn
    ctx.close()
    return pos


def test_device_fire_equals_host_fire(setup):
    m, structs = setup
    ref = fire_host(m, structs, gapx.GRAD, 25, 1e-6, 0.2)
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    out = ctx.relax(gapx.GRAD, max_steps=25, fmax=1e-6, maxstep=0.2)
    assert out["steps"] == 25
    d = np.abs(ctx.positions() - ref).max()
    print("device vs host FIRE after 25 steps: max |dx| %.2e A" % d)
    assert d <= 1e-9
    ctx.close()


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_relaxation_lowers_energy_and_converges(setup, mixed):
    m, structs = setup
    flags = gapx.GRAD | (gapx.MIXED if mixed else 0)
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    ctx.run(flags)
    e0 = ctx.results()["energy"].copy()
    ctx.use_graph(True)
    out = ctx.relax(flags, max_steps=300, fmax=0.05, maxstep=0.2)
    r = ctx.results()
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    fm = np.array([np.sqrt((r["forces"][offs[k]:offs[k + 1]] ** 2).sum(1)).max() for k in range(len(structs))])
    print("relax (%s): %d steps, %d/%d converged, dE/structure min %.3e max %.3e eV, max|F| %.3e"
          % ("mixed" if mixed else "fp64", out["steps"], out["converged"], len(structs),
             (r["energy"] - e0).min(), (r["energy"] - e0).max(), fm.max()))
    assert np.all(r["energy"] < e0)
    assert out["converged"] >= len(structs) // 2
    assert np.sum(fm <= 0.05) >= out["converged"]
    ctx.close()


def lbfgs_trajectory(m, structs, steps, memory):
    """Device positions after 0, 1, .., steps LBFGS steps (a fresh context per
    length: every run is bit for bit the same, so these are one trajectory),
    and the GPU forces at each of them."""
    xs, fs = [], []
    for k in range(steps + 1):
        ctx = gapx.Context(m)
        ctx.load(structs, 6.0)
        if k:
            ctx.relax(gapx.GRAD, max_steps=k, fmax=1e-6, optimizer="lbfgs", memory=memory)
        else:
            ctx.run(gapx.GRAD)
        xs.append(ctx.positions())
        fs.append(ctx.results()["forces"].copy())
        ctx.close()
    return xs, fs


def check_lbfgs_steps(m, structs, steps, memory):
    xs, fs = lbfgs_trajectory(m, structs, steps, memory)
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    worst = 0.0
    for k in range(steps):
        for j in range(len(structs)):
            sl = slice(offs[j], offs[j + 1])
            want = lbfgs_step_host([x[sl] for x in xs[:k + 1]], [f[sl] for f in fs[:k + 1]], memory=memory)
            worst = max(worst, float(np.abs(xs[k + 1][sl] - want).max()))
    return worst


def test_device_lbfgs_steps_equal_ase_steps(setup):
    """Every one of 25 device LBFGS steps moves every structure where ASE's
    LBFGS (the reference's optimiser, memory 100) moves it from the same
    history of positions and forces: within 1e-9 A.  The steps are compared
    one at a time along the device's own trajectory, because on this
    synthetic energy surface most updates have y . s < 0 (negative
    curvature, which ASE's LBFGS does not guard against) and two whole
    trajectories that differ in the last bit drift apart."""
    m, structs = setup
    worst = check_lbfgs_steps(m, structs, 25, 100)
    print("device vs ASE LBFGS step, worst over 25 steps: %.2e A" % worst)
    assert worst <= 1e-9


def test_lbfgs_history_beyond_its_memory(setup):
    """With a history of 4 steps the ring buffer wraps several times in 25
    steps; every device step still equals ASE's LBFGS with memory = 4."""
    m, structs = setup
    worst = check_lbfgs_steps(m, structs[:4], 25, 4)
    print("memory 4: worst step difference %.2e A" % worst)
    assert worst <= 1e-9


def test_lbfgs_and_fire_relax_a_batch(setup):
    """Both device optimisers run the same batch for up to 300 steps at
    fmax = 0.05 eV/A: FIRE lowers every energy; LBFGS (the reference's
    choice) stops every structure it reports converged at max |F_i| < fmax.
    On this synthetic surface ASE's LBFGS itself climbs to higher energies
    on structures whose updates have negative curvature (measured on the
    CPU with the oracle's forces: 6 of these 12 structures end above their
    start), so the energy is not asserted for LBFGS."""
    m, structs = setup
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    res = {}
    for opt in ("fire", "lbfgs"):
        ctx = gapx.Context(m)
        ctx.load(structs, 6.0)
        ctx.run(gapx.GRAD)
        e0 = ctx.results()["energy"].copy()
        ctx.use_graph(True)
        out = ctx.relax(gapx.GRAD, max_steps=300, fmax=0.05, optimizer=opt)
        r = ctx.results()
        fm = np.array([np.sqrt((r["forces"][offs[k]:offs[k + 1]] ** 2).sum(1)).max() for k in range(len(structs))])
        res[opt] = dict(out=out, e=r["energy"].copy(), e0=e0, fm=fm)
        ctx.close()
        print("%s: %s, energies %s -> %s" % (opt, out, np.round(e0, 4), np.round(r["energy"], 4)))
    assert np.all(res["fire"]["e"] < res["fire"]["e0"])
    lb = res["lbfgs"]
    assert lb["out"]["steps"] <= 300 and lb["out"]["converged"] >= 1
    assert np.sum(lb["fm"] < 0.05) >= lb["out"]["converged"]
