"""Batched relaxation on the device (FIRE at fixed cell).

* The device loop equals a host restatement of the same FIRE steps driven by
  the same GAP forces (positions after 25 steps within 1e-9 A).
* Relaxing a population lowers every structure's energy, and every structure
  reported converged has max |F_i| <= fmax at its final positions.
"""
import numpy as np
import pytest

import gapx
from gapx import synth
from helpers import workload_model

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
