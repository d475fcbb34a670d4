"""Batched NVT MD on the device (gapx_ctx_md: ASE NVTBerendsen steps).

* The device loop equals a host restatement of ASE 3.18.1's NVTBerendsen.step
  (Berendsen scaling clipped to [0.9, 1.1], half kick, total-momentum removal,
  drift, forces, half kick) driven by the same GPU forces: positions within
  1e-9 A and momenta within 1e-9 amu A/t after 20 steps, log rows equal.
* With fixcm the total momentum of every structure stays zero.
* Without the thermostat (plain velocity Verlet) the total-energy error is
  second order in dt: halving dt divides the largest excursion of
  E_pot + E_kin from its start by about 4 (bar: between 2.5 and 6).
* The same with a 5 fs taut from rest, where every scale factor is clipped
  (to 1.1 while T << T0, the T = 0 start included).
"""
import numpy as np
import pytest

import gapx
from gapx import synth
from helpers import workload_model

pytestmark = pytest.mark.gpu

MASS = {1: 1.008, 8: 15.999, 22: 47.867}


@pytest.fixture(scope="module")
def setup():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    m = gapx.Model.from_model(workload_model("ternary4k"))
    structs = synth.random_ternary(6, nmin=16, nmax=40, seed=91)
    yield m, structs
    m.close()


def masses_of(structs):
    return np.concatenate([[MASS[int(z)] for z in s["species"]] for s in structs])


def maxwell_boltzmann(masses, T, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(len(masses), 3)) * np.sqrt(masses * T * gapx.KB)[:, None]


def berendsen_host(m, structs, masses, p0, steps, dt, T0, taut, fixcm=True, flags=gapx.GRAD):
    """ASE NVTBerendsen (3.18.1) restated with numpy, forces from GPU runs."""
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    x = np.concatenate([s["pos"] for s in structs]).copy()
    p = p0.copy()
    ctx.run(flags)
    f = ctx.results()["forces"]
    for _ in range(steps):
        for k in range(len(structs)):
            a, b = offs[k], offs[k + 1]
            pk, mk = p[a:b], masses[a:b]
            if taut > 0:
                T = np.sum(pk ** 2 / mk[:, None]) / (3 * (b - a) * gapx.KB)
                with np.errstate(divide="ignore"):
                    s = np.sqrt(1.0 + (T0 / T - 1.0) * dt / taut)
                s = min(max(s, 0.9), 1.1)
            else:
                s = 1.0
            pk = s * pk + 0.5 * dt * f[a:b]
            if fixcm:
                pk = pk - pk.sum(axis=0) / (b - a)
            x[a:b] += dt * pk / mk[:, None]
            p[a:b] = pk
        ctx.update(x)
        ctx.run(flags)
        f = ctx.results()["forces"]
        p = p + 0.5 * dt * f
    ctx.close()
    return x, p


@pytest.mark.parametrize("taut_fs,start", [(500.0, "mb300"), (5.0, "rest")])
def test_md_matches_host_restatement(setup, taut_fs, start):
    m, structs = setup
    masses = masses_of(structs)
    p0 = maxwell_boltzmann(masses, 300.0, 5) if start == "mb300" else np.zeros((len(masses), 3))
    dt, T0, taut = 1.0 * gapx.FS, 300.0, taut_fs * gapx.FS
    x_h, p_h = berendsen_host(m, structs, masses, p0, 20, dt, T0, taut)
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    log = ctx.md(masses, 20, dt_fs=1.0, temperature=T0, taut_fs=taut_fs, momenta=p0 if start == "mb300" else None)
    x_d, p_d = ctx.positions(), ctx.momenta()
    dx, dp = np.abs(x_d - x_h).max(), np.abs(p_d - p_h).max()
    print("device vs host after 20 steps: |dx| %.2e A, |dp| %.2e" % (dx, dp))
    assert dx <= 1e-9 and dp <= 1e-9
    # log: the last row is the energy and kinetic energy at the final state
    r = ctx.results()
    offs = np.concatenate([[0], np.cumsum([len(s["species"]) for s in structs])])
    for k in range(len(structs)):
        a, b = offs[k], offs[k + 1]
        ek = 0.5 * np.sum(p_d[a:b] ** 2 / masses[a:b, None])
        assert log[-1, k, 0] == pytest.approx(r["energy"][k], rel=1e-14, abs=1e-12)
        assert log[-1, k, 1] == pytest.approx(ek, rel=1e-12)
        assert log[-1, k, 2] == pytest.approx(2 * ek / (3 * (b - a) * gapx.KB), rel=1e-12)
        assert np.abs(p_d[a:b].sum(axis=0)).max() <= 1e-10
    ctx.close()


def energy_excursion(m, structs, masses, p0, dt_fs, t_total_fs):
    ctx = gapx.Context(m)
    ctx.load(structs, 6.0)
    log = ctx.md(masses, int(round(t_total_fs / dt_fs)), dt_fs=dt_fs, temperature=0.0, taut_fs=0.0, fixcm=False,
                 momenta=p0)
    ctx.close()
    etot = log[:, :, 0] + log[:, :, 1]
    return np.abs(etot - etot[0]).max(axis=0)


def test_velocity_verlet_energy_error_is_second_order(setup):
    m, structs = setup
    masses = masses_of(structs)
    p0 = maxwell_boltzmann(masses, 300.0, 6)
    e1 = energy_excursion(m, structs, masses, p0, 0.5, 20.0)
    e2 = energy_excursion(m, structs, masses, p0, 0.25, 20.0)
    ratio = e1 / e2
    print("energy excursion dt 0.5 fs %s, dt 0.25 fs %s, ratio %s" % (e1, e2, ratio))
    assert np.all(e1 > 0) and np.all((ratio > 2.5) & (ratio < 6.0))


def converging_cluster(n, radius, box, seed, shrink=0.25, steps=48):
    """n atoms at least 2.4 A apart inside a sphere at the centre of a cubic
    box, with velocities -k (r - c) that would shrink every distance by the
    factor (1 - shrink) in `steps` fs of free flight: the cluster contracts
    and every atom's neighbour count grows step by step."""
    rng = np.random.default_rng(seed)
    c = np.full(3, box / 2)
    pos = []
    while len(pos) < n:
        r = (rng.random(3) * 2 - 1) * radius
        if np.linalg.norm(r) <= radius and all(np.linalg.norm(c + r - q) >= 2.4 for q in pos):
            pos.append(c + r)
    pos = np.array(pos)
    species = np.array([(1, 8, 22)[k % 3] for k in range(n)], np.int32)
    s = dict(species=species, lat=np.eye(3) * box, pos=pos)
    masses = masses_of([s])
    v = -(shrink / (steps * gapx.FS)) * (pos - c)
    return s, masses, masses[:, None] * v


def test_md_outgrowing_its_capacities_matches_host_restatement(setup):
    """A contracting cluster: the neighbour lists outgrow the capacities (and
    their headroom) set at the start, the device notices at a block's check,
    goes back to the block's start and runs it with larger lists, and the
    trajectory equals the host restatement's, which checks every step."""
    m, _ = setup
    s, masses, p0 = converging_cluster(120, 11.0, 40.0, 12)
    dt = 1.0 * gapx.FS
    ctx = gapx.Context(m)
    ctx.load([s], 6.0)
    ctx.run(gapx.GRAD)
    ctx.results()
    pairs0, _ = ctx.counts()
    log = ctx.md(masses, 48, dt_fs=1.0, temperature=0.0, taut_fs=0.0, momenta=p0)
    x_d, p_d = ctx.positions(), ctx.momenta()
    pairs1, _ = ctx.counts()
    ctx.close()
    x_h, p_h = berendsen_host(m, [s], masses, p0, 48, dt, 0.0, 0.0)
    dx, dp = np.abs(x_d - x_h).max(), np.abs(p_d - p_h).max()
    print("contracting cluster: pairs %d -> %d; device vs host |dx| %.2e A, |dp| %.2e" % (pairs0, pairs1, dx, dp))
    assert pairs1 > 1.6 * pairs0
    assert dx <= 1e-9 and dp <= 1e-9
    assert np.all(np.isfinite(log))
