"""The CPU oracle pinned against the reference's own outputs and against
what the mathematics fixes (closed forms, finite differences, invariants,
brute force)."""
import itertools
import math
import os
import tempfile

import numpy as np
import pytest

import oracle
from conftest import CAR_CASE_NAMES, GOLDEN_CASE_NAMES, car_case, golden_cases, golden_model
from helpers import fitted_small_model, ternary_structure

# parity bar of the FP64 build: 1e-9 eV/atom, 1e-7 eV/A, 1e-6 GPa; for the
# stress also 1e-8 relative to the largest component, the FP64 rounding level
# of a virial summed over some 10^4 terms in cells of about 1 A^3 per atom
# (stresses of 1e4-6e5 GPa; measured 0.6e-9 between the two oracle modes)
TOL_E, TOL_F, TOL_S = 1e-9, 1e-7, 1e-6
TOL_S_REL = 1e-8


@pytest.mark.parametrize("mode", ["faithful", "sparse"])
@pytest.mark.parametrize("frame", range(11))
def test_ase_traj_frames(frames, example_model, mode, frame):
    """ase.traj (reference output, 11 MD frames of C64 diamond)."""
    r = oracle.gap_calc(example_model, frames["numbers"], frames["cells"][frame], frames["positions"][frame], 6.0,
                        mode=mode, nthreads=2)
    n = len(frames["numbers"])
    assert abs(r["energy"] - frames["energies"][frame]) / n < TOL_E
    assert np.abs(r["forces"] - frames["forces"][frame]).max() < TOL_F
    assert np.abs(r["stress"] - frames["stresses"][frame]).max() < TOL_S


@pytest.mark.parametrize("mode", ["faithful", "sparse"])
@pytest.mark.parametrize("case", GOLDEN_CASE_NAMES)
def test_reference_vectors(ref_vectors, mode, case):
    """Vectors written by the compiled reference FGAP_CALC (tests/golden)."""
    c = [x for x in golden_cases(ref_vectors) if x["name"] == case][0]
    m = golden_model(ref_vectors, c["model"])
    r = oracle.gap_calc(m, c["species"], c["lat"], c["pos"], c["rcut"], lgrad=c["lgrad"], mode=mode, nthreads=2)
    n = len(c["species"])
    assert abs(r["energy"] - c["energy"]) / n < TOL_E
    assert np.abs(r["forces"] - c["forces"]).max() < TOL_F
    assert np.abs(r["stress"] - c["stress"]).max() < max(TOL_S, TOL_S_REL * np.abs(c["stress"]).max())
    if not c["lgrad"]:
        assert not np.any(r["forces"]) and not np.any(r["stress"])


def test_model_file_roundtrip(example_model):
    m = example_model
    assert (m.nsf, m.des_len, m.nsparse) == (33, 66, 129)
    assert list(m.z) == [5, 6] and list(m.w) == [-1.0, 4.0]
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "gap_parameters")
        oracle.write_model(p, m)
        m2 = oracle.parse_model(p)
    for k in ("z", "w", "sf_type", "sf_alpha", "sf_cut", "theta", "mm", "coeff"):
        np.testing.assert_array_equal(getattr(m, k), getattr(m2, k))


# --------------------------------------------------------------------------
# models and structures for the mathematical checks
# --------------------------------------------------------------------------

# --------------------------------------------------------------------------
# finite differences (forces and stress), including the zeta > 1 extension
# --------------------------------------------------------------------------

@pytest.mark.parametrize("mode", ["faithful", "sparse"])
def test_forces_are_minus_energy_gradient(mode):
    rng = np.random.default_rng(3)
    s = ternary_structure(12, 12.0, 31)
    m = fitted_small_model(rng, s)
    r = oracle.gap_calc(m, s["species"], s["lat"], s["pos"], 6.0, mode=mode)
    h = 1e-5
    for i, c in [(0, 0), (3, 1), (7, 2), (11, 0)]:
        p1 = s["pos"].copy()
        p1[i, c] += h
        p2 = s["pos"].copy()
        p2[i, c] -= h
        e1 = oracle.gap_calc(m, s["species"], s["lat"], p1, 6.0, lgrad=False, mode=mode)["energy"]
        e2 = oracle.gap_calc(m, s["species"], s["lat"], p2, 6.0, lgrad=False, mode=mode)["energy"]
        fd = -(e1 - e2) / (2 * h)
        assert abs(fd - r["forces"][i, c]) < 1e-6 * max(1.0, abs(fd)), (i, c, fd, r["forces"][i, c])


@pytest.mark.parametrize("mode", ["faithful", "sparse"])
def test_stress_is_strain_derivative(mode):
    """libgap stress (GPa, VASP sign): sigma_ac = -(1/V) dE/d(eps_ac) / 6.24219e-3."""
    rng = np.random.default_rng(4)
    s = ternary_structure(10, 13.0, 41)
    m = fitted_small_model(rng, s)
    r = oracle.gap_calc(m, s["species"], s["lat"], s["pos"], 6.0, mode=mode)
    vol = abs(np.linalg.det(s["lat"]))
    h = 1e-6
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]
    for t, (a, c) in enumerate(order):
        es = []
        for sgn in (1, -1):
            F = np.eye(3)
            F[a, c] += sgn * h       # x'_a = x_a + h x_c
            lat = s["lat"] @ F.T
            pos = s["pos"] @ F.T
            es.append(oracle.gap_calc(m, s["species"], lat, pos, 6.0, lgrad=False, mode=mode)["energy"])
        dedeps = (es[0] - es[1]) / (2 * h)
        expect = -dedeps / vol / oracle.GPA
        assert abs(expect - r["stress"][t]) < 1e-5 * max(1.0, abs(expect)), (t, expect, r["stress"][t])


# --------------------------------------------------------------------------
# invariants
# --------------------------------------------------------------------------

def _rot(ax, ang):
    ax = np.asarray(ax, float) / np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def test_invariants_translation_rotation_permutation():
    rng = np.random.default_rng(5)
    s = ternary_structure(14, 12.0, 51)
    m = fitted_small_model(rng, s)
    base = oracle.gap_calc(m, s["species"], s["lat"], s["pos"], 6.0, mode="sparse")
    assert np.abs(base["forces"].sum(axis=0)).max() < 1e-10
    # translation (positions only, cell unchanged)
    t = oracle.gap_calc(m, s["species"], s["lat"], s["pos"] + np.array([0.37, -1.2, 2.9]), 6.0, mode="sparse")
    assert abs(t["energy"] - base["energy"]) < 1e-10
    np.testing.assert_allclose(t["forces"], base["forces"], atol=1e-10)
    # rotation of cell and positions
    R = _rot([1, 2, 3], 0.7)
    r = oracle.gap_calc(m, s["species"], s["lat"] @ R.T, s["pos"] @ R.T, 6.0, mode="sparse")
    assert abs(r["energy"] - base["energy"]) < 1e-10
    np.testing.assert_allclose(r["forces"], base["forces"] @ R.T, atol=1e-9)

    def full(v):
        return np.array([[v[0], v[3], v[5]], [v[3], v[1], v[4]], [v[5], v[4], v[2]]])
    np.testing.assert_allclose(full(r["stress"]), R @ full(base["stress"]) @ R.T, atol=1e-8)
    # permutation
    perm = rng.permutation(len(s["species"]))
    p = oracle.gap_calc(m, s["species"][perm], s["lat"], s["pos"][perm], 6.0, mode="sparse")
    assert abs(p["energy"] - base["energy"]) < 1e-10
    np.testing.assert_allclose(p["forces"], base["forces"][perm], atol=1e-10)


def test_supercell_replication():
    rng = np.random.default_rng(6)
    s = ternary_structure(9, 14.0, 61)
    m = fitted_small_model(rng, s)
    base = oracle.gap_calc(m, s["species"], s["lat"], s["pos"], 6.0, mode="sparse", want_details=True)
    lat2 = s["lat"].copy()
    lat2[0] *= 2
    pos2 = np.concatenate([s["pos"], s["pos"] + s["lat"][0]])
    sp2 = np.concatenate([s["species"], s["species"]])
    big = oracle.gap_calc(m, sp2, lat2, pos2, 6.0, mode="sparse", want_details=True)
    assert abs(big["energy"] - 2 * base["energy"]) < 1e-9
    np.testing.assert_allclose(big["x"][:9], base["x"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(big["forces"][9:], base["forces"], atol=1e-9)
    np.testing.assert_allclose(big["stress"], base["stress"], atol=1e-8)


# --------------------------------------------------------------------------
# closed forms
# --------------------------------------------------------------------------

def _fc(r, rc):
    return 0.5 * (math.cos(oracle.PI_TRUNC * r / rc) + 1)


def _iso_model(types, alphas, cuts, zetas=None):
    nsf = len(types)
    return oracle.Model(z=np.array([1, 8], np.int32), w=np.array([1.0, 3.0]), sf_type=np.array(types, np.int32),
                        sf_alpha=np.array(alphas, float), sf_cut=np.array(cuts, float), theta=np.ones(2 * nsf),
                        mm=np.zeros((1, 2 * nsf)), coeff=np.ones(1),
                        sf_zeta=None if zetas is None else np.array(zetas, float))


def test_dimer_radial_closed_form():
    lat = np.eye(3) * 40.0
    r = 2.3
    pos = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]])
    sp = np.array([1, 8], np.int32)
    m = _iso_model([1, 3], [0.3, 2.0], [6.0, 5.0])
    x = oracle.descriptors(m, sp, lat, pos, 6.0)
    g2 = math.exp(-0.3 * r * r) * _fc(r, 6.0)
    g3 = math.exp(-4.0 * (r - 2.0) ** 2) * _fc(r, 5.0)
    np.testing.assert_allclose(x[0], [g2, g3, 3.0 * g2, 3.0 * g3], rtol=1e-14)
    np.testing.assert_allclose(x[1], [g2, g3, 1.0 * g2, 1.0 * g3], rtol=1e-14)
    # GPR with M = 1: E = c * exp(-1/2 sum((x - m)/theta)^2) per atom
    m.mm = np.array([[0.1, 0.2, 0.3, 0.4]])
    m.theta = np.array([0.5, 0.6, 0.7, 0.8])
    m.coeff = np.array([1.7])
    e = oracle.gap_calc(m, sp, lat, pos, 6.0, mode="faithful")["energy"]
    expect = sum(1.7 * math.exp(-0.5 * np.sum(((xi - m.mm[0]) / m.theta) ** 2)) for xi in x)
    assert abs(e - expect) < 1e-14


@pytest.mark.parametrize("angle", [60.0, 90.0, 137.0])
def test_trimer_angular_closed_form(angle):
    lat = np.eye(3) * 40.0
    a, b = 1.9, 2.4
    th = math.radians(angle)
    pos = np.array([[10.0, 10, 10], [10 + a, 10, 10], [10 + b * math.cos(th), 10 + b * math.sin(th), 10]])
    sp = np.array([1, 8, 8], np.int32)
    e = math.dist(pos[1], pos[2])
    eta = 0.07
    zetas = [1, 1, 2, 4]
    m = _iso_model([2, 4, 2, 4], [eta] * 4, [6.0] * 4, zetas=zetas)
    x = oracle.descriptors(m, sp, lat, pos, 6.0)
    base = math.exp(-eta * (a * a + b * b + e * e)) * _fc(a, 6) * _fc(b, 6) * _fc(e, 6)
    cth = math.cos(th)
    expect = [(1 + cth) * base, (1 - cth) * base, 0.5 * (1 + cth) ** 2 * base, 2 ** -3 * (1 - cth) ** 4 * base]
    np.testing.assert_allclose(x[0, :4], expect, rtol=1e-12)
    np.testing.assert_allclose(x[0, 4:], 9.0 * np.array(expect), rtol=1e-12)


def test_norms_minus_dot_identity():
    """The expanded distance used on the GPU equals the direct sum (FP64)."""
    rng = np.random.default_rng(7)
    X = rng.normal(size=(50, 30))
    Mm = rng.normal(size=(40, 30))
    th = rng.uniform(0.5, 2, 30)
    direct = (((X[:, None, :] - Mm[None, :, :]) / th) ** 2).sum(-1)
    xt, mt = X / th, Mm / th
    mu = mt.mean(axis=0)
    xt, mt = xt - mu, mt - mu
    expanded = (xt ** 2).sum(1)[:, None] + (mt ** 2).sum(1)[None, :] - 2 * xt @ mt.T
    np.testing.assert_allclose(expanded, direct, rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------
# neighbour search: cell list == brute force over images
# --------------------------------------------------------------------------

def brute_force_pairs(lat, pos, rc):
    inv = np.linalg.inv(lat)
    spacing = 1.0 / np.linalg.norm(inv, axis=0)
    frac = pos @ inv
    span = frac.max(axis=0) - frac.min(axis=0)
    n = np.ceil(rc / spacing + span).astype(int) + 1
    out = []
    for i in range(len(pos)):
        for j in range(len(pos)):
            for img in itertools.product(*(range(-k, k + 1) for k in n)):
                if i == j and img == (0, 0, 0):
                    continue
                d = pos[j] + np.array(img) @ lat - pos[i]
                if np.linalg.norm(d) <= rc:
                    out.append((i, j) + img)
    return sorted(out)


@pytest.mark.parametrize("n,vpa,seed", [(5, 12.0, 1), (7, 30.0, 2), (20, 11.0, 3), (3, 6.0, 4)])
def test_cell_list_equals_brute_force(n, vpa, seed):
    from gapx import synth
    rng = np.random.default_rng(seed)
    s = synth.structure(n, n * vpa, (1,), rng, min_dist=1.0)
    pos = s["pos"] + rng.normal(0, 3.0, size=s["pos"].shape)  # unwrapped positions
    rows, _ = oracle.neighbor_set(s["lat"], pos, 6.0)
    assert [tuple(int(v) for v in r) for r in rows] == brute_force_pairs(s["lat"], pos, 6.0)


@pytest.mark.parametrize("case", CAR_CASE_NAMES)
def test_car2acsf_restatement_matches_reference(ref_vectors, case):
    """CAR2ACSF restated (dense dxdy/strs) against the compiled reference's
    outputs on FGAP_CALC's own neighbour tables (tests/golden), and the
    neighbour-table builder against the stored table."""
    c = car_case(ref_vectors, case)
    m = golden_model(ref_vectors, c["model"])
    xx, dxdy, strs = oracle.car2acsf(m, c["pos"], c["neighbor"], c["count"])
    for got, ref in ((xx, c["xx"]), (dxdy, c["dxdy"]), (strs, c["strs"])):
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))
    g = [x for x in golden_cases(ref_vectors) if x["name"] == case][0]
    nb, cnt = oracle.neighbor_table(m, g["species"], g["lat"], g["pos"], g["rcut"])
    assert np.array_equal(cnt, c["count"])
    k = c["neighbor"].shape[1]
    np.testing.assert_array_equal(nb[:, :k, :], c["neighbor"])


@pytest.mark.parametrize("case", GOLDEN_CASE_NAMES)
def test_min_bond_restatement_matches_reference(ref_vectors, case):
    """FGET_BOND restated (numpy, the reference's image range and order of
    operations) equals the compiled reference's min_bond bit for bit, at the
    evaluation cutoff and at 0.9 A."""
    c = [x for x in golden_cases(ref_vectors) if x["name"] == case][0]
    want = ref_vectors["case_%s_min_bond" % case]
    assert oracle.min_bond(c["lat"], c["pos"], c["rcut"]) == want[0]
    assert oracle.min_bond(c["lat"], c["pos"], 0.9) == want[1]
