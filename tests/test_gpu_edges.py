"""Edge cases of the GPU path against the oracle (FP64 bar: 1e-9 eV/atom,
1e-7 eV/A, 1e-6 GPa) and error paths that must leave the context usable.

* An isolated atom (no neighbour within rcut), one atom in a cell smaller
  than rcut (only its own periodic images as neighbours, the reference keeps
  them: gap_calc.f90:97-101), a close pair inside the reference's
  d < 0.5 A warning range (gap_calc.f90:102-105), and a ragged batch of 1-,
  2- and 64-atom structures.
* Empty input (no atoms, a structure with no atoms) is an argument error.
* A load that fails part-way (an atomic number the model does not list)
  leaves the context empty: a run then reports "no structures loaded", and
  the next good load works as in a fresh context.
* A rejected lattice update (singular cell) keeps the loaded cells.
* Lists above the reference's 1000 (gap_calc.f90:107-111) stop FGAP_CALC's
  flags like the reference and are evaluated whole without them; lists
  longer than the descriptor kernels take (about 2000) are reported as
  GAPX_ERR_NEIGHBORS.  Lists of about 920 (under the reference's limit) are
  pinned by the reference's own vectors (tests/test_gpu_parity.py).
* A cell too small for rcut / 48 is refused; lists above 256 slots (the
  triplet builder's other path) match the oracle.
* Strongly sheared cells, where the list kernel's image-bin distance bound is
  loosest: pair counts against a test of every image, descriptors and
  results against the oracle.
* A 0.4 A-thin slab whose image search reaches 15 bins (compact 32-bit
  reverse-slot keys) and 18 bins (64-bit keys): pair counts and results
  against the oracle.
"""
import numpy as np
import pytest

import gapx
import oracle
from helpers import fitted_small_model, small_model, ternary_structure

pytestmark = pytest.mark.gpu

TOL_E, TOL_F, TOL_S = 1e-9, 1e-7, 1e-6


@pytest.fixture(scope="module")
def model():
    if gapx.device_count() == 0:
        pytest.skip("no GPU visible")
    rng = np.random.default_rng(17)
    s = ternary_structure(48, 12.0, 171)
    om = fitted_small_model(rng, s)
    m = gapx.Model.from_model(om)
    yield om, m
    m.close()


def check(om, m, species, lat, pos, rcut=6.0):
    e, f, st, _ = m.calc(species, lat, pos, rcut)
    r = oracle.gap_calc(om, species, lat, pos, rcut)
    n = len(species)
    assert abs(e - r["energy"]) / n <= TOL_E
    assert np.abs(f - r["forces"]).max() <= TOL_F
    assert np.abs(st - r["stress"]).max() <= TOL_S
    return e, f


def test_isolated_atom(model):
    om, m = model
    e, f = check(om, m, np.array([8], np.int32), np.eye(3) * 20.0, np.array([[3.0, 4.0, 5.0]]))
    assert not np.any(f)


def test_single_atom_in_cell_smaller_than_rcut(model):
    om, m = model
    lat = np.array([[2.6, 0.0, 0.0], [0.4, 2.8, 0.0], [0.3, -0.2, 3.1]])
    check(om, m, np.array([22], np.int32), lat, np.array([[0.5, 0.7, 0.9]]))


def test_close_pair(model):
    om, m = model
    pos = np.array([[5.0, 5.0, 5.0], [5.3, 5.0, 5.0], [6.4, 5.9, 5.2]])
    check(om, m, np.array([1, 1, 8], np.int32), np.eye(3) * 10.0, pos)


def test_ragged_batch_equals_single_structures(model):
    om, m = model
    structs = [dict(species=np.array([8], np.int32), lat=np.eye(3) * 9.0, pos=np.array([[1.0, 2.0, 3.0]])),
               dict(species=np.array([1, 22], np.int32), lat=np.eye(3) * 7.5, pos=np.array([[0.0, 0.0, 0.0],
                                                                                              [1.9, 0.3, 0.1]])),
               ternary_structure(64, 12.0, 172)]
    ene, forces, stress = m.calc_batch(structs, 6.0)
    for k, s in enumerate(structs):
        e, f = check(om, m, s["species"], s["lat"], s["pos"])
        assert abs(ene[k] - e) / len(s["species"]) <= 1e-12
        assert np.abs(forces[k] - f).max() <= 1e-12


def test_empty_inputs_are_argument_errors(model):
    om, m = model
    with pytest.raises(gapx.GapxError) as e:
        m.calc(np.zeros(0, np.int32), np.eye(3) * 5.0, np.zeros((0, 3)), 6.0)
    assert e.value.code == 1
    ctx = gapx.Context(m)
    with pytest.raises(gapx.GapxError) as e:
        ctx.load([dict(species=np.zeros(0, np.int32), lat=np.eye(3) * 5.0, pos=np.zeros((0, 3)))], 6.0)
    assert e.value.code == 1
    with pytest.raises(gapx.GapxError) as e:
        ctx.run(gapx.GRAD)
    assert e.value.code == 1
    ctx.close()


def test_failed_load_leaves_context_empty(model):
    om, m = model
    good = [ternary_structure(40, 12.0, 173)]
    ctx = gapx.Context(m)
    ctx.load(good, 6.0)
    ctx.run(gapx.GRAD)
    r0 = ctx.results()
    bad = [ternary_structure(300, 12.0, 174)]
    bad[0]["species"] = bad[0]["species"].copy()
    bad[0]["species"][7] = 79  # not in the model
    with pytest.raises(gapx.GapxError) as e:
        ctx.load(bad, 6.0)
    assert e.value.code == 4
    for call in (lambda: ctx.run(gapx.GRAD), lambda: ctx.update(good[0]["pos"])):
        with pytest.raises(gapx.GapxError) as e:
            call()
        assert e.value.code == 1
    ctx.load(good, 6.0)
    ctx.run(gapx.GRAD)
    r1 = ctx.results()
    assert np.array_equal(r0["forces"], r1["forces"]) and r0["energy"][0] == r1["energy"][0]
    ctx.close()


def test_rejected_lattice_update_keeps_cells(model):
    om, m = model
    s = ternary_structure(40, 12.0, 175)
    ctx = gapx.Context(m)
    ctx.load([s], 6.0)
    ctx.run(gapx.GRAD)
    r0 = ctx.results()
    with pytest.raises(gapx.GapxError) as e:
        ctx.update(s["pos"], lat=np.zeros((1, 3, 3)))
    assert e.value.code == 1
    ctx.run(gapx.GRAD)
    r1 = ctx.results()
    assert np.array_equal(r0["forces"], r1["forces"]) and np.array_equal(r0["stress"], r1["stress"])
    ctx.close()


def test_neighbour_limit_is_reported(model):
    """About 4500 neighbours per atom (40 atoms in a 4 A cube, rcut 12): above
    the list length the descriptor kernels take (about 2000 in FP64 for this
    table, kept within the device's LDS with the cutoff tables in global
    memory), so GAPX_ERR_NEIGHBORS in both precisions."""
    om, m = model
    lat = np.eye(3) * 4.0
    pos = np.random.default_rng(2).random((40, 3)) * 4.0
    for flags in (0, gapx.MIXED):
        with pytest.raises(gapx.GapxError) as e:
            m.calc(np.ones(40, np.int32), lat, pos, 12.0, flags=flags)
        assert e.value.code == 5


def dense_cell(vpa, seed):
    from gapx import synth
    return synth.structure(8, 8 * vpa, (1, 8, 22), np.random.default_rng(seed), cell="cubic", min_dist=0.9)


def dense_model(s, mixed):
    """A model fitted on the dense cell's own descriptors (kernel values
    O(1)).  theta = 1.5 std over the cell's 8 atoms makes some features 4e4
    times their length scale; with mixed=True theta also carries 2^-10 of the
    feature's mean, which keeps the FP32 features within the mixed bar."""
    om = fitted_small_model(np.random.default_rng(42), s)
    if mixed:
        x = oracle.descriptors(om, s["species"], s["lat"], s["pos"], 6.0)
        om.theta = x.std(axis=0) * 1.5 + 2.0 ** -10 * np.abs(x.mean(axis=0)) + 1e-3
    return om


@pytest.mark.parametrize("mixed", [False, True], ids=["fp64", "mixed"])
def test_above_the_reference_cap_without_it(mixed):
    """About 1090 neighbours per atom (8 atoms at 0.83 A^3 each) with a model
    fitted on the cell's own descriptors: FGAP_CALC's flags
    (GAPX_REF_NEIGHBOR_CAP) stop above 1000 like the reference
    (gap_calc.f90:107-111); without the flag the lists are taken whole and
    match the oracle's sparse mode (FP64 bar, the stress to 1e-8 of its
    5.6e5 GPa: a virial of some 10^4 terms, measured 2.5e-9; mixed, on the
    16-bit split contraction with FP32 features, within 1 meV/atom and
    20 meV/A of FP64)."""
    s = dense_cell(0.83, 41)
    om = dense_model(s, mixed)
    m = gapx.Model.from_model(om)
    if mixed:  # the rescaled model keeps the 16-bit path (the as-fitted one is routed to FP64)
        assert not m.precision()["mixed_runs_fp64"]
    ctx = gapx.Context(m)
    ctx.load([s], 6.0)
    ctx.run(gapx.GRAD | (gapx.MIXED if mixed else 0))
    ctx.results()
    pairs, _ = ctx.counts()
    ctx.close()
    assert pairs / 8 > 1000
    with pytest.raises(gapx.GapxError) as e:
        m.calc(s["species"], s["lat"], s["pos"], 6.0, flags=gapx.REF_NEIGHBOR_CAP | (gapx.MIXED if mixed else 0))
    assert e.value.code == 5
    e64, f64, s64, _ = m.calc(s["species"], s["lat"], s["pos"], 6.0)
    if mixed:
        r = m.calc(s["species"], s["lat"], s["pos"], 6.0, mixed=True)
        print("dense 1090 mixed vs fp64: dE/atom %.3e dF %.3e" % (abs(r[0] - e64) / 8, np.abs(r[1] - f64).max()))
        assert abs(r[0] - e64) / 8 <= 1e-3 and np.abs(r[1] - f64).max() <= 2e-2
    else:
        ref = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 6.0, mode="sparse", nthreads=8)
        print("dense 1090: E %.6f |F| %.4f S %.1f" % (e64, np.abs(f64).max(), s64[0]))
        assert abs(e64 - ref["energy"]) / 8 <= TOL_E
        assert np.abs(f64 - ref["forces"]).max() <= TOL_F
        assert np.abs(s64 - ref["stress"]).max() <= max(TOL_S, 1e-8 * np.abs(ref["stress"]).max())
    m.close()


def test_mixed_runs_of_a_badly_scaled_model_are_fp64_runs():
    """The same cell under the model whose features reach 4e4 times theta:
    its FP32 features would move E by about 1.4 meV/atom (measured before the
    probe routed it), so the model is recognised at load (B 2^-24 = 3.4e-4
    eV) and GAPX_MIXED runs equal FP64 runs bit for bit."""
    s = dense_cell(0.83, 41)
    m = gapx.Model.from_model(dense_model(s, False))
    prec = m.precision()
    assert prec["mixed_runs_fp64"] and prec["feature_sensitivity"] * 2.0 ** -24 > 1e-4, prec
    e64, f64, s64, _ = m.calc(s["species"], s["lat"], s["pos"], 6.0)
    for flags in (gapx.MIXED, gapx.MIXED | gapx.BF16):
        r = m.calc(s["species"], s["lat"], s["pos"], 6.0, flags=flags)
        assert r[0] == e64 and np.array_equal(r[1], f64) and np.array_equal(r[2], s64)
    m.close()


def test_too_many_periodic_images_is_an_argument_error(model):
    """A lattice plane spacing below rcut / 48 would need more periodic
    images than the list kernel indexes exactly; it is refused at load."""
    om, m = model
    with pytest.raises(gapx.GapxError) as e:
        m.calc(np.array([8], np.int32), np.eye(3) * 0.1, np.array([[0.01, 0.02, 0.03]]), 6.0)
    assert e.value.code == 1
    check(om, m, np.array([8], np.int32), np.eye(3) * 2.0, np.array([[0.5, 0.7, 0.9]]), rcut=3.0)


def test_more_than_256_neighbours_match_oracle(model):
    """Lists longer than 256 slots take the triplet builder's path that reads
    both records of a pair from LDS (shorter lists keep each lane's own records
    in registers); FP64 against the oracle, mixed within its bar."""
    om, m = model
    s = ternary_structure(40, 12.0, 176)
    e, f, st, _ = m.calc(s["species"], s["lat"], s["pos"], 10.0)
    r = oracle.gap_calc(om, s["species"], s["lat"], s["pos"], 10.0)
    n = len(s["species"])
    assert abs(e - r["energy"]) / n <= TOL_E
    assert np.abs(f - r["forces"]).max() <= TOL_F
    assert np.abs(st - r["stress"]).max() <= TOL_S
    ctx = gapx.Context(m)
    ctx.load([s], 10.0)
    ctx.run(gapx.GRAD)
    pairs, _ = ctx.counts()
    ctx.close()
    assert pairs / n > 256
    mix = m.calc(s["species"], s["lat"], s["pos"], 10.0, mixed=True)
    assert abs(mix[0] - e) / n <= 1e-3 and np.abs(mix[1] - f).max() <= 2e-2


def brute_pairs(lat, pos, rcut):
    """Ordered pairs (i, j, image) with 0 < r <= rcut, every image tested."""
    h = 1.0 / np.linalg.norm(np.linalg.inv(lat), axis=0)  # plane spacings
    rng3 = [np.arange(-k, k + 1) for k in np.ceil(rcut / h).astype(int) + 1]
    shifts = np.stack(np.meshgrid(*rng3, indexing="ij"), -1).reshape(-1, 3) @ lat
    d = pos[None, :, None, :] + shifts[None, None, :, :] - pos[:, None, None, :]
    r = np.sqrt((d ** 2).sum(-1))
    return int(((r <= rcut) & (r > 0)).sum())


@pytest.mark.parametrize("cell", [
    [[9.0, 0.0, 0.0], [7.5, 4.0, 0.0], [-6.0, 3.0, 5.0]],
    [[14.0, 0.0, 0.0], [3.0, 13.0, 0.0], [5.0, -4.0, 3.2]],
    [[7.0, 0.0, 0.0], [-5.5, 6.0, 0.0], [3.0, -2.5, 6.5]],
    [[20.0, 0.0, 0.0], [12.0, 15.0, 0.0], [-8.0, 6.0, 14.0]],
])
def test_sheared_cells_match_oracle(model, cell):
    """The list kernel skips image bins that a lower bound on their distance
    (through the bin metric's Cholesky factor) puts beyond rcut; in strongly
    sheared cells that bound is furthest from the true distance.  The pair
    count against every image tested, the descriptors against the oracle's,
    and energy, forces and stress."""
    from gapx import synth
    om, m = model
    lat = np.array(cell)
    rng = np.random.default_rng(int(abs(np.linalg.det(lat))))
    n = max(8, int(0.065 * abs(np.linalg.det(lat))))
    pos = synth.random_positions(lat, n, rng, min_dist=1.5)
    species = rng.choice(np.array([1, 8, 22], np.int32), size=len(pos)).astype(np.int32)
    ctx = gapx.Context(m)
    ctx.load([dict(species=species, lat=lat, pos=pos)], 6.0)
    ctx.run(gapx.GRAD)
    pairs, _ = ctx.counts()
    ctx.close()
    assert pairs == brute_pairs(lat, pos, 6.0)
    x = m.descriptors(species, lat, pos, 6.0)
    xo = oracle.descriptors(om, species, lat, pos, 6.0)
    np.testing.assert_allclose(x, xo, rtol=1e-11, atol=1e-13)
    check(om, m, species, lat, pos)


def test_wide_image_search_matches_oracle():
    """A cell with one 0.4 A lattice axis needs a search half-width of 15
    bins along it at rcut 6 for the 16-atom, 0.4 A-thin slab below (the list
    kernel's compact 32-bit reverse-slot keys hold up to 15), and 18 at rcut 7
    (64-bit keys).  Both against the oracle with a model fitted on the cell's
    own descriptors (kernel values O(1), so a wrong reverse slot would show
    in the forces), and the pair counts against a test of every image."""
    rng = np.random.default_rng(31)
    lat = np.array([[9.0, 0.0, 0.0], [2.0, 8.5, 0.0], [0.0, 0.0, 0.4]])
    pos = np.concatenate([rng.random((16, 2)) * 8.0, rng.random((16, 1)) * 0.4], axis=1)
    species = rng.choice(np.array([1, 8, 22], np.int32), size=16).astype(np.int32)
    s = dict(species=species, lat=lat, pos=pos)
    om = fitted_small_model(rng, s)
    m = gapx.Model.from_model(om)
    try:
        for rcut in (6.0, 7.0):
            ctx = gapx.Context(m)
            ctx.load([s], rcut)
            ctx.run(gapx.GRAD)
            pairs, _ = ctx.counts()
            ctx.close()
            assert pairs == brute_pairs(lat, pos, rcut)
            check(om, m, species, lat, pos, rcut=rcut)
    finally:
        m.close()
