"""Synthetic structures, symmetry-function tables and GAP models.

Everything here is seeded and deterministic (numpy PCG64), so the same call
gives the same structures on every machine with this image.

Workloads (sizes as used by the benchmark and the tests):

* ``c64``      -- 64 carbon atoms, cubic cell at 0.06 atoms/A^3, rc 6 A,
                  24 radial + 24 angular lines (D = 96), M = 500.
* ``ternary4k``-- 4096 atoms of Z in {1, 8, 22} (weights = Z), triclinic cell
                  (angles 80-100 deg) at 0.07 atoms/A^3, rc 6 A, 32 radial +
                  32 angular lines (4 eta x 4 zeta x 2 lambda, D = 128),
                  M = 4000, alpha ~ N(0, 0.05).
* ``ternary32k``-- 32768 atoms at the same density, rc 6.5 A, 64 + 64 lines
                  (D = 256), M = 8000.
* ``population``-- 512 structures of 8-256 atoms, 10-20 A^3/atom, 1-4
                  species from {1, 6, 8, 22}, D = 128, M = 4000, rc 6 A.

Models: theta_d = standard deviation of descriptor column d over a pool of
extra seeded structures; the sparse set MM is drawn from that pool without
replacement; alpha ~ N(0, sigma).  The model's descriptor pool is computed by
a caller-supplied function (the GPU descriptor entry point in the product).
"""
import math

import numpy as np

MIN_DIST = 1.8


# --------------------------------------------------------------------------
# structures
# --------------------------------------------------------------------------

def triclinic_cell(volume, rng, angle_range=(80.0, 100.0)):
    """Equal edges, angles uniform in angle_range (degrees), scaled to volume."""
    al, be, ga = np.radians(rng.uniform(*angle_range, size=3))
    a = np.array([1.0, 0.0, 0.0])
    b = np.array([math.cos(ga), math.sin(ga), 0.0])
    cx = math.cos(be)
    cy = (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz = math.sqrt(max(1.0 - cx * cx - cy * cy, 1e-6))
    c = np.array([cx, cy, cz])
    lat = np.stack([a, b, c])
    v1 = abs(np.linalg.det(lat))
    return lat * (volume / v1) ** (1.0 / 3.0)


def cubic_cell(volume):
    return np.eye(3) * volume ** (1.0 / 3.0)


def random_positions(lat, n, rng, min_dist=MIN_DIST, max_tries=200000):
    """Sequential random insertion with minimum-image rejection at min_dist.

    A bin grid in fractional space (bins at least min_dist wide) keeps each
    test local; periodic images are covered by wrapping bin indices and
    testing the +-1 lattice translations of each candidate neighbour.
    """
    inv = np.linalg.inv(lat)
    spacing = 1.0 / np.linalg.norm(inv, axis=0)
    nb = np.maximum(1, np.floor(spacing / min_dist).astype(int))
    grid = {}
    frac = np.zeros((n, 3))
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], float)
    shift_cart = shifts @ lat
    placed = 0
    tries = 0
    md2 = min_dist * min_dist
    if n <= 600:
        # small cells: draw candidates in batches (32, doubling while few are
        # accepted) and test each batch
        # against every placed atom at once; the survivors are then accepted
        # in draw order, each tested against those accepted before it.
        def clear(cands, ref):
            df = cands[:, None, :] - ref[None, :, :]
            df -= np.round(df)
            dc = df @ lat
            d2 = ((dc[:, :, None, :] + shift_cart[None, None, :, :]) ** 2).sum(-1).min(axis=2)
            return np.all(d2 >= md2, axis=1)

        batch = 32
        while placed < n:
            tries += batch
            if tries > max_tries * 10:
                raise RuntimeError("could not place %d atoms at min distance %.2f" % (n, min_dist))
            cands = rng.random((batch, 3))
            if placed:
                cands = cands[clear(cands, frac[:placed])]
            start = placed
            for f in cands:
                if placed > start and not clear(f[None, :], frac[start:placed])[0]:
                    continue
                frac[placed] = f
                placed += 1
                if placed == n:
                    break
            # grow the batch as the acceptance rate falls
            if placed - start < 4 and batch < 1024:
                batch *= 2
        return frac @ lat
    while placed < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place %d atoms at min distance %.2f" % (n, min_dist))
        f = rng.random(3)
        b = np.floor(f * nb).astype(int)
        cand = []
        rng_k = [range(-1, 2) if nb[k] > 2 else range(nb[k]) for k in range(3)]
        for o0 in rng_k[0]:
            for o1 in rng_k[1]:
                for o2 in rng_k[2]:
                    if nb[0] > 2:
                        q0 = (b[0] + o0) % nb[0]
                    else:
                        q0 = o0
                    if nb[1] > 2:
                        q1 = (b[1] + o1) % nb[1]
                    else:
                        q1 = o1
                    if nb[2] > 2:
                        q2 = (b[2] + o2) % nb[2]
                    else:
                        q2 = o2
                    cand.extend(grid.get((q0, q1, q2), ()))
        ok = True
        if cand:
            df = frac[cand] - f
            df -= np.round(df)
            dc = df @ lat
            d2 = ((dc[:, None, :] + shift_cart[None, :, :]) ** 2).sum(-1).min(axis=1)
            ok = bool(np.all(d2 >= md2))
        if ok:
            frac[placed] = f
            grid.setdefault(tuple(int(v) for v in b), []).append(placed)
            placed += 1
    return frac @ lat


def structure(n, volume, species_pool, rng, cell="triclinic", species_probs=None,
              min_dist=MIN_DIST):
    lat = triclinic_cell(volume, rng) if cell == "triclinic" else cubic_cell(volume)
    pos = random_positions(lat, n, rng, min_dist=min_dist)
    sp = np.asarray(species_pool, dtype=np.int32)
    if species_probs is None:
        species = sp[np.arange(n) % len(sp)]
        rng.shuffle(species)
    else:
        species = rng.choice(sp, size=n, p=species_probs).astype(np.int32)
    return dict(species=species.astype(np.int32), lat=lat, pos=pos)


# --------------------------------------------------------------------------
# symmetry-function tables
# --------------------------------------------------------------------------

def sf_table(n_radial, etas_ang, zetas, rc, alpha_range=(0.01, 4.0)):
    """Radial type-1 lines with log-spaced alpha, then angular lines ordered
    by eta, lambda (type 2 = +1, type 4 = -1), zeta."""
    types, alphas, cuts, zs = [], [], [], []
    lo, hi = alpha_range
    for k in range(n_radial):
        types.append(1)
        alphas.append(lo * (hi / lo) ** (k / max(n_radial - 1, 1)))
        cuts.append(rc)
        zs.append(1.0)
    for eta in etas_ang:
        for t in (2, 4):
            for z in zetas:
                types.append(t)
                alphas.append(float(eta))
                cuts.append(rc)
                zs.append(float(z))
    return (np.array(types, np.int32), np.array(alphas), np.array(cuts), np.array(zs))


def logspace(lo, hi, n):
    return [lo * (hi / lo) ** (k / max(n - 1, 1)) for k in range(n)]


WORKLOADS = {
    # name: atoms, density (atoms/A^3), species (weights = Z), cell, rc,
    #       n_radial, angular etas, zetas, M, alpha sigma, seed
    "c64": dict(n=64, density=0.06, species=(6,), cell="cubic", rc=6.0, n_radial=24,
                etas=logspace(0.005, 0.5, 3), zetas=(1, 2, 4, 16), M=500, sigma=0.1, seed=1001,
                pool=16),
    "c64_zeta1": dict(n=64, density=0.06, species=(6,), cell="cubic", rc=6.0, n_radial=24,
                      etas=logspace(0.005, 0.5, 12), zetas=(1,), M=500, sigma=0.1, seed=1001,
                      pool=16),
    "ternary4k": dict(n=4096, density=0.07, species=(1, 8, 22), cell="triclinic", rc=6.0,
                      n_radial=32, etas=logspace(0.005, 0.5, 4), zetas=(1, 2, 4, 16), M=4000,
                      sigma=0.05, seed=1002, pool=2),
    "ternary32k": dict(n=32768, density=0.07, species=(1, 8, 22), cell="triclinic", rc=6.5,
                       n_radial=64, etas=logspace(0.005, 0.5, 8), zetas=(1, 2, 4, 16), M=8000,
                       sigma=0.05, seed=1003, pool=1),
}


def workload_sf(name):
    w = WORKLOADS[name]
    return sf_table(w["n_radial"], w["etas"], w["zetas"], w["rc"])


def workload_structure(name, seed_offset=0):
    w = WORKLOADS[name]
    rng = np.random.default_rng(w["seed"] + 7919 * seed_offset)
    return structure(w["n"], w["n"] / w["density"], w["species"], rng, cell=w["cell"])


def make_model(species_z, species_w, sf, pool_x, M, sigma, rng):
    """GAP model from a descriptor pool: theta = column std, MM = M pool rows."""
    types, alphas, cuts, zetas = sf
    pool_x = np.asarray(pool_x)
    theta = pool_x.std(axis=0)
    theta = np.where(theta > 0, theta, 1.0)
    idx = rng.choice(len(pool_x), size=M, replace=False)
    mm = pool_x[idx].copy()
    coeff = rng.normal(0.0, sigma, size=M)
    return dict(z=np.asarray(species_z, np.int32), w=np.asarray(species_w, float), sf_type=types,
                sf_alpha=alphas, sf_cut=cuts, sf_zeta=zetas, theta=theta, mm=mm, coeff=coeff)


def workload_model(name, descriptor_fn):
    """Model for a workload; descriptor_fn(model_header, species, lat, pos, rc) -> X."""
    w = WORKLOADS[name]
    sf = workload_sf(name)
    header = dict(z=np.array(w["species"], np.int32), w=np.array(w["species"], float), sf=sf)
    rng = np.random.default_rng(w["seed"] + 100000)
    xs = []
    for p in range(w["pool"]):
        s = workload_structure(name, seed_offset=p + 1)
        xs.append(descriptor_fn(header, s["species"], s["lat"], s["pos"], w["rc"]))
    pool = np.concatenate(xs)
    return make_model(w["species"], w["species"], sf, pool, w["M"], w["sigma"], rng)


def population(n_struct=512, nmin=8, nmax=256, seed=1004, species_set=(1, 6, 8, 22),
               min_dist=1.5):
    """Uneven structure population: atoms U{nmin..nmax}, 10-20 A^3/atom,
    triclinic cells, 1-4 species drawn from species_set.  The minimum
    interatomic distance is 1.5 A: at 10 A^3/atom a 1.8 A exclusion is a
    packing fraction of 0.31, close enough to random-sequential jamming that
    seeded insertion takes minutes per population."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_struct):
        n = int(rng.integers(nmin, nmax + 1))
        vpa = rng.uniform(10.0, 20.0)
        ns = int(rng.integers(1, len(species_set) + 1))
        sp = rng.choice(np.asarray(species_set), size=ns, replace=False)
        out.append(structure(n, n * vpa, sp, rng, cell="triclinic", min_dist=min_dist))
    return out


# Batched form: a population of uneven periodic structures, and the D = 128,
# M = 4000 model it is evaluated with (species 1, 6, 8, 22, weights = Z; the
# SF table is the ternary4k one; the sparse set is drawn from the descriptors
# of a separately seeded population).
POPULATION = dict(n_struct=512, nmin=8, nmax=256, seed=1004, species=(1, 6, 8, 22), rc=6.0, n_radial=32,
                  etas=logspace(0.005, 0.5, 4), zetas=(1, 2, 4, 16), M=4000, sigma=0.05, pool_structs=48,
                  pool_seed=2004)


def population_sf():
    p = POPULATION
    return sf_table(p["n_radial"], p["etas"], p["zetas"], p["rc"])


def population_model(descriptor_fn, M=None):
    """GAP model for the population workload; descriptor_fn as for workload_model."""
    p = POPULATION
    sf = population_sf()
    header = dict(z=np.array(p["species"], np.int32), w=np.array(p["species"], float), sf=sf)
    pool = population(p["pool_structs"], seed=p["pool_seed"], species_set=p["species"])
    xs = [descriptor_fn(header, s["species"], s["lat"], s["pos"], p["rc"]) for s in pool]
    rng = np.random.default_rng(p["seed"] + 100000)
    return make_model(p["species"], p["species"], sf, np.concatenate(xs), M or p["M"], p["sigma"], rng)


# Surrogate form: many small ternary structures drawn like the model's own
# descriptor pool (random sequential insertion with a 1.8 A exclusion at
# 0.07 atoms/A^3).  rsa_structure draws 4 n candidates at once, finds every
# close pair among them with one vectorised minimum-image test (exact for a
# threshold below half the smallest interplanar spacing, which holds for the
# cells used here), and accepts candidates in draw order unless they clash
# with one accepted before: sequential insertion, about 5x faster than
# random_positions.  Structure k of a set uses seed (seed, k), so a set is the
# same for any number of worker processes.
SURROGATE = dict(n_train=200000, nmin=16, nmax=128, density=0.07, species=(1, 8, 22), seed=1005,
                 candidates=10000, top_fraction=0.05, min_dist=1.8)


def rsa_structure(n, volume, species, rng, min_dist=MIN_DIST):
    lat = triclinic_cell(volume, rng)
    sp = np.asarray(species, np.int32)
    inv = np.linalg.inv(lat)
    half_spacing = 0.5 / np.linalg.norm(inv, axis=0).max()
    if min_dist >= half_spacing:
        return dict(species=rng.choice(sp, size=n).astype(np.int32), lat=lat,
                    pos=random_positions(lat, n, rng, min_dist=min_dist))
    md2 = min_dist * min_dist
    frac = np.zeros((0, 3))
    while len(frac) < n:
        cand = rng.random((4 * n, 3))
        allf = np.concatenate([frac, cand])
        df = cand[:, None, :] - allf[None, :, :]
        df -= np.round(df)
        d2 = ((df @ lat) ** 2).sum(-1)
        clash = d2 < md2                       # [candidate, accepted + candidate]
        ok_old = ~clash[:, :len(frac)].any(axis=1)
        cc = clash[:, len(frac):]
        blocked = np.zeros(len(cand), bool)
        take = []
        for j in np.flatnonzero(ok_old):
            if blocked[j]:
                continue
            take.append(j)
            blocked |= cc[j]
            if len(frac) + len(take) == n:
                break
        frac = np.concatenate([frac, cand[take]])
    return dict(species=rng.choice(sp, size=n).astype(np.int32), lat=lat, pos=frac[:n] @ lat)


def _rsa_one(args):
    k, seed, nmin, nmax, density, species, min_dist = args
    rng = np.random.default_rng([seed, k])
    n = int(rng.integers(nmin, nmax + 1))
    return rsa_structure(n, n / density, species, rng, min_dist=min_dist)


def random_ternary(n_struct, nmin=16, nmax=128, density=0.07, species=(1, 8, 22), seed=1005, min_dist=MIN_DIST,
                   workers=1, start=0):
    """Structures start .. start + n_struct - 1 of the seeded set."""
    args = [(k, seed, nmin, nmax, density, tuple(species), min_dist) for k in range(start, start + n_struct)]
    if workers > 1 and n_struct >= 4 * workers:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            return pool.map(_rsa_one, args, chunksize=max(1, n_struct // (8 * workers)))
    return [_rsa_one(a) for a in args]


def jittered_structures(n_struct, nmin, nmax, density, species, seed, jitter=0.2):
    """Jittered lattice sites (no rejection; about 0.1 ms per structure):
    with jitter 0.2 of the site spacing the closest pair is at least 0.6
    spacings.  Fast, but its environments are far from a random-insertion
    pool's."""
    rng = np.random.default_rng(seed)
    out = []
    for n in rng.integers(nmin, nmax + 1, size=n_struct):
        n = int(n)
        m = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
        lat = triclinic_cell(n / density, rng)
        idx = rng.choice(m ** 3, size=n, replace=False)
        g = np.stack(np.unravel_index(idx, (m, m, m)), axis=1).astype(float)
        frac = (g + 0.5 + rng.uniform(-jitter, jitter, size=(n, 3))) / m
        out.append(dict(species=rng.choice(np.asarray(species, np.int32), size=n).astype(np.int32), lat=lat,
                        pos=frac @ lat))
    return out
