"""Benchmark: atom energy+force+stress evaluations per second on one MI355X.

Headline workload (default, ``--workload population``): the batched form, 512
uneven periodic structures of 8-256 atoms (10-20 A^3/atom, triclinic cells,
1-4 species from {1, 6, 8, 22}, weights = Z), 32 radial + 32 angular wACSF
lines (zeta in {1, 2, 4, 16}, D = 128), M = 4000 sparse points, rcut 6 A; a
synthetic model built from GPU descriptors of a separately seeded population.
All structures are one device-resident batch, autotuned and replayed from a
captured HIP graph.  A step is one full E+F+S evaluation of every structure
from positions already in HBM: neighbour lists, descriptors, contraction,
backward pass, forces and stress.  ``value`` is the mixed-precision rate of
the whole population; the FP64 rate, their ratio and the largest
per-structure mixed errors against FP64 are in the same line.

Extra records in the same line: ``ternary32k`` (the 32768-atom cell, rcut
6.5 A, D = 256, M = 8000) and ``ternary4k`` (the 4096-atom cell, D = 128,
M = 4000), both precisions each; ``relax`` and ``md`` (batched FIRE
relaxation and NVT MD of the population); ``surrogate`` (the MLP pre-screen,
trained on GPU GAP labels); ``variance`` (predictive variance of the
4096-atom cell).  Skip them with --no-large, --no-4k, --no-dynamics,
--no-surrogate.  ``cpu_baseline`` is the reference FGAP_CALC compiled from its
own sources (oracle/_ref) on the first structures of the population, one
core, about --cpu-seconds of work (--no-cpu skips it).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload population]
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "calypso-gap_amd"))

import numpy as np  # noqa: E402

import gapx  # noqa: E402
from gapx import synth  # noqa: E402

PEAK = {"fp64": 78.6, "fp32": 157.3, "fp16": 2500.0, "bf16": 2500.0}   # TFLOP/s, MI355X dense


def gpu_descriptor_fn(header, species, lat, pos, rc):
    t, a, c, z = header["sf"]
    D = 2 * len(t)
    m = gapx.Model.from_arrays(header["z"], header["w"], t, a, c, np.ones(D), np.zeros((1, D)), np.zeros(1), z)
    x = m.descriptors(species, lat, pos, rc)
    m.close()
    return x


# Algorithmic work per unit, following the per-unit figures of the reference
# path's cost model: a (pair, radial line) costs about 30 flops and a
# (triplet, angular line) about 80 flops for forward plus backward (geometry,
# exponential, cutoff products and the three distance derivatives, counted per
# line as the reference computes them); split 1/3 forward, 2/3 backward.  The
# contraction is 4 D M flops per atom (S = xt mt^T and G = W mt); the other
# kernels are byte-bound and counted in bytes.
FLOPS_RAD_FWD, FLOPS_RAD_BWD = 10, 20          # per (pair, radial line)
FLOPS_ANG_FWD, FLOPS_ANG_BWD = 27, 53          # per (triplet, angular line)
BYTES_NEIGHBOR = 4 + 8 + 32 + 4                # per pair: index, key, (dx,dy,dz,r), reverse slot
# the headline metric this record reports (value: the mixed-precision rate;
# the FP64 rate, their ratio, the mixed errors and the roofline fraction
# are in the same record)
METRIC = "atom E+F+stress evals/sec, mixed vs FP64 GPU at ≤1 meV/atom; roofline frac"
SINGLE_KERNEL_STAGES = ("descriptors_fwd", "gpr_contraction", "gpr_finish", "descriptors_bwd")


def zeta_1_2_4_16(header):
    """True when the angular lines' zeta list is {1, 2, 4, 16}: the grid
    descriptor kernels then run their specialisation for that list."""
    ang = np.isin(header["sf_type"], (2, 4))
    return sorted(set(np.round(header["sf_zeta"][ang]).astype(int).tolist())) == [1, 2, 4, 16]


def contraction_grid(n_atoms, des_len, nsparse, mixed, tune):
    """Threads of one contraction launch over n_atoms rows (the host's split
    of the sparse points over grid.y for the tuned workgroup target)."""
    dp = (des_len + 31) // 32 * 32
    mp = (nsparse + 63) // 64 * 64
    rows = (128 if dp > 128 else tune["gpr_rows"]) if mixed else 64
    tiles = (n_atoms + 127) // 128 * 128 // rows
    chunk = 64 if dp <= 128 else 32
    splits = min(max(1, -(-tune["gpr_blocks"] // tiles)), mp // chunk)
    per = -(-(-(-mp // splits)) // chunk) * chunk
    splits = -(-mp // per)
    return tiles * splits * (512 if (mixed and dp > 128) else 256)


def stage_kernel(stage, mixed, tune, z16, des_len, n_atoms, nsparse, c64=False):
    """The one kernel of a single-kernel stage under the given launch shapes,
    as rocprofv3 names it, and its grid in threads for a one-piece launch over
    n_atoms (grid-form tables with 4 zetas; z16: the zeta list is {1, 2, 4,
    16}; c64: the mixed run contracts in FP64).  The contraction is
    instantiated for the descriptor length padded to 32; the mixed one as
    <bf16, D, row blocks per wave, waves>: four waves of one or two 16-row
    blocks up to 128 features, eight waves of one block above."""
    t = "float" if mixed else "double"
    z = "true" if z16 else "false"
    if stage == "descriptors_fwd":
        w = min(tune["waves_fwd"], 2)
        return "k_desc_fwd_grid<%s, 4, %d, %s, false>" % (t, w, z), n_atoms * 64 * w
    if stage == "descriptors_bwd":
        w = tune["waves_bwd"]
        return "k_desc_bwd<%s, %d, 4, %s, false>" % (t, w, z), n_atoms * 64 * w
    if stage == "gpr_contraction":
        dp = (des_len + 31) // 32 * 32
        grid = contraction_grid(n_atoms, des_len, nsparse, mixed and not c64, tune)
        if not mixed or c64:
            return "k_gpr_f64<%d>" % dp, grid
        if dp > 128:
            return "k_gpr_mixed<false, %d, 1, 8>" % dp, grid
        return "k_gpr_mixed<false, %d, %d, 4>" % (dp, tune["gpr_rows"] // 64), grid
    return "k_gpr_finish", n_atoms * 128


def stage_rates(st, N, D, M, pairs, triplets, m, mixed):
    hdr = m.header()
    n_ang = int(np.sum((hdr["sf_type"] == 2) | (hdr["sf_type"] == 4)))
    n_rad = int(np.sum((hdr["sf_type"] == 1) | (hdr["sf_type"] == 3)))
    vpeak = PEAK["fp32"] if mixed else PEAK["fp64"]
    mpeak = PEAK["fp16"] if (mixed and not m.precision()["mixed_contraction_fp64"]) else PEAK["fp64"]
    work = {
        "neighbor_list": ("hbm", pairs * BYTES_NEIGHBOR * 2 + N * 64, 8000.0, "GB/s", 1e9),
        "descriptors_fwd": ("valu", pairs * n_rad * FLOPS_RAD_FWD + triplets * n_ang * FLOPS_ANG_FWD, vpeak,
                            "TFLOP/s", 1e12),
        "gpr_contraction": ("mfma", 4.0 * D * M * N, mpeak, "TFLOP/s", 1e12),
        "gpr_finish": ("hbm", N * D * 8 * 3.0, 8000.0, "GB/s", 1e9),
        "descriptors_bwd": ("valu", pairs * n_rad * FLOPS_RAD_BWD + triplets * n_ang * FLOPS_ANG_BWD, vpeak,
                            "TFLOP/s", 1e12),
        "forces_stress": ("hbm", pairs * 24 * 2 + N * 80, 8000.0, "GB/s", 1e9),
    }
    out = {}
    for k, ms in st.items():
        bound, amount, peak, unit, scale = work[k]
        ach = amount / (ms * 1e-3) / scale if ms > 0 else 0.0
        out[k] = dict(bound=bound, achieved=round(ach, 3), peak=peak, unit=unit, frac=round(ach / peak, 4),
                      work=float(amount))
    return out


def _profiled(workload, kernel, grid):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("workloads", {}).get(workload, {}).get("kernels", {}).get("%s grid %d" % (kernel, grid))


def measured_traffic(workload, kernel, grid):
    """HBM bytes per launch of a kernel at a grid (threads) from the committed
    counter summary (profiles/traffic.json: two rocprofv3 passes per workload,
    scaled by a copy of known size), or None when it does not cover them."""
    v = _profiled(workload, kernel, grid)
    return None if v is None else v["traffic_bytes"]


def measured_counters(workload, kernel, grid):
    """The same summary's busy fractions for the kernel (VALU busy and lane
    utilisation, MFMA busy, LDS active and bank-conflict cycles, VALU
    instructions per launch), from the counter passes beside the traffic
    ones; None when not covered."""
    v = _profiled(workload, kernel, grid)
    if v is None or "valu_busy" not in v:
        return None
    return {k: v[k] for k in ("valu_busy", "valu_lane_util", "mfma_busy", "lds_active", "lds_bank_conflict",
                              "valu_instructions") if k in v}


def time_runs(ctx, flags, steps, warmup):
    for _ in range(warmup):
        ctx.run(flags)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.run(flags)
    ctx.sync()
    return (time.perf_counter() - t0) / steps


def stage_times(ctx, flags, steps):
    """Per-stage milliseconds from HIP events on the context's stream (eager
    one-piece launches: each single-kernel stage is one launch over all atoms)."""
    ctx.timing(True)
    for _ in range(steps):
        ctx.run(flags)
    ctx.sync()
    st = ctx.stages()
    ctx.timing(False)
    return {k: (v[0] / max(v[1], 1)) for k, v in st.items()}


def measure(m, structs, rc, steps, warmup, label):
    """Both precisions on one loaded batch: autotuned launch shapes, graph
    replay for the rate, eager per-stage events for the stage times.  Returns
    the context (still loaded) and the record."""
    ctx = gapx.Context(m)
    ctx.load(structs, rc)
    natoms = np.array([len(s["species"]) for s in structs])
    N = int(natoms.sum())
    runs = {}
    for name, flags in (("fp64", gapx.GRAD), ("mixed", gapx.GRAD | gapx.MIXED)):
        tune = ctx.autotune(flags)  # launch shapes for this batch (cached per shape class)
        ctx.use_graph(True)  # replay the captured launch sequence
        sec = time_runs(ctx, flags, steps, warmup)
        ctx.use_graph(False)  # per-stage events need eager launches
        st = stage_times(ctx, flags, max(3, steps // 4))
        ctx.use_graph(True)
        ctx.run(flags)
        runs[name] = dict(sec=sec, stages=st, res=ctx.results(), tune=tune)
        print("bench: %s %s %.3f ms/step tuning %s stages %s" % (label, name, sec * 1e3, tune,
                                                                 {k: round(v, 4) for k, v in st.items()}),
              file=sys.stderr, flush=True)
    pairs, triplets = ctx.counts()
    r64, rmx = runs["fp64"]["res"], runs["mixed"]["res"]
    offs = np.concatenate([[0], np.cumsum(natoms)])
    dF = [float(np.abs(rmx["forces"][offs[k]:offs[k + 1]] - r64["forces"][offs[k]:offs[k + 1]]).max())
          for k in range(len(structs))]
    err = dict(energy_meV_per_atom=float(np.abs((rmx["energy"] - r64["energy"]) / natoms).max()) * 1e3,
               force_meV_per_A=max(dF) * 1e3,
               stress_GPa=float(np.abs(rmx["stress"] - r64["stress"]).max()))
    D, M = m.des_len, m.nsparse
    c64 = m.precision()["mixed_contraction_fp64"]
    st = runs["mixed"]["stages"]
    rates = stage_rates(st, N, D, M, pairs, triplets, m, mixed=True)
    # the dominant kernel: the longest of the stages that are one kernel each
    dom = max(SINGLE_KERNEL_STAGES, key=lambda k: st[k])
    kname, grid = stage_kernel(dom, True, runs["mixed"]["tune"], zeta_1_2_4_16(m.header()), D, N, M, c64)
    roof = {kk: rates[dom][kk] for kk in ("bound", "achieved", "peak", "unit", "frac")}
    roof["traffic"] = measured_traffic(label, kname, grid)
    roof["counters"] = measured_counters(label, kname, grid)
    roof.update(kernel=kname, grid=grid, stage=dom, ms_per_launch=round(st[dom], 4), work_per_launch=rates[dom]["work"])
    roof["stages"] = {k: dict(ms=round(st[k], 4), **{kk: vv for kk, vv in rates[k].items()
                                                     if kk in ("achieved", "frac", "bound", "unit")}) for k in st}
    for k in SINGLE_KERNEL_STAGES:
        kn, g = stage_kernel(k, True, runs["mixed"]["tune"], zeta_1_2_4_16(m.header()), D, N, M, c64)
        roof["stages"][k].update(kernel=kn, traffic=measured_traffic(label, kn, g),
                                 counters=measured_counters(label, kn, g))
    sec = runs["mixed"]["sec"]
    rec = dict(
        value=round(N / sec, 1), ms_per_step=round(sec * 1e3, 4),
        config=dict(workload=label, structures=len(structs), atoms=N, des_len=D, nsparse=M, rcut=rc,
                    pairs=pairs, triplets=triplets, mixed_contraction_fp64=bool(c64),
                    tuning={k: runs[k]["tune"] for k in runs}),
        fp64=dict(value=round(N / runs["fp64"]["sec"], 1), ms_per_step=round(runs["fp64"]["sec"] * 1e3, 4),
                  stages_ms={k: round(v, 4) for k, v in runs["fp64"]["stages"].items()}),
        mixed_stages_ms={k: round(v, 4) for k, v in st.items()},
        mixed_over_fp64=round(runs["fp64"]["sec"] / sec, 3),
        mixed_error_vs_fp64_max={k: float("%.4g" % v) for k, v in err.items()},
        roofline=roof)
    return ctx, rec


def dynamics_bench(ctx, structs, rc):
    """Batched FIRE relaxation and NVT MD of the loaded population (mixed
    precision, graph replay)."""
    natoms = np.array([len(s["species"]) for s in structs])
    ctx.use_graph(True)
    t0 = time.perf_counter()
    rel = ctx.relax(gapx.GRAD | gapx.MIXED, max_steps=200, fmax=0.05, maxstep=0.2)
    t_rel = time.perf_counter() - t0
    relax = dict(steps=rel["steps"], converged=rel["converged"], seconds=round(t_rel, 3),
                 structure_steps_per_s=round(len(structs) * rel["steps"] / t_rel, 1), fmax=0.05, max_steps=200)
    print("bench: population relax %s" % relax, file=sys.stderr, flush=True)
    # NVTBerendsen as the reference's MD driver configures it (1 fs, 300 K,
    # taut 500 fs), from the starting positions with Maxwell-Boltzmann momenta
    ctx.load(structs, rc)
    mass_of = {1: 1.008, 6: 12.011, 8: 15.999, 22: 47.867}
    masses = np.concatenate([[mass_of[int(z)] for z in s["species"]] for s in structs])
    p0 = np.random.default_rng(3).normal(size=(len(masses), 3)) * np.sqrt(masses * 300.0 * gapx.KB)[:, None]
    md_steps = 100
    t0 = time.perf_counter()
    log = ctx.md(masses, md_steps, dt_fs=1.0, temperature=300.0, taut_fs=500.0, momenta=p0,
                 flags=gapx.GRAD | gapx.MIXED)
    t_md = time.perf_counter() - t0
    md = dict(steps=md_steps, seconds=round(t_md, 3), atom_steps_per_s=round(float(natoms.sum()) * md_steps / t_md, 1),
              ms_per_step=round(t_md / md_steps * 1e3, 4), mean_T_end=round(float(log[-1, :, 2].mean()), 1))
    print("bench: population md %s" % md, file=sys.stderr, flush=True)
    return relax, md


def surrogate_bench(m, n_train, n_cand, steps, batch=4096):
    """Surrogate form: label random ternary structures with the GPU GAP (mixed
    precision), train the MLP on the GPU, then pre-screen candidates with it
    and judge its top 5% with the GAP."""
    w = synth.SURROGATE
    workers = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
    t0 = time.perf_counter()
    train = synth.random_ternary(n_train, w["nmin"], w["nmax"], w["density"], w["species"], w["seed"],
                                 workers=workers)
    t_gen = time.perf_counter() - t0
    net = gapx.MLP(seed=0)
    ctx = gapx.Context(m)
    t0 = time.perf_counter()
    s_e = s_e2 = s_g2 = 0.0
    rows = 0
    theta = m.arrays()[0]
    for k in range(0, len(train), 2000):
        ctx.load(train[k:k + 2000], 6.0)
        ctx.run(gapx.GRAD | gapx.MIXED)
        ctx.export_labels(net)
        e = ctx.results()["eatom"]
        g = ctx.features(dedx=True)[1] * theta
        s_e += e.sum()
        s_e2 += (e * e).sum()
        s_g2 += (g * g).sum()
        rows += len(e)
    t_label = time.perf_counter() - t0
    var_e = s_e2 / rows - (s_e / rows) ** 2
    cE, cG = 1.0 / var_e, rows / s_g2
    net.train(1, batch=batch, lr=0.0, cE=cE, cG=cG)  # allocate the work buffers
    t0 = time.perf_counter()
    last = None
    for k in range(0, steps, 500):
        last = net.train(min(500, steps - k), batch=batch, lr=1e-3, cE=cE, cG=cG, seed=k)
    t_train = time.perf_counter() - t0
    # per row: 2 forward + 2 input-gradient + 3 second-order + 4 weight-gradient products of 128 x 128
    train_flops = 11 * 2 * 128 * 128 * batch * steps
    cands = synth.random_ternary(n_cand, w["nmin"], w["nmax"], w["density"], w["species"], w["seed"] + 1,
                                 workers=workers)
    n = np.array([len(c["species"]) for c in cands])
    ctx.load(cands, 6.0)

    def timed(fn, reps=3):
        fn()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t) / reps

    sec_sur = timed(lambda: ctx.run_surrogate(net, gapx.MIXED))
    e_sur = ctx.results()["energy"] / n
    sec_gap = timed(lambda: ctx.run(gapx.MIXED))
    e_gap = ctx.results()["energy"] / n
    sec_gap64 = timed(lambda: ctx.run(gapx.GRAD), reps=1)
    # where a screening run spends its time (per-stage HIP events, eager)
    stages = {}
    for name, fn in (("surrogate", lambda: ctx.run_surrogate(net, gapx.MIXED)), ("gap_mixed_energy", lambda: ctx.run(gapx.MIXED))):
        ctx.timing(True)  # resets the stage sums
        fn()
        ctx.sync()
        stages[name] = {k: round(v[0] / max(v[1], 1), 3) for k, v in ctx.stages().items() if v[1] > 0}
    ctx.timing(False)
    ctx.close()
    net.close()
    k = max(1, int(w["top_fraction"] * len(cands)))
    pick = np.argsort(e_sur)[:k]
    recall = len(set(pick) & set(np.argsort(e_gap)[:k])) / k
    rank = lambda v: np.argsort(np.argsort(v))
    return dict(
        train_structures=n_train, train_atoms=int(rows), generate_s=round(t_gen, 2), label_s=round(t_label, 2),
        train_steps=steps, batch=batch, train_s=round(t_train, 3),
        train_rows_per_s=round(steps * batch / t_train, 1), train_tflops=round(train_flops / t_train / 1e12, 2),
        last_batch_mse_E=float("%.4g" % last[0]), label_var_E=float("%.4g" % var_e),
        candidates=n_cand, candidate_atoms=int(n.sum()),
        surrogate_cands_per_s=round(n_cand / sec_sur, 1), gap_mixed_energy_cands_per_s=round(n_cand / sec_gap, 1),
        gap_fp64_efs_cands_per_s=round(n_cand / sec_gap64, 1),
        speedup_vs_gap_mixed_energy=round(sec_gap / sec_sur, 2), speedup_vs_gap_fp64_efs=round(sec_gap64 / sec_sur, 2),
        target_comparator="the 20x target is read against speedup_vs_gap_fp64_efs: the GAP evaluation the reference "
                          "screens with (FP64 energy, forces and stress, here on the GPU); speedup_vs_gap_mixed_energy "
                          "is against the fastest GAP screen this library has (mixed precision, energy only)",
        stages_ms=stages,
        judged=k, recall_of_gap_top=round(recall, 3),
        spearman=round(float(np.corrcoef(rank(e_gap), rank(e_sur))[0, 1]), 3),
        judged_mean_abs_err_eV_per_atom=float("%.4g" % np.abs(e_sur[pick] - e_gap[pick]).mean()))


def variance_bench(m, ctx, N, reps=10):
    """Predictive variance after an FP64 run of the loaded structure: Q is a
    synthetic symmetric M x M matrix (timing does not depend on its values)."""
    M, D = m.nsparse, m.des_len
    rng = np.random.default_rng(7)
    A = rng.normal(0, 1e-3, (M, M))
    m.set_qmm(0.5 * (A + A.T))
    ctx.run(gapx.GRAD)
    ctx.results()
    ctx.variance()  # workspace
    t = time.perf_counter()
    for _ in range(reps):
        ctx.variance()
    sec = (time.perf_counter() - t) / reps
    m.set_qmm(None)
    Mp, Dp = -(-M // 64) * 64, -(-D // 32) * 32
    flops = 2.0 * N * Mp * Dp + 2.0 * N * Mp * Mp
    return dict(ms=round(sec * 1e3, 4), atoms_per_s=round(N / sec, 1), gemm_tflops=round(flops / sec / 1e12, 2),
                frac_fp64_peak=round(flops / sec / 1e12 / PEAK["fp64"], 3),
                note="wall clock per gapx_ctx_variance call (enqueue, kernels, sync, copy-back); "
                     "flops = 2 N Mp Dp + 2 N Mp^2")


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def _oracle_model(oracle, model):
    return oracle.Model(z=model["z"], w=model["w"], sf_type=model["sf_type"], sf_alpha=model["sf_alpha"],
                        sf_cut=model["sf_cut"], theta=model["theta"], mm=model["mm"], coeff=model["coeff"],
                        sf_zeta=model["sf_zeta"])


def cpu_baseline(model, structs, rc, seconds):
    """The reference FGAP_CALC compiled from its own sources (oracle/_ref;
    one core: its OpenMP directives are commented out) on the first
    structures of the same population until about `seconds` of work, the
    model written as ./gap_parameters in a temporary folder (the reference
    reads its header there); beside it the oracle's sparse FP64 path on the
    same sample with all host cores (up to 16).  Without the compiled
    reference the oracle's rate is the baseline (kind "port")."""
    oracle = _oracle()
    om = _oracle_model(oracle, model)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    ref = None
    k, n_ref, t_ref = 0, 0, 0.0
    if oracle.ref_available():
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "gap_parameters")
            oracle.write_model(path, om)
            while k < len(structs) and t_ref < seconds:
                s = structs[k]
                t0 = time.perf_counter()
                oracle.ref_gap_calc(path, s["species"], s["lat"], s["pos"], om.theta, om.mm, om.coeff, rc, True)
                t_ref += time.perf_counter() - t0
                n_ref += len(s["species"])
                k += 1
        ref = dict(value=n_ref / t_ref, structures=k, atoms=n_ref, seconds=round(t_ref, 2))
    # the oracle (sparse mode, OpenMP over centre atoms) on the same sample,
    # or on the first structures until about `seconds` without the reference
    kk, t_port, n_port = 0, 0.0, 0
    while kk < len(structs) and (kk < k if ref else t_port < seconds):
        s = structs[kk]
        t0 = time.perf_counter()
        oracle.gap_calc(om, s["species"], s["lat"], s["pos"], rc, mode="sparse", nthreads=cores)
        t_port += time.perf_counter() - t0
        n_port += len(s["species"])
        kk += 1
    port = dict(value=n_port / t_port, cores=cores, structures=kk, atoms=n_port, seconds=round(t_port, 2))
    if ref:
        return dict(value=round(ref["value"], 1), unit="atom E+F+S evals/s", cores=1, kind="reference",
                    sample="FGAP_CALC (gappy/libgap/gap_calc.f90, wacsf.f90; amdflang -O3) on the first %d of %d "
                           "population structures (%d atoms), %.1f s; it reads zeta = 1 for every angular line "
                           "(the zeta column is ignored by its list-directed read), the same number of lines and "
                           "so the same cost" % (k, len(structs), n_ref, t_ref),
                    port=dict(value=round(port["value"], 1), cores=cores, kind="port",
                              sample="the oracle's sparse FP64 path on the same structures, %.2f s" % t_port))
    return dict(value=round(port["value"], 1), unit="atom E+F+S evals/s", cores=cores, kind="port",
                sample="the oracle's sparse FP64 path on the first %d population structures (%d atoms), %.1f s"
                       % (kk, n_port, t_port))


def device_copy(nbytes):
    """One device-to-device copy of nbytes through the HIP runtime (the
    runtime's plain copy kernel), for calibrating HBM byte counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:  # 3: device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def load_workload(name):
    """(model dict, structures, rcut) of a workload; its model is built from
    GPU descriptors of separately seeded structures."""
    if name == "population":
        return synth.population_model(gpu_descriptor_fn), synth.population(synth.POPULATION["n_struct"]), \
            synth.POPULATION["rc"]
    return synth.workload_model(name, gpu_descriptor_fn), [synth.workload_structure(name)], synth.WORKLOADS[name]["rc"]


def probe(args):
    """Copy of known size, then evaluations (for rocprofv3 counter passes)."""
    device_copy(1 << 30)
    print("probe: copied %d bytes" % (1 << 30), flush=True)
    model, structs, rc = load_workload(args.workload)
    m = gapx.Model.from_model(model)
    ctx = gapx.Context(m)
    ctx.load(structs, rc)
    # one-piece launches (per-stage timing on), the launches whose per-stage
    # times the benchmark's roofline uses, so the counters match them by grid
    ctx.timing(True)
    for flags, name in ((gapx.GRAD, "fp64"), (gapx.GRAD | gapx.MIXED, "mixed")):
        # every launch shape the autotuner can choose (the descriptor kernels'
        # waves one at a time from the defaults, the contraction's row tile and
        # workgroup target together, as it searches), `--probe` eager
        # evaluations each; no graphs, whose replay the counter passes do not need
        base = dict(waves_fwd=1, waves_bwd=1, gpr_rows=64, gpr_blocks=512)
        shapes = [dict(base, waves_fwd=v) for v in (1, 2, 4)] + [dict(base, waves_bwd=v) for v in (2, 4)]
        for r in ((64, 128) if flags & gapx.MIXED else (64,)):
            shapes += [dict(base, gpr_rows=r, gpr_blocks=b) for b in (256, 512, 768, 1024, 1536, 2048, 3072)]
        for t in shapes:
            ctx.set_tuning(flags, **t)
            for _ in range(args.probe):
                ctx.run(flags)
            ctx.results()
        print("probe: %s launch shapes done" % name, flush=True)
    ctx.close()
    m.close()
    print("probe: done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="population", choices=["population", "ternary4k", "ternary32k", "c64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 32768-atom record")
    ap.add_argument("--no-4k", action="store_true", help="skip the 4096-atom record (and variance, surrogate)")
    ap.add_argument("--no-dynamics", action="store_true", help="skip the population's relaxation and MD")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--surrogate-train", type=int, default=200000,
                    help="training structures (the surrogate form's 200000 by default)")
    ap.add_argument("--surrogate-steps", type=int, default=12000)
    ap.add_argument("--probe", type=int, default=0,
                    help="profiling probe: a 1 GiB device copy (calibrates HBM byte counters), then per precision "
                         "this many eager evaluations at every launch shape the autotuner can choose; no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this evaluator runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if gapx.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.probe:
        return probe(args)
    t0 = time.perf_counter()
    # measured vector-FMA throughput (register-only microkernels)
    ceil = {k: round(v, 1) for k, v in gapx.probe_fma_rates().items()}
    model, structs, rc = load_workload(args.workload)
    print("bench: workload %s ready in %.1f s; measured FMA rates %s TFLOP/s" % (
        args.workload, time.perf_counter() - t0, ceil), file=sys.stderr, flush=True)
    m = gapx.Model.from_model(model)
    ctx, head = measure(m, structs, rc, args.steps, args.warmup, args.workload)
    relax = md = None
    if args.workload == "population" and not args.no_dynamics:
        relax, md = dynamics_bench(ctx, structs, rc)
    ctx.close()
    m.close()

    extra = {}
    for name, skip in (("ternary32k", args.no_large), ("ternary4k", args.no_4k)):
        if skip or name == args.workload:
            continue
        t0 = time.perf_counter()
        emodel, estructs, erc = load_workload(name)
        print("bench: workload %s ready in %.1f s" % (name, time.perf_counter() - t0), file=sys.stderr, flush=True)
        em = gapx.Model.from_model(emodel)
        ectx, rec = measure(em, estructs, erc, args.steps, args.warmup, name)
        if name == "ternary4k":
            rec["variance"] = variance_bench(em, ectx, rec["config"]["atoms"])
            print("bench: variance %s" % rec["variance"], file=sys.stderr, flush=True)
        ectx.close()
        if name == "ternary4k" and not args.no_surrogate:
            extra["surrogate"] = surrogate_bench(em, args.surrogate_train, synth.SURROGATE["candidates"],
                                                 args.surrogate_steps)
            print("bench: surrogate %s" % extra["surrogate"], file=sys.stderr, flush=True)
        em.close()
        extra[name] = rec

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(model, structs, rc, args.cpu_seconds)
        print("bench: cpu baseline %s" % cpu, file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "atom E+F+S evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 descriptors + f16 hi/lo MFMA (f32 accumulate) + f64 norms/reductions",
        "data": "synthetic (seeded structures and model)",
        "config": dict(head["config"], precision="mixed"),
        "fp64": head["fp64"],
        "mixed_over_fp64": head["mixed_over_fp64"],
        "mixed_error_vs_fp64": head["mixed_error_vs_fp64_max"],
        "mixed_stages_ms": head["mixed_stages_ms"],
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "relax": relax,
        "md": md,
        "fma_rates_tflops": dict(ceil, note="measured by gapx_probe_fma_rates on this box: register-only FMA "
                                            "microkernels, FP32 without packed instructions as the kernels "
                                            "are compiled; the roofline peaks stay the published ones"),
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
