"""Benchmark: atom energy+force+stress evaluations per second on one MI355X.

Workload (default): 4096-atom ternary structure (Z 1/8/22, weights = Z),
triclinic cell at 0.07 atoms/A^3, rcut 6 A, 32 radial + 32 angular wACSF
lines (D = 128), M = 4000 sparse points, synthetic seeded data and a
synthetic model built from GPU descriptors of seeded pool structures.

A step is one full E+F+S evaluation of the structure from positions already
in HBM: neighbour list, descriptors, contraction, backward pass, forces and
stress.  Prints one JSON line (see the fields below).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload ternary4k]

The JSON line also carries "population": the batched form (512 uneven
structures, 8-256 atoms, D = 128, M = 4000), autotuned and replayed from a
HIP graph, in mixed precision and in FP64, with the per-structure mixed
error against FP64 (skip with --no-population).
"""
import argparse
import json
import os
import sys
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


# Algorithmic work per stage (documented in DESIGN.md).  Descriptor flops are
# counted per (pair, radial line) and (triplet, angular line) from the
# arithmetic each term needs, a transcendental counted as one flop.
FLOPS_RAD_FWD, FLOPS_RAD_BWD = 8, 12          # per pair and radial line
FLOPS_ANG_FWD, FLOPS_ANG_BWD = 10, 17         # per triplet and angular line
FLOPS_TRIP_FWD, FLOPS_TRIP_BWD = 30, 60       # per triplet (geometry, cutoff products)
BYTES_NEIGHBOR = 4 + 12 + 32 + 4              # per pair: index, image, (dx,dy,dz,r), reverse slot


def stage_rates(st, N, D, M, pairs, triplets, m, mixed):
    hdr = m.header()
    n_ang = int(np.sum((hdr["sf_type"] == 2) | (hdr["sf_type"] == 4)))
    n_rad = int(np.sum((hdr["sf_type"] == 1) | (hdr["sf_type"] == 3)))
    vpeak = PEAK["fp32"] if mixed else PEAK["fp64"]
    mpeak = PEAK["fp16"] if mixed else PEAK["fp64"]
    work = {
        "neighbor_list": ("hbm", pairs * BYTES_NEIGHBOR * 2 + N * 64, 8000.0, "GB/s", 1e9),
        "descriptors_fwd": ("valu", pairs * n_rad * FLOPS_RAD_FWD + triplets * (n_ang * FLOPS_ANG_FWD + FLOPS_TRIP_FWD),
                            vpeak, "TFLOP/s", 1e12),
        "gpr_contraction": ("mfma", 4.0 * D * M * N, mpeak, "TFLOP/s", 1e12),
        "gpr_finish": ("hbm", N * D * 8 * 3.0, 8000.0, "GB/s", 1e9),
        "descriptors_bwd": ("valu", pairs * n_rad * FLOPS_RAD_BWD + triplets * (n_ang * FLOPS_ANG_BWD + FLOPS_TRIP_BWD),
                            vpeak, "TFLOP/s", 1e12),
        "forces_stress": ("hbm", pairs * 24 * 2 + N * 80, 8000.0, "GB/s", 1e9),
    }
    out = {}
    for k, ms in st.items():
        bound, amount, peak, unit, scale = work[k]
        ach = amount / (ms * 1e-3) / scale if ms > 0 else 0.0
        out[k] = dict(bound=bound, achieved=round(ach, 3), peak=peak, unit=unit, frac=round(ach / peak, 4),
                      traffic=None, work=float(amount))
    return out


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
    ctx.timing(True)
    for _ in range(steps):
        ctx.run(flags)
    ctx.sync()
    st = ctx.stages()
    ctx.timing(False)
    return {k: (v[0] / max(v[1], 1)) for k, v in st.items()}


def population_bench(steps, warmup):
    """Batched population in both precisions: autotuned, graph replay."""
    t0 = time.perf_counter()
    structs = synth.population(synth.POPULATION["n_struct"])
    model = synth.population_model(gpu_descriptor_fn)
    print("bench: population ready in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    m = gapx.Model.from_model(model)
    ctx = gapx.Context(m)
    ctx.load(structs, synth.POPULATION["rc"])
    natoms = np.array([len(s["species"]) for s in structs])
    out, res = {}, {}
    for name, flags in (("fp64", gapx.GRAD), ("mixed", gapx.GRAD | gapx.MIXED)):
        tune = ctx.autotune(flags)
        ctx.use_graph(True)
        sec = time_runs(ctx, flags, steps, warmup)
        ctx.use_graph(False)
        st = stage_times(ctx, flags, max(3, steps // 4))
        ctx.use_graph(True)
        ctx.run(flags)
        res[name] = ctx.results()
        out[name] = dict(value=round(float(natoms.sum()) / sec, 1), ms_per_step=round(sec * 1e3, 4), tuning=tune,
                         stages_ms={k: round(v, 4) for k, v in st.items()})
        print("bench: population %s %.3f ms/step tuning %s" % (name, sec * 1e3, tune), file=sys.stderr, flush=True)
    ctx.close()
    m.close()
    r64, rmx = res["fp64"], res["mixed"]
    err = dict(energy_meV_per_atom=float(np.abs((rmx["energy"] - r64["energy"]) / natoms).max()) * 1e3,
               force_meV_per_A=float(np.abs(rmx["forces"] - r64["forces"]).max()) * 1e3,
               stress_GPa=float(np.abs(rmx["stress"] - r64["stress"]).max()))
    return dict(structures=len(structs), atoms=int(natoms.sum()), unit="atom E+F+S evals/s", **out,
                mixed_over_fp64=round(out["mixed"]["value"] / out["fp64"]["value"], 3),
                mixed_error_vs_fp64_max={k: float("%.4g" % v) for k, v in err.items()})


def cpu_baseline(model, s, rc, seconds):
    """The CPU oracle (sparse FP64 path, OpenMP over centre atoms) on a
    bounded sample: the first n centre atoms of the same structure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    om = oracle.Model(z=model["z"], w=model["w"], sf_type=model["sf_type"], sf_alpha=model["sf_alpha"],
                      sf_cut=model["sf_cut"], theta=model["theta"], mm=model["mm"], coeff=model["coeff"],
                      sf_zeta=model["sf_zeta"])
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 16))
    n = 64
    t = 0.0
    while True:
        t0 = time.perf_counter()
        oracle.gap_calc_range(om, s["species"], s["lat"], s["pos"], rc, 0, n, nthreads=cores)
        t = time.perf_counter() - t0
        if t > seconds / 4 or n >= len(s["species"]):
            break
        n = min(len(s["species"]), int(n * max(2.0, seconds / 4 / max(t, 1e-3))))
    return dict(value=n / t, unit="atom E+F+S evals/s", cores=cores, kind="port",
                sample="first %d of %d centre atoms of the same structure, %.1f s" % (n, len(s["species"]), t))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ternary4k")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-population", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this evaluator runs on one GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if gapx.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    w = synth.WORKLOADS[args.workload]
    rc = w["rc"]
    t0 = time.perf_counter()
    model = synth.workload_model(args.workload, gpu_descriptor_fn)
    s = synth.workload_structure(args.workload)
    print("bench: workload %s ready in %.1f s" % (args.workload, time.perf_counter() - t0), file=sys.stderr,
          flush=True)
    m = gapx.Model.from_model(model)
    ctx = gapx.Context(m)
    ctx.load([s], rc)
    N = len(s["species"])
    D, M = m.des_len, m.nsparse

    runs = {}
    for name, flags in (("fp64", gapx.GRAD), ("mixed", gapx.GRAD | gapx.MIXED)):
        ctx.use_graph(True)  # replay the captured launch sequence
        sec = time_runs(ctx, flags, args.steps, args.warmup)
        ctx.use_graph(False)  # per-stage events need eager launches
        st = stage_times(ctx, flags, max(3, args.steps // 4))
        ctx.run(flags)
        res = ctx.results()
        runs[name] = dict(sec=sec, stages=st, res=res)
        print("bench: %s %.3f ms/step; stages %s" % (name, sec * 1e3, {k: round(v, 4) for k, v in st.items()}),
              file=sys.stderr, flush=True)
    pairs, triplets = ctx.counts()
    r64, rmx = runs["fp64"]["res"], runs["mixed"]["res"]
    err = dict(energy_meV_per_atom=abs(rmx["energy"][0] - r64["energy"][0]) / N * 1e3,
               force_meV_per_A=float(np.abs(rmx["forces"] - r64["forces"]).max()) * 1e3,
               stress_GPa=float(np.abs(rmx["stress"] - r64["stress"]).max()))

    sec = runs["mixed"]["sec"]
    value = N / sec
    st = runs["mixed"]["stages"]
    rates = stage_rates(st, N, D, M, pairs, triplets, m, mixed=True)
    dom = max(st, key=lambda k: st[k])
    roof = dict(rates[dom])
    roof["kernel"] = dom
    roof["stages"] = {k: dict(ms=round(st[k], 4), **{kk: vv for kk, vv in rates[k].items()
                                                     if kk in ("achieved", "frac", "bound")}) for k in st}
    ctx.close()
    m.close()
    popr = None
    if not args.no_population:
        popr = population_bench(args.steps, args.warmup)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(model, s, rc, args.cpu_seconds)
    out = {
        "metric": "atom E+F+stress evals/sec",
        "value": round(value, 1),
        "unit": "atom E+F+S evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 descriptors + f16 hi/lo MFMA (f32 accumulate) + f64 norms/reductions",
        "data": "synthetic (seeded structures and model)",
        "config": {"workload": args.workload, "atoms": N, "des_len": D, "nsparse": M, "rcut": rc,
                   "precision": "mixed", "pairs": pairs, "triplets": triplets},
        "fp64": {"value": round(N / runs["fp64"]["sec"], 1), "ms_per_step": round(runs["fp64"]["sec"] * 1e3, 4),
                 "stages_ms": {k: round(v, 4) for k, v in runs["fp64"]["stages"].items()}},
        "mixed_over_fp64": round(runs["fp64"]["sec"] / sec, 3),
        "mixed_error_vs_fp64": {k: float("%.4g" % v) for k, v in err.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "population": popr,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
