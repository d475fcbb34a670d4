"""Benchmark: atom energy+force+stress evaluations per second on one MI355X.

Workload (default): 4096-atom ternary structure (Z 1/8/22, weights = Z),
triclinic cell at 0.07 atoms/A^3, rcut 6 A, 32 radial + 32 angular wACSF
lines (D = 128), M = 4000 sparse points, synthetic seeded data and a
synthetic model built from GPU descriptors of seeded pool structures.

A step is one full E+F+S evaluation of the structure from positions already
in HBM: neighbour list, descriptors, contraction, backward pass, forces and
stress.  Prints one JSON line (see the fields below).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--workload ternary4k]
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

    flags64 = gapx.GRAD
    sec64 = time_runs(ctx, flags64, args.steps, args.warmup)
    st64 = stage_times(ctx, flags64, max(3, args.steps // 4))
    ctx.run(flags64)
    r64 = ctx.results()
    print("bench: fp64 %.3f ms/step; stages %s" % (sec64 * 1e3, {k: round(v, 4) for k, v in st64.items()}),
          file=sys.stderr, flush=True)

    value = N / sec64
    ms_per_step = sec64 * 1e3
    # roofline of the contraction kernel: 4 D M flops per atom (S = xt mt^T and W mt)
    flops = 4.0 * D * M * N
    t_con = st64["gpr_contraction"] * 1e-3
    con_tf = flops / t_con / 1e12
    roof = dict(bound="mfma", achieved=round(con_tf, 3), peak=PEAK["fp64"], unit="TFLOP/s",
                frac=round(con_tf / PEAK["fp64"], 4), traffic=None, kernel="k_gpr_f64 (gpr_contraction)",
                stage_ms={k: round(v, 4) for k, v in st64.items()})
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
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded structures and model)",
        "config": {"workload": args.workload, "atoms": N, "des_len": D, "nsparse": M, "rcut": rc,
                   "precision": "fp64"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "energy_eV": float(r64["energy"][0]),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
