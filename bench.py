#!/usr/bin/env python3
"""Benchmark of the fitting core on one MI355X.

Workload (the Stage One random search at real scale): the self-consistent
Lommel-Seeliger ellipsoid with torque (P_PSI, TORQUE, INTERP; 11 parameters),
one synthetic 400-observation single-filter light curve over 3 days,
TIME_STEP 1e-2 d, L limits 48pi/10..48pi/0.1 rad/day, Tmax 10, P_psi range
2..4800 h, and 2^18 independent simplexes (1024 reseed groups of 256), each
stopped after 4800 chi^2 evaluations (or size convergence).  One step is one
launch cycle of the reference's host loop over all 2^18 simplexes
(chi2_gpu + sync, asteroid.c:675-688) with inputs already resident in HBM.

Prints one JSON line: model-point evaluations per second (evaluations times
observations) over the timed steps, fits/s beside it, the roofline of the
fitting kernel (algorithmic FLOPs from asteroid_amd/flops.py over its
HIP-event-timed launch durations) and the CPU oracle timed on a bounded
sample of the same workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (and dense MFMA) peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)

N_OBS, SPAN_DAYS, CURVE_SEED = 400, 3.0, 2024
PPSI_HOURS = (2.0, 4800.0)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(precision):
    from asteroid_amd import synth, variants
    v = variants.ls_torque(precision, time_step=1e-2)
    curve, truth = synth.make_curve(CURVE_SEED, v, N_OBS, SPAN_DAYS, n_filters=1, sigma_range=(0.05, 0.3))
    space = variants.default_model_space(v, ppsi_hours=PPSI_HOURS, reopt=0)
    return v, curve, truth, space


def traffic_from_profile():
    """HBM bytes per launch of the fitting kernel from the committed PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "fit_kernel_stage_one_2p18_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch="):
                return float(line.split("=", 1)[1])
    return None


def cpu_baseline(precision, budget, seconds):
    """The oracle (canonical form, C++) on this host's cores, bounded sample."""
    import oracle
    v, curve, truth, space = make_workload(precision)
    pb = oracle.Problem(space, curve)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    # calibrate with one simplex per core, then size the sample to ~`seconds`
    t0 = time.perf_counter()
    r = pb.run(cores, 1, nstages=1, n_steps=5000, eval_budget=budget, seed=7, form=oracle.CANON, n_threads=cores)
    t1 = time.perf_counter() - t0
    reps = max(1, int(round(seconds / max(t1, 1e-3))))
    nfit = cores * reps
    t0 = time.perf_counter()
    r = pb.run(nfit, 1, nstages=1, n_steps=5000, eval_budget=budget, seed=11, form=oracle.CANON, n_threads=cores)
    dt = time.perf_counter() - t0
    mp = float(r["lane_nev"].sum()) * curve.n_data
    return {"value": mp / dt, "unit": "model-point evals/s", "cores": cores, "kind": "port",
            "sample": f"{nfit} simplexes x {budget} chi2 evals on the same 400-point curve "
                      f"({int(r['lane_nev'].sum())} evals, {dt:.1f} s wall, oracle canonical form, "
                      f"{cores} threads)",
            "fits_per_s": nfit / dt}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "f32simplex"], default="fp64")
    ap.add_argument("--fits", type=int, default=1 << 18)
    ap.add_argument("--budget", type=int, default=4800)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the fits are independent and shard across GPUs "
              "without communication (see DESIGN.md); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from asteroid_amd import abi, api, flops
    prec = {"fp64": abi.PREC_FP64, "f32simplex": abi.PREC_F32SIMPLEX}[args.precision]
    v, curve, truth, space = make_workload(prec)
    group = 256
    n_groups = args.fits // group
    ctx = api.Context(v)
    ctx.set_model_space(space)
    ctx.upload_curve(curve)
    ctx.rng_seed(1)

    def step():
        ctx.reset_best()
        ctx.run(n_groups, group, nstages=1, n_steps=5000, eval_budget=args.budget)

    for w in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"warmup {w + 1}/{args.warmup} done in {time.perf_counter() - t:.2f} s")
    tot_points = tot_evals = 0
    kern_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        e, p = ctx.last_counts()
        ms, nl = ctx.last_timing()
        tot_evals += e
        tot_points += p
        kern_ms += ms
        launches += nl
        log(f"step {k + 1}/{args.steps}: {e} evals, kernel {ms / 1e3:.2f} s in {nl} launches")
    wall = time.perf_counter() - t0
    f, params, dv = ctx.fetch()
    ctx.close()

    value = tot_points / wall
    flop_per_eval = flops.eval_flops(v, curve.mjd)
    flop_total = flop_per_eval * tot_evals
    achieved = flop_total / (kern_ms / 1e3) / 1e12
    traffic = traffic_from_profile()
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
            "kernel": "fit_kernel<double,TORQUE,LS,NF=1>",
            "flop_per_model_point": round(flop_per_eval / curve.n_data, 2),
            "avg_launch_ms": round(kern_ms / max(launches, 1), 3), "launches": launches}
    out = {
        "metric": "model-point evals/s, fits/s at 1e-6 rel chi^2 vs FP64; p50/p99 fit latency",
        "value": value,
        "unit": "model-point evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.PREC_FP64 else "f64 model, f32 simplex",
        "data": "synthetic",
        "config": {"workload": "stage-one random search (LS ellipsoid + torque, P_PSI, INTERP)",
                   "n_obs": N_OBS, "span_days": SPAN_DAYS, "filters": 1, "simplexes": args.fits,
                   "evals_per_simplex": args.budget, "group_size": group, "time_step_days": 1e-2,
                   "precision": args.precision},
        "fits_per_s": args.fits * args.steps / wall,
        "evals_per_s": tot_evals / wall,
        "best_chi2": float(np.min(f)),
        "roofline": roof,
    }
    if not args.no_cpu:
        log("timing the CPU oracle sample ...")
        out["cpu_baseline"] = cpu_baseline(prec, args.budget, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
