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

Both precision forms run the same workload: the mixed form (FP64 Euler-angle
integration and chi^2 sums, FP32 geometry, brightness and simplex) is the
headline, the plain FP64 form is reported beside it.  The mixed run's
per-group best models are re-evaluated with the FP64 form to state its
relative chi^2 error against the 1e-6 gate.

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
# Both forms stop on the evaluation budget: the simplex-size test uses the
# FP64 threshold (SIZE_MIN of the ACC build, asteroid.h:233-237) for both.
SIZE_MIN = 1e-10
MIXED = 2
PREC_NAMES = {0: "fp64", 1: "f32simplex", 2: "mixed"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(precision):
    from asteroid_amd import synth, variants
    v = variants.ls_torque(precision, time_step=1e-2)
    curve, truth = synth.make_curve(CURVE_SEED, v, N_OBS, SPAN_DAYS, n_filters=1, sigma_range=(0.05, 0.3))
    space = variants.default_model_space(v, ppsi_hours=PPSI_HOURS, reopt=0)
    return v, curve, truth, space


def traffic_from_profile(form):
    """HBM bytes per launch of the fitting kernel from the committed PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", f"fit_kernel_stage_one_2p18_{form}_traffic.txt")
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


def run_form(prec, args):
    """Warm up, then time args.steps Stage One steps of one precision form."""
    from asteroid_amd import api, flops
    v, curve, truth, space = make_workload(prec)
    group = 256
    n_groups = args.fits // group
    ctx = api.Context(v)
    ctx.set_model_space(space)
    ctx.upload_curve(curve)
    ctx.rng_seed(1)

    def step():
        ctx.reset_best()
        ctx.run(n_groups, group, nstages=1, n_steps=5000, eval_budget=args.budget, size_min=SIZE_MIN)

    name = PREC_NAMES[prec]
    for w in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"{name}: warmup {w + 1}/{args.warmup} done in {time.perf_counter() - t:.2f} s")
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
        log(f"{name}: step {k + 1}/{args.steps}: {e} evals, kernel {ms / 1e3:.2f} s in {nl} launches")
    wall = time.perf_counter() - t0
    f, params, dv = ctx.fetch()
    ctx.close()
    flop_per_eval = flops.eval_flops(v, curve.mjd)
    achieved = flop_per_eval * tot_evals / (kern_ms / 1e3) / 1e12
    if prec == MIXED:
        f64, f32 = flops.eval_flops_mixed(v, curve.mjd)
        # peak of this op mix: FP64 FLOPs at the FP64 rate, FP32 FLOPs at the FP32 rate
        peak = (f64 + f32) / (f64 / FP64_PEAK_TFLOPS + f32 / FP32_PEAK_TFLOPS)
        kernel = "fit_kernel<float,TORQUE,LS,NF=1,mixed>"
        split = {"fp64_flop_per_model_point": round(f64 / curve.n_data, 2),
                 "fp32_flop_per_model_point": round(f32 / curve.n_data, 2)}
    else:
        peak = FP64_PEAK_TFLOPS
        kernel = "fit_kernel<double,TORQUE,LS,NF=1>"
        split = {}
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic_from_profile(name),
            "kernel": kernel, "flop_per_model_point": round(flop_per_eval / curve.n_data, 2), **split,
            "avg_launch_ms": round(kern_ms / max(launches, 1), 3), "launches": launches}
    return {"value": tot_points / wall, "wall": wall, "ms_per_step": wall / args.steps * 1e3,
            "fits_per_s": args.fits * args.steps / wall, "evals_per_s": tot_evals / wall,
            "evals_per_fit": tot_evals / (args.fits * args.steps), "best_f": f, "best_params": params,
            "roofline": roof}


def gate_at_best_fits(res_mx):
    """FP64 chi^2 (GPU FP64 form, bitwise the oracle's) of the mixed run's
    per-group best models against the chi^2 the mixed run reported."""
    from asteroid_amd import abi, api
    v, curve, truth, space = make_workload(abi.PREC_FP64)
    ok = res_mx["best_f"] < 1e29
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        c64, _ = ctx.chi2_batch(res_mx["best_params"][ok])
    rel = np.abs(res_mx["best_f"][ok] - c64) / c64
    return {"max_rel_chi2_err": float(rel.max()), "median_rel_chi2_err": float(np.median(rel)),
            "n_models": int(ok.sum()), "tolerance": 1e-6, "pass": bool(rel.max() <= 1e-6 + 2.0 ** -23)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64", "f32simplex"], default="both")
    ap.add_argument("--fits", type=int, default=1 << 18)
    ap.add_argument("--budget", type=int, default=4800)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the fits are independent and shard across GPUs "
              "without communication (see DESIGN.md); run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from asteroid_amd import abi
    forms = {"both": [MIXED, abi.PREC_FP64], "mixed": [MIXED], "fp64": [abi.PREC_FP64],
             "f32simplex": [abi.PREC_F32SIMPLEX]}[args.precision]
    results = {prec: run_form(prec, args) for prec in forms}
    head_prec = forms[0]
    head = results[head_prec]
    dtype = {abi.PREC_FP64: "f64", abi.PREC_F32SIMPLEX: "f64 model, f32 simplex",
             MIXED: "f64 Euler-angle integration and chi^2 sums, f32 geometry and brightness, f32 simplex"}
    out = {
        "metric": "model-point evals/s, fits/s at 1e-6 rel chi^2 vs FP64; p50/p99 fit latency",
        "value": head["value"],
        "unit": "model-point evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype[head_prec],
        "data": "synthetic",
        "config": {"workload": "stage-one random search (LS ellipsoid + torque, P_PSI, INTERP)",
                   "n_obs": N_OBS, "span_days": SPAN_DAYS, "filters": 1, "simplexes": args.fits,
                   "evals_per_simplex": args.budget, "size_min": SIZE_MIN, "group_size": 256,
                   "time_step_days": 1e-2, "precision": PREC_NAMES[head_prec]},
        "fits_per_s": head["fits_per_s"],
        "evals_per_s": head["evals_per_s"],
        "evals_per_fit": round(head["evals_per_fit"], 1),
        "best_chi2": float(np.min(head["best_f"])),
        "roofline": head["roofline"],
    }
    if MIXED in results:
        out["precision_gate"] = gate_at_best_fits(results[MIXED])
    for prec, r in results.items():
        if prec == head_prec:
            continue
        out[PREC_NAMES[prec]] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                                 "fits_per_s": r["fits_per_s"], "evals_per_fit": round(r["evals_per_fit"], 1),
                                 "best_chi2": float(np.min(r["best_f"])), "roofline": r["roofline"]}
    if not args.no_cpu:
        log("timing the CPU oracle sample ...")
        out["cpu_baseline"] = cpu_baseline(abi.PREC_FP64, args.budget, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
