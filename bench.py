#!/usr/bin/env python3
"""Benchmark of the fitting core on one MI355X.

Default workload (the Stage One random search at real scale): the
self-consistent Lommel-Seeliger ellipsoid with torque (P_PSI, TORQUE, INTERP;
11 parameters), one synthetic 400-observation single-filter light curve over
3 days, TIME_STEP 1e-2 d, L limits 48pi/10..48pi/0.1 rad/day, Tmax 10, P_psi
range 2..4800 h, and 2^18 simplexes in flight (1024 workgroups of 256), each
fit stopped after 4800 chi^2 evaluations; a lane whose fit ends starts the
next unstarted fit from the device queue.  One step is one launch of the
fitting kernel (--ticks evaluation rounds) over the 2^18 resident simplexes,
the pass of the reference's host loop (chi2_gpu + sync, asteroid.c:675-688),
with inputs already resident in HBM.  One complete search of 2^18 fits is
run after the steps for the fits/s of a whole search and the per-group best
models.

Both precision forms run the same workload: the mixed form (FP64 Euler-angle
integration, axes and chi^2 sums, FP32 brightness and simplex) is the
headline, the plain FP64 form is reported beside it, and the mixed run's
per-group best models are re-evaluated with the FP64 form to state its
relative chi^2 error against the 1e-6 gate.  A short serving run (Poisson
arrivals of batched-style fit requests, mixed form) adds the p50/p99 fit
latency.

Prints one JSON line: model-point evaluations per second (evaluations times
observations) over the timed steps, fits/s beside it, the roofline of the
fitting kernel (algorithmic FLOPs from asteroid_amd/flops.py over its
HIP-event-timed launch durations) and the CPU oracle timed on a bounded
sample of the same workload on this host's cores.

--workload small|reopt|batched|serving|plot|rmsd|variants measures the other
configurations and paths, one JSON line each: the plain FP64 small search
with its full CPU run, the Stage Two BC reoptimisation, batched fits of many
curves, the serving sweep, the -plot model curves, the RMSD intervals, and
the Stage One search for every built model variant.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector peak
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
METRIC = "model-point evals/s, fits/s at 1e-6 rel chi^2 vs FP64; p50/p99 fit latency"
# Both forms stop on the evaluation budget: the simplex-size test uses the
# FP64 threshold (SIZE_MIN of the ACC build, asteroid.h:233-237) for both.
SIZE_MIN = 1e-10
FP64, F32SIMPLEX, MIXED = 0, 1, 2
PREC_NAMES = {FP64: "fp64", F32SIMPLEX: "f32simplex", MIXED: "mixed"}
DTYPE = {FP64: "f64", F32SIMPLEX: "f64 model, f32 simplex",
         MIXED: "f64 Euler-angle integration, axes and chi^2 sums; f32 brightness and simplex"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


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


def peak_of(prec, variant, mjd, seg_t=None):
    """Vector peak of the form's op mix: FP64 FLOPs at the FP64 rate, FP32 FLOPs at the FP32 rate."""
    from asteroid_amd import flops
    if prec != MIXED:
        return FP64_PEAK_TFLOPS, {}
    f64, f32 = flops.eval_flops_mixed(variant, mjd, seg_t=seg_t)
    return ((f64 + f32) / (f64 / FP64_PEAK_TFLOPS + f32 / FP32_PEAK_TFLOPS),
            {"fp64_flop_per_model_point": round(f64 / len(mjd), 2),
             "fp32_flop_per_model_point": round(f32 / len(mjd), 2)})


def base_line(value, steps, warmup, ms_per_step, dtype, config):
    return {"metric": METRIC, "value": value, "unit": "model-point evals/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


# ------------------------------------------------------------- stage one
def cpu_baseline(budget, seconds):
    """The oracle (canonical form, C++) on this host's cores, bounded sample."""
    import oracle
    from asteroid_amd import workloads
    v, curve, truth, space = workloads.stage_one(FP64)
    pb = oracle.Problem(space, curve)
    n = cores()
    # calibrate with one simplex per core, then size the sample to ~`seconds`
    t0 = time.perf_counter()
    pb.run(n, 1, nstages=1, n_steps=5000, eval_budget=budget, seed=7, form=oracle.CANON, n_threads=n)
    t1 = time.perf_counter() - t0
    nfit = n * max(1, int(round(seconds / max(t1, 1e-3))))
    t0 = time.perf_counter()
    r = pb.run(nfit, 1, nstages=1, n_steps=5000, eval_budget=budget, seed=11, form=oracle.CANON, n_threads=n)
    dt = time.perf_counter() - t0
    mp = float(r["lane_nev"].sum()) * curve.n_data
    return {"value": mp / dt, "unit": "model-point evals/s", "cores": n, "kind": "port",
            "sample": f"{nfit} simplexes x {budget} chi2 evals on the same 400-point curve "
                      f"({int(r['lane_nev'].sum())} evals, {dt:.1f} s wall, oracle canonical form, {n} threads)",
            "fits_per_s": nfit / dt}


def run_stage_one(prec, args):
    """Warm up, then time args.steps Stage One steps of one precision form.

    The search keeps 2^18 simplexes in flight (1024 workgroups of 256 lanes);
    each fit has the 4800-evaluation budget, and a lane whose fit ends starts
    the job's next unstarted fit from the device queue.  One step is one
    launch of the fitting kernel over the 2^18 resident simplexes
    (args.ticks evaluation rounds).  Afterwards one complete search of 2^18
    fits (every fit started at once, run to its end) gives the fits/s of a
    whole search and the per-group best models for the precision checks."""
    from asteroid_amd import api, flops, workloads
    v, curve, truth, space = workloads.stage_one(prec)
    group = 256
    n_groups = args.fits // group
    ctx = api.Context(v)
    ctx.set_model_space(space)
    ctx.upload_curve(curve)
    if args.waves:
        ctx.set_waves(v, 1, args.waves)
        tune = [{"waves": args.waves, "set_by_hand": True}]
    elif args.autotune:
        # one launch per register budget and pass, of args.ticks x n_obs
        # model-points of work (about 50 evaluation rounds here: the launch size
        # also counts each RK4 step as ~1.9 observations)
        tune = ctx.autotune([dict(space=space, curve=curve, n_fits=args.fits, n_steps=5000, eval_budget=args.budget,
                                  size_min=SIZE_MIN, seed=1)],
                            work_points_per_launch=float(args.ticks * len(curve.mjd)), cycles=1)
    else:
        tune = []
    ctx.rng_seed(1)
    name = PREC_NAMES[prec]
    # enough fits in the queue that it cannot run dry during the steps
    rounds = int(np.ceil((args.warmup + args.steps) * args.ticks / (0.8 * args.budget))) + 2
    run = dict(nstages=1, n_steps=5000, eval_budget=args.budget, size_min=SIZE_MIN, use_graphs=args.graphs,
               ticks_per_launch=args.ticks)
    ctx.run_begin(n_groups * rounds, group, resident_groups=n_groups, **run)
    for w in range(args.warmup):
        t = time.perf_counter()
        ctx.run_continue(1)
        log(f"{name}: warmup {w + 1}/{args.warmup} done in {time.perf_counter() - t:.2f} s")
    e0, _ = ctx.last_counts()
    ms0, nl0 = ctx.last_timing()
    _, done0, _ = ctx.run_status()
    t0 = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        if ctx.run_continue(1):
            raise RuntimeError("the fit queue ran dry inside the timed steps")
        log(f"{name}: step {k + 1}/{args.steps} in {time.perf_counter() - t:.3f} s")
    wall = time.perf_counter() - t0
    e1, _ = ctx.last_counts()
    ms1, nl1 = ctx.last_timing()
    _, done1, _ = ctx.run_status()
    evals, kern_ms, launches = e1 - e0, ms1 - ms0, nl1 - nl0
    # one complete search: 2^18 fits started together and run to their end
    ctx.run_cancel()
    ctx.rng_seed(2)
    t1 = time.perf_counter()
    ctx.run(n_groups, group, **run)
    full_wall = time.perf_counter() - t1
    fe, fp = ctx.last_counts()
    fms, fnl = ctx.last_timing()
    f, params, dv = ctx.fetch()
    ctx.close()
    flop_per_eval = flops.eval_flops(v, curve.mjd)
    achieved = flop_per_eval * evals / (kern_ms / 1e3) / 1e12
    peak, split = peak_of(prec, v, curve.mjd)
    kernel = {MIXED: "fit_kernel<float,TORQUE,LS,NF=1,mixed>", FP64: "fit_kernel<double,TORQUE,LS,NF=1>",
              F32SIMPLEX: "fit_kernel<float,TORQUE,LS,NF=1>"}[prec]
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic_from_profile(name), "kernel": kernel,
            "flop_per_model_point": round(flop_per_eval / curve.n_data, 2), **split,
            "avg_launch_ms": round(kern_ms / max(launches, 1), 3), "launches": launches,
            "evals_per_launch": int(evals / max(launches, 1)),
            "waves_per_simd": tune[0]["waves"] if tune else 4}
    full = {"wall_s": round(full_wall, 3), "fits_per_s": args.fits / full_wall,
            "model_points_per_s": fp / full_wall, "evals_per_fit": round(fe / args.fits, 1),
            "kernel_ms": round(fms, 1), "launches": fnl,
            "roofline_frac": round(flop_per_eval * fe / (fms / 1e3) / 1e12 / peak, 4),
            "best_chi2": float(np.min(f))}
    log(f"{name}: {evals * curve.n_data / wall:.3e} model-points/s over the steps, "
        f"{achieved:.2f} TFLOP/s ({achieved / peak:.1%}); complete search {full_wall:.2f} s")
    # in steady state every lane is inside some fit, so fits end at the
    # evaluation rate over the evaluations a fit takes (the complete search's
    # mean); the fits that happen to end inside the few timed launches are
    # reported beside it
    return {"value": evals * curve.n_data / wall, "wall": wall, "ms_per_step": wall / args.steps * 1e3,
            "fits_per_s": evals / wall / (fe / args.fits), "fits_ended_in_steps": int(done1 - done0),
            "evals_per_s": evals / wall, "best_f": f, "best_params": params, "roofline": roof, "autotune": tune,
            "complete_search": full}


def complete_search(prec, args, seed=2):
    """One complete Stage One search of 2^18 fits in one precision form (no
    timed steps): per-group best models for the precision comparisons."""
    from asteroid_amd import api, workloads
    v, curve, truth, space = workloads.stage_one(prec)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(seed)
        t0 = time.perf_counter()
        ctx.run(args.fits // 256, 256, nstages=1, n_steps=5000, eval_budget=args.budget, size_min=SIZE_MIN,
                use_graphs=args.graphs)
        wall = time.perf_counter() - t0
        f, params, dv = ctx.fetch()
    return {"best_f": f, "best_params": params, "wall_s": wall}


def gate_stage_one(res_mx):
    """FP64 chi^2 (GPU FP64 form, bitwise the oracle's) of the mixed run's
    per-group best models against the chi^2 the mixed run reported."""
    from asteroid_amd import api, workloads
    v, curve, truth, space = workloads.stage_one(FP64)
    ok = res_mx["best_f"] < 1e29
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        c64, _ = ctx.chi2_batch(res_mx["best_params"][ok])
    rel = np.abs(res_mx["best_f"][ok] - c64) / c64
    return {"max_rel_chi2_err": float(rel.max()), "median_rel_chi2_err": float(np.median(rel)),
            "n_models": int(ok.sum()), "tolerance": 1e-6, "pass": bool(rel.max() <= 1e-6 + 2.0 ** -23)}


def fit_errors(res_mx, res_64, truth):
    """The fits of two forms side by side: the best model of each complete
    search (chi^2 and parameters), and per group the difference of the two
    forms' best chi^2.  Against the FP64 form (double simplex) the mixed
    form's float simplex takes other paths from the first steps on, so the
    two searches are compared by their answers; against the float-simplex
    form (the same float simplex, FP64 evaluator) the paths are the same until
    a comparison of two nearly equal chi^2 flips."""
    b_mx, b_64 = int(np.argmin(res_mx["best_f"])), int(np.argmin(res_64["best_f"]))
    p_mx, p_64 = res_mx["best_params"][b_mx], res_64["best_params"][b_64]
    rel = np.abs(p_mx - p_64) / np.maximum(np.abs(p_64), 1e-300)
    ok = (res_mx["best_f"] < 1e29) & (res_64["best_f"] < 1e29)
    g = np.abs(res_mx["best_f"][ok] - res_64["best_f"][ok]) / res_64["best_f"][ok]
    return {"best_chi2_mixed": float(res_mx["best_f"][b_mx]), "best_chi2_fp64": float(res_64["best_f"][b_64]),
            "best_rel_chi2_diff": float(abs(res_mx["best_f"][b_mx] - res_64["best_f"][b_64]) / res_64["best_f"][b_64]),
            "best_params_max_rel_diff": float(rel.max()), "best_params_median_rel_diff": float(np.median(rel)),
            "best_params_mixed": [float(x) for x in p_mx], "best_params_fp64": [float(x) for x in p_64],
            "truth": [float(x) for x in truth],
            "per_group_rel_chi2_diff_median": float(np.median(g)), "per_group_rel_chi2_diff_p90": float(np.percentile(g, 90)),
            "per_group_same_model_1e-4": float(np.mean(np.all(
                np.abs(res_mx["best_params"][ok] - res_64["best_params"][ok])
                <= 1e-4 * np.maximum(np.abs(res_64["best_params"][ok]), 1e-3), axis=1)))}


def request_stream(prec, rate, duration, budget):
    """A Poisson request stream.  Generated before the process touches the
    GPU: the curve generator forks worker processes."""
    from asteroid_amd import workloads
    times, jobs, made = workloads.requests(prec, rate, duration, eval_budget=budget)
    return rate, times, jobs


def serving_leg(prec, stream, max_resident_fits=1 << 18):
    """One Poisson request stream through the serving API; latency percentiles."""
    from asteroid_amd import api, variants
    rate, times, jobs = stream
    ctx = api.Context(variants.ls_torque(prec))
    ctx.serve_start(max_resident_fits=max_resident_fits)
    ids = []
    t0 = time.perf_counter()
    for t, j in zip(times, jobs):
        dt = t0 + t - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(ctx.submit(**j))
    res = ctx.wait_all(ids, timeout_s=900)
    wall = time.perf_counter() - t0
    st = ctx.serve_stats()
    ctx.serve_stop()
    ctx.close()
    lat = np.array([res[i]["latency_ms"] for i in ids])
    fits = sum(j["n_fits"] for j in jobs)
    mp = sum(res[i]["model_points"] for i in ids)
    return {"rate_req_per_s": rate, "requests": len(ids), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "fits_per_s": fits / wall, "model_points_per_s": mp / wall,
            "wall_s": wall, "launches": st["launches"], "ok": all(res[i]["status"] == 0 for i in ids),
            "lat_ms": lat}, jobs, times


def main_stage_one(args):
    t_start = time.perf_counter()

    def time_left():
        return args.deadline - (time.perf_counter() - t_start)

    forms = {"both": [MIXED, FP64], "mixed": [MIXED], "fp64": [FP64], "f32simplex": [F32SIMPLEX]}[args.precision]
    stream = None if args.no_serving else request_stream(forms[0], 30.0, 4.0, 200)
    results = {}
    skipped = []
    for prec in forms:
        if results and time_left() < 120:
            skipped.append(PREC_NAMES[prec])
            continue
        results[prec] = run_stage_one(prec, args)
    head_prec = forms[0]
    head = results[head_prec]
    out = base_line(head["value"], args.steps, args.warmup, head["ms_per_step"], DTYPE[head_prec],
                    {"workload": "stage-one random search (LS ellipsoid + torque, P_PSI, INTERP): 2^18 simplexes "
                                 "in flight, 4800-evaluation budget per fit, a lane starting the next fit from "
                                 "the device queue when its fit ends; one step = one fitting-kernel launch of "
                                 f"{args.ticks} evaluation rounds over the 2^18 resident simplexes",
                     "n_obs": 400, "span_days": 3.0, "filters": 1, "simplexes": args.fits,
                     "evals_per_simplex": args.budget, "size_min": SIZE_MIN, "group_size": 256,
                     "ticks_per_step": args.ticks, "time_step_days": 1e-2, "precision": PREC_NAMES[head_prec],
                     "hip_graphs": args.graphs})
    out.update({"fits_per_s": head["fits_per_s"], "fits_ended_in_steps": head["fits_ended_in_steps"],
                "evals_per_s": head["evals_per_s"],
                "best_chi2": float(np.min(head["best_f"])), "roofline": head["roofline"],
                "complete_search": head["complete_search"], "autotune": head["autotune"]})
    if MIXED in results:
        out["precision_gate"] = gate_stage_one(results[MIXED])
    if MIXED in results and FP64 in results:
        from asteroid_amd import workloads
        truth = workloads.stage_one(FP64)[2]
        out["fit_vs_fp64"] = fit_errors(results[MIXED], results[FP64], truth)
        if time_left() > 90:
            out["fit_vs_f32simplex"] = fit_errors(results[MIXED], complete_search(F32SIMPLEX, args), truth)
        else:
            skipped.append("fit_vs_f32simplex")
    for prec, r in results.items():
        if prec != head_prec:
            out[PREC_NAMES[prec]] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                                     "fits_per_s": r["fits_per_s"], "best_chi2": float(np.min(r["best_f"])),
                                     "roofline": r["roofline"], "complete_search": r["complete_search"]}
    if not args.no_serving and time_left() > 60:
        log("serving leg ...")
        lat, _, _ = serving_leg(head_prec, stream)
        del lat["lat_ms"]
        out["latency"] = dict(lat, workload="Poisson 30 req/s for 4 s, batched-style curves, "
                                            "256-4096 starts x 200 chi2 evals each")
    elif not args.no_serving:
        skipped.append("serving")
    if not args.no_cpu and time_left() > 2 * args.cpu_seconds + 10:
        log("timing the CPU oracle sample ...")
        out["cpu_baseline"] = cpu_baseline(args.budget, args.cpu_seconds)
    elif not args.no_cpu:
        skipped.append("cpu_baseline")
    if skipped:
        out["skipped_for_time"] = skipped
    out["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    return out


# ------------------------------------------------- other configurations
def run_batch(jobs, prec, work_points=16000.0, autotune=True, graphs=True):
    from asteroid_amd import api, flops, variants
    with api.Context(variants.ls_torque(prec)) as ctx:
        tune = ctx.autotune(jobs, work_points_per_launch=work_points, cycles=3) if autotune else []
        t0 = time.perf_counter()
        res = ctx.run_jobs(jobs, work_points_per_launch=work_points, use_graphs=graphs)
        wall = time.perf_counter() - t0
        ms, launches = ctx.last_timing()
    mp = sum(r["model_points"] for r in res)
    fits = sum(j["n_fits"] for j in jobs)
    # algorithmic FLOPs of the evaluations done, FP64 and FP32 apart
    f64 = f32 = 0.0
    for j, r in zip(jobs, res):
        v, mjd = j["space"].variant, j["curve"].mjd
        if prec == MIXED:
            a, b = flops.eval_flops_mixed(v, mjd)
        else:
            a, b = flops.eval_flops(v, mjd), 0.0
        f64 += a * r["n_evals"]
        f32 += b * r["n_evals"]
    achieved = (f64 + f32) / (ms / 1e3) / 1e12
    peak = (f64 + f32) / (f64 / FP64_PEAK_TFLOPS + f32 / FP32_PEAK_TFLOPS)
    return res, {"value": mp / wall, "fits_per_s": fits / wall, "wall_s": wall, "kernel_ms": ms,
                 "launches": launches, "evals": int(sum(r["n_evals"] for r in res)),
                 "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2),
                              "unit": "TFLOP/s", "frac": round(achieved / peak, 4)},
                 "waves_per_simd": {f"tq{e['torque']}bc{e['bc']}bw{e['bw_ball']}nf{e['filter_capacity']}": e["waves"]
                                    for e in tune}}


def gate_jobs(jobs, res_mx):
    """FP64 chi^2 of each mixed job's reported best model."""
    from asteroid_amd import api, workloads
    jobs64 = workloads.with_precision(jobs, FP64)
    rel = []
    for j, r in zip(jobs64, res_mx):
        if r["best_f"] >= 1e29:
            continue
        with api.Context(j["space"].variant) as ctx:
            ctx.set_model_space(j["space"])
            ctx.upload_curve(j["curve"])
            c64, _ = ctx.chi2_batch(r["best_params"])
        rel.append(abs(r["best_f"] - c64[0]) / c64[0])
    rel = np.array(rel)
    return {"max_rel_chi2_err": float(rel.max()), "median_rel_chi2_err": float(np.median(rel)),
            "n_models": int(len(rel)), "tolerance": 1e-6, "pass": bool(rel.max() <= 1e-6 + 2.0 ** -23)}


def main_small(args):
    """4096 random starts x 200 simplex steps on the 400-point curve, FP64,
    on the GPU and on the host CPU (the oracle, all cores)."""
    import oracle
    from asteroid_amd import api, workloads
    v, curve, truth, space = workloads.stage_one(FP64)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(3)
        ctx.run(16, 256, nstages=1, n_steps=200)  # warm-up
        ctx.rng_seed(3)
        ctx.reset_best()
        t0 = time.perf_counter()
        ctx.run(16, 256, nstages=1, n_steps=200)
        wall = time.perf_counter() - t0
        e, p = ctx.last_counts()
        f, params, dv = ctx.fetch()
    log(f"GPU: {e} evals in {wall:.3f} s")
    pb = oracle.Problem(space, curve)
    n = cores()
    t0 = time.perf_counter()
    want = pb.run(16, 256, nstages=1, n_steps=200, seed=3, form=oracle.CANON, n_threads=n)
    cpu_wall = time.perf_counter() - t0
    same = bool(np.array_equal(want["f"], f) and np.array_equal(want["params"], params))
    out = base_line(p / wall, 1, 1, wall * 1e3, DTYPE[FP64],
                    {"workload": "plain FP64 small: 4096 random starts x 200 simplex steps, 400-point curve",
                     "n_obs": 400, "span_days": 3.0, "simplexes": 4096, "n_steps": 200})
    out.update({"fits_per_s": 4096 / wall, "evals": e, "best_chi2": float(np.min(f)),
                "identical_to_oracle": same,
                "cpu_baseline": {"value": float(want["lane_nev"].sum()) * curve.n_data / cpu_wall,
                                 "unit": "model-point evals/s", "cores": n, "kind": "port",
                                 "sample": f"the whole run: 4096 simplexes x 200 steps ({cpu_wall:.1f} s wall)",
                                 "fits_per_s": 4096 / cpu_wall}})
    return out


def jobs_vs_fp64(res_mx, res_64):
    """Each job's answer in the mixed form beside its answer in the FP64 form:
    the relative difference of the two best chi^2 and the largest relative
    difference of the two best parameter vectors, summarised over the jobs
    (the two forms' simplexes take different paths, so the searches are
    compared by their answers)."""
    dchi, dpar = [], []
    for a, b in zip(res_mx, res_64):
        if a["best_f"] >= 1e29 or b["best_f"] >= 1e29:
            continue
        dchi.append(abs(a["best_f"] - b["best_f"]) / b["best_f"])
        dpar.append(float(np.max(np.abs(a["best_params"] - b["best_params"]) /
                                 np.maximum(np.abs(b["best_params"]), 1e-3))))
    dchi, dpar = np.array(dchi), np.array(dpar)
    q = lambda x: {"median": float(np.median(x)), "p90": float(np.percentile(x, 90)), "max": float(x.max())}
    return {"jobs": len(dchi), "best_rel_chi2_diff": q(dchi), "best_params_max_rel_diff": q(dpar),
            "jobs_same_model_1e-3": float(np.mean(dpar <= 1e-3)),
            "jobs_mixed_lower_chi2": float(np.mean([a["best_f"] < b["best_f"] for a, b in zip(res_mx, res_64)]))}


def main_reopt(args):
    from asteroid_amd import workloads
    jobs_mx, truth = workloads.reopt_bc(MIXED)
    jobs64 = workloads.with_precision(jobs_mx, FP64)
    res_mx, m_mx = run_batch(jobs_mx, MIXED, autotune=args.autotune, graphs=args.graphs)
    log(f"mixed: {m_mx}")
    res_64, m_64 = run_batch(jobs64, FP64, autotune=args.autotune, graphs=args.graphs)
    log(f"fp64: {m_64}")
    b_mx = min(res_mx, key=lambda r: r["best_f"])
    b_64 = min(res_64, key=lambda r: r["best_f"])
    dpar = np.abs(b_mx["best_params"] - b_64["best_params"]) / np.maximum(np.abs(b_64["best_params"]), 1e-300)
    out = base_line(m_mx["value"], 1, 0, m_mx["wall_s"] * 1e3, DTYPE[MIXED],
                    {"workload": "stage-two reoptimisation, BC no torque, 3 filters, 818 points over 5 days, "
                                 "TIME_STEP 5e-3, 20 starts (truth +-5%) x 3328 simplexes", "simplexes": 66560,
                     "n_steps": 5000})
    out.update({"fits_per_s": m_mx["fits_per_s"], "roofline": m_mx["roofline"], "mixed": m_mx, "fp64": m_64,
                "best_chi2_mixed": b_mx["best_f"], "best_chi2_fp64": b_64["best_f"],
                "best_rel_chi2_diff": abs(b_mx["best_f"] - b_64["best_f"]) / b_64["best_f"],
                "best_param_max_rel_diff": float(dpar.max()), "per_job_vs_fp64": jobs_vs_fp64(res_mx, res_64),
                "precision_gate": gate_jobs(jobs_mx, res_mx)})
    return out


def main_batched(args):
    from asteroid_amd import workloads
    t = time.perf_counter()
    jobs_mx, made = workloads.batched(MIXED, eval_budget=args.batch_budget)
    log(f"512 curves generated in {time.perf_counter() - t:.1f} s")
    jobs64 = workloads.with_precision(jobs_mx, FP64)
    res_mx, m_mx = run_batch(jobs_mx, MIXED, autotune=args.autotune, graphs=args.graphs)
    log(f"mixed: {m_mx}")
    res_64, m_64 = run_batch(jobs64, FP64, autotune=args.autotune, graphs=args.graphs)
    log(f"fp64: {m_64}")
    pts = sum(j["curve"].n_data for j in jobs_mx)
    out = base_line(m_mx["value"], 1, 0, m_mx["wall_s"] * 1e3, DTYPE[MIXED],
                    {"workload": "batched uneven: 512 curves, 50-2000 points (sum %d), 1-4 filters, 1-8 days, "
                                 "60%% LS+torque / 30%% BC / 10%% BW_BALL+torque, 1024 simplexes each" % pts,
                     "simplexes": 512 * 1024, "evals_per_simplex": args.batch_budget})
    out.update({"fits_per_s": m_mx["fits_per_s"], "roofline": m_mx["roofline"], "mixed": m_mx, "fp64": m_64,
                "per_job_vs_fp64": jobs_vs_fp64(res_mx, res_64), "precision_gate": gate_jobs(jobs_mx, res_mx)})
    return out


def _models_near(truth, n, seed, scale=1e-3):
    """truth plus n-1 models with every parameter shifted by a relative
    normal deviate of `scale` (candidate models around a fit)."""
    r = np.random.default_rng(seed).standard_normal((n - 1, len(truth)))
    return np.vstack([truth[None], truth[None] * (1 + scale * r)])


def main_plot(args):
    """The -plot path: model light curves on the reference's 20000-point grid
    (chi2_plot) for a batch of models on the Stage One curve, the reference's
    single-model case, and the minima periodogram of one curve."""
    import oracle
    from asteroid_amd import api, flops, workloads
    prec = F32SIMPLEX if args.precision in ("both", "f32simplex") else {"fp64": FP64, "mixed": MIXED}[args.precision]
    v, curve, truth, space = workloads.stage_one(prec)
    n_models, n_plot = args.plot_models, 20000
    params = _models_near(truth, n_models, 5)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.model_curve(params[:64], n_plot=n_plot)  # warm-up
        t0 = time.perf_counter()
        t, vm, chi2, dv = ctx.model_curve(params, n_plot=n_plot)
        wall = time.perf_counter() - t0
        ms, _ = ctx.last_timing()
        t0 = time.perf_counter()
        t1, vm1, _, _ = ctx.model_curve(truth, n_plot=n_plot)
        one_wall = time.perf_counter() - t0
        one_ms, _ = ctx.last_timing()
        t0 = time.perf_counter()
        fr, H, nmin = ctx.periodogram(t1, vm1[0], 5)
        pg_wall = time.perf_counter() - t0
        pg_ms, _ = ctx.last_timing()
    pts = n_models * (curve.n_data + n_plot)
    f_model = flops.eval_flops(v, curve.mjd) + flops.eval_flops(v, t)
    achieved = f_model * n_models / (ms / 1e3) / 1e12
    log(f"{n_models} models x {n_plot}: kernel {ms:.1f} ms, call {wall * 1e3:.1f} ms; one model {one_ms:.2f} ms")
    out = base_line(pts / wall, 1, 1, wall * 1e3, DTYPE[prec],
                    {"workload": f"plot: {n_models} models (the truth and models within 0.1% of it) x "
                                 f"{n_plot}-point grid on the Stage One 400-point curve, chi2_plot per model",
                     "models": n_models, "n_plot": n_plot, "precision": PREC_NAMES[prec]})
    out.update({"kernel_ms": ms, "model_points_per_s_kernel": pts / (ms / 1e3),
                "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                             "kernel": "model_curve_kernel", "flop_per_model": round(f_model, 1),
                             "lanes": n_models},
                "single_model": {"kernel_ms": one_ms, "call_ms": one_wall * 1e3,
                                 "periodogram_ms": pg_wall * 1e3, "periodogram_kernel_ms": pg_ms,
                                 "minima": nmin, "clusters_per_day": [float(x) for x in fr]},
                "cpu_baseline": None})
    if args.no_cpu:
        return out
    pb = oracle.Problem(space, curve)
    n = cores()
    # calibrate on one model per core, then size the sample to ~cpu_seconds
    t0 = time.perf_counter()
    pb.model_curve(params[:n], n_plot=n_plot, n_threads=n)
    t1 = time.perf_counter() - t0
    sample = params[: min(n_models, n * max(1, int(round(args.cpu_seconds / max(t1, 1e-3)))))]
    t0 = time.perf_counter()
    want = pb.model_curve(sample, n_plot=n_plot, n_threads=n)
    cpu_dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    pb.model_curve(truth, n_plot=n_plot)
    cpu_one = time.perf_counter() - t0
    same = bool(np.array_equal(want[1], vm[: len(sample)]) and np.array_equal(want[2], chi2[: len(sample)]))
    out["single_model"]["cpu_oracle_ms"] = cpu_one * 1e3
    out.update({"identical_to_oracle": same,
                "cpu_baseline": {"value": len(sample) * (curve.n_data + n_plot) / cpu_dt, "unit": "model-point evals/s",
                                 "cores": n, "kind": "port",
                                 "sample": f"{len(sample)} of the models ({cpu_dt:.2f} s wall, oracle canonical form, "
                                           f"{n} threads)"}})
    return out


def main_rmsd(args):
    """RMSD confidence intervals (chi2_gpu_rms) around the Stage One truth
    model: 2^18 lanes (1024 groups of 256), args.rmsd_samples shifted models
    each, dx = args.dx; float-simplex form (the reference's default) with the
    mixed form beside it."""
    import oracle
    from asteroid_amd import api, flops, variants, workloads
    res = {}
    for prec in (F32SIMPLEX, MIXED):
        v, curve, truth, _ = workloads.stage_one(prec)
        space = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0), reopt=1)
        with api.Context(v) as ctx:
            ctx.set_model_space(space)
            ctx.upload_curve(curve)
            ctx.rng_seed(21)
            ctx.set_start(truth)
            ctx.rmsd(1024, 256, 4, args.dx)  # warm-up
            t0 = time.perf_counter()
            r = ctx.rmsd(1024, 256, args.rmsd_samples, args.dx)
            wall = time.perf_counter() - t0
            ms, _ = ctx.last_timing()
        f_eval = flops.eval_flops(v, curve.mjd) if prec != MIXED else sum(flops.eval_flops_mixed(v, curve.mjd))
        peak = peak_of(prec, v, curve.mjd)[0]
        achieved = f_eval * r["n_tot"] / (ms / 1e3) / 1e12
        res[prec] = dict(value=r["n_tot"] * curve.n_data / wall, wall_s=wall, kernel_ms=ms, n_tot=r["n_tot"],
                         n_bad=r["n_bad"], f0=r["f0"], f1=r["f1"],
                         ranges=[[float(a), float(b)] for a, b in zip(r["par_min"].min(0), r["par_max"].max(0))],
                         roofline={"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2),
                                   "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "kernel": "rmsd_kernel"})
        log(f"{PREC_NAMES[prec]}: {res[prec]['value']:.3e} model-pts/s, kernel {ms:.0f} ms, "
            f"accepted {r['n_tot'] - r['n_bad']} of {r['n_tot']}")
    a = res[F32SIMPLEX]
    out = base_line(a["value"], 1, 1, a["wall_s"] * 1e3, DTYPE[F32SIMPLEX],
                    {"workload": f"RMSD confidence intervals around the Stage One truth model: 1024 groups x 256 "
                                 f"lanes x {args.rmsd_samples} models, dx {args.dx}, 400-point curve",
                     "lanes": 262144, "samples_per_lane": args.rmsd_samples, "dx": args.dx})
    out.update({"roofline": a["roofline"], "f32simplex": a, "mixed": res[MIXED], "cpu_baseline": None})
    if args.no_cpu:
        return out
    v, curve, truth, _ = workloads.stage_one(F32SIMPLEX)
    space = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0), reopt=1)
    pb = oracle.Problem(space, curve)
    n = cores()
    # calibrate on one group of 16 lanes x 8 models per core, then size the sample to ~cpu_seconds
    t0 = time.perf_counter()
    pb.rmsd(n, 16, 8, args.dx, truth, seed=21, n_threads=n)
    t1 = (time.perf_counter() - t0) / 8
    S = max(1, int(round(args.cpu_seconds / max(t1, 1e-4))))
    t0 = time.perf_counter()
    w = pb.rmsd(n, 16, S, args.dx, truth, seed=21, n_threads=n)
    dt = time.perf_counter() - t0
    out.update({"cpu_baseline": {"value": float(w["counts"][:, 0].sum()) * curve.n_data / dt,
                                 "unit": "model-point evals/s", "cores": n, "kind": "port",
                                 "sample": f"{n} groups x 16 lanes x {S} models ({dt:.2f} s wall, oracle "
                                           f"canonical form, {n} threads)"}})
    return out


def fifo_latency(arrivals, service):
    """Latencies (ms) of a single FIFO server with these arrival and service times (s)."""
    free, lat = 0.0, []
    for a, s in zip(arrivals, service):
        start = max(a, free)
        free = start + s
        lat.append(free - a)
    return np.array(lat) * 1e3


def main_serving(args):
    """Continuous batching (mixed form) over swept Poisson rates, beside the
    plain FP64 form run one request at a time (its own launch cycles per
    request, FIFO), the latter timed on each stream's first requests."""
    from asteroid_amd import api, variants, workloads
    streams = [request_stream(MIXED, rate, min(args.duration, args.max_requests / rate), args.serve_budget)
               for rate in args.rates]
    rows = []
    for rate, stream in zip(args.rates, streams):
        lat, jobs, times = serving_leg(MIXED, stream)
        log(f"rate {rate}: {lat['requests']} requests, p50 {lat['p50_ms']:.0f} ms, p99 {lat['p99_ms']:.0f} ms, "
            f"{lat['fits_per_s']:.0f} fits/s")
        k = min(args.baseline_requests, len(jobs))
        jobs64 = workloads.with_precision(jobs[:k], FP64)
        svc = []
        with api.Context(variants.ls_torque(FP64)) as ctx:
            for j in jobs64:
                t0 = time.perf_counter()
                ctx.run_jobs([j])
                svc.append(time.perf_counter() - t0)
        l64 = fifo_latency(times[:k], svc)
        lat["first_k"] = {"k": k, "p50_ms": float(np.percentile(lat["lat_ms"][:k], 50)),
                          "p99_ms": float(np.percentile(lat["lat_ms"][:k], 99))}
        lat["fp64_one_at_a_time_first_k"] = {"p50_ms": float(np.percentile(l64, 50)),
                                             "p99_ms": float(np.percentile(l64, 99)),
                                             "mean_service_ms": float(np.mean(svc) * 1e3)}
        del lat["lat_ms"]
        rows.append(lat)
    best = max(rows, key=lambda r: r["fits_per_s"])
    out = base_line(best["model_points_per_s"], 1, 0, best["wall_s"] * 1e3, DTYPE[MIXED],
                    {"workload": "serving: Poisson arrivals, batched-style curves, 256-4096 starts x "
                                 f"{args.serve_budget} chi2 evals per request", "rates": args.rates})
    out.update({"fits_per_s": best["fits_per_s"], "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
                "sweep": rows})
    return out


VARIANT_RUNS = {
    # name: (factory(precision, time_step), n_filters)
    "ls_torque": (lambda p, ts: variants_mod().ls_torque(p, time_step=ts), 1),
    "ls": (lambda p, ts: variants_mod().make_variant(False, False, False, True, True, p, ts), 1),
    "bc": (lambda p, ts: variants_mod().make_variant(False, True, False, True, True, p, ts), 1),
    "bc_torque": (lambda p, ts: variants_mod().make_variant(True, True, False, True, True, p, ts), 1),
    "bw": (lambda p, ts: variants_mod().make_variant(False, False, True, True, True, p, ts), 1),
    "bw_torque": (lambda p, ts: variants_mod().bw_ball_torque(p, time_step=ts), 1),
    "trend_torque": (lambda p, ts: variants_mod().ls_torque_trend(p, time_step=ts), 1),
    "torque2": (lambda p, ts: variants_mod().ls_torque2(p, time_step=ts), 1),
    "bc_rotate": (lambda p, ts: variants_mod().bc_rotate(p, time_step=ts), 1),
    "bc_rotate_torque": (lambda p, ts: variants_mod().bc_rotate(p, time_step=ts, torque=True), 1),
    "rect": (lambda p, ts: variants_mod().rect(p, time_step=ts), 1),
    "rect_torque": (lambda p, ts: variants_mod().rect(p, time_step=ts, torque=True), 1),
    "nudge_torque": (lambda p, ts: variants_mod().ls_torque_nudge(p, time_step=ts), 1),
    "segment3": (lambda p, ts: variants_mod().ls_segments(p, time_step=ts), 1),
}


def minima_of(v, curve, truth, k=4):
    """k observed minima for a NUDGE run: local maxima of the ground truth's
    magnitudes (the synthetic model, not the fitted one) at the observations."""
    from asteroid_amd import synth
    vm = synth.truth_vmod(v, truth, curve.mjd, curve.E, curve.S)
    idx = [i - 1 for i in range(2, len(vm)) if vm[i - 1] > vm[i - 2] and vm[i - 1] >= vm[i]]
    pick = idx[::max(1, len(idx) // k)][:k]
    return (np.array([curve.mjd[i] for i in pick], np.float32), np.array([vm[i] for i in pick], np.float32))


def variants_mod():
    from asteroid_amd import variants
    return variants


def main_variants(args):
    """The Stage One random search (400 points over 3 days, TIME_STEP 0.01,
    P_psi 2-4800 h, args.budget evaluations per simplex) for every built
    variant, args.fits simplexes each, in one precision form: model-point
    evals/s and the fit kernel's fraction of the op-mix peak per variant."""
    from asteroid_amd import api, flops, synth, variants
    prec = {"both": MIXED, "mixed": MIXED, "fp64": FP64, "f32simplex": F32SIMPLEX}[args.precision]
    group = 256
    rows = {}
    names = args.variant_list.split(",") if args.variant_list else list(VARIANT_RUNS)
    for name in names:
        mk, nf = VARIANT_RUNS[name]
        v = mk(prec, 1e-2)
        if v.nudge and prec == MIXED:
            continue  # NUDGE runs in the FP64 and float-simplex forms only
        seg_t = None
        if v.n_seg > 1:
            # SEGMENT: the LS curve split in thirds; each further segment's
            # Euler angles and psi_0 a little away from the first one's
            v1 = variants.make_variant(False, False, False, True, True, prec, 1e-2)
            curve, t8 = synth.make_curve(2024, v1, 400, 3.0, n_filters=nf, sigma_range=(0.05, 0.3))
            seg_t = curve.mjd[0] + np.array([0.0, 1.0, 2.0])
            p8, off = variants.param_positions(v1), np.random.default_rng(7)
            truth = np.array([t8[p8[k.split(":")[0]]] + (off.uniform(-0.2, 0.2) if ":" in k else 0.0)
                              for k in variants.param_positions(v)])
        else:
            curve, truth = synth.make_curve(2024, v, 400, 3.0, n_filters=nf, sigma_range=(0.05, 0.3))
        # NUDGE is Stage Three (README.md:87-101): a reoptimisation from a model
        # (here the truth perturbed by 2%), not a random search
        stage3 = bool(v.nudge)
        space = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0), reopt=1 if stage3 else 0)
        with api.Context(v) as ctx:
            ctx.set_model_space(space)
            if seg_t is not None:
                ctx.set_segments(seg_t)
            ctx.upload_curve(curve)
            if stage3:
                ctx.set_minima(*minima_of(v, curve, truth))
                ctx.set_start(truth * (1.0 + np.random.default_rng(3).uniform(-0.02, 0.02, truth.shape)))
            ctx.rng_seed(5)
            run = dict(nstages=1, n_steps=5000, eval_budget=args.budget, size_min=SIZE_MIN, use_graphs=args.graphs)
            ctx.run(args.fits // group // 4, group, **run)  # warm-up
            ctx.reset_best()
            t0 = time.perf_counter()
            ctx.run(args.fits // group, group, **run)
            wall = time.perf_counter() - t0
            e, pts = ctx.last_counts()
            ms, nl = ctx.last_timing()
            # the evaluator alone: one chi^2 per lane for 2^17 models near the
            # truth (no simplex, every lane busy to the end)
            near = truth * (1.0 + np.random.default_rng(1).uniform(-0.02, 0.02, (1 << 17, truth.size)))
            ctx.chi2_batch(near[:1024])
            ctx.chi2_batch(near)
            ev_ms, _ = ctx.last_timing()
        split = truth[variants.param_positions(v)["Tt"]] if v.torque2 else None
        fl = flops.eval_flops(v, curve.mjd, split, seg_t)
        peak, mix = peak_of(prec, v, curve.mjd, seg_t)
        achieved = fl * e / (ms / 1e3) / 1e12
        rows[name] = {"n_params": space.n_params, "stage": "three (reopt)" if stage3 else "one", "value": pts / wall, "fits_per_s": args.fits / wall,
                      "evals_per_fit": round(e / args.fits, 1), "kernel_ms": round(ms, 2), "launches": nl,
                      "flop_per_model_point": round(fl / curve.n_data, 2), "achieved_tflops": round(achieved, 3),
                      "chi2_batch_2p17_ms": round(ev_ms, 3),
                      "chi2_batch_model_points_per_s": (1 << 17) * curve.n_data / (ev_ms / 1e3),
                      "peak_tflops": round(peak, 2), "frac": round(achieved / peak, 4)}
        log(f"{name}: {pts / wall:.3e} model-points/s, {achieved:.2f} TFLOP/s ({achieved / peak:.1%} of {peak:.1f})")
    head = rows[names[0]]
    out = base_line(head["value"], 1, 1, None, DTYPE[prec],
                    {"workload": "stage-one random search per variant (400 points, 3 days, TIME_STEP 0.01)",
                     "simplexes": args.fits, "evals_per_simplex": args.budget, "group_size": group,
                     "precision": PREC_NAMES[prec]})
    out["variants"] = rows
    return out


def main_counters(args):
    """One short pass over every kept kernel, for rocprofv3 counter runs
    (each PMC pass reruns the program): three 192-round launches of the
    Stage One fitting kernel over 2^18 resident simplexes (mixed and FP64),
    a 2^17-model chi^2 batch in each form, 4096 model curves of 20000 grid
    points, 2^18 RMSD lanes x 16 models (float-simplex and mixed), and a
    plain copy kernel reading and writing 1 GiB whose FETCH_SIZE and
    WRITE_SIZE calibrate the traffic counters."""
    from asteroid_amd import api, variants, workloads
    t0 = time.perf_counter()
    out = {}
    for prec in (MIXED, FP64):
        v, curve, truth, space = workloads.stage_one(prec)
        with api.Context(v) as ctx:
            ctx.set_model_space(space)
            ctx.upload_curve(curve)
            ctx.set_waves(v, 1, 3)
            ctx.rng_seed(1)
            ctx.run_begin(4096, 256, nstages=1, n_steps=5000, eval_budget=args.budget, size_min=SIZE_MIN,
                          ticks_per_launch=args.ticks, use_graphs=False, resident_groups=1024)
            ctx.run_continue(3)
            e, _ = ctx.last_counts()
            ms, nl = ctx.last_timing()
            ctx.run_cancel()
            near = truth * (1.0 + np.random.default_rng(1).uniform(-0.02, 0.02, (1 << 17, truth.size)))
            ctx.chi2_batch(near)
            bms, _ = ctx.last_timing()
        out[PREC_NAMES[prec]] = {"fit_evals": e, "fit_ms": ms, "fit_launches": nl, "chi2_batch_ms": bms}
        log(f"{PREC_NAMES[prec]}: fit {nl} launches {ms:.1f} ms, chi2 batch {bms:.1f} ms")
    v, curve, truth, space = workloads.stage_one(F32SIMPLEX)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.model_curve(truth * (1.0 + np.random.default_rng(5).uniform(-0.001, 0.001, (4096, truth.size))),
                        n_plot=20000)
        out["model_curve_ms"], _ = ctx.last_timing()
    for prec in (F32SIMPLEX, MIXED):
        v, curve, truth, _ = workloads.stage_one(prec)
        sp = variants.default_model_space(v, ppsi_hours=(2.0, 4800.0), reopt=1)
        with api.Context(v) as ctx:
            ctx.set_model_space(sp)
            ctx.upload_curve(curve)
            ctx.rng_seed(21)
            ctx.set_start(truth)
            r = ctx.rmsd(1024, 256, 16, args.dx)
            out[f"rmsd_{PREC_NAMES[prec]}_ms"], _ = ctx.last_timing()
            out[f"rmsd_{PREC_NAMES[prec]}_models"] = r["n_tot"]
    # A kernel of known traffic for the traffic counters: the library's plain
    # copy of 1 GiB (1 GiB read and 1 GiB written per launch).
    with api.Context(variants.ls_torque(MIXED)) as ctx:
        out["copy_bytes_per_s"] = ctx.copy_bandwidth(1 << 30, 1)
    out["copy_kernel_bytes_read"] = out["copy_kernel_bytes_written"] = 1 << 30
    line = base_line(None, 1, 0, None, "mixed", {"workload": "counters: one pass over every kept kernel"})
    line.update(out)
    line["bench_wall_s"] = round(time.perf_counter() - t0, 1)
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["stage1", "small", "reopt", "batched", "serving", "plot", "rmsd",
                                           "variants", "counters"],
                    default="stage1")
    ap.add_argument("--plot-models", type=int, default=4096)
    ap.add_argument("--variant-list", default="", help="comma-separated subset of the variants workload")
    ap.add_argument("--rmsd-samples", type=int, default=64)
    ap.add_argument("--dx", type=float, default=0.003)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64", "f32simplex"], default="both")
    ap.add_argument("--fits", type=int, default=1 << 18)
    ap.add_argument("--budget", type=int, default=4800)
    ap.add_argument("--ticks", type=int, default=192, help="evaluation rounds per fitting-kernel launch (one step)")
    ap.add_argument("--deadline", type=float, default=360.0,
                    help="seconds after which the optional legs of the default workload are skipped")
    ap.add_argument("--batch-budget", type=int, default=1000)
    ap.add_argument("--serve-budget", type=int, default=200)
    ap.add_argument("--rates", type=float, nargs="+", default=[10, 30, 100, 300, 1000, 2000])
    ap.add_argument("--duration", type=float, default=6.0)
    ap.add_argument("--max-requests", type=float, default=200.0)
    ap.add_argument("--baseline-requests", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--waves", type=int, choices=[0, 2, 3, 4], default=0,
                    help="Stage One: run the fitting kernel's build for this many waves per SIMD (0: autotune)")
    ap.add_argument("--no-graphs", dest="graphs", action="store_false")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the fits are independent and shard across GPUs "
              "without communication; run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    fn = {"stage1": main_stage_one, "small": main_small, "reopt": main_reopt, "batched": main_batched,
          "serving": main_serving, "plot": main_plot, "rmsd": main_rmsd, "variants": main_variants,
          "counters": main_counters}[args.workload]
    print(json.dumps(fn(args)), flush=True)


if __name__ == "__main__":
    main()
