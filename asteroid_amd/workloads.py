"""Synthetic workloads of the benchmark configurations.

Every light curve comes from asteroid_amd.synth (a seeded ground-truth model
integrated with an independent high-order ODE solver, a 3-point ephemeris
with the asteroid ~1.3 AU from the Sun and ~0.3 AU from the Earth, Gaussian
noise); nothing is read from disk.

  stage_one   one 400-point single-filter curve over 3 days, LS ellipsoid
              with torque (P_PSI, TORQUE, INTERP), TIME_STEP 1e-2 d.
  reopt_bc    BC without torque, 3 filters, 818 points over 5 days, INTERP,
              TIME_STEP 5e-3 d, noise 0.02-0.2 mag; a Stage Two
              reoptimisation around 20 starting models, each the truth
              perturbed by +-5% in every parameter.
  batched     512 curves, N log-uniform in 50..2000 points, 1-4 filters,
              spans 1-8 days; variants 60% LS+torque, 30% BC, 10%
              BW_BALL+torque.
  requests    Poisson arrivals of batched-style curves, each with 256..4096
              simplex starts (multiples of 256).
"""
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from . import abi, synth, variants

STAGE_ONE = dict(n_obs=400, span=3.0, seed=2024, ppsi_hours=(2.0, 4800.0))


def stage_one(precision):
    v = variants.ls_torque(precision, time_step=1e-2)
    curve, truth = synth.make_curve(STAGE_ONE["seed"], v, STAGE_ONE["n_obs"], STAGE_ONE["span"], n_filters=1,
                                    sigma_range=(0.05, 0.3))
    space = variants.default_model_space(v, ppsi_hours=STAGE_ONE["ppsi_hours"], reopt=0)
    return v, curve, truth, space


def _workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def reopt_bc(precision, n_starts=20, fits_per_start=3328, seed=77, n_steps=5000):
    """Jobs (dicts for Context.run_jobs) of the Stage Two reoptimisation."""
    v = variants.bc_free(precision, time_step=5e-3)
    curve, truth = synth.make_curve(seed, v, 818, 5.0, n_filters=3, sigma_range=(0.02, 0.2))
    space = variants.default_model_space(v, reopt=1)
    rng = np.random.default_rng(seed + 1)
    jobs = []
    for k in range(n_starts):
        p0 = truth * (1.0 + rng.uniform(-0.05, 0.05, size=truth.shape))
        jobs.append(dict(space=space, curve=curve, n_fits=fits_per_start, nstages=1, n_steps=n_steps,
                         seed=seed * 1000 + k, params0=p0, tag=k))
    return jobs, truth


_MIX = ((0.6, "ls_torque"), (0.3, "bc"), (0.1, "bw_torque"))


def _variant(name, precision):
    if name == "ls_torque":
        return variants.ls_torque(precision, time_step=1e-2)
    if name == "bc":
        return variants.bc_free(precision, time_step=5e-3)
    return variants.bw_ball_torque(precision, time_step=1e-2)


def _curve_spec(rng):
    u = rng.uniform()
    name = _MIX[0][1] if u < _MIX[0][0] else (_MIX[1][1] if u < _MIX[0][0] + _MIX[1][0] else _MIX[2][1])
    n_obs = int(round(np.exp(rng.uniform(np.log(50), np.log(2000)))))
    return dict(name=name, n_obs=n_obs, n_filters=int(rng.integers(1, 5)), span=float(rng.uniform(1.0, 8.0)),
                seed=int(rng.integers(1 << 30)))


def _make(spec_prec):
    spec, precision = spec_prec
    v = _variant(spec["name"], precision)
    curve, truth = synth.make_curve(spec["seed"], v, spec["n_obs"], spec["span"], n_filters=spec["n_filters"])
    return curve, truth


def curves(specs, precision):
    """Light curves of `specs` (dicts of _curve_spec), generated in parallel."""
    with ProcessPoolExecutor(max_workers=_workers()) as ex:
        return list(ex.map(_make, [(s, precision) for s in specs], chunksize=4))


def batched(precision, n_curves=512, fits=1024, seed=4, eval_budget=1000, n_steps=5000, made=None):
    """Jobs of the batched uneven workload; `made` reuses curves generated
    for the same specs (the curves do not depend on the precision)."""
    rng = np.random.default_rng(seed)
    specs = [_curve_spec(rng) for _ in range(n_curves)]
    made = made if made is not None else curves(specs, precision)
    jobs = []
    for k, (spec, (curve, truth)) in enumerate(zip(specs, made)):
        v = _variant(spec["name"], precision)
        space = variants.default_model_space(v)
        jobs.append(dict(space=space, curve=curve, n_fits=fits, nstages=1, n_steps=n_steps,
                         eval_budget=eval_budget, seed=1000 + k, tag=k))
    return jobs, made


def requests(precision, rate, duration_s, seed=5, eval_budget=200, n_steps=5000, max_requests=4000, made=None):
    """Poisson request stream: (arrival offsets in s, jobs).  Each request is a
    batched-style curve with U{1..16} x 256 simplex starts."""
    rng = np.random.default_rng(seed)
    t, times = 0.0, []
    while len(times) < max_requests:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        times.append(t)
    specs = [_curve_spec(rng) for _ in times]
    fits = [256 * int(rng.integers(1, 17)) for _ in times]
    made = made if made is not None else curves(specs, precision)
    jobs = []
    for k, (spec, (curve, truth), nf) in enumerate(zip(specs, made, fits)):
        v = _variant(spec["name"], precision)
        space = variants.default_model_space(v)
        jobs.append(dict(space=space, curve=curve, n_fits=nf, nstages=1, n_steps=n_steps, eval_budget=eval_budget,
                         seed=2000 + k, tag=k))
    return np.array(times), jobs, made


def with_precision(jobs, precision):
    """The same jobs with every variant switched to `precision`."""
    out = []
    for j in jobs:
        v = j["space"].variant
        v2 = abi.Variant(v.torque, v.bc, v.bw_ball, v.p_psi, v.interp, int(precision), v.time_step)
        space = variants.default_model_space(v2, reopt=j["space"].x2().reopt)
        out.append(dict(j, space=space))
    return out
