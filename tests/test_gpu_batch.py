"""Batched jobs (ast_run_jobs): many light curves of uneven length, filter
count, model variant and precision fitted in shared launch cycles (one
launch per kernel class).  Each job's results equal the same job run alone,
and for the FP64 and float-simplex forms they equal the oracle's run of that
job bitwise (best chi^2, model and delta_V per group, evaluation counts)."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu


def _jobs():
    specs = [  # (variant, n_filters, n_obs, span, n_fits, nstages, n_steps, reopt)
        (variants.ls_torque(abi.PREC_FP64), 1, 90, 1.5, 512, 2, 30, 0),
        (variants.bc_free(abi.PREC_FP64), 3, 140, 2.0, 256, 1, 40, 0),
        (variants.bw_ball_torque(abi.PREC_F32SIMPLEX), 2, 60, 1.0, 768, 1, 30, 0),
        (variants.ls_torque(abi.PREC_MIXED), 1, 200, 3.0, 256, 1, 30, 0),
        (variants.ls_torque(abi.PREC_FP64), 4, 50, 1.0, 256, 1, 25, 1),
        (variants.bc_free(abi.PREC_MIXED), 2, 120, 2.5, 512, 1, 30, 0),
        (variants.ls_torque_trend(abi.PREC_FP64), 2, 80, 1.5, 256, 1, 25, 0),
        (variants.ls_torque2(abi.PREC_F32SIMPLEX), 1, 70, 1.2, 256, 1, 25, 0),
        (variants.bc_rotate(abi.PREC_FP64, torque=True), 1, 90, 1.5, 256, 1, 25, 1),
        (variants.rect(abi.PREC_F32SIMPLEX), 2, 60, 1.0, 256, 1, 25, 0),
    ]
    jobs = []
    for k, (v, nf, n_obs, span, n_fits, nst, nsteps, reopt) in enumerate(specs):
        curve, truth = synth.make_curve(100 + k, v, n_obs, span, n_filters=nf)
        space = variants.default_model_space(v, reopt=reopt)
        jobs.append(dict(space=space, curve=curve, n_fits=n_fits, nstages=nst, n_steps=nsteps, seed=500 + k,
                         params0=truth if reopt else None, tag=k))
    return jobs


def _same(a, b):
    for key in ("status", "best_f", "best_group", "n_evals", "model_points", "tag"):
        assert a[key] == b[key], key
    assert np.array_equal(a["best_params"], b["best_params"])
    assert np.array_equal(a["best_dV"], b["best_dV"])


def test_batch_equals_each_job_alone():
    jobs = _jobs()
    with api.Context(variants.ls_torque()) as ctx:
        batch = ctx.run_jobs(jobs, work_points_per_launch=3000)
        alone = [ctx.run_jobs([j], work_points_per_launch=5000)[0] for j in jobs]
    for k, (b, a) in enumerate(zip(batch, alone)):
        print(f"job {k}: best {b['best_f']:.6g} evals {b['n_evals']} latency {b['latency_ms']:.1f} ms")
        _same(b, a)
    assert all(r["status"] == 0 for r in batch)


def test_batch_matches_oracle():
    jobs = _jobs()
    with api.Context(variants.ls_torque()) as ctx:
        res = ctx.run_jobs(jobs)
    for k, (j, r) in enumerate(zip(jobs, res)):
        v = j["space"].variant
        if v.precision == abi.PREC_MIXED:
            continue  # tolerance-checked in test_gpu_mixed
        pb = oracle.Problem(j["space"], j["curve"])
        G = j["n_fits"] // 256
        want = pb.run(G, 256, nstages=j["nstages"], n_steps=j["n_steps"], seed=j["seed"],
                      reopt=j["space"].x2().reopt, params0=j["params0"], n_threads=8)
        g = int(np.argmin(want["f"]))
        assert r["best_f"] == want["f"][g] and r["best_group"] == g, k
        assert np.array_equal(r["best_params"], want["params"][g]), k
        assert np.array_equal(r["best_dV"], want["dV"][g]), k
        assert r["n_evals"] == want["lane_nev"].sum(), k


def test_batch_one_launch_per_class():
    """Eight jobs of one kernel class (different curves and seeds) share one
    launch per cycle: the batch takes about as many launches as its slowest
    job alone, not eight times as many."""
    v = variants.ls_torque(abi.PREC_FP64)
    space = variants.default_model_space(v)
    jobs = []
    for k in range(8):
        curve, _ = synth.make_curve(300 + k, v, 80, 1.5)
        jobs.append(dict(space=space, curve=curve, n_fits=256, n_steps=40, seed=k))
    with api.Context(v) as ctx:  # schedule 1: the same ticks per launch alone and in the batch
        ctx.run_jobs(jobs, work_points_per_launch=2000, schedule=1)
        _, launches = ctx.last_timing()
        alone = []
        for j in jobs:
            ctx.run_jobs([j], work_points_per_launch=2000, schedule=1)
            alone.append(ctx.last_timing()[1])
    print("batch launches", launches, "alone", alone)
    assert launches == max(alone)


def test_batch_schedules_agree():
    """The three launch schedules (heaviest first with equal finish, equal
    work per launch in job order, equal work heaviest first) change only how
    the work is cut into launches: every result is the same bit for bit."""
    jobs = _jobs()
    with api.Context(variants.ls_torque()) as ctx:
        runs = [ctx.run_jobs(jobs, work_points_per_launch=4000, schedule=s) for s in (0, 1, 2)]
    for k in range(len(jobs)):
        _same(runs[0][k], runs[1][k])
        _same(runs[0][k], runs[2][k])


def test_batch_nudge_job_matches_alone_and_oracle():
    """A Stage Three NUDGE job carries its observed minima in the batch; it
    equals the same job alone and the oracle's run, beside jobs of other
    variants."""
    v = variants.ls_torque_nudge(abi.PREC_FP64)
    curve, truth = synth.make_curve(77, v, 100, 1.5)
    space = variants.default_model_space(v, reopt=1)
    vm = synth.truth_vmod(v, truth, curve.mjd, curve.E, curve.S)
    idx = [i - 1 for i in range(2, len(vm)) if vm[i - 1] > vm[i - 2] and vm[i - 1] >= vm[i]][:3]
    minima = (np.array([curve.mjd[i] for i in idx], np.float32), np.array([vm[i] for i in idx], np.float32))
    nj = dict(space=space, curve=curve, n_fits=256, n_steps=25, seed=9, params0=truth * 1.01, tag=99,
              minima=minima)
    jobs = _jobs()[:3] + [nj]
    with api.Context(variants.ls_torque()) as ctx:
        batch = ctx.run_jobs(jobs, work_points_per_launch=3000)
        alone = ctx.run_jobs([nj], work_points_per_launch=5000)[0]
    _same(batch[3], alone)
    pb = oracle.Problem(space, curve, minima=minima)
    want = pb.run(1, 256, nstages=1, n_steps=25, seed=9, reopt=1, params0=truth * 1.01, n_threads=8)
    r = batch[3]
    assert r["status"] == 0 and r["best_f"] == want["f"][0]
    assert np.array_equal(r["best_params"], want["params"][0]) and np.array_equal(r["best_dV"], want["dV"][0])
    assert r["n_evals"] == want["lane_nev"].sum()


def test_batch_segment_job_matches_alone_and_oracle():
    """A SEGMENT job carries its start times in the batch (observations before
    the first start left out); it equals the same job alone and the oracle's
    run, beside jobs of other variants."""
    v1 = variants.make_variant(False, False, False, True, True, abi.PREC_FP64)
    curve, t8 = synth.make_curve(81, v1, 150, 2.0)
    v = variants.ls_segments(abi.PREC_FP64)
    space = variants.default_model_space(v)
    t_seg = np.array([curve.mjd[4], curve.mjd[0] + 0.7, curve.mjd[0] + 1.3])
    sj = dict(space=space, curve=curve, n_fits=256, n_steps=30, seed=13, tag=7, t_segments=t_seg)
    jobs = _jobs()[:3] + [sj]
    with api.Context(variants.ls_torque()) as ctx:
        batch = ctx.run_jobs(jobs, work_points_per_launch=3000)
        alone = ctx.run_jobs([sj], work_points_per_launch=5000)[0]
    _same(batch[3], alone)
    want = oracle.Problem(space, curve, t_start=t_seg).run(1, 256, nstages=1, n_steps=30, seed=13, n_threads=8)
    r = batch[3]
    assert r["status"] == 0 and r["best_f"] == want["f"][0]
    assert np.array_equal(r["best_params"], want["params"][0]) and np.array_equal(r["best_dV"], want["dV"][0])
    assert r["n_evals"] == want["lane_nev"].sum()


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_F32SIMPLEX])
def test_full_size_curves_match_oracle(prec):
    """Curves at the largest size of the batched workload (2000 points over 8
    days, four filters; and a 1000-point three-filter BC curve over 5 days at
    TIME_STEP 5e-3, the Stage Two shape), a few simplexes each, in one batch
    beside a small job: every job equals the oracle's run bitwise."""
    specs = [(variants.ls_torque(prec), 4, 2000, 8.0, 128, 1, 40, 0),
             (variants.bc_free(prec, time_step=5e-3), 3, 1000, 5.0, 128, 2, 20, 0),
             (variants.ls_torque(prec), 1, 60, 1.0, 256, 1, 30, 0)]
    jobs = []
    for k, (v, nf, n_obs, span, n_fits, nst, nsteps, reopt) in enumerate(specs):
        curve, truth = synth.make_curve(300 + k, v, n_obs, span, n_filters=nf)
        jobs.append(dict(space=variants.default_model_space(v), curve=curve, n_fits=n_fits, nstages=nst,
                         n_steps=nsteps, seed=700 + k, tag=k))
    with api.Context(variants.ls_torque(prec)) as ctx:
        res = ctx.run_jobs(jobs, group_size=64, work_points_per_launch=4000)
    for j, r in zip(jobs, res):
        print("job", r["tag"], "points", j["curve"].n_data, "best", r["best_f"], flush=True)
        want = oracle.Problem(j["space"], j["curve"]).run(j["n_fits"] // 64, 64, nstages=j["nstages"],
                                                          n_steps=j["n_steps"], seed=j["seed"], n_threads=8)
        g = int(np.argmin(want["f"]))
        assert r["status"] == 0 and r["best_f"] == want["f"][g] and r["best_group"] == g
        assert np.array_equal(r["best_params"], want["params"][g]) and np.array_equal(r["best_dV"], want["dV"][g])
        assert r["n_evals"] == want["lane_nev"].sum()
