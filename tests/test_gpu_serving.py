"""Serving (continuous batching) on the GPU: a job's result equals the
oracle's run of the same job, is identical alone and in a mix of other jobs
(other curves, variants and filter counts admitted while it runs), and
cancellation and admission behave."""
import time

import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu


def _job(seed, v, n_obs, span, nf=1, reopt=0):
    curve, truth = synth.make_curve(seed, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v, reopt=reopt)
    return space, curve, truth


def test_job_matches_oracle():
    v = variants.ls_torque()
    space, curve, truth = _job(21, v, 70, 1.2)
    with api.Context(v) as ctx:
        ctx.serve_start(max_resident_fits=4096)
        jid = ctx.submit(space, curve, n_fits=512, nstages=2, n_steps=30, seed=5)
        r = ctx.wait_all([jid])[jid]
        ctx.serve_stop()
    want = oracle.Problem(space, curve).run(2, 256, nstages=2, n_steps=30, seed=5, n_threads=8)
    g = int(np.argmin(want["f"]))
    assert r["status"] == 0
    assert r["best_f"] == want["f"][g] and r["best_group"] == g
    assert np.array_equal(r["best_params"], want["params"][g])
    assert np.array_equal(r["best_dV"], want["dV"][g])
    assert r["n_evals"] == want["lane_nev"].sum()


def test_result_same_alone_and_in_the_mix():
    vA = variants.ls_torque()
    A = _job(31, vA, 80, 1.5)
    others = [(_job(32, variants.bc_free(), 120, 2.0, nf=3), 1024),
              (_job(33, variants.bw_ball_torque(abi.PREC_F32SIMPLEX), 60, 1.0, nf=2), 768),
              (_job(34, variants.ls_torque(abi.PREC_F32SIMPLEX), 200, 3.0), 512),
              (_job(35, vA, 40, 0.8), 256)]
    with api.Context(vA) as ctx:
        ctx.serve_start(max_resident_fits=8192)
        a1 = ctx.submit(*A[:2], n_fits=512, nstages=2, n_steps=40, seed=9)
        alone = ctx.wait_all([a1])[a1]
        ids = [ctx.submit(*o[0][:2], n_fits=o[1], nstages=1, n_steps=60, seed=11 + k) for k, o in enumerate(others[:2])]
        a2 = ctx.submit(*A[:2], n_fits=512, nstages=2, n_steps=40, seed=9)
        time.sleep(0.01)
        ids += [ctx.submit(*o[0][:2], n_fits=o[1], nstages=1, n_steps=60, seed=21 + k) for k, o in enumerate(others[2:])]
        res = ctx.wait_all(ids + [a2])
        ctx.serve_stop()
    mixed = res[a2]
    for k in ("best_f", "best_group", "n_evals"):
        assert alone[k] == mixed[k], k
    assert np.array_equal(alone["best_params"], mixed["best_params"])
    assert np.array_equal(alone["best_dV"], mixed["best_dV"])
    assert all(res[i]["status"] == 0 for i in ids)


def test_cancel_and_queueing():
    v = variants.ls_torque()
    space, curve, truth = _job(41, v, 100, 2.0)
    with api.Context(v) as ctx:
        ctx.serve_start(max_resident_fits=512)   # room for one 512-fit job at a time
        big = ctx.submit(space, curve, n_fits=512, nstages=1, n_steps=100000, eval_budget=100000, seed=1)
        queued = [ctx.submit(space, curve, n_fits=512, nstages=1, n_steps=20, seed=2 + k) for k in range(3)]
        time.sleep(0.2)
        st = ctx.serve_stats()
        assert st["resident"] == 1 and st["pending"] == 3
        ctx.cancel(big)
        res = ctx.wait_all([big] + queued)
        ctx.serve_stop()
    assert res[big]["status"] == abi.E_CANCELLED
    for q in queued:
        assert res[q]["status"] == 0 and res[q]["queue_ms"] > 0
