"""The autotuner's register budgets (4, 3, 2 waves per SIMD) compute the
same results bit for bit, and ast_autotune keeps the fastest it measured."""
import numpy as np
import pytest

from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu


def _jobs():
    out = []
    for k, (v, nf) in enumerate(((variants.ls_torque(abi.PREC_FP64), 1), (variants.bc_free(abi.PREC_MIXED), 3),
                                 (variants.bw_ball_torque(abi.PREC_F32SIMPLEX), 2))):
        curve, _ = synth.make_curve(40 + k, v, 80, 1.5, n_filters=nf)
        out.append(dict(space=variants.default_model_space(v), curve=curve, n_fits=512, nstages=2, n_steps=30,
                        seed=9 + k))
    return out


def test_budgets_give_identical_results():
    jobs = _jobs()
    res = {}
    with api.Context(variants.ls_torque()) as ctx:
        for w in (4, 3, 2):
            for j in jobs:
                ctx.set_waves(j["space"].variant, j["curve"].n_filters, w)
            res[w] = ctx.run_jobs(jobs)
    for w in (3, 2):
        for a, b in zip(res[4], res[w]):
            assert a["best_f"] == b["best_f"] and a["n_evals"] == b["n_evals"], w
            assert np.array_equal(a["best_params"], b["best_params"]), w
            assert np.array_equal(a["best_dV"], b["best_dV"]), w


def test_autotune_keeps_the_fastest():
    jobs = _jobs()
    with api.Context(variants.ls_torque()) as ctx:
        rep = ctx.autotune(jobs, cycles=3)
        after = ctx.run_jobs(jobs)
    assert len(rep) == 3
    for e in rep:
        ms = e["kernel_ms"]
        print(e)
        assert all(m > 0 for m in ms)
        assert e["waves"] == (4, 3, 2)[int(np.argmin(ms))]
    assert all(r["status"] == 0 for r in after)
