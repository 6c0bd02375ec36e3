"""Single-stage runs on fewer workgroups than they have groups (the device
work queue): a lane whose fit ends starts the job's next unstarted fit.
Each fit's trajectory depends only on its own index (random stream, start),
so per-fit results, evaluation and step counts and the per-group best models
equal the oracle's run of the same job bitwise, however the fits were
spread over lanes; a run driven cycle by cycle (ast_run_begin /
ast_run_continue) equals one ast_run; serving jobs admitted on part of
their groups' slots equal the same jobs alone."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu


def _case(v, nf=1, n_obs=60, span=1.0, reopt=0, seed=11):
    curve, truth = synth.make_curve(seed, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v, reopt=reopt)
    return curve, truth, space


@pytest.mark.parametrize("name,mk,nf,reopt,resident", [
    ("ls_torque_fp64", lambda: variants.ls_torque(abi.PREC_FP64), 1, 0, 1),
    ("bc_f32_reopt", lambda: variants.bc_free(abi.PREC_F32SIMPLEX), 3, 1, 2),
    ("bw_torque_fp64", lambda: variants.bw_ball_torque(abi.PREC_FP64), 2, 0, 3),
    ("torque2_f32", lambda: variants.ls_torque2(abi.PREC_F32SIMPLEX), 1, 0, 1),
])
def test_refill_matches_oracle(name, mk, nf, reopt, resident):
    v = mk()
    curve, truth, space = _case(v, nf=nf, reopt=reopt)
    G, gs, n_steps, seed = 5, 64, 40, 4321
    want = oracle.Problem(space, curve).run(G, gs, nstages=1, n_steps=n_steps, seed=seed, reopt=reopt,
                                            params0=truth, form=oracle.CANON, n_threads=8)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(seed)
        ctx.set_start(truth)
        ctx.run(G, gs, nstages=1, n_steps=n_steps, ticks_per_launch=23, resident_groups=resident)
        f, p, dv = ctx.fetch()
        lf, lne, lns = ctx.fetch_lanes()
    print(name, "groups", f, "oracle", want["f"])
    assert np.array_equal(lne, want["lane_nev"])
    assert np.array_equal(lns, want["lane_steps"])
    assert np.array_equal(lf, want["lane_f"])
    assert np.array_equal(f, want["f"])
    assert np.array_equal(p, want["params"])
    assert np.array_equal(dv, want["dV"])


@pytest.mark.parametrize("prec", [abi.PREC_MIXED, abi.PREC_FP64])
def test_run_in_cycles_equals_one_run(prec):
    v = variants.ls_torque(prec)
    curve, truth, space = _case(v, n_obs=80, span=1.5, seed=12)
    G, gs, budget, seed = 6, 64, 300, 99
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(seed)
        ctx.run(G, gs, nstages=1, n_steps=5000, eval_budget=budget, ticks_per_launch=50)
        f0, p0, d0 = ctx.fetch()
        l0 = ctx.fetch_lanes()
        e0, _ = ctx.last_counts()
        ctx.rng_seed(seed)
        ctx.run_begin(G, gs, nstages=1, n_steps=5000, eval_budget=budget, ticks_per_launch=50, resident_groups=2)
        cycles, last_done = 0, 0
        while not ctx.run_continue(1):
            started, done, fin = ctx.run_status()
            assert not fin and done >= last_done and started <= G * gs
            last_done = done
            cycles += 1
            assert cycles < 1000
        started, done, fin = ctx.run_status()
        assert fin and started == done == G * gs
        f1, p1, d1 = ctx.fetch()
        l1 = ctx.fetch_lanes()
        e1, _ = ctx.last_counts()
    assert cycles > 3
    assert e0 == e1 == l0[1].sum()
    assert np.array_equal(f0, f1) and np.array_equal(p0, p1) and np.array_equal(d0, d1)
    for a, b in zip(l0, l1):
        assert np.array_equal(a, b)


def test_run_continue_refuses_conflicting_calls():
    v = variants.ls_torque(abi.PREC_FP64)
    curve, truth, space = _case(v)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(1)
        ctx.run_begin(4, 64, nstages=1, n_steps=30, resident_groups=1, ticks_per_launch=5)
        ctx.run_continue(1)
        with pytest.raises(RuntimeError, match="AST_E_STATE"):
            ctx.rng_seed(2)
        with pytest.raises(RuntimeError, match="AST_E_STATE"):
            ctx.run_jobs([dict(space=space, curve=curve, n_fits=64, nstages=1, n_steps=10, seed=1)])
        assert ctx.run_continue(0)
        ctx.rng_seed(2)  # allowed again once the run has ended
        ctx.run_begin(4, 64, nstages=1, n_steps=30, resident_groups=1, ticks_per_launch=5)
        ctx.run_continue(2)
        ctx.run_cancel()  # or once it is cancelled
        ctx.rng_seed(3)


def test_serving_partial_admission_equals_alone():
    # a pool of 5 slots: the 2-group job takes 2; the 8-group job A waits for
    # half its groups' worth (4) and the second 8-group job starts on what its
    # predecessors free; each takes the rest of its fits from its own queue
    vA = variants.ls_torque(abi.PREC_FP64)
    vB = variants.ls_torque(abi.PREC_MIXED)
    A = _case(vA, n_obs=70, span=1.2, seed=31)
    B = _case(vB, n_obs=90, span=1.5, seed=32)
    with api.Context(vA) as ctx:
        ctx.serve_start(max_resident_fits=5 * 256)
        a_alone = ctx.wait_all([ctx.submit(A[2], A[0], n_fits=8 * 256, nstages=1, n_steps=40, seed=3)])
        ids = [ctx.submit(B[2], B[0], n_fits=2 * 256, nstages=1, n_steps=40, seed=5),
               ctx.submit(A[2], A[0], n_fits=8 * 256, nstages=1, n_steps=40, seed=3),
               ctx.submit(B[2], B[0], n_fits=8 * 256, nstages=1, n_steps=40, seed=4)]
        res = ctx.wait_all(ids)
        ctx.serve_stop()
    a_alone = list(a_alone.values())[0]
    mixed = res[ids[1]]
    for k in ("status", "best_f", "best_group", "n_evals"):
        assert a_alone[k] == mixed[k], k
    assert np.array_equal(a_alone["best_params"], mixed["best_params"])
    want = oracle.Problem(A[2], A[0]).run(8, 256, nstages=1, n_steps=40, seed=3, n_threads=8)
    g = int(np.argmin(want["f"]))
    assert mixed["best_f"] == want["f"][g] and np.array_equal(mixed["best_params"], want["params"][g])
    assert mixed["n_evals"] == want["lane_nev"].sum()
    # the mixed-precision jobs: their reported chi^2 is within 1e-6 of the
    # FP64 oracle's chi^2 of the model they report
    for i in (ids[0], ids[2]):
        r = res[i]
        assert r["status"] == 0
        space64 = variants.default_model_space(variants.ls_torque(abi.PREC_FP64))
        c64, _ = oracle.Problem(space64, B[0]).chi2(r["best_params"][None])
        assert abs(r["best_f"] - c64[0]) <= 1e-6 * c64[0] + 2.0 ** -23 * c64[0]
