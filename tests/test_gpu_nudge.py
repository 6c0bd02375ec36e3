"""NUDGE on the GPU (Stage Three: chi^2 nudged by model minima near the
observed ones) against the oracle's canonical form, bit for bit: chi^2 and
delta_V of the evaluator (rewarded, penalised and failed models, the
failures' delta_V NaN), whole fits (per-lane chi^2, evaluation and step
counts, per-group best with the last evaluation's delta_V), the -plot curve
(including a grid pass that fails and leaves the rest of the curve NaN) and
the RMSD intervals; the call order and error rules of ast_set_minima."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu

PRECS = {"fp64": abi.PREC_FP64, "f32simplex": abi.PREC_F32SIMPLEX}


def _case(prec, bc=False, n=300, span=3.0, seed=5, ppsi=(6.0, 9.0), n_min=4, v1s=False):
    v = variants.ls_torque_nudge(prec, bc=bc, v1s=v1s)
    truth = synth.random_truth(np.random.default_rng(seed + 100), v, ppsi_hours=ppsi)
    curve, truth = synth.make_curve(seed, v, n, span, truth=truth)
    space = variants.default_model_space(v)
    # observed minima near the truth's model minima (local V maxima among the observations)
    vb = variants.make_variant(True, bc, False, True, True, prec)
    _, dv, vm = oracle.Problem(variants.default_model_space(vb), curve).chi2(truth, want_vmod=True)
    vm = vm[0]
    idx = [i - 1 for i in range(2, n) if vm[i - 1] > vm[i - 2] and vm[i - 1] >= vm[i]]
    pick = idx[::max(1, len(idx) // n_min)][:n_min]
    minima = (np.array([curve.mjd[i] + 0.01 for i in pick], np.float32),
              np.array([vm[i] + dv[0, 0] + 0.03 for i in pick], np.float32))
    return v, curve, truth, space, minima


def _ctx(v, space, curve, minima):
    ctx = api.Context(v)
    ctx.set_model_space(space)
    ctx.upload_curve(curve)
    ctx.set_minima(*minima)
    return ctx


@pytest.mark.parametrize("bc", [False, True], ids=["ls", "bc"])
@pytest.mark.parametrize("prec", list(PRECS))
def test_nudge_chi2_batch_bitwise(prec, bc):
    v, curve, truth, space, minima = _case(PRECS[prec], bc=bc)
    pb = oracle.Problem(space, curve, minima=minima)
    x = np.random.default_rng(3).uniform(0.02, 0.98, (400, space.n_params))
    p, rc = pb.x2params(x, reopt=0)
    near = truth * (1 + np.random.default_rng(4).uniform(-0.003, 0.003, (64, space.n_params)))
    params = np.vstack([truth, near, p[rc == 0]])
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    with _ctx(v, space, curve, minima) as ctx:
        got_c, got_dv = ctx.chi2_batch(params)
    plain, _ = oracle.Problem(variants.default_model_space(variants.make_variant(True, bc, False, True, True,
                                                                                 PRECS[prec])), curve).chi2(params)
    changed = np.sum(want_c != plain)
    print(prec, bc, "models", len(params), "nudged", changed, "failed", np.sum(want_c >= 1e29))
    assert changed > 10
    np.testing.assert_array_equal(np.isnan(got_c), np.isnan(want_c))
    ok = ~np.isnan(want_c)
    assert np.array_equal(got_c[ok], want_c[ok])
    assert np.array_equal(np.isnan(got_dv), np.isnan(want_dv))
    assert np.array_equal(got_dv[~np.isnan(want_dv)], want_dv[~np.isnan(want_dv)])


def test_nudge_failed_models_bitwise():
    # a fast rotator densely observed with ten observed minima across the data:
    # more than 30 model minima near them fail most models
    v, curve, truth, space, _ = _case(abi.PREC_FP64, n=3000, span=3.0, seed=9, ppsi=(1.2, 1.4))
    minima = (np.linspace(0.2, 2.8, 10).astype(np.float32), np.full(10, 20.0, np.float32))
    pb = oracle.Problem(space, curve, minima=minima)
    near = truth * (1 + np.random.default_rng(5).uniform(-0.01, 0.01, (40, space.n_params)))
    params = np.vstack([truth, near])
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    with _ctx(v, space, curve, minima) as ctx:
        got_c, got_dv = ctx.chi2_batch(params)
    assert np.sum(want_c == 1e30) > 10
    assert np.array_equal(got_c, want_c)
    assert np.array_equal(np.isnan(got_dv), np.isnan(want_dv)) and np.isnan(got_dv[want_c == 1e30]).all()


@pytest.mark.parametrize("prec", list(PRECS))
def test_nudge_fit_trajectories_identical(prec):
    # Stage Three: reoptimisation from a model (README.md:87-101)
    v, curve, truth, space, minima = _case(PRECS[prec], n=120, span=1.5, seed=11, n_min=2)
    space1 = space.with_reopt(1)
    pb = oracle.Problem(space1, curve, minima=minima)
    start = truth * 1.01
    G, gs, n_steps, seed = 2, 64, 40, 77
    want = pb.run(G, gs, nstages=1, n_steps=n_steps, seed=seed, reopt=1, params0=start, form=oracle.CANON,
                  n_threads=8)
    with _ctx(v, space1, curve, minima) as ctx:
        ctx.rng_seed(seed)
        ctx.set_start(start)
        ctx.run(G, gs, nstages=1, n_steps=n_steps, ticks_per_launch=29)
        f, p, dv = ctx.fetch()
        lf, lne, lns = ctx.fetch_lanes()
    print(prec, "group f", f, "oracle", want["f"])
    assert np.array_equal(lne, want["lane_nev"]) and np.array_equal(lns, want["lane_steps"])
    assert np.array_equal(lf, want["lane_f"])
    assert np.array_equal(f, want["f"]) and np.array_equal(p, want["params"]) and np.array_equal(dv, want["dV"])


@pytest.mark.parametrize("ppsi,failing", [((6.0, 9.0), False), ((1.2, 1.4), True)], ids=["ok", "grid_fails"])
def test_nudge_model_curve_bitwise(ppsi, failing):
    # failing: 60 observations of a fast rotator pass (fewer than 31 model
    # minima among them), the 20000-point grid does not: its curve is cut
    n = 60 if failing else 300
    v, curve, truth, space, minima = _case(abi.PREC_F32SIMPLEX, n=n, span=3.0, seed=13, ppsi=ppsi, n_min=2)
    if failing:
        minima = (np.linspace(0.2, 2.8, 10).astype(np.float32), np.full(10, 20.0, np.float32))
    pb = oracle.Problem(space, curve, minima=minima)
    t, vm, c2, dv = pb.model_curve(truth, n_plot=20000)
    with _ctx(v, space, curve, minima) as ctx:
        gt, gvm, gc2, gdv = ctx.model_curve(truth, n_plot=20000)
    nan = np.isnan(vm)
    print("NaN grid points:", nan.sum(), "chi2", c2)
    assert c2[0] < 1e29 and nan.any() == failing
    if failing:  # a tail of NaN after the point where the grid pass failed
        first = np.argmax(nan[0])
        assert 0 < first and nan[0, first:].all()
    assert np.array_equal(gt, t) and np.array_equal(np.isnan(gvm), nan)
    assert np.array_equal(gvm[~nan], vm[~nan])
    assert np.array_equal(gc2, c2) and np.array_equal(np.isnan(gdv), np.isnan(dv))
    assert np.array_equal(gdv[~np.isnan(dv)], dv[~np.isnan(dv)])


def test_nudge_rmsd_bitwise():
    v, curve, truth, space, minima = _case(abi.PREC_F32SIMPLEX, n=150, span=2.0, seed=15)
    space1 = variants.default_model_space(v, reopt=1)
    with _ctx(v, space1, curve, minima) as ctx:
        ctx.rng_seed(21)
        ctx.set_start(truth)
        got = ctx.rmsd(2, 64, 4, 3e-4)
    want = oracle.Problem(space1, curve, minima=minima).rmsd(2, 64, 4, 3e-4, truth, seed=21, n_threads=8)
    assert np.array_equal(got["par_min"], want["par_min"]) and np.array_equal(got["par_max"], want["par_max"])
    assert got["n_tot"] == want["counts"][:, 0].sum() and got["n_bad"] == want["counts"][:, 1].sum()
    assert got["f0"] == want["f0"] and got["f1"] == want["f1"]


def test_set_minima_rules():
    v, curve, truth, space, minima = _case(abi.PREC_FP64, n=120, span=1.5, seed=17)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        with pytest.raises(api.AsteroidError):  # no minima yet
            ctx.chi2_batch(truth[None])
        with pytest.raises(api.AsteroidError):  # more than AST_MAX_MINIMA
            ctx.set_minima(np.zeros(11), np.zeros(11))
        ctx.set_minima(*minima)
        a, _ = ctx.chi2_batch(truth[None])
    # minima before the curve: the same result
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.set_minima(*minima)
        ctx.upload_curve(curve)
        b, _ = ctx.chi2_batch(truth[None])
    assert a[0] == b[0]
    with api.Context(variants.ls_torque()) as ctx:  # not a NUDGE context
        with pytest.raises(api.AsteroidError):
            ctx.set_minima(*minima)
