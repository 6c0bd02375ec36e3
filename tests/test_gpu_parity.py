"""GPU FP64 path against the CPU oracle's canonical form: bitwise chi^2 and
delta_V of the evaluator, and identical simplex trajectories (per-simplex
final chi^2, evaluation and step counts, per-group best model) for a fixed
seed.  Runs through the C ABI (libasteroid_hip.so)."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu

VARIANTS = {
    "ls_torque": (lambda prec: variants.ls_torque(prec), 1, (2.0, 4800.0)),
    "bc": (lambda prec: variants.bc_free(prec), 3, (2.0, 4800.0)),
    "bw_torque": (lambda prec: variants.bw_ball_torque(prec), 2, (2.0, 4800.0)),
    "ls_plain": (lambda prec: variants.make_variant(False, False, False, True, True, prec, 1e-2), 1, (2.0, 4800.0)),
    "ls_noninterp": (lambda prec: variants.ls_torque(prec, interp=False), 2, (2.0, 4800.0)),
}
PRECS = {"fp64": abi.PREC_FP64, "f32simplex": abi.PREC_F32SIMPLEX}


def _setup(name, prec, n_obs=80, span=1.5, reopt=0, seed=11):
    mk, nf, pp = VARIANTS[name]
    v = mk(PRECS[prec])
    curve, truth = synth.make_curve(seed, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v, ppsi_hours=pp, reopt=reopt)
    return v, curve, truth, space


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("name", list(VARIANTS))
def test_chi2_batch_bitwise(name, prec):
    v, curve, truth, space = _setup(name, prec)
    pb = oracle.Problem(space, curve)
    rng = np.random.default_rng(3)
    # random points of the scale-free cube mapped through the oracle's x2params
    x = rng.uniform(0.02, 0.98, size=(300, space.n_params))
    params, rc = pb.x2params(x, reopt=0)
    params = np.vstack([truth, params[rc == 0]])
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got_c, got_dv = ctx.chi2_batch(params)
    print(name, prec, "n", len(params), "max |d chi2|", np.nanmax(np.abs(got_c - want_c)))
    np.testing.assert_array_equal(np.isnan(got_c), np.isnan(want_c))
    ok = ~np.isnan(want_c)
    assert np.array_equal(got_c[ok], want_c[ok])
    assert np.array_equal(got_dv[ok], want_dv[ok])


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("name,nstages,reopt", [("ls_torque", 2, 0), ("bc", 1, 1), ("bw_torque", 1, 0),
                                                ("ls_plain", 2, 0)])
def test_fit_trajectories_identical(name, prec, nstages, reopt):
    v, curve, truth, space = _setup(name, prec, n_obs=60, span=1.0, reopt=reopt)
    pb = oracle.Problem(space, curve)
    G, gs, n_steps, seed = 3, 64, 40, 1234
    want = pb.run(G, gs, nstages=nstages, n_steps=n_steps, seed=seed, reopt=reopt, params0=truth,
                  form=oracle.CANON, n_threads=8)
    with api.Context(v) as ctx:
        ctx.set_model_space(space.with_reopt(reopt))
        ctx.upload_curve(curve)
        ctx.rng_seed(seed)
        ctx.set_start(truth)
        ctx.run(G, gs, nstages=nstages, n_steps=n_steps, ticks_per_launch=37)
        f, p, dv = ctx.fetch()
        lf, lne, lns = ctx.fetch_lanes()
    print(name, prec, "group f", f, "oracle", want["f"])
    assert np.array_equal(lne, want["lane_nev"])
    assert np.array_equal(lns, want["lane_steps"])
    assert np.array_equal(lf, want["lane_f"])
    assert np.array_equal(f, want["f"])
    assert np.array_equal(p, want["params"])
    assert np.array_equal(dv, want["dV"])


@pytest.mark.parametrize("name", ["ls_torque_fp64", "bc_fp64_3f", "bw_torque_f32", "ls_torque_noninterp_f32"])
def test_gpu_matches_golden_vectors(name):
    from tests.test_golden import G, _problem, fx
    case = G[name]
    pb, space = _problem(case)
    params = fx(case["params"], (-1, space.n_params))
    with api.Context(space.variant) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(pb.curve)
        c, dv = ctx.chi2_batch(params)
        assert np.array_equal(c, fx(case["chi2_canon"])) and np.array_equal(dv.ravel(), fx(case["dV_canon"]))
