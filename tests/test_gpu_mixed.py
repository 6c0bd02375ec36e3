"""The mixed-precision form (AST_PREC_MIXED) against the FP64 oracle.

The mixed form integrates the Euler angles and angular velocity in FP64
exactly as the FP64 form does, evaluates each observation's geometry and
brightness in FP32, and keeps the chi^2 sums in FP64.  Its gate is a
relative chi^2 error of at most 1e-6 against the FP64 oracle, for every
evaluation, on light curves of the benchmark configurations (400 points over
3 days with torque; 818 points in 3 filters over 5 days, TIME_STEP 5e-3, for
BC) and on the other variants.  The evaluated models are random points of
the whole model space (slow to fast rotators, all shapes) plus the
ground-truth model and models near it.
"""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu

REL_TOL = 1e-6  # relative chi^2 error of the mixed form against FP64

CASES = {
    # name: (variant factory, n_filters, n_obs, span_days)
    "ls_torque_400x3d": (lambda p: variants.ls_torque(p), 1, 400, 3.0),
    "bc_818x5d_3f": (lambda p: variants.bc_free(p), 3, 818, 5.0),
    "bw_torque": (lambda p: variants.bw_ball_torque(p), 2, 200, 2.0),
    "ls_plain": (lambda p: variants.make_variant(False, False, False, True, True, p, 1e-2), 1, 200, 2.0),
    "ls_noninterp": (lambda p: variants.ls_torque(p, interp=False), 2, 150, 1.5),
    "trend": (lambda p: variants.ls_torque_trend(p), 2, 200, 2.0),
    "torque2": (lambda p: variants.ls_torque2(p), 1, 200, 2.0),
    "bc_rotate_torque": (lambda p: variants.bc_rotate(p, torque=True), 2, 200, 2.0),
    "rect_torque": (lambda p: variants.rect(p, torque=True), 1, 200, 2.0),
}


def _models(space, truth, pb, n_random, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.02, 0.98, size=(n_random, space.n_params))
    params, rc = pb.x2params(x, reopt=0)
    near = truth * (1.0 + rng.uniform(-0.01, 0.01, size=(64, space.n_params)))
    return np.vstack([truth, near, params[rc == 0]])


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_chi2_within_tolerance(name):
    mk, nf, n_obs, span = CASES[name]
    v = mk(abi.PREC_MIXED)
    curve, truth = synth.make_curve(7, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    params = _models(space, truth, pb, 1500, 5)
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got_c, got_dv = ctx.chi2_batch(params)
    ok = np.isfinite(want_c)
    assert np.array_equal(np.isfinite(got_c), ok)
    rel = np.abs(got_c[ok] - want_c[ok]) / want_c[ok]
    ddv = np.abs(got_dv[ok] - want_dv[ok])
    print(f"{name}: n {ok.sum()} rel chi2 err median {np.median(rel):.2e} p99 {np.quantile(rel, 0.99):.2e} "
          f"max {rel.max():.2e}; max |d dV| {ddv.max():.2e} mag; truth rel {rel[0]:.2e}")
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_chi2_within_tolerance_whole_cube(name):
    """The gate over the whole scale-free cube, its faces (x = 0, x = 1)
    included: the extreme shapes and spins there still stay within 1e-6, and
    the models whose chi^2 is not finite are the same in both forms."""
    mk, nf, n_obs, span = CASES[name]
    v = mk(abi.PREC_MIXED)
    curve, truth = synth.make_curve(7, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    rng = np.random.default_rng(12)
    x = rng.uniform(0.0, 1.0, size=(2048, space.n_params))
    faces = rng.random(size=x.shape)
    x[faces < 0.05] = 0.0
    x[faces > 0.95] = 1.0
    params, rc = pb.x2params(x, reopt=0)
    params = params[rc == 0]
    want_c, _ = pb.chi2(params, form=oracle.CANON)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got_c, _ = ctx.chi2_batch(params)
    ok = np.isfinite(want_c)
    assert ok.sum() > 1000
    assert np.array_equal(np.isfinite(got_c), ok)
    rel = np.abs(got_c[ok] - want_c[ok]) / want_c[ok]
    print(f"{name}: n {ok.sum()} of {len(params)} finite, max rel chi2 err {rel.max():.2e}")
    assert rel.max() <= REL_TOL


@pytest.mark.parametrize("name", ["ls_torque_400x3d", "bc_818x5d_3f"])
def test_mixed_fit_best_matches_fp64_chi2(name):
    """A mixed-precision fit's reported best chi^2 is the FP64 chi^2 of its
    reported best model within the gate (plus the float rounding of the
    reported value)."""
    mk, nf, n_obs, span = CASES[name]
    v = mk(abi.PREC_MIXED)
    curve, truth = synth.make_curve(8, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(77)
        ctx.run(8, 128, nstages=2, n_steps=150)
        f, p, dv = ctx.fetch()
    # A group can end without a result (f = 1e30): when the stage-one best
    # point, converted from P_psi to L (cuda.c:1791-1798), has L outside its
    # hard limits, every stage-two vertex is rejected.  The oracle does the same.
    ok = f < 1e29
    assert ok.sum() >= len(f) // 2
    want, _ = pb.chi2(p[ok], form=oracle.CANON)
    rel = np.abs(f[ok] - want) / want
    print(f"{name}: best-of-group rel err max {rel.max():.2e} over {ok.sum()} groups; best chi2 {f.min():.6g}")
    assert rel.max() <= REL_TOL + 2.0 ** -23
