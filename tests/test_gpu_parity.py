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
    "trend": (lambda prec: variants.ls_torque_trend(prec), 2, (2.0, 4800.0)),
    "torque2": (lambda prec: variants.ls_torque2(prec), 1, (2.0, 4800.0)),
    "bc_rotate": (lambda prec: variants.bc_rotate(prec), 2, (2.0, 4800.0)),
    "bc_rotate_torque": (lambda prec: variants.bc_rotate(prec, torque=True), 1, (2.0, 4800.0)),
    "rect": (lambda prec: variants.rect(prec), 2, (2.0, 4800.0)),
    "rect_torque": (lambda prec: variants.rect(prec, torque=True), 1, (2.0, 4800.0)),
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


@pytest.mark.parametrize("name", ["ls_torque", "bc", "bw_torque", "rect_torque", "torque2"])
def test_chi2_batch_bitwise_wide(name):
    """A wider sweep for the evaluator's hardware-refined division, square root
    and log: 4096 models over the whole scale-free cube (its faces x = 0 and
    x = 1 included), a 400-point curve over 3 days, FP64 form, bitwise."""
    v, curve, truth, space = _setup(name, "fp64", n_obs=400, span=3.0, seed=29)
    pb = oracle.Problem(space, curve)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.0, 1.0, size=(4096, space.n_params))
    faces = rng.random(size=x.shape)
    x[faces < 0.05] = 0.0
    x[faces > 0.95] = 1.0
    params, rc = pb.x2params(x, reopt=0)
    params = np.vstack([truth, params[rc == 0]])
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got_c, got_dv = ctx.chi2_batch(params)
    print(name, "n", len(params), "nan", int(np.isnan(want_c).sum()))
    assert len(params) > 1000
    np.testing.assert_array_equal(np.isnan(got_c), np.isnan(want_c))
    ok = ~np.isnan(want_c)
    assert np.array_equal(got_c[ok], want_c[ok])
    assert np.array_equal(got_dv[ok], want_dv[ok])


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("name,nstages,reopt", [("ls_torque", 2, 0), ("bc", 1, 1), ("bw_torque", 1, 0),
                                                ("ls_plain", 2, 0), ("trend", 1, 0), ("torque2", 1, 0),
                                                ("bc_rotate_torque", 1, 1), ("rect", 1, 0)])
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


EDGE = {
    # name: (variant, n_filters, n_obs, span_days)
    "min_dof_ls": (lambda: variants.ls_torque(abi.PREC_FP64), 1, 13, 0.5),       # 11 params + 1 filter + 1
    "nf8_class_bc": (lambda: variants.bc_free(abi.PREC_FP64), 6, 300, 2.0),     # the 8-filter kernels
    "long_6000": (lambda: variants.ls_torque(abi.PREC_F32SIMPLEX), 2, 6000, 10.0),
    "sparse_rk4": (lambda: variants.bw_ball_torque(abi.PREC_FP64), 1, 30, 8.0),  # ~27 RK4 steps a point
}


@pytest.mark.parametrize("name", list(EDGE))
def test_chi2_edge_shapes_bitwise(name):
    mk, nf, n_obs, span = EDGE[name]
    v = mk()
    curve, truth = synth.make_curve(17, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    x = np.random.default_rng(4).uniform(0.02, 0.98, size=(200, space.n_params))
    params, rc = pb.x2params(x, reopt=0)
    params = np.vstack([truth, params[rc == 0]])
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got_c, got_dv = ctx.chi2_batch(params)
    ok = ~np.isnan(want_c)
    assert np.array_equal(np.isnan(got_c), ~ok)
    assert np.array_equal(got_c[ok], want_c[ok]) and np.array_equal(got_dv[ok], want_dv[ok])


@pytest.mark.parametrize("name,G,gs", [("nf8_class_bc", 2, 64), ("min_dof_ls", 1, 1024), ("sparse_rk4", 3, 64)])
def test_fit_edge_shapes_identical(name, G, gs):
    mk, nf, n_obs, span = EDGE[name]
    v = mk()
    curve, truth = synth.make_curve(18, v, n_obs, span, n_filters=nf)
    space = variants.default_model_space(v)
    pb = oracle.Problem(space, curve)
    want = pb.run(G, gs, nstages=2, n_steps=25, seed=77, n_threads=8)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        ctx.rng_seed(77)
        ctx.run(G, gs, nstages=2, n_steps=25, ticks_per_launch=11)
        f, p, dv = ctx.fetch()
        lf, ln, ls = ctx.fetch_lanes()
    assert np.array_equal(lf, want["lane_f"]) and np.array_equal(ln, want["lane_nev"])
    assert np.array_equal(ls, want["lane_steps"])
    assert np.array_equal(f, want["f"]) and np.array_equal(p, want["params"]) and np.array_equal(dv, want["dV"])
