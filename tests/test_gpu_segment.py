"""SEGMENT on the GPU against the oracle: for -DSEGMENT -DN_SEG=3 -DONE_LE
(LS), chi^2 and delta_V of the evaluator bit for bit (FP64 and float-simplex
forms; the mixed form within its 1e-6 relative tolerance), whole fits lane
by lane, observations before the first segment and an observation past two
start times, the -plot curve (grid points before the first start NaN), the
call rules of ast_set_segments and the refusals; for all eight SEGMENT
builds (three or four segments, with and without ONE_LE, TORQUE and BC:
16 to 40 parameters), chi^2, delta_V and two-stage fits bit for bit and the mixed
form within tolerance; slot pools cut for the largest job's parameter
count."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu

PRECS = {"fp64": abi.PREC_FP64, "f32simplex": abi.PREC_F32SIMPLEX, "mixed": abi.PREC_MIXED}


def _case(prec, n=240, span=3.0, seed=6, skip=0):
    v8 = variants.make_variant(False, False, False, True, True, prec)
    curve, truth8 = synth.make_curve(seed, v8, n, span)
    v = variants.ls_segments(prec)
    space = variants.default_model_space(v)
    t0 = curve.mjd[skip] if skip else curve.mjd[0] - 0.01
    t_start = np.array([t0, curve.mjd[0] + span / 3, curve.mjd[0] + 2 * span / 3])
    pos8 = variants.param_positions(v8)
    rng = np.random.default_rng(seed)
    truth = np.array([truth8[pos8[name.split(":")[0]]] + (rng.uniform(-0.2, 0.2) if ":" in name else 0.0)
                      for name in variants.param_positions(v)])
    return v, curve, truth, space, t_start


def _ctx(v, space, curve, t_start):
    ctx = api.Context(v)
    ctx.set_model_space(space)
    ctx.set_segments(t_start)
    ctx.upload_curve(curve)
    return ctx


def _models(pb, space, truth, n=300, seed=3):
    x = np.random.default_rng(seed).uniform(0.02, 0.98, (n, space.n_params))
    p, rc = pb.x2params(x, reopt=0)
    near = truth * (1 + np.random.default_rng(seed + 1).uniform(-0.003, 0.003, (32, space.n_params)))
    return np.vstack([truth, near, p[rc == 0]])


@pytest.mark.parametrize("skip", [0, 7], ids=["all", "first7_out"])
@pytest.mark.parametrize("prec", ["fp64", "f32simplex"])
def test_segment_chi2_batch_bitwise(prec, skip):
    v, curve, truth, space, t_start = _case(PRECS[prec], skip=skip)
    pb = oracle.Problem(space, curve, t_start=t_start)
    params = _models(pb, space, truth)
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    print(prec, skip, "models", len(params), flush=True)
    with _ctx(v, space, curve, t_start) as ctx:
        got_c, got_dv = ctx.chi2_batch(params)
    assert np.array_equal(got_c, want_c)
    assert np.array_equal(got_dv, want_dv)


def test_segment_mixed_within_tolerance():
    v, curve, truth, space, t_start = _case(PRECS["mixed"])
    pb = oracle.Problem(variants.default_model_space(variants.ls_segments(abi.PREC_FP64)), curve, t_start=t_start)
    params = _models(pb, space, truth)
    want_c, _ = pb.chi2(params, form=oracle.CANON)
    with _ctx(v, space, curve, t_start) as ctx:
        got_c, _ = ctx.chi2_batch(params)
    ok = want_c < 1e29
    rel = np.abs(got_c[ok] - want_c[ok]) / want_c[ok]
    print("mixed max rel", rel.max(), flush=True)
    assert rel.max() < 1e-6


def test_segment_observation_past_two_starts():
    v, curve, truth, space, _ = _case(abi.PREC_FP64, seed=8)
    i = 100
    t_start = np.array([curve.mjd[0], curve.mjd[i] - 1e-4, curve.mjd[i] - 5e-5])  # segment 2 opens at i + 1
    pb = oracle.Problem(space, curve, t_start=t_start)
    params = _models(pb, space, truth, n=64)
    want_c, want_dv = pb.chi2(params)
    with _ctx(v, space, curve, t_start) as ctx:
        got_c, got_dv = ctx.chi2_batch(params)
    assert np.array_equal(got_c, want_c) and np.array_equal(got_dv, want_dv)


@pytest.mark.parametrize("prec", ["fp64", "f32simplex"])
def test_segment_fit_trajectories_identical(prec):
    v, curve, truth, space, t_start = _case(PRECS[prec], n=150, span=2.0, seed=12, skip=3)
    pb = oracle.Problem(space, curve, t_start=t_start)
    G, gs, n_steps, seed = 2, 64, 60, 31
    want = pb.run(G, gs, nstages=2, n_steps=n_steps, seed=seed, form=oracle.CANON, n_threads=8)
    print(prec, "oracle f", want["f"], flush=True)
    with _ctx(v, space, curve, t_start) as ctx:
        ctx.rng_seed(seed)
        ctx.run(G, gs, nstages=2, n_steps=n_steps, ticks_per_launch=37)
        f, p, dv = ctx.fetch()
        lf, lne, lns = ctx.fetch_lanes()
    assert np.array_equal(lne, want["lane_nev"]) and np.array_equal(lns, want["lane_steps"])
    assert np.array_equal(lf, want["lane_f"])
    assert np.array_equal(f, want["f"]) and np.array_equal(p, want["params"]) and np.array_equal(dv, want["dV"])


def test_segment_call_rules():
    v, curve, truth, space, t_start = _case(abi.PREC_FP64, n=60)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        with pytest.raises(api.AsteroidError):
            ctx.upload_curve(curve)  # start times first
        with pytest.raises(api.AsteroidError):
            ctx.set_segments(t_start[:2])  # three segments
        with pytest.raises(api.AsteroidError):
            ctx.set_segments([t_start[0], t_start[1], curve.mjd[-1] + 1.0])
            ctx.upload_curve(curve)  # the third segment has no observation
        ctx.set_segments(t_start)
        ctx.upload_curve(curve)
        with pytest.raises(api.AsteroidError):
            ctx.rmsd(1, 64, 4, 0.003)
    with api.Context(variants.ls_torque()) as ctx:
        with pytest.raises(api.AsteroidError):
            ctx.set_segments(t_start)  # not a SEGMENT context


@pytest.mark.parametrize("skip", [0, 9], ids=["all", "first9_out"])
@pytest.mark.parametrize("prec", ["fp64", "f32simplex"])
def test_segment_model_curve_bitwise(prec, skip):
    v, curve, truth, space, t_start = _case(PRECS[prec], skip=skip)
    pb = oracle.Problem(space, curve, t_start=t_start)
    params = _models(pb, space, truth, n=16)[:8]
    t_w, vm_w, c_w, dv_w = pb.model_curve(params, n_plot=3000, form=oracle.CANON, n_threads=4)
    with _ctx(v, space, curve, t_start) as ctx:
        t_g, vm_g, c_g, dv_g = ctx.model_curve(params, n_plot=3000)
    g0 = int(np.searchsorted(t_w, t_start[0]))
    print(prec, skip, "grid points before the first start", g0, flush=True)
    assert (g0 > 0) == (skip > 0)
    assert np.array_equal(t_g, t_w) and np.array_equal(c_g, c_w) and np.array_equal(dv_g, dv_w)
    assert np.all(np.isnan(vm_g[:, :g0])) and np.all(np.isnan(vm_w[:, :g0]))
    assert np.array_equal(vm_g[:, g0:], vm_w[:, g0:])


@pytest.mark.parametrize("prec", ["fp64", "f32simplex"])
def test_segment_starts_by_index_bitwise(prec):
    # ast_set_segment_starts: the observations split at given indices (here
    # each segment opening two observations before its start time, as a split
    # on the observer's clock can), the grid at the start times
    v, curve, truth, space, t_start = _case(PRECS[prec], skip=5)
    by_time = [int(np.argmax(curve.mjd >= t)) for t in t_start]
    first = [by_time[0]] + [k - 2 for k in by_time[1:]]
    pb = oracle.Problem(space, curve, t_start=t_start, seg_first=first)
    params = _models(pb, space, truth, n=120)
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    assert not np.array_equal(want_c, oracle.Problem(space, curve, t_start=t_start).chi2(params)[0])
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.set_segment_starts(first, t_start)
        ctx.upload_curve(curve)
        got_c, got_dv = ctx.chi2_batch(params)
        res = ctx.run_jobs([dict(space=space, curve=curve, n_fits=256, nstages=1, n_steps=20, seed=3,
                                 t_segments=t_start, seg_first_obs=first)])
    assert np.array_equal(got_c, want_c) and np.array_equal(got_dv, want_dv)
    want = pb.run(1, 256, nstages=1, n_steps=20, seed=3, n_threads=8)
    assert res[0]["best_f"] == want["f"][0] and np.array_equal(res[0]["best_params"], want["params"][0])


SEG_BUILDS = [(n, le, tq, bc) for n in (3, 4) for le in (True, False) for tq in (False, True) for bc in (False, True)]


def _build_case(prec, n_seg, one_le, torque, bc=False, seed=9, n=150, span=4.0):
    v0 = variants.make_variant(torque, bc, False, True, True, prec)
    v = variants.ls_segments(prec, n_seg=n_seg, one_le=one_le, torque=torque, bc=bc)
    curve, truth0 = synth.make_curve(seed, v0, n, span, n_filters=2)
    t_start = curve.mjd[0] + 0.3 + np.arange(n_seg) * (span - 0.3) / n_seg  # the first observations left out
    pos0, pos = variants.param_positions(v0), variants.param_positions(v)
    rng = np.random.default_rng(seed)
    truth = np.array([truth0[pos0[k.split(":")[0]]] for k in pos])
    for k, i in pos.items():
        if ":" in k:
            truth[i] *= 1.0 + rng.uniform(-0.02, 0.02)
    return v, curve, truth, variants.default_model_space(v), t_start


def _bid(b):
    return f"seg{b[0]}{'_onele' if b[1] else ''}{'_torque' if b[2] else ''}{'_bc' if b[3] else ''}"


@pytest.mark.parametrize("build", SEG_BUILDS, ids=_bid)
@pytest.mark.parametrize("prec", ["fp64", "f32simplex"])
def test_segment_builds_chi2_and_fits_bitwise(prec, build):
    v, curve, truth, space, t_start = _build_case(PRECS[prec], *build)
    print(prec, _bid(build), "params", space.n_params, flush=True)
    pb = oracle.Problem(space, curve, t_start=t_start)
    params = _models(pb, space, truth, n=200)
    want_c, want_dv = pb.chi2(params, form=oracle.CANON)
    want = pb.run(2, 64, nstages=2, n_steps=25, seed=13, n_threads=8)
    with _ctx(v, space, curve, t_start) as ctx:
        got_c, got_dv = ctx.chi2_batch(params)
        ctx.rng_seed(13)
        ctx.run(2, 64, nstages=2, n_steps=25, ticks_per_launch=19)
        f, p, dv = ctx.fetch()
        lf, ln, ls = ctx.fetch_lanes()
    ok = ~np.isnan(want_c)
    assert np.array_equal(np.isnan(got_c), ~ok)
    assert np.array_equal(got_c[ok], want_c[ok]) and np.array_equal(got_dv[ok], want_dv[ok])
    assert np.array_equal(lf, want["lane_f"]) and np.array_equal(ln, want["lane_nev"])
    assert np.array_equal(ls, want["lane_steps"])
    assert np.array_equal(f, want["f"]) and np.array_equal(p, want["params"]) and np.array_equal(dv, want["dV"])


@pytest.mark.parametrize("build", SEG_BUILDS, ids=_bid)
def test_segment_builds_mixed_within_tolerance(build):
    v, curve, truth, space, t_start = _build_case(abi.PREC_MIXED, *build)
    v64, _, _, space64, _ = _build_case(abi.PREC_FP64, *build)
    pb = oracle.Problem(space64, curve, t_start=t_start)
    params = _models(pb, space64, truth, n=300)
    want_c, _ = pb.chi2(params, form=oracle.CANON)
    with _ctx(v, space, curve, t_start) as ctx:
        got_c, _ = ctx.chi2_batch(params)
    ok = np.isfinite(want_c) & (want_c < 1e6)
    rel = np.abs(got_c[ok] - want_c[ok]) / want_c[ok]
    print(_bid(build), "max rel", rel.max(), flush=True)
    assert rel.max() <= 1e-6


def test_pool_sized_for_largest_job():
    """A batch holding a 10-parameter job and a 40-parameter one cuts its slots
    for 40 parameters; each job equals the same job alone (on a pool cut for
    its own count) and the 40-parameter job equals the oracle.  Serving slots
    cut for fewer parameters than a job refuse it; a job that fits runs as
    alone."""
    v, curve, truth, space, t_start = _build_case(abi.PREC_FP64, 4, False, True, True)
    assert space.n_params == abi.MAX_PARAMS == 40
    v10 = variants.ls_torque(abi.PREC_FP64)
    c10, _ = synth.make_curve(5, v10, 80, 1.5)
    s10 = variants.default_model_space(v10)
    j10 = dict(space=s10, curve=c10, n_fits=256, n_steps=30, seed=3)
    j40 = dict(space=space, curve=curve, n_fits=256, n_steps=20, seed=4, t_segments=t_start)
    with api.Context(v10) as ctx:
        a10 = ctx.run_jobs([j10])[0]
        a40 = ctx.run_jobs([j40])[0]
        both = ctx.run_jobs([j10, j40])
        ctx.serve_start(max_resident_fits=512, max_params=s10.n_params)
        with pytest.raises(RuntimeError, match="AST_E_ARG"):
            ctx.submit(space, curve, n_fits=256, nstages=1, n_steps=20, seed=4, t_segments=t_start)
        sid = ctx.submit(s10, c10, n_fits=256, nstages=1, n_steps=30, seed=3)
        served = ctx.wait_all([sid])[sid]
        ctx.serve_stop()
    for got, want in ((both[0], a10), (both[1], a40), (served, a10)):
        for key in ("status", "best_f", "best_group", "n_evals"):
            assert got[key] == want[key], key
        assert np.array_equal(got["best_params"], want["best_params"])
        assert np.array_equal(got["best_dV"], want["best_dV"])
    want = oracle.Problem(space, curve, t_start=t_start).run(1, 256, nstages=1, n_steps=20, seed=4, n_threads=8)
    assert a40["status"] == 0 and a40["best_f"] == want["f"][0]
    assert np.array_equal(a40["best_params"], want["params"][0]) and np.array_equal(a40["best_dV"], want["dV"][0])
    assert a40["n_evals"] == want["lane_nev"].sum()
