"""SEGMENT (-DSEGMENT -DN_SEG=3 -DONE_LE) in the oracle: the split of the
curve at the segment start times (read_data.c:244-250), the per-segment
parameters (Types[type][iseg], asteroid.c:128-168) and the sums over all
segments with the full N_data in the degrees of freedom (cuda.c:162-268,
884-906).

The pins: every segment's model magnitudes equal the plain LS model's on
that segment's observations alone, with the segment's parameters picked by
name; chi^2 and delta_V equal a brute-force weighted least-squares fit of
those magnitudes; observations before the first start are not visited; an
observation at or after two start times opens only one segment."""
import numpy as np
import pytest

import oracle
from asteroid_amd import synth, variants


def _sub(curve, i1, i2):
    return synth.Curve(curve.mjd[i1:i2], curve.V[i1:i2], curve.w[i1:i2], curve.filter[i1:i2], curve.n_filters,
                       curve.E[i1:i2], curve.S[i1:i2], curve.eph)


def _seg_params(v_seg, truth8, rng):
    """A 16-parameter SEGMENT model around an LS model: each further segment
    gets its own Euler angles and psi_0."""
    pos8 = variants.param_positions(variants.make_variant(False, False, False, True, True))
    pos = variants.param_positions(v_seg)
    p = np.zeros(len(pos))
    for name, i in pos.items():
        base = name.split(":")[0]
        p[i] = truth8[pos8[base]]
        if ":" in name:
            p[i] += rng.uniform(-0.3, 0.3)
    return p


def _setup(n=90, span=3.0, nf=2, seed=4, t_start=None):
    v8 = variants.make_variant(False, False, False, True, True)
    curve, truth8 = synth.make_curve(seed, v8, n, span, n_filters=nf)
    vs = variants.ls_segments()
    space = variants.default_model_space(vs)
    if t_start is None:
        t_start = [curve.mjd[0] - 0.1, curve.mjd[0] + span / 3, curve.mjd[0] + 2 * span / 3]
    pb = oracle.Problem(space, curve, t_start=t_start)
    pb8 = oracle.Problem(variants.default_model_space(v8), curve)
    params = _seg_params(vs, truth8, np.random.default_rng(seed))
    return pb, pb8, curve, truth8, params, np.asarray(t_start)


def _starts(mjd, t_start):
    s, k = [], 0
    for i, t in enumerate(mjd):  # restated from read_data.c:244-250
        if k < len(t_start) and t >= t_start[k]:
            s.append(i)
            k += 1
    return s


def _segment_vmod(pb8, curve, params, vs, starts, form):
    """Magnitudes of each segment from the plain LS model on its observations."""
    pos8 = variants.param_positions(pb8.space.variant)
    pos = variants.param_positions(vs)
    vm = np.zeros(curve.n_data)
    for k, i1 in enumerate(starts):
        i2 = starts[k + 1] if k + 1 < len(starts) else curve.n_data
        p8 = np.zeros(8)
        for name, j in pos8.items():
            p8[j] = params[pos[name if k == 0 or variants._shared(vs, name) else f"{name}:{k}"]]
        sub = oracle.Problem(pb8.space, _sub(curve, i1, i2))
        _, _, v = sub.chi2(p8, form=form, want_vmod=True)
        vm[i1:i2] = v[0]
    return vm


@pytest.mark.parametrize("form", [oracle.LITERAL, oracle.CANON])
def test_segments_are_plain_models_on_their_observations(form):
    pb, pb8, curve, _, params, t_start = _setup()
    starts = _starts(curve.mjd, t_start)
    assert len(starts) == 3 and starts[0] == 0 and 0 < starts[1] < starts[2] < curve.n_data
    _, _, vm = pb.chi2(params, form=form, want_vmod=True)
    ref = _segment_vmod(pb8, curve, params, pb.space.variant, starts, form)
    assert np.array_equal(vm[0], ref)
    # each segment restarts: the segment model differs from one long integration
    _, _, v_one = pb8.chi2(params[:8], form=form, want_vmod=True)
    assert np.array_equal(vm[0][:starts[1]], v_one[0][:starts[1]])
    assert not np.allclose(vm[0][starts[1]:], v_one[0][starts[1]:])


def test_chi2_is_the_least_squares_fit_over_all_segments():
    pb, pb8, curve, _, params, t_start = _setup(t_start=None, seed=7)
    # the first observations lie before the first start: left out of the sums
    t_start = np.array([curve.mjd[5], curve.mjd[0] + 1.0, curve.mjd[0] + 2.0])
    pb = oracle.Problem(pb.space, curve, t_start=t_start)
    starts = _starts(curve.mjd, t_start)
    assert starts[0] == 5
    c, dV, vm = pb.chi2(params, want_vmod=True)
    assert np.all(vm[0][:5] == 0)  # never visited
    ref = _segment_vmod(pb8, curve, params, pb.space.variant, starts, oracle.CANON)
    assert np.array_equal(vm[0][5:], ref[5:])
    inc = np.arange(curve.n_data) >= 5
    V, w, f = curve.V.astype(float), curve.w.astype(float), curve.filter
    chi2 = 0.0
    for m in range(curve.n_filters):
        s = inc & (f == m)
        d = V[s] - ref[s]
        dv = np.sum(w[s] * d) / np.sum(w[s])  # the weighted mean minimises sum w (d - dv)^2
        assert dV[0][m] == pytest.approx(dv, rel=1e-12, abs=1e-12)
        chi2 += np.sum(w[s] * (d - dv) ** 2)
    dof = curve.n_data - pb.space.n_params - curve.n_filters  # all observations count
    assert c[0] == pytest.approx(chi2 / dof, rel=1e-10)


def test_one_observation_past_two_starts_opens_one_segment():
    pb, pb8, curve, _, params, _ = _setup(seed=9)
    i = 30
    t_start = np.array([curve.mjd[0], curve.mjd[i] - 1e-4, curve.mjd[i] - 5e-5])
    pb = oracle.Problem(pb.space, curve, t_start=t_start)
    starts = _starts(curve.mjd, t_start)
    assert starts == [0, i, i + 1]  # the third segment opens at the next observation
    _, _, vm = pb.chi2(params, want_vmod=True)
    assert np.array_equal(vm[0], _segment_vmod(pb8, curve, params, pb.space.variant, starts, oracle.CANON))


def test_a_segment_without_observations_is_refused():
    pb, _, curve, _, params, _ = _setup()
    bad = oracle.Problem(pb.space, curve, t_start=[curve.mjd[0], curve.mjd[-1], curve.mjd[-1] + 1.0])
    with pytest.raises(AssertionError):
        bad.chi2(params)


def test_fit_runs_on_segments():
    pb, _, curve, _, params, _ = _setup(n=60)
    r = pb.run(2, 16, n_steps=60, seed=3)
    assert np.all(np.isfinite(r["f"])) and np.all(r["f"] < 1e29)
    c, _ = pb.chi2(r["params"][0])
    assert c[0] == r["f"][0]


def test_plot_curve_restarts_at_each_segment():
    # the -plot grid splits by the same rule (read_data.c:357-363): with every
    # segment given the first segment's parameters, the curve equals the plain
    # LS curve up to the second start and differs after it; grid points before
    # the first start are never computed
    pb, pb8, curve, truth8, params, _ = _setup(seed=11)
    t_start = np.array([curve.mjd[4], curve.mjd[0] + 1.0, curve.mjd[0] + 2.0])
    pb = oracle.Problem(pb.space, curve, t_start=t_start)
    same = params.copy()
    same[8:12] = same[[0, 1, 2, 7]]
    same[12:16] = same[[0, 1, 2, 7]]
    t, vm, c, dv = pb.model_curve(same, n_plot=2000)
    t8, vm8, c8, dv8 = pb8.model_curve(same[:8], n_plot=2000)
    assert np.array_equal(t, t8)  # the grid spans every observation
    g = np.searchsorted(t, t_start)
    assert np.all(np.isnan(vm[0][:g[0]]))
    # segment 0 starts at its first grid point, and every segment restarts
    assert np.all(np.isfinite(vm[0][g[0]:]))
    assert not np.allclose(vm[0][g[1]:] - dv[0][0], vm8[0][g[1]:] - dv8[0][0])
    c_obs, dv_obs = pb.chi2(same)
    assert c[0] == c_obs[0] and np.array_equal(dv[0], dv_obs[0])


SEG_BUILDS = [(n, le, tq, bc) for n in (3, 4) for le in (True, False) for tq in (False, True) for bc in (False, True)]


def _ids(b):
    return f"seg{b[0]}{'_onele' if b[1] else ''}{'_torque' if b[2] else ''}{'_bc' if b[3] else ''}"


def _build(n_seg, one_le, torque, bc=False, seed=7, n=120, span=4.0):
    v0 = variants.make_variant(torque, bc, False, True, True)
    vs = variants.ls_segments(n_seg=n_seg, one_le=one_le, torque=torque, bc=bc)
    curve, truth0 = synth.make_curve(seed, v0, n, span, n_filters=2)
    t_start = curve.mjd[0] - 0.05 + np.arange(n_seg) * span / n_seg
    pos0, pos = variants.param_positions(v0), variants.param_positions(vs)
    rng = np.random.default_rng(seed)
    params = np.zeros(len(pos))
    for name, i in pos.items():
        b = name.split(":")[0]
        params[i] = truth0[pos0[b]]
        if ":" in name:  # each further segment its own rotation (and L, Es)
            params[i] *= 1.0 + rng.uniform(-0.02, 0.02) if b in ("L", "Es", "Ti", "Ts", "Tl") else 1.0
            if b not in ("L", "Es", "Ti", "Ts", "Tl"):
                params[i] += rng.uniform(-0.3, 0.3)
    return v0, vs, curve, params, t_start


@pytest.mark.parametrize("build", SEG_BUILDS, ids=_ids)
def test_every_segment_build_is_plain_models_on_their_observations(build):
    n_seg, one_le, torque, bc = build
    v0, vs, curve, params, t_start = _build(n_seg, one_le, torque, bc)
    space = variants.default_model_space(vs)
    assert space.n_params == 8 + 3 * torque + 2 * bc + (6 + 3 * torque - 2 * one_le) * (n_seg - 1)  # asteroid.h:104-106
    pb = oracle.Problem(space, curve, t_start=t_start)
    starts = _starts(curve.mjd, t_start)
    assert len(starts) == n_seg
    pos0, pos = variants.param_positions(v0), variants.param_positions(vs)
    space0 = variants.default_model_space(v0)
    vm = np.zeros(curve.n_data)
    for k, i1 in enumerate(starts):
        i2 = starts[k + 1] if k + 1 < n_seg else curve.n_data
        p0 = np.array([params[pos[n if k == 0 or variants._shared(vs, n) else f"{n}:{k}"]] for n in pos0])
        _, _, v = oracle.Problem(space0, _sub(curve, i1, i2)).chi2(p0, form=oracle.CANON, want_vmod=True)
        vm[i1:i2] = v[0]
    c, dv, got = pb.chi2(params, form=oracle.CANON, want_vmod=True)
    assert np.array_equal(got[0][starts[0]:], vm[starts[0]:])
    # chi^2 and delta_V: the weighted least squares of those magnitudes, with
    # every observation in the degrees of freedom
    sel = np.arange(curve.n_data) >= starts[0]
    chi2, dvs = 0.0, []
    for m in range(curve.n_filters):
        f = sel & (curve.filter == m)
        y, w = curve.V[f].astype(float) - vm[f], curve.w[f].astype(float)
        d = np.sum(w * y) / np.sum(w)
        dvs.append(d)
        chi2 += np.sum(w * (y - d) ** 2)
    chi2 /= curve.n_data - space.n_params - curve.n_filters
    assert c[0] == pytest.approx(chi2, rel=1e-10) and np.allclose(dv[0], dvs, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("build", SEG_BUILDS, ids=_ids)
def test_segment_x2params_maps_each_segment_like_the_plain_model(build):
    # with every Es on one side of 1/2, segment k's parameters are the plain
    # model's map of segment k's coordinates (L from 1/P_psi with the
    # segment's own Es, cuda.c:1263)
    n_seg, one_le, torque, bc = build
    v0, vs, curve, params, t_start = _build(n_seg, one_le, torque, bc)
    space, space0 = variants.default_model_space(vs), variants.default_model_space(v0)
    pb = oracle.Problem(space, curve, t_start=t_start)
    pb0 = oracle.Problem(space0, curve)
    pos0, pos = variants.param_positions(v0), variants.param_positions(vs)
    rng = np.random.default_rng(n_seg * 10 + one_le * 2 + torque + 100 * bc)
    for side in (0.25, 0.75):
        x = rng.uniform(0.05, 0.95, (50, space.n_params))
        for name, i in pos.items():
            if name.split(":")[0] == "Es":
                x[:, i] = side + rng.uniform(-0.2, 0.2, 50)
        p, rc = pb.x2params(x, reopt=0)
        for k in range(n_seg):
            idx = [pos[n if k == 0 or variants._shared(vs, n) else f"{n}:{k}"] for n in pos0]
            p0, rc0 = pb0.x2params(x[:, idx], reopt=0)
            ok = (rc == 0) & (rc0 == 0)
            assert ok.sum() > 20
            assert np.array_equal(p[ok][:, idx], p0[ok])


def test_segment_lam_is_the_last_segments():
    # the reference's x2params takes LAM from the last Es in the parameter
    # order (its first loop, cuda.c:1128-1147) and maps every segment's Es
    # with it (cuda.c:1170-1184): a SAM coordinate of the first segment is
    # mapped on the LAM branch when the last segment's Es is LAM
    v0, vs, curve, params, t_start = _build(3, False, False)
    space = variants.default_model_space(vs)
    pb = oracle.Problem(space, curve, t_start=t_start)
    pos = variants.param_positions(vs)
    x = np.full(space.n_params, 0.5)
    x[pos["c_tumb"]], x[pos["b_tumb"]] = 0.4, 0.6
    x[pos["Es"]], x[pos["Es:1"]], x[pos["Es:2"]] = 0.2, 0.3, 0.8
    p, rc = pb.x2params(x, reopt=1)
    assert rc[0] == 0
    c, b = p[0, pos["c_tumb"]], p[0, pos["b_tumb"]]
    Ii = (1 + c * c) / (b * b + c * c)
    for name, xe in (("Es", 0.2), ("Es:1", 0.3), ("Es:2", 0.8)):
        assert p[0, pos[name]] == pytest.approx(2 * (xe - 0.5) * (1 - 1 / Ii) + 1 / Ii, rel=1e-14)
