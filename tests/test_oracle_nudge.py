"""The oracle's NUDGE merit (Stage Three, cuda.c:716-758 and 912-992),
pinned by what its definition fixes rather than by a second copy of it:

* no observed minimum near any model minimum leaves chi^2 unchanged, bit
  for bit;
* one observed minimum placed on a model minimum (the vertex of the
  parabola through the three model magnitudes around it, computed here in
  numpy, and its magnitude plus delta_V) multiplies chi^2 by P_MIN = 0.01
  when chi^2 < CHI2_0 and leaves it alone when chi^2 > CHI2_1;
* ten copies of one observed minimum far off in magnitude count every model
  minimum within DT_MAX ten times: S_M = 10 k gives the penalty 1 for k <= 2
  and exactly 4 for k >= 3, with no reward;
* more than M_MAX = 30 model minima near the observed ones fail the model
  (1e30, delta_V untouched);
* V1S zeroes positive magnitude distances only;
* the canonical form equals the literal one to rounding.
The model minima the merit sees are located independently here from the
oracle's magnitudes (local maxima of V among the observations).
"""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, synth, variants

DT_MAX, DT_MAX2 = 0.12, np.float32(1.5 * np.float64(np.float32(0.12)))


def _setup(prec=abi.PREC_FP64, n=300, span=3.0, seed=5, ppsi=(4.0, 10.0), v1s=False, truth=None):
    v = variants.ls_torque_nudge(prec, v1s=v1s)
    rng = np.random.default_rng(seed + 100)
    if truth is None:
        truth = synth.random_truth(rng, v, ppsi_hours=ppsi)
    curve, truth = synth.make_curve(seed, v, n, span, truth=truth)
    space = variants.default_model_space(v)
    plain = oracle.Problem(variants.default_model_space(variants.ls_torque(prec)), curve)
    return v, curve, truth, space, plain


def _model_minima(plain, params, curve):
    """(t_vertex, V_vertex, i) of every local V maximum among the observations."""
    _, dv, vm = plain.chi2(params, want_vmod=True)
    vm = vm[0]
    t = curve.mjd
    out = []
    for i in range(2, len(t)):
        if vm[i - 1] > vm[i - 2] and vm[i - 1] >= vm[i]:
            a, b, c = np.polyfit(t[i - 2:i + 1], vm[i - 2:i + 1], 2)
            tv = -b / (2 * a)
            out.append((tv, a * tv * tv + b * tv + c, i - 1))
    return out, dv[0, 0]


@pytest.mark.parametrize("form", [oracle.LITERAL, oracle.CANON])
def test_far_minima_leave_chi2_unchanged(form):
    v, curve, truth, space, plain = _setup()
    pb = oracle.Problem(space, curve, minima=(np.array([-50.0, 80.0], np.float32), np.array([20.0, 20.0], np.float32)))
    c, d = pb.chi2(truth, form=form)
    c0, d0 = plain.chi2(truth, form=form)
    assert c[0] == c0[0] and np.array_equal(d, d0)


def _isolated(mins, gap):
    return [m for k, m in enumerate(mins) if all(abs(m[0] - o[0]) > gap for j, o in enumerate(mins) if j != k)]


@pytest.mark.parametrize("form", [oracle.LITERAL, oracle.CANON])
def test_minimum_on_model_minimum_is_rewarded_by_p_min(form):
    v, curve, truth, space, plain = _setup(ppsi=(20.0, 30.0))
    mins, dv0 = _model_minima(plain, truth, curve)
    iso = _isolated(mins, 2 * DT_MAX2)
    assert iso, "no isolated model minimum"
    tv, Vv, _ = iso[len(iso) // 2]
    pb = oracle.Problem(space, curve, minima=(np.array([tv], np.float32), np.array([Vv + dv0], np.float32)))
    c, _ = pb.chi2(truth, form=form)
    c0, _ = plain.chi2(truth, form=form)
    assert c0[0] < 10.0
    assert abs(c[0] / c0[0] - 0.01) < 1e-5 * 0.01
    # the same model with chi^2 above CHI2_1 (weights scaled up) gets no reward
    heavy = synth.Curve(curve.mjd, curve.V, curve.w * 100.0, curve.filter, 1, curve.E, curve.S, curve.eph)
    pbh = oracle.Problem(space, heavy, minima=pb.minima)
    ch, _ = pbh.chi2(truth, form=form)
    ch0, _ = oracle.Problem(plain.space, heavy).chi2(truth, form=form)
    assert ch0[0] > 30.0 and ch[0] == ch0[0]


@pytest.mark.parametrize("form", [oracle.LITERAL, oracle.CANON])
def test_too_many_model_minima_penalty(form):
    # ten copies of an observed minimum 79 mag off (no reward): every model
    # minimum within DT_MAX adds 10 to S_M, one in the DT_MAX..DT_MAX2 band less
    fast = _setup(n=2000, span=2.0, ppsi=(2.0, 2.5), seed=7)
    slow = _setup(ppsi=(20.0, 30.0))
    checked = {1.0: 0, 4.0: 0}
    for v, curve, truth, space, plain in (fast, slow):
        mins, dv0 = _model_minima(plain, truth, curve)
        c0, _ = plain.chi2(truth, form=form)
        for T in np.arange(0.2, 1.9, 0.1):
            d = np.array([abs(np.float32(m[0]) - np.float32(T)) for m in mins])
            k_close, k_near = int(np.sum(d <= DT_MAX)), int(np.sum(d < DT_MAX2))
            if 10 * k_close >= 30:
                want = 4.0
            elif k_near <= 1:
                want = 1.0
            else:
                continue
            pb = oracle.Problem(space, curve, minima=(np.full(10, T, np.float32), np.full(10, 99.0, np.float32)))
            c, _ = pb.chi2(truth, form=form)
            assert c[0] == c0[0] * want, (T, k_close, k_near, c[0] / c0[0])
            checked[want] += 1
    assert checked[1.0] > 2 and checked[4.0] > 2, checked


def test_more_than_m_max_minima_fail_the_model():
    v, curve, truth, space, plain = _setup(n=3000, span=3.0, ppsi=(1.2, 1.4), seed=9)
    mins, _ = _model_minima(plain, truth, curve)
    ts = np.linspace(0.2, 2.8, 10).astype(np.float32)
    near = sum(1 for m in mins if np.min(np.abs(np.float32(curve.mjd[m[2]]) - ts)) < DT_MAX2)
    assert near > 30
    pb = oracle.Problem(space, curve, minima=(ts, np.full(10, 20.0, np.float32)))
    c, d = pb.chi2(truth)
    assert c[0] == 1e30 and np.isnan(d).all()
    # with fewer observed minima the count stays below the limit
    pb2 = oracle.Problem(space, curve, minima=(ts[:1], np.full(1, 20.0, np.float32)))
    c2, d2 = pb2.chi2(truth)
    assert c2[0] < 1e30 and np.isfinite(d2).all()


@pytest.mark.parametrize("form", [oracle.LITERAL, oracle.CANON])
def test_v1s_zeroes_positive_distances_only(form):
    v, curve, truth, space, plain = _setup(ppsi=(20.0, 30.0))
    mins, dv0 = _model_minima(plain, truth, curve)
    tv, Vv, _ = _isolated(mins, 2 * DT_MAX2)[0]
    for shift, same in ((-0.5, True), (0.5, False)):
        # observed minimum brighter (smaller V) than the model's: dV > 0, zeroed under V1S
        mn = (np.array([tv], np.float32), np.array([Vv + dv0 + shift], np.float32))
        c1, _ = oracle.Problem(variants.default_model_space(variants.ls_torque_nudge(v1s=True)), curve,
                               minima=mn).chi2(truth, form=form)
        c0, _ = oracle.Problem(space, curve, minima=mn).chi2(truth, form=form)
        cz, _ = oracle.Problem(space, curve, minima=(mn[0], np.array([Vv + dv0], np.float32))).chi2(truth, form=form)
        if shift < 0:
            assert c1[0] != c0[0] and abs(c1[0] / cz[0] - 1) < 1e-5
        else:
            assert c1[0] == c0[0]


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_F32SIMPLEX])
def test_canonical_form_equals_literal(prec):
    v, curve, truth, space, plain = _setup(prec=prec, ppsi=(6.0, 9.0))
    mins, dv0 = _model_minima(plain, truth, curve)
    pick = mins[::max(1, len(mins) // 4)][:4]
    mn = (np.array([m[0] + 0.01 for m in pick], np.float32), np.array([m[1] + dv0 + 0.05 for m in pick], np.float32))
    pb = oracle.Problem(space, curve, minima=mn)
    rng = np.random.default_rng(3)
    x = rng.uniform(0.02, 0.98, (150, space.n_params))
    p, rc = pb.x2params(x, reopt=0, form=oracle.LITERAL)
    L = p[:, variants.param_positions(v)["L"]]
    p = np.vstack([truth, p[(rc == 0) & (L * v.time_step < 2.0)]])
    cl, _ = pb.chi2(p, form=oracle.LITERAL)
    cc, _ = pb.chi2(p, form=oracle.CANON)
    ok = np.isfinite(cl) & (cl < 1e29)
    assert ok.sum() > 20 and cl[0] < cc[0] * (1 + 1e-5) and np.array_equal(cl >= 1e29, cc >= 1e29)
    rel = np.abs(cl[ok] - cc[ok]) / cl[ok]
    # a merit factor is a float, and a model minimum right at a window edge can
    # flip between the forms; all but a few models agree to float rounding
    assert np.mean(rel < 1e-5) > 0.95
