"""The oracle's -plot path (chi2_plot, cuda.c:1865-1966; its grid,
read_data.c:304-360) and minima periodogram (minima(), misc.c:65-276),
pinned on the CPU by what the mathematics fixes:

* the grid runs evenly from the first to the last observation;
* on a curve whose observations sit on the grid, the plot magnitudes are the
  evaluator's magnitudes plus delta_V of the first filter, bit for bit;
* the plot magnitudes agree with the independent numpy model of the same
  physics (scipy DOP853 integration, rotation matrices);
* interior grid points take E,S from the ephemeris whatever INTERP says;
* the periodogram of equally spaced minima peaks at the harmonics k/P, the
  left edge of each full-count plateau where the closed form puts it, with
  equal heights; a time shift changes nothing; the histogram is normalised
  to mean 0 and unit sample deviation; plateau tops count once.
"""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, synth, variants


def _setup(v, n=120, span=3.0, nf=2, seed=4):
    curve, truth = synth.make_curve(seed, v, n, span, n_filters=nf)
    space = variants.default_model_space(v)
    return oracle.Problem(space, curve), curve, truth, space


def test_grid_runs_from_first_to_last_observation():
    v = variants.ls_torque(abi.PREC_FP64)
    pb, curve, truth, _ = _setup(v)
    t, vm, chi2, dv = pb.model_curve(truth, n_plot=2001)
    assert t[0] == curve.mjd[0] and t[-1] == curve.mjd[-1]
    h = (curve.mjd[-1] - curve.mjd[0]) / 2000
    np.testing.assert_allclose(np.diff(t), h, rtol=1e-9)
    c2, d2 = pb.chi2(truth)
    assert chi2[0] == c2[0] and np.array_equal(dv[0], d2[0])  # step one is the evaluator's chi^2


@pytest.mark.parametrize("form", [oracle.LITERAL, oracle.CANON])
@pytest.mark.parametrize("vname", ["ls_torque", "bc", "bw_torque"])
def test_plot_on_the_observation_times_is_the_evaluator(form, vname):
    v = {"ls_torque": variants.ls_torque(abi.PREC_FP64), "bc": variants.bc_free(abi.PREC_FP64),
         "bw_torque": variants.bw_ball_torque(abi.PREC_FP64)}[vname]
    pb, curve, truth, space = _setup(v, nf=2)
    n_plot = 900
    t, _, _, _ = pb.model_curve(truth, n_plot=n_plot, form=form)
    # a curve observed at exactly the grid's times (so its own grid is itself), two filters
    filt = (np.arange(n_plot) % 2).astype(np.int32)
    grid_curve = synth.Curve(t, np.full(n_plot, 15.0) + 0.3 * filt, np.ones(n_plot), filt, 2,
                             np.zeros((n_plot, 3)), np.zeros((n_plot, 3)), curve.eph)
    pg = oracle.Problem(space, grid_curve)
    t2, vm, _, dv = pg.model_curve(truth, n_plot=n_plot, form=form)
    assert np.array_equal(t2, t)
    _, dv_e, vm_e = pg.chi2(truth, form=form, want_vmod=True)
    assert np.array_equal(dv, dv_e)
    assert dv[0, 0] != dv[0, 1]
    assert np.array_equal(vm[0], vm_e[0] + dv[0, 0])  # delta_V of the first filter, nothing else


@pytest.mark.parametrize("vname", ["ls_torque", "bc", "bw_torque"])
def test_plot_matches_independent_physics(vname):
    v = {"ls_torque": variants.ls_torque(abi.PREC_FP64), "bc": variants.bc_free(abi.PREC_FP64),
         "bw_torque": variants.bw_ball_torque(abi.PREC_FP64)}[vname]
    pb, curve, truth, _ = _setup(v, nf=1)
    t, vm, _, dv = pb.model_curve(truth, n_plot=3000, form=oracle.LITERAL)
    E = np.array([synth.quad_interp(curve.eph, x)[0] for x in t])
    S = np.array([synth.quad_interp(curve.eph, x)[1] for x in t])
    E /= np.linalg.norm(E, axis=1)[:, None]
    S /= np.linalg.norm(S, axis=1)[:, None]
    want = synth.truth_vmod(v, truth, t, E, S)
    assert np.max(np.abs(vm[0] - dv[0, 0] - want)) < 1e-7


def test_interior_grid_uses_the_ephemeris_without_interp():
    v_i = variants.ls_torque(abi.PREC_FP64, interp=True)
    v_n = variants.ls_torque(abi.PREC_FP64, interp=False)
    curve, truth = synth.make_curve(8, v_n, 150, 2.0)
    pi = oracle.Problem(variants.default_model_space(v_i), curve)
    pn = oracle.Problem(variants.default_model_space(v_n), curve)
    _, vi, _, dvi = pi.model_curve(truth, n_plot=500)
    _, vn, _, dvn = pn.model_curve(truth, n_plot=500)
    # the ends come from the observations (their own E,S without INTERP)
    a = vi[0, 1:-1] - dvi[0, 0]
    b = vn[0, 1:-1] - dvn[0, 0]
    assert np.max(np.abs(a - b)) < 1e-12


def _periodic_minima(P, T=3.0, n=20000, phase=0.7, shift=0.0):
    t = np.linspace(0.0, T, n) + shift
    vm = -np.cos(2 * np.pi * (t - shift) / P + phase)
    return t, vm


def test_periodogram_peaks_at_the_harmonics():
    P, T, n = 0.3, 3.0, 20000
    t, vm = _periodic_minima(P, T, n)
    fr, H, nmin, Hn = oracle.periodogram(t, vm, 5)
    assert nmin == 10
    mmax = nmin - 1
    binw = (24.0 - 24.0 / 80.0) / 9999
    for j in range(5):
        # the first bin where every interval m*P lies within 0.1 of a whole number of periods
        edge = (j + 1) / P * (1 - 0.1 / ((j + 1) * mmax))
        assert abs(fr[j] - edge) <= 3 * binw, (j, fr[j], edge)
    assert np.all(H == H[0]) and H[0] >= 3.0   # every harmonic reaches the full count
    assert np.max(Hn) == H[0]


def test_periodogram_normalisation_and_shift():
    t, vm = _periodic_minima(0.37, 2.5, 12000, phase=0.1)
    fr, H, nmin, Hn = oracle.periodogram(t, vm, 5)
    assert abs(Hn.mean()) < 1e-12 and abs(Hn.std(ddof=1) - 1) < 1e-12
    fr2, H2, nmin2, Hn2 = oracle.periodogram(t + 100.0, vm, 5)
    assert nmin2 == nmin
    np.testing.assert_allclose(fr2, fr, rtol=0, atol=0)
    np.testing.assert_allclose(H2, H, rtol=1e-9)


def test_periodogram_edge_cases():
    t = np.arange(8, dtype=float)
    # a flat top counts once (V[i] > V[i-1] and V[i] >= V[i+1])
    _, _, nmin, _ = oracle.periodogram(t, np.array([0, 1, 1, 0, 2, 2, 2, 0.0]), 5)
    assert nmin == 2
    # fewer than two minima: no intervals, no clusters
    fr, H, nmin, _ = oracle.periodogram(t, np.array([0, 1, 0, 0, 0, 0, 0, 0.0]), 5)
    assert nmin == 1 and np.all(fr == -1.0) and np.all(H == 0.0)
    # more than 10000 minima is refused (misc.c:80-84)
    n = 20004
    assert oracle.periodogram(np.arange(n, dtype=float), np.tile([0.0, 1.0], n // 2), 5) is None
