"""The oracle's RMSD confidence intervals (chi2_gpu_rms, cuda.c:2148-2389),
pinned on the CPU by what the definitions fix:

* the RMSD of a model is sqrt(chi^2 * dof / sum of weights), linking it to
  the pinned chi^2;
* with no shift every draw is the start model: all accepted, the range is
  the start model's round trip through the scale-free coordinates;
* small shifts keep the start model inside every range;
* every draw takes 2 stream draws per coordinate, accepted or not, and two
  calls continue the streams exactly as one call of twice the draws;
* periodic angles stay unwrapped, and the spherical torque components
  bracket the start model's.
"""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, synth, variants


def _setup(v, n=120, span=2.0, nf=2, seed=31, reopt=1):
    curve, truth = synth.make_curve(seed, v, n, span, n_filters=nf)
    space = variants.default_model_space(v, reopt=reopt)
    return oracle.Problem(space, curve), curve, truth, space


def test_rmsd_is_chi2_times_dof_over_weights():
    v = variants.ls_torque(abi.PREC_FP64)
    pb, curve, truth, space = _setup(v)
    r = pb.rmsd(1, 1, 0, 0.0, truth)
    chi2, _ = pb.chi2(truth)
    dof = curve.n_data - space.n_params - curve.n_filters
    sw = float(np.sum(curve.w.astype(np.float64)))
    assert abs(r["f0"] ** 2 - chi2[0] * dof / sw) <= 1e-12 * r["f0"] ** 2
    assert r["f1"] == r["f0"] + r["f0"] / np.sqrt(dof)


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_F32SIMPLEX])
def test_no_shift_draws_the_start_model(prec):
    v = variants.bc_free(prec)
    pb, curve, truth, space = _setup(v)
    G, S = 8, 5
    r = pb.rmsd(2, G, S, 0.0, truth)
    assert r["counts"][:, 0].sum() == 2 * G * S and r["counts"][:, 1].sum() == 0
    x = pb.params2x(truth)
    p, rc = pb.x2params(x, reopt=1)
    assert rc[0] == 0
    want = p[0].astype(np.float32)
    assert np.array_equal(r["par_min"], np.tile(want, (2, 1)))
    assert np.array_equal(r["par_max"], np.tile(want, (2, 1)))
    assert np.all(r["rng"] == S * 2 * space.n_params)  # two draws per coordinate and model


def test_small_shifts_bracket_the_start_model():
    v = variants.ls_torque(abi.PREC_FP64)
    pb, curve, truth, space = _setup(v)
    r = pb.rmsd(2, 32, 8, 1e-5, truth)
    nt, nb = r["counts"].sum(0)
    assert nt == 2 * 32 * 8 and nb == 0
    lo, hi = r["par_min"].min(0), r["par_max"].max(0)
    assert np.all(lo <= truth.astype(np.float32)) and np.all(truth.astype(np.float32) <= hi)
    assert np.all(hi - lo < 0.05 * np.maximum(np.abs(truth), 1.0))


def test_streams_continue_across_calls():
    v = variants.bw_ball_torque(abi.PREC_F32SIMPLEX)
    pb, curve, truth, space = _setup(v, nf=1)
    rng = np.zeros(64, np.uint32)
    a = pb.rmsd(2, 32, 3, 0.01, truth, rng=rng, seed=7)
    b = pb.rmsd(2, 32, 3, 0.01, truth, rng=rng, seed=7)
    one = pb.rmsd(2, 32, 6, 0.01, truth, seed=7)
    assert np.array_equal(np.minimum(a["par_min"], b["par_min"]), one["par_min"])
    assert np.array_equal(np.maximum(a["par_max"], b["par_max"]), one["par_max"])
    assert np.array_equal(a["counts"] + b["counts"], one["counts"])
    assert np.array_equal(rng, one["rng"])


def test_periodic_angles_are_not_wrapped():
    # PERIODIC coordinates map to x * 2 pi without wrapping (cuda.c:1208-1211): a start
    # azimuth near 0 gives accepted models on both sides of 0, and a narrow range
    v = variants.ls_torque(abi.PREC_FP64)
    pb, curve, truth, space = _setup(v)
    ids = variants.param_positions(v)
    start = truth.copy()
    start[ids["phi_0"]] = 1e-6
    r = pb.rmsd(1, 64, 8, 1e-6, start)
    i = ids["phi_0"]
    assert r["par_min"][0, i] < 0 < r["par_max"][0, i] < 1e-3


def test_spherical_torque_components():
    v = variants.ls_torque(abi.PREC_FP64)
    pb, curve, truth, space = _setup(v)
    ids = variants.param_positions(v)
    r = pb.rmsd(1, 64, 8, 1e-6, truth, spherical_k=True)
    bt, ct = truth[ids["b_tumb"]], truth[ids["c_tumb"]]
    Is = (1 + bt * bt) / (bt * bt + ct * ct)
    Ii = (1 + ct * ct) / (bt * bt + ct * ct)
    K = np.array([Ii * truth[ids["Ti"]], Is * truth[ids["Ts"]], truth[ids["Tl"]]])
    k0 = np.linalg.norm(K)
    want = {"Ti": k0, "Ts": np.arccos(K[2] / k0), "Tl": np.arctan2(K[1], K[0])}
    r0 = pb.rmsd(1, 8, 2, 0.0, truth, spherical_k=True)
    for name, w in want.items():
        j = ids[name]
        tol = 1e-3 * abs(w)
        assert r["par_min"][0, j] <= w + tol and r["par_max"][0, j] >= w - tol, name
        assert r["par_max"][0, j] - r["par_min"][0, j] < 1e-2 * abs(w), name
        # no shift: the start model's own components
        assert abs(r0["par_min"][0, j] - w) <= 1e-6 * abs(w) and r0["par_max"][0, j] == r0["par_min"][0, j], name
