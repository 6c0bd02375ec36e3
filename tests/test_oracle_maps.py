"""The oracle's parameter maps (x2params, cuda.c:1120-1319; params2x,
cuda.c:1016-1115) pinned by what the maps are, not by their own code:

* closed forms at chosen points: c_tumb = exp(lerp of its log limits),
  b_tumb = c_tumb^lerp, Es spanning [1/Is, 1/Ii] (SAM, x < 1/2) and
  [1/Ii, 1] (LAM), psi_0 spanning +-atan(sqrt(Ii (Is - 1/Es) / Is /
  (1/Es - Ii))) (SAM) or [0, 2 pi] (LAM), periodic angles 2 pi x, the BC
  photometric c and b as c_tumb and b_tumb, BW_BALL's kappa = exp(lerp),
  linear torque and L limits;
* every hard-limit code: which side of [0, 1] is rejected (HARD_BOTH,
  HARD_LEFT, HARD_RIGHT, PERIODIC_LAM under SAM only, never SOFT_BOTH or
  PERIODIC), and BC's |log c - log c_tumb| > 100 and |log b - log b_tumb| >
  100 rejections;
* the two maps as inverses: params2x(x2params(x)) = x on random interior x,
  and x2params(params2x(p)) = p, for every built variant.

The expected values are computed here with numpy from the formulas above,
with the variant's limits read from the model space."""
import math

import numpy as np
import pytest

import oracle
from asteroid_amd import abi, synth, variants

PI = math.pi


def _problem(v, limits=None, reopt=1):
    curve, truth = synth.make_curve(3, v, 40, 1.0)
    space = variants.default_model_space(v, reopt=reopt, limits=limits)
    return oracle.Problem(space, curve), space


def _x2p(pb, x, reopt=1):
    """x2params with reopt = 1 (L mapped linearly); the P_PSI mapping of L
    under reopt = 0 is pinned by the psi-period test of test_oracle_physics."""
    return pb.x2params(x, reopt=reopt)


def _pos(space):
    return {n: i for i, n in enumerate(space.param_names())}


def _lim(space, name):
    t = space.ids[name]
    return space.limits[t], space.limits[space.n_types + t]


def _x(space, **kw):
    """A point of the cube at 0.5 in every coordinate except the given ones."""
    x = np.full(space.n_params, 0.5)
    pos = _pos(space)
    for k, val in kw.items():
        x[pos[k]] = val
    return x


def _inertia(c, b):
    Is = (1 + b * b) / (b * b + c * c)
    Ii = (1 + c * c) / (b * b + c * c)
    return Is, Ii


FP64_VARIANTS = {
    "ls_torque": lambda: variants.ls_torque(abi.PREC_FP64),
    "bc": lambda: variants.bc_free(abi.PREC_FP64),
    "bw_torque": lambda: variants.bw_ball_torque(abi.PREC_FP64),
    "bc_rotate_torque": lambda: variants.bc_rotate(abi.PREC_FP64, torque=True),
    "torque2": lambda: variants.ls_torque2(abi.PREC_FP64),
    "trend": lambda: variants.ls_torque_trend(abi.PREC_FP64),
    "rect": lambda: variants.rect(abi.PREC_FP64),
    "segment3": lambda: variants.ls_segments(abi.PREC_FP64),
}


@pytest.mark.parametrize("xc,xb", [(0.0, 0.0), (1.0, 1.0), (0.3, 0.7), (0.9, 0.2)])
def test_tumbling_shape_closed_form(xc, xb):
    pb, sp = _problem(variants.ls_torque())
    p, rc = _x2p(pb, _x(sp, c_tumb=xc, b_tumb=xb))
    assert rc[0] == 0
    pos = _pos(sp)
    lo, hi = _lim(sp, "c_tumb")
    c = math.exp(lo + xc * (hi - lo))
    blo, bhi = _lim(sp, "b_tumb")
    b = c ** (blo + xb * (bhi - blo))
    assert p[0, pos["c_tumb"]] == pytest.approx(c, rel=1e-14)
    assert p[0, pos["b_tumb"]] == pytest.approx(b, rel=1e-13)
    if xc == 0.0:
        assert p[0, pos["c_tumb"]] == pytest.approx(0.01, rel=1e-14)
    if xb == 0.0:
        assert p[0, pos["b_tumb"]] == 1.0


@pytest.mark.parametrize("xe", [0.0, 0.1, 0.25, 0.4999, 0.5, 0.6, 0.9, 1.0])
def test_energy_and_psi0_ranges_closed_form(xe):
    pb, sp = _problem(variants.ls_torque())
    pos = _pos(sp)
    for xpsi in (0.0, 0.3, 1.0):
        p, rc = _x2p(pb, _x(sp, c_tumb=0.4, b_tumb=0.6, Es=xe, psi_0=xpsi))
        assert rc[0] == 0
        c, b = p[0, pos["c_tumb"]], p[0, pos["b_tumb"]]
        Is, Ii = _inertia(c, b)
        if xe >= 0.5:  # LAM: Es in [1/Ii, 1], psi_0 in [0, 2 pi]
            es = 2 * (xe - 0.5) * (1 - 1 / Ii) + 1 / Ii
            pmin, pmax = 0.0, 2 * PI
        else:          # SAM: Es in [1/Is, 1/Ii], psi_0 within the separatrix
            es = 2 * xe * (1 / Ii - 1 / Is) + 1 / Is
            pmax = math.atan(math.sqrt(Ii * (Is - 1 / es) / Is / (1 / es - Ii)))
            pmin = -pmax
        assert p[0, pos["Es"]] == pytest.approx(es, rel=1e-14)
        if xe in (0.5, 1.0):
            assert p[0, pos["Es"]] == pytest.approx(1 / Ii if xe == 0.5 else 1.0, rel=1e-14)
        if xe == 0.0:
            assert p[0, pos["Es"]] == pytest.approx(1 / Is, rel=1e-14)
        assert p[0, pos["psi_0"]] == pytest.approx(pmin + xpsi * (pmax - pmin), rel=1e-13, abs=1e-15)
        if 0 < xe < 0.5 and xpsi == 0.0:
            assert p[0, pos["psi_0"]] < 0  # the SAM range is symmetric about 0 (a point at Es = 1/Is)


def test_angles_torque_and_L_closed_form():
    pb, sp = _problem(variants.ls_torque(), reopt=1)
    pos = _pos(sp)
    x = _x(sp, theta_M=0.2, phi_M=0.7, phi_0=0.95, Ti=0.0, Ts=1.0, Tl=0.25, L=0.6)
    p, rc = _x2p(pb, x)
    assert rc[0] == 0
    lo, hi = _lim(sp, "theta_M")
    assert p[0, pos["theta_M"]] == pytest.approx(lo + 0.2 * (hi - lo), rel=1e-15)
    assert p[0, pos["phi_M"]] == pytest.approx(0.7 * 2 * PI, rel=1e-15)
    assert p[0, pos["phi_0"]] == pytest.approx(0.95 * 2 * PI, rel=1e-15)
    tlo, thi = _lim(sp, "Ti")
    assert p[0, pos["Ti"]] == pytest.approx(tlo, rel=1e-15)
    assert p[0, pos["Ts"]] == pytest.approx(thi, rel=1e-15)
    assert p[0, pos["Tl"]] == pytest.approx(tlo + 0.25 * (thi - tlo), rel=1e-15)
    llo, lhi = _lim(sp, "L")
    assert p[0, pos["L"]] == pytest.approx(llo + 0.6 * (lhi - llo), rel=1e-15)


def test_bc_photometric_shape_and_kappa_closed_form():
    pb, sp = _problem(variants.bc_free())
    pos = _pos(sp)
    p, rc = _x2p(pb, _x(sp, c_tumb=0.5, b_tumb=0.5, c=0.2, b=0.8))
    assert rc[0] == 0
    lo, hi = _lim(sp, "c")
    c = math.exp(lo + 0.2 * (hi - lo))
    blo, bhi = _lim(sp, "b")
    assert p[0, pos["c"]] == pytest.approx(c, rel=1e-14)
    assert p[0, pos["b"]] == pytest.approx(c ** (blo + 0.8 * (bhi - blo)), rel=1e-13)
    pb, sp = _problem(variants.bw_ball_torque())
    pos = _pos(sp)
    p, rc = _x2p(pb, _x(sp, kappa=0.3))
    lo, hi = _lim(sp, "kappa")
    assert rc[0] == 0 and p[0, pos["kappa"]] == pytest.approx(math.exp(lo + 0.3 * (hi - lo)), rel=1e-14)


HARD = [  # (parameter, x, rejected)
    ("theta_M", -1e-6, True), ("theta_M", 1 + 1e-6, True),            # HARD_BOTH
    ("L", -1e-6, True), ("L", 1.3, False),                            # HARD_LEFT
    ("c_tumb", 1 + 1e-6, True), ("c_tumb", -0.2, False),              # HARD_RIGHT
    ("Ti", -0.5, False), ("Ts", 1.5, False),                          # SOFT_BOTH
    ("phi_M", -0.3, False), ("phi_0", 1.7, False),                    # PERIODIC
    ("b_tumb", -1e-6, True), ("Es", 1 + 1e-6, True),                  # HARD_BOTH (dependent)
]


@pytest.mark.parametrize("name,xv,rejected", HARD)
def test_hard_limit_codes(name, xv, rejected):
    pb, sp = _problem(variants.ls_torque())
    p, rc = _x2p(pb, _x(sp, **{name: xv}))
    assert (rc[0] != 0) == rejected


def test_psi0_limits_only_in_sam():
    pb, sp = _problem(variants.ls_torque())
    for xe, rejected in ((0.3, True), (0.7, False)):   # PERIODIC_LAM: hard only for SAM
        for xpsi in (-0.01, 1.01):
            p, rc = _x2p(pb, _x(sp, Es=xe, psi_0=xpsi))
            assert (rc[0] != 0) == rejected, (xe, xpsi)


def test_bc_deviation_rejections():
    v = variants.bc_free()
    T = abi.type_ids(v)
    # c may reach log c = -300: more than 100 below log c_tumb >= -log(100)
    pb, sp = _problem(v, limits={T["c"]: (-300.0, 0.0)})
    assert _x2p(pb, _x(sp, c=0.0, c_tumb=1.0))[1][0] != 0
    assert _x2p(pb, _x(sp, c=0.99, c_tumb=1.0))[1][0] == 0
    # log c = -150, log c_tumb = -60 (90 apart, accepted); b = c^1 against
    # b_tumb = c_tumb^0: 150 apart, rejected; b = c^0.3 (45 apart) accepted
    pb, sp = _problem(v, limits={T["c"]: (-150.0, 0.0), T["c_tumb"]: (-60.0, 0.0)})
    assert _x2p(pb, _x(sp, c=0.0, c_tumb=0.0, b=1.0, b_tumb=0.0))[1][0] != 0
    assert _x2p(pb, _x(sp, c=0.0, c_tumb=0.0, b=0.3, b_tumb=0.0))[1][0] == 0


@pytest.mark.parametrize("name", list(FP64_VARIANTS))
def test_maps_are_inverse(name):
    v = FP64_VARIANTS[name]()
    pb, sp = _problem(v, reopt=1)
    if v.n_seg > 1:
        pb = oracle.Problem(sp, pb.curve, t_start=pb.curve.mjd[0] + np.array([0.0, 0.3, 0.6]))
    rng = np.random.default_rng(5)
    x = rng.uniform(0.02, 0.98, (400, sp.n_params))
    p, rc = _x2p(pb, x)
    ok = rc == 0
    assert ok.mean() > 0.9
    x2 = pb.params2x(p[ok])
    assert np.max(np.abs(x2 - x[ok])) <= 1e-12
    p2, rc2 = _x2p(pb, x2)
    assert np.all(rc2 == 0)
    assert np.max(np.abs(p2 - p[ok]) / np.maximum(np.abs(p[ok]), 1e-300)) <= 1e-12


def test_maps_inverse_float_simplex():
    # the reference's default build keeps x in float (CHI_FLOAT): the round
    # trip holds to float resolution
    v = variants.ls_torque(abi.PREC_F32SIMPLEX)
    pb, sp = _problem(v, reopt=1)
    x = np.random.default_rng(6).uniform(0.02, 0.98, (300, sp.n_params)).astype(np.float32).astype(np.float64)
    p, rc = _x2p(pb, x)
    x2 = pb.params2x(p[rc == 0])
    assert np.max(np.abs(x2 - x[rc == 0])) <= 4e-7
