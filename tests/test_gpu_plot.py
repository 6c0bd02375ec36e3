"""The -plot path on the GPU (ast_model_curve, ast_periodogram) against the
oracle: the grid, chi^2, delta_V and every plotted magnitude bit for bit in
the FP64 and float-simplex forms; the mixed form within its tolerance; a
batch of models equal to each model alone; the periodogram's clusters bit
for bit (its histogram counts are exact integers)."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu

VARIANTS = {
    "ls_torque": lambda p: variants.ls_torque(p),
    "bc": lambda p: variants.bc_free(p),
    "bw_torque": lambda p: variants.bw_ball_torque(p),
    "ls_plain": lambda p: variants.make_variant(False, False, False, True, True, p, 1e-2),
    "ls_noninterp": lambda p: variants.ls_torque(p, interp=False),
    "trend": lambda p: variants.ls_torque_trend(p),
    "torque2": lambda p: variants.ls_torque2(p),
    "bc_rotate": lambda p: variants.bc_rotate(p),
    "rect_torque": lambda p: variants.rect(p, torque=True),
}


def _case(vname, prec, nf=2, n=150, seed=21):
    v = VARIANTS[vname](prec)
    curve, truth = synth.make_curve(seed, v, n, 2.5, n_filters=nf)
    space = variants.default_model_space(v)
    return v, curve, truth, space


def _models(space, curve, truth, n, seed):
    """truth plus n-1 random models of the space (physical units, inside its limits)."""
    pb = oracle.Problem(space, curve)
    x = np.random.default_rng(seed).uniform(0.05, 0.95, (4 * n, space.n_params))
    p, rc = pb.x2params(x)
    return np.vstack([truth[None], p[rc == 0][: n - 1]])


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_F32SIMPLEX], ids=["fp64", "f32simplex"])
@pytest.mark.parametrize("vname", list(VARIANTS))
def test_model_curve_bitwise(vname, prec):
    v, curve, truth, space = _case(vname, prec)
    params = _models(space, curve, truth, 5, 3)
    want = oracle.Problem(space, curve).model_curve(params, n_plot=3000)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got = ctx.model_curve(params, n_plot=3000)
    for name, g, w in zip(("t", "vmod", "chi2", "dV"), got, want):
        assert np.array_equal(g, w), (name, np.max(np.abs(g - w)))


def test_model_curve_full_grid_one_model():
    """The reference's case: one model on the 20000-point grid."""
    v, curve, truth, space = _case("ls_torque", abi.PREC_F32SIMPLEX, nf=1, n=400, seed=5)
    want = oracle.Problem(space, curve).model_curve(truth, n_plot=20000)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        got = ctx.model_curve(truth, n_plot=20000)
        ms, launches = ctx.last_timing()
    print(f"one model x 20000 grid points: {ms:.2f} ms")
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_model_curve_batch_equals_alone():
    v, curve, truth, space = _case("bc", abi.PREC_FP64, nf=3)
    params = _models(space, curve, truth, 300, 9)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        t, vm, c2, dv = ctx.model_curve(params, n_plot=1500)
        for k in (0, 77, 299):
            t1, vm1, c1, dv1 = ctx.model_curve(params[k], n_plot=1500)
            assert np.array_equal(vm1[0], vm[k]) and c1[0] == c2[k] and np.array_equal(dv1[0], dv[k])


@pytest.mark.parametrize("vname", ["ls_torque", "bc", "bw_torque"])
def test_model_curve_mixed_within_tolerance(vname):
    v64, curve, truth, space64 = _case(vname, abi.PREC_FP64, nf=2)
    vmx = VARIANTS[vname](abi.PREC_MIXED)
    spacemx = variants.default_model_space(vmx)
    params = _models(space64, curve, truth, 8, 4)
    t, vm64, c64, _ = oracle.Problem(space64, curve).model_curve(params, n_plot=3000)
    with api.Context(vmx) as ctx:
        ctx.set_model_space(spacemx)
        ctx.upload_curve(curve)
        tm, vmm, cmx, _ = ctx.model_curve(params, n_plot=3000)
    assert np.array_equal(t, tm)
    # the FP32 brightness tail moves a magnitude by a few 1e-7; chi^2 within the mixed form's gate
    err = np.max(np.abs(vmm - vm64))
    print("max |dVmod| mixed vs FP64:", err)
    assert err < 2e-5
    assert np.max(np.abs(cmx - c64) / c64) < 1e-6 + 2.0 ** -23


def test_periodogram_bitwise_on_model_curves():
    v, curve, truth, space = _case("ls_torque", abi.PREC_F32SIMPLEX, nf=1, n=300, seed=12)
    params = _models(space, curve, truth, 3, 6)
    with api.Context(v) as ctx:
        ctx.set_model_space(space)
        ctx.upload_curve(curve)
        t, vm, _, _ = ctx.model_curve(params, n_plot=20000)
        for k in range(len(params)):
            fr, H, nmin = ctx.periodogram(t, vm[k], 5)
            wfr, wH, wn, _ = oracle.periodogram(t, vm[k], 5)
            print(k, nmin, fr, H)
            assert nmin == wn and np.array_equal(fr, wfr) and np.array_equal(H, wH)


def test_periodogram_bitwise_many_minima():
    # 400 minima: 79800 intervals, every histogram bin counted on the GPU
    n, P = 40000, 0.05
    t = np.linspace(0.0, 20.0, n)
    vm = -np.cos(2 * np.pi * t / P + 0.3) + 0.01 * np.sin(2 * np.pi * t / 7.3)
    wfr, wH, wn, _ = oracle.periodogram(t, vm, 5)
    v = variants.ls_torque(abi.PREC_FP64)
    with api.Context(v) as ctx:
        fr, H, nmin = ctx.periodogram(t, vm, 5)
        ms, _ = ctx.last_timing()
    print(f"{nmin} minima: histogram kernel {ms:.3f} ms; clusters {fr}")
    assert nmin == wn == 400
    assert np.array_equal(fr, wfr) and np.array_equal(H, wH)


def test_periodogram_too_many_minima_is_refused():
    n = 20004
    with api.Context(variants.ls_torque()) as ctx:
        with pytest.raises(api.AsteroidError):
            ctx.periodogram(np.arange(n, dtype=float), np.tile([0.0, 1.0], n // 2), 5)
