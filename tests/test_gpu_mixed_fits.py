"""Mixed-precision fits of every built variant (NUDGE has no mixed form):
a small Stage One search per variant (2 groups of 64, 40 simplex steps,
launch cycles replayed from HIP graphs), whose every group's reported chi^2
is within 1e-6 of the FP64 oracle's chi^2 of the model it reports, and whose
evaluation counts equal those of the same search run eagerly."""
import numpy as np
import pytest

import oracle
from asteroid_amd import abi, api, synth, variants

pytestmark = pytest.mark.gpu

MX = abi.PREC_MIXED
VARIANTS = {
    "ls_torque": lambda p: variants.ls_torque(p),
    "ls": lambda p: variants.make_variant(False, False, False, True, True, p),
    "bc": lambda p: variants.make_variant(False, True, False, True, True, p),
    "bc_torque": lambda p: variants.make_variant(True, True, False, True, True, p),
    "bw": lambda p: variants.make_variant(False, False, True, True, True, p),
    "bw_torque": lambda p: variants.bw_ball_torque(p),
    "trend_torque": lambda p: variants.ls_torque_trend(p),
    "torque2": lambda p: variants.ls_torque2(p),
    "bc_rotate": lambda p: variants.bc_rotate(p),
    "bc_rotate_torque": lambda p: variants.bc_rotate(p, torque=True),
    "rect": lambda p: variants.rect(p),
    "rect_torque": lambda p: variants.rect(p, torque=True),
    "segment3": lambda p: variants.ls_segments(p),
}


@pytest.mark.parametrize("name", list(VARIANTS))
def test_mixed_fit_within_tolerance_of_fp64(name):
    v = VARIANTS[name](MX)
    v64 = VARIANTS[name](abi.PREC_FP64)
    seg = None
    if v.n_seg > 1:
        curve, _ = synth.make_curve(61, variants.make_variant(False, False, False, True, True, MX), 90, 1.5)
        seg = curve.mjd[0] + np.array([0.0, 0.5, 1.0])
    else:
        curve, _ = synth.make_curve(61, v, 90, 1.5)
    space = variants.default_model_space(v)
    print(name, "params", space.n_params, flush=True)
    res = {}
    for graphs in (True, False):
        with api.Context(v) as ctx:
            ctx.set_model_space(space)
            if seg is not None:
                ctx.set_segments(seg)
            ctx.upload_curve(curve)
            ctx.rng_seed(8)
            ctx.run(2, 64, nstages=1, n_steps=40, use_graphs=graphs, ticks_per_launch=17)
            f, p, dv = ctx.fetch()
            e, _ = ctx.last_counts()
        res[graphs] = (f, p, e)
    f, p, e = res[True]
    assert np.array_equal(f, res[False][0]) and np.array_equal(p, res[False][1]) and e == res[False][2]
    pb = oracle.Problem(variants.default_model_space(v64), curve, t_start=seg)
    ok = f < 1e29
    assert ok.any()
    c64, _ = pb.chi2(p[ok], form=oracle.CANON)
    rel = np.abs(f[ok] - c64) / c64
    print(name, "rel chi2 err", rel, flush=True)
    assert rel.max() <= 1e-6 + 2.0 ** -23
