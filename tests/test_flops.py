"""The frozen operation table (asteroid_amd/flops.py) equals the oracle's
operation-counting build on curves of different lengths and spans."""
import ctypes as C

import numpy as np
import pytest

import oracle
from asteroid_amd import flops, synth, variants


def _count(v, curve, params):
    L = oracle.lib()
    L.orc_count_ops.argtypes = [C.POINTER(oracle.OrcProblem), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    pb = oracle.Problem(variants.default_model_space(v), curve)
    out = np.zeros(6)
    assert L.orc_count_ops(C.byref(pb.pb), params.ctypes.data_as(C.POINTER(C.c_double)),
                           out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return out


@pytest.mark.parametrize("v", [variants.ls_torque(), variants.bc_free(), variants.bw_ball_torque(),
                               variants.make_variant(False, False, False, True, True),
                               variants.make_variant(False, False, True, True, True), variants.ls_torque_trend(),
                               variants.ls_torque2(), variants.bc_rotate(), variants.bc_rotate(torque=True),
                               variants.rect(), variants.rect(torque=True)])
@pytest.mark.parametrize("n,span", [(37, 0.7), (400, 3.0), (250, 8.0)])
def test_table_matches_counting_oracle(v, n, span):
    curve, truth = synth.make_curve(17, v, n, span)
    split = truth[variants.param_positions(v)["Tt"]] if v.torque2 else None
    got, table = _count(v, curve, truth), flops.eval_ops(v, curve.mjd, split)
    # The per-point and per-step work is branch-free; the once-per-model
    # asin(...) of theta_0 goes through atan2, whose octant reduction costs
    # up to two adds, one division and two adds more on some arguments.
    d = np.abs(got - table)
    assert np.all(d <= np.array([4, 0, 0, 1, 0, 0])), (got, table)


def test_mixed_table_consistent():
    """The FP64 brightness tail the mixed form replaces holds every division
    and square root of an observation (the rest of the point is fma/mul/add
    work), and the mixed counts are the FP64 counts minus that tail plus the
    mixed tail (its FP64 and FP32 parts)."""
    from asteroid_amd import abi, synth, variants
    for torque, kind in ((True, "ls"), (False, "ls"), (True, "bw")):
        per_point = np.array(flops.TABLE[(torque, kind)]["per_point"], float)
        assert np.array_equal(flops._TAIL64[kind][3:5], per_point[3:5])
        assert np.all(per_point - flops._TAIL64[kind] >= 0)
    v64, vmx = variants.ls_torque(abi.PREC_FP64), variants.ls_torque(abi.PREC_MIXED)
    curve, _ = synth.make_curve(3, v64, 50, 1.0)
    f64, f32 = flops.eval_flops_mixed(vmx, curve.mjd)
    w = np.array([1, 1, 2, 1, 1, 1], float)
    expect64 = flops.eval_flops(v64, curve.mjd) - 50 * (flops._TAIL64["ls"] - flops._MX_F64_EXTRA["ls"]) @ w
    assert f64 == expect64 and f32 == 50 * 77 and flops.eval_flops(vmx, curve.mjd) == f64 + f32
    # RECT keeps its FP64 tail, and the body axes are the FP64 form's: the same count
    r64, rmx = variants.rect(abi.PREC_FP64), variants.rect(abi.PREC_MIXED)
    f64, f32 = flops.eval_flops_mixed(rmx, curve.mjd)
    assert f32 == 0 and f64 == flops.eval_flops(r64, curve.mjd)


def test_segment_counts_are_the_sum_over_segments():
    """SEGMENT: the counting oracle on each segment's observations (with that
    segment's parameters) summed equals the table's count for the split
    curve; observations before the first start and the RK4 steps into each
    segment's first observation are not counted."""
    v1 = variants.make_variant(False, False, False, True, True)
    curve, truth = synth.make_curve(5, v1, 300, 3.0)
    vs = variants.ls_segments()
    t_start = curve.mjd[0] + np.array([0.2, 1.1, 2.05])
    parts = flops.segments(vs, curve.mjd, t_start)
    assert [len(m) for _, m in parts] == [int(np.sum((curve.mjd >= a) & (curve.mjd < b)))
                                          for a, b in zip(t_start, list(t_start[1:]) + [np.inf])]
    starts = np.searchsorted(curve.mjd, t_start)
    total = np.zeros(6)
    for k, (a, b) in enumerate(zip(starts, list(starts[1:]) + [curve.n_data])):
        sub = synth.Curve(curve.mjd[a:b], curve.V[a:b], curve.w[a:b], curve.filter[a:b], 1, curve.E[a:b],
                          curve.S[a:b], curve.eph)
        total += _count(v1, sub, truth)
    d = np.abs(total - flops.eval_ops(vs, curve.mjd, seg_t=t_start))
    assert np.all(d <= 3 * np.array([4, 0, 0, 1, 0, 0])), (total, flops.eval_ops(vs, curve.mjd, seg_t=t_start))
