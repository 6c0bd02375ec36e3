"""Algorithmic operation counts of one chi^2 evaluation (the roofline's work).

Counts are of the canonical evaluation order, i.e. exactly the operations
the GPU evaluator executes, measured by running the oracle's evaluator on an
operation-counting scalar type and fitted (exactly, zero residual) to

    ops(eval) = per_model + n_points * per_point + n_rk4_steps * per_rk4

where n_rk4_steps = sum over intervals of int(dt / TIME_STEP + 1)
(cuda.c:321).  Columns: add/sub, mul, fma, div, sqrt, rint/floor.  FLOP
convention: every operation is 1 FLOP except fma, which is 2.  Selects,
sign flips and the simplex bookkeeping (O(N^2) per evaluation) are not
counted.  tests/test_flops.py re-derives these numbers from the oracle.
"""
import numpy as np

COLUMNS = ("add", "mul", "fma", "div", "sqrt", "rint")
_W = np.array([1, 1, 2, 1, 1, 1], float)

# key: (torque, brightness) with brightness "ls" (LS ellipsoid, BC or not),
# "bw", "rect", and the LS forms with "+trend" (TREND) or "+rot" (ROTATE).
# TORQUE2 is the torque row: its split interval runs as two RK4 stretches,
# counted by rk4_steps when the split fraction is given.
TABLE = {
    (True, "ls"): dict(per_model=(16, 33, 61, 11, 3, 0), per_point=(19, 66, 69, 5, 4, 0),
                       per_rk4=(14, 44, 148, 4, 0, 0)),
    (False, "ls"): dict(per_model=(15, 27, 61, 11, 3, 0), per_point=(20, 69, 80, 5, 4, 0),
                        per_rk4=(15, 44, 110, 0, 0, 0)),
    (True, "bw"): dict(per_model=(18, 38, 83, 9, 3, 0), per_point=(11, 42, 50, 1, 0, 0),
                       per_rk4=(14, 44, 148, 4, 0, 0)),
    (False, "bw"): dict(per_model=(17, 32, 83, 9, 3, 0), per_point=(12, 45, 61, 1, 0, 0),
                        per_rk4=(15, 44, 110, 0, 0, 0)),
    # TREND: the LS torque row plus the de-trending fma (the phase angle is a
    # property of the point, computed once on the host)
    (True, "ls+trend"): dict(per_model=(16, 33, 61, 11, 3, 0), per_point=(19, 66, 70, 5, 4, 0),
                             per_rk4=(14, 44, 148, 4, 0, 0)),
    (True, "ls+rot"): dict(per_model=(19, 42, 94, 11, 3, 0), per_point=(19, 84, 87, 5, 4, 0),
                           per_rk4=(14, 44, 148, 4, 0, 0)),
    (False, "ls+rot"): dict(per_model=(18, 36, 94, 11, 3, 0), per_point=(20, 87, 98, 5, 4, 0),
                            per_rk4=(15, 44, 110, 0, 0, 0)),
    (True, "rect"): dict(per_model=(16, 33, 61, 11, 3, 0), per_point=(11, 61, 68, 1, 3, 0),
                         per_rk4=(14, 44, 148, 4, 0, 0)),
    (False, "rect"): dict(per_model=(15, 27, 61, 11, 3, 0), per_point=(12, 64, 79, 1, 3, 0),
                          per_rk4=(15, 44, 110, 0, 0, 0)),
}


def table_key(variant):
    torque = bool(variant.torque or variant.torque2)
    if variant.bw_ball:
        return torque, "bw"
    if variant.rect:
        return torque, "rect"
    return torque, "ls+trend" if variant.trend else "ls+rot" if variant.rotate else "ls"


def rk4_steps(mjd, time_step, split=None):
    """RK4 steps of one evaluation (cuda.c:321); split = TORQUE2's Tt, whose
    interval runs as two stretches (cuda.c:288-323)."""
    mjd = np.asarray(mjd, float)
    d = np.diff(mjd)
    n = int(np.sum((d / time_step + 1).astype(np.int64)))
    if split is not None:
        ts = split * (mjd[-1] - mjd[0]) + mjd[0]
        i = int(np.searchsorted(mjd, ts, side="right"))
        if 0 < i < len(mjd):
            t1, t2 = mjd[i - 1], mjd[i]
            n += int((ts - t1) / time_step + 1) + int((t2 - ts) / time_step + 1) - int((t2 - t1) / time_step + 1)
    return n


def segments(variant, mjd, seg_t):
    """SEGMENT: (one-segment variant, times) of each segment, by the split of
    read_data.c:244-250; every segment is a whole evaluation of its own
    (cuda.c:163-252 run per segment)."""
    import copy
    plain = copy.copy(variant)
    plain.n_seg, plain.one_le = 0, 0
    mjd = np.asarray(mjd, float)
    start, k = [], 0
    for i, t in enumerate(mjd):
        if k < variant.n_seg and t >= seg_t[k]:
            start.append(i)
            k += 1
    ends = start[1:] + [len(mjd)]
    return [(plain, mjd[a:b]) for a, b in zip(start, ends)]


def eval_ops(variant, mjd, split=None, seg_t=None):
    """Operation counts (6 columns) of one evaluation on a curve with times mjd
    (TORQUE2: split = Tt, its split interval counted as two stretches;
    SEGMENT: seg_t = the segment start times)."""
    if variant.n_seg > 1:
        return sum(eval_ops(pv, m) for pv, m in segments(variant, mjd, seg_t))
    t = TABLE[table_key(variant)]
    split = split if variant.torque2 else None
    return (np.array(t["per_model"], float) + len(mjd) * np.array(t["per_point"], float)
            + rk4_steps(mjd, variant.time_step, split) * np.array(t["per_rk4"], float))


def eval_flops(variant, mjd, split=None, seg_t=None):
    """FLOPs of one evaluation; for the mixed form, its FP64 plus FP32 FLOPs.
    TORQUE2 without `split` counts its split interval as one stretch (at most
    one RK4 step short)."""
    if variant.precision == MIXED:
        f64, f32 = eval_flops_mixed(variant, mjd, split, seg_t)
        return f64 + f32
    return float(eval_ops(variant, mjd, split, seg_t) @ _W)


# ------------------------------------------------------------ mixed form
# The mixed-precision evaluator runs the per-model set-up, the RK4
# integration, the scaled Sun/Earth directions, the chi^2 sums and the body
# axes exactly as the FP64 form (the rows above) and replaces
# the FP64 brightness tail by its own (RECT keeps its FP64 tail).
# Counted from ast_device.h and astmath.h:
#   FP64 Lommel-Seeliger tail (ls_tail_f64, from |u_S|^2 to the magnitude,
#     with the log and log10 of astmath.h): 16 add, 32 mul, 27 fma, 5 div,
#     4 sqrt; the Lambert + ball tail (log10 and the -2.5 factor): 6 add,
#     8 mul, 8 fma, 1 div.
#   Mixed Lommel-Seeliger tail (ls_tail_mx): in FP64 the norms, the unit
#     vectors and their sum and differences, 12 add, 20 mul, 12 fma and 3
#     hardware rsq (1 FLOP each, in the sqrt column); in FP32, 3 add, 25 mul,
#     21 fma and 7 hardware rsq/rcp/sqrt/log2 ("special", 1 FLOP each).  The
#     Lambert + ball tail: 1 FP32 log2.  Both add 1 FP64 add and 1 FP64 mul
#     (the magnitude from the exponent and the log2 of the mantissa).
MIXED = 2
_TAIL64 = {"ls": np.array([16, 32, 27, 5, 4, 0], float), "bw": np.array([6, 8, 8, 1, 0, 0], float)}
_TAIL32 = {"ls": np.array([3, 25, 21, 7], float), "bw": np.array([0, 0, 0, 1], float)}  # add mul fma special
_MX_F64_EXTRA = {"ls": np.array([13, 21, 12, 0, 3, 0], float), "bw": np.array([1, 1, 0, 0, 0, 0], float)}
_W4 = np.array([1, 1, 2, 1], float)


def eval_flops_mixed(variant, mjd, split=None, seg_t=None):
    """(FP64 FLOPs, FP32 FLOPs) of one mixed-form evaluation."""
    if variant.n_seg > 1:
        parts = [eval_flops_mixed(pv, m) for pv, m in segments(variant, mjd, seg_t)]
        return sum(a for a, _ in parts), sum(b for _, b in parts)
    key = table_key(variant)
    kind = key[1]
    t = TABLE[key]
    n = len(mjd)
    tail = "rect" if kind == "rect" else "bw" if kind == "bw" else "ls"
    f64_pt = np.array(t["per_point"], float)
    f32_pt = 0.0
    if tail != "rect":
        f64_pt = f64_pt - _TAIL64[tail] + _MX_F64_EXTRA[tail]
        f32_pt = _TAIL32[tail] @ _W4
    split = split if variant.torque2 else None
    f64 = (np.array(t["per_model"], float) @ _W
           + rk4_steps(mjd, variant.time_step, split) * (np.array(t["per_rk4"], float) @ _W) + n * (f64_pt @ _W))
    return float(f64), float(n * f32_pt)
