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

# key: (torque, brightness) with brightness "ls" (LS ellipsoid, BC or not) or "bw".
TABLE = {
    (True, "ls"): dict(per_model=(28, 41, 77, 11, 3, 4), per_point=(22, 68, 73, 5, 4, 1),
                       per_rk4=(38, 60, 180, 4, 0, 8)),
    (False, "ls"): dict(per_model=(27, 35, 77, 11, 3, 4), per_point=(26, 73, 88, 5, 4, 2),
                        per_rk4=(39, 60, 142, 0, 0, 8)),
    (True, "bw"): dict(per_model=(36, 50, 107, 9, 3, 6), per_point=(14, 44, 54, 1, 0, 1),
                       per_rk4=(38, 60, 180, 4, 0, 8)),
    (False, "bw"): dict(per_model=(35, 44, 107, 9, 3, 6), per_point=(18, 49, 69, 1, 0, 2),
                        per_rk4=(39, 60, 142, 0, 0, 8)),
}


def rk4_steps(mjd, time_step):
    d = np.diff(np.asarray(mjd, float))
    return int(np.sum((d / time_step + 1).astype(np.int64)))


def eval_ops(variant, mjd):
    """Operation counts (6 columns) of one evaluation on a curve with times mjd."""
    t = TABLE[(bool(variant.torque), "bw" if variant.bw_ball else "ls")]
    return (np.array(t["per_model"], float) + len(mjd) * np.array(t["per_point"], float)
            + rk4_steps(mjd, variant.time_step) * np.array(t["per_rk4"], float))


def eval_flops(variant, mjd):
    """FLOPs of one evaluation; for the mixed form, its FP64 plus FP32 FLOPs."""
    if variant.precision == MIXED:
        f64, f32 = eval_flops_mixed(variant, mjd)
        return f64 + f32
    return float(eval_ops(variant, mjd) @ _W)


# ------------------------------------------------------------ mixed form
# The mixed-precision evaluator runs the per-model set-up, the RK4
# integration, the body axes, the scaled Sun/Earth directions and the chi^2
# sums exactly as the FP64 form (the rows above), and replaces the FP64
# brightness tail by its own.  Counted from ast_device.h:
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


def eval_flops_mixed(variant, mjd):
    """(FP64 FLOPs, FP32 FLOPs) of one mixed-form evaluation."""
    kind = "bw" if variant.bw_ball else "ls"
    t = TABLE[(bool(variant.torque), kind)]
    n = len(mjd)
    f64_pt = np.array(t["per_point"], float) - _TAIL64[kind] + _MX_F64_EXTRA[kind]
    f64 = (np.array(t["per_model"], float) @ _W + rk4_steps(mjd, variant.time_step) * (np.array(t["per_rk4"], float) @ _W)
           + n * (f64_pt @ _W))
    return float(f64), float(n * (_TAIL32[kind] @ _W4))
