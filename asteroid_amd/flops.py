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
    return float(eval_ops(variant, mjd) @ _W)


def flops_per_model_point(variant, mjd):
    return eval_flops(variant, mjd) / len(mjd)
