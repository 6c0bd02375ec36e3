"""Model variants and their default parameter tables.

Python mirror of what the reference's main() builds before the fit
(asteroid.c:86-169 Property and Types, asteroid.c:473-601 hard-coded limits,
-f/-l overrides and the P_psi mapping).  The C++ copy of the same tables is
ast_default_model_space() in libasteroid_hip.so (used by the host driver); a
test checks that the two agree.
"""
import math

import numpy as np

from . import abi

PI = 3.141592653589793238          # asteroid.h:16
RAD = 180.0 / PI                   # asteroid.h:17
TMAX = 10.0                        # asteroid.c:507


def make_variant(torque=True, bc=False, bw_ball=False, p_psi=True, interp=True,
                 precision=abi.PREC_FP64, time_step=1e-2, trend=False, torque2=False, rotate=False, rect=False,
                 nudge=False, v1s=False, n_seg=0, one_le=False):
    """The reference's macros; TORQUE2 implies TORQUE and RECT implies BC (asteroid.h:32-38)."""
    return abi.Variant(int(torque or torque2), int(bc or rect), int(bw_ball), int(p_psi), int(interp),
                       int(precision), float(time_step), int(trend), int(torque2), int(rotate), int(rect),
                       int(nudge), int(v1s), int(n_seg), int(one_le))


def param_positions(variant):
    """Position of each named parameter in the parameter vector (the Property
    order, asteroid.c:86-125); type ids (abi.type_ids) are a different order."""
    order = ["theta_M", "phi_M", "phi_0"]
    if variant.trend:
        order += ["A"]
    if variant.torque or variant.torque2:
        order += ["Ti", "Ts", "Tl"]
    if variant.torque2:
        order += ["T2i", "T2s", "T2l", "Tt"]
    order += ["c_tumb", "b_tumb", "Es", "L", "psi_0"]
    if variant.bc or variant.rect:
        order += ["c", "b"]
    if variant.rotate or variant.bw_ball:
        order += ["theta_R", "phi_R"]
    if variant.rotate:
        order += ["psi_R"]
    if variant.bw_ball:
        order += ["kappa"]
    # SEGMENT: segment k > 0 repeats the parameters that are not shared, as
    # "name:k", after the first segment's (asteroid.c:128-150)
    for k in range(1, variant.n_seg):
        order += [f"{n}:{k}" for n in order[:len(order)] if ":" not in n and not _shared(variant, n)]
    return {n: i for i, n in enumerate(order)}


def _shared(variant, name):
    """P_multi_segment of a parameter (asteroid.c:86-125; Es and L under ONE_LE)."""
    if name in ("Es", "L"):
        return bool(variant.one_le)
    return name in ("A", "c_tumb", "b_tumb", "c", "b", "theta_R", "phi_R", "psi_R", "kappa")


# Named variants of the benchmark workloads.
def ls_torque(precision=abi.PREC_FP64, time_step=1e-2, interp=True):
    """Self-consistent LS ellipsoid with torque (-DP_PSI -DTORQUE), 11 parameters."""
    return make_variant(True, False, False, True, interp, precision, time_step)


def bc_free(precision=abi.PREC_FP64, time_step=5e-3, interp=True):
    """Relaxed brightness ellipsoid without torque (-DBC), 10 parameters."""
    return make_variant(False, True, False, True, interp, precision, time_step)


def bw_ball_torque(precision=abi.PREC_FP64, time_step=1e-2, interp=True):
    """Black-and-white ball with torque (-DBW_BALL -DTORQUE -DP_PSI), 14 parameters."""
    return make_variant(True, False, True, True, interp, precision, time_step)


class ModelSpace:
    """Property / Types / Limits / x2 of one variant (the inputs of ast_set_model_space)."""

    def __init__(self, variant, property_rows, types, limits_lo, limits_hi, reopt, ppsi1, ppsi2):
        self.variant = variant
        self.ids = abi.type_ids(variant)
        self.property = np.array(property_rows, dtype=np.int32)
        self.types = np.array(types, dtype=np.int32)
        self.limits = np.concatenate([np.asarray(limits_lo, float), np.asarray(limits_hi, float)])
        self.reopt = int(reopt)
        self.ppsi1 = float(np.float32(ppsi1))
        self.ppsi2 = float(np.float32(ppsi2))

    @property
    def n_params(self):
        return self.property.shape[0]

    @property
    def n_types(self):
        return len(self.types)

    def x2(self):
        return abi.X2(self.reopt, self.ppsi1, self.ppsi2)

    def param_names(self):
        inv = {v: k for k, v in self.ids.items() if k not in ("n_types", "n_params")}
        return [inv[int(t)] for t in self.property[:, abi.P_TYPE]]

    def with_reopt(self, reopt):
        ms = ModelSpace(self.variant, self.property.copy(), self.types.copy(),
                        self.limits[:self.n_types].copy(), self.limits[self.n_types:].copy(),
                        reopt, self.ppsi1, self.ppsi2)
        return ms


def default_model_space(variant, ppsi_hours=(2.0, 4800.0), reopt=0, frozen=None, limits=None,
                        free_random=()):
    """The reference's tables for `variant`.

    frozen: {type_id: value} as the -f option (asteroid.c:566-581);
    limits: {type_id: (lo, hi)} as the -l option (asteroid.c:583-589);
    free_random: parameter positions given as "v" in -m (P_frozen = -1, asteroid.c:308-313).
    """
    T = abi.type_ids(variant)
    rows = [
        (T["theta_M"], 1, 0, 0, 0, abi.HARD_BOTH),
        (T["phi_M"], 1, 0, 0, 0, abi.PERIODIC),
        (T["phi_0"], 1, 0, 0, 0, abi.PERIODIC),
    ]
    if variant.trend:
        rows += [(T["A"], 1, 0, 0, 1, abi.HARD_BOTH)]
    if variant.torque or variant.torque2:
        rows += [(T["Ti"], 1, 0, 0, 0, abi.SOFT_BOTH), (T["Ts"], 1, 0, 0, 0, abi.SOFT_BOTH),
                 (T["Tl"], 1, 0, 0, 0, abi.SOFT_BOTH)]
    if variant.torque2:
        rows += [(T["T2i"], 1, 0, 0, 0, abi.SOFT_BOTH), (T["T2s"], 1, 0, 0, 0, abi.SOFT_BOTH),
                 (T["T2l"], 1, 0, 0, 0, abi.SOFT_BOTH), (T["Tt"], 1, 0, 0, 0, abi.HARD_BOTH)]
    rows += [
        (T["c_tumb"], 1, 0, 0, 1, abi.HARD_RIGHT),
        (T["b_tumb"], 0, 0, 0, 1, abi.HARD_BOTH),
        (T["Es"], 0, 0, 0, 0, abi.HARD_BOTH),
        (T["L"], 1, 0, 0, 0, abi.HARD_LEFT),
        (T["psi_0"], 0, 0, 0, 0, abi.PERIODIC_LAM),
    ]
    if variant.bc or variant.rect:
        rows += [(T["c"], 1, 0, 0, 1, abi.HARD_RIGHT), (T["b"], 0, 0, 0, 1, abi.HARD_BOTH)]
    if variant.rotate or variant.bw_ball:
        rows += [(T["theta_R"], 1, 0, 0, 1, abi.HARD_BOTH), (T["phi_R"], 1, 0, 0, 1, abi.PERIODIC)]
    if variant.rotate:
        rows += [(T["psi_R"], 1, 0, 0, 1, abi.PERIODIC)]
    if variant.bw_ball:
        rows += [(T["kappa"], 1, 0, 0, 1, abi.SOFT_BOTH)]
    rows = [list(r) for r in rows]
    if variant.one_le:  # ONE_LE: Es and L shared by the segments (asteroid.c:68-72, 108-109)
        for r in rows:
            if r[abi.P_TYPE] in (T["Es"], T["L"]):
                r[abi.P_MULTI_SEGMENT] = 1
    n0 = len(rows)
    for k in range(1, variant.n_seg):  # asteroid.c:128-150
        for j in range(n0):
            if rows[j][abi.P_MULTI_SEGMENT] == 0:
                rows.append(list(rows[j]))
                rows[-1][abi.P_ISEG] = k
    assert len(rows) == T["n_params"]
    nt = T["n_types"]
    types = [-1] * nt
    for i, r in enumerate(rows[:n0]):
        types[r[abi.P_TYPE]] = i
    lo = [0.0] * nt
    hi = [0.0] * nt
    lo[T["theta_M"]], hi[T["theta_M"]] = 0.001 / RAD, 179.999 / RAD
    lo[T["L"]], hi[T["L"]] = 48.0 * PI / 10, 48.0 * PI / 0.1
    if variant.trend:
        lo[T["A"]], hi[T["A"]] = -5.0, 5.0                   # asteroid.c:498-503
    if variant.torque or variant.torque2:
        for n in ("Ti", "Ts", "Tl"):
            lo[T[n]], hi[T[n]] = -TMAX, TMAX
    if variant.torque2:                                      # asteroid.c:515-525
        for n in ("T2i", "T2s", "T2l"):
            lo[T[n]], hi[T[n]] = -TMAX, TMAX
        lo[T["Tt"]], hi[T["Tt"]] = 0.001, 0.999
    lo[T["c_tumb"]], hi[T["c_tumb"]] = math.log(0.01), math.log(1.0)
    lo[T["b_tumb"]], hi[T["b_tumb"]] = 0.0, 1.0
    if variant.bc or variant.rect:
        lo[T["c"]], hi[T["c"]] = lo[T["c_tumb"]], hi[T["c_tumb"]]
        lo[T["b"]], hi[T["b"]] = 0.0, 1.0
    if variant.rotate or variant.bw_ball:
        lo[T["theta_R"]], hi[T["theta_R"]] = 0.001 / RAD, 179.999 / RAD
    if variant.bw_ball:
        lo[T["kappa"]], hi[T["kappa"]] = math.log(0.01), math.log(0.1)
    for itype, value in (frozen or {}).items():
        for r in rows:
            if r[abi.P_TYPE] == itype:
                r[abi.P_FROZEN] = 1
                lo[itype] = hi[itype] = float(value)
    for itype, (a, b) in (limits or {}).items():
        lo[itype], hi[itype] = float(a), float(b)
    for k in free_random:
        rows[k][abi.P_FROZEN] = -1
    ppsi1 = np.float32(24.0 / ppsi_hours[1])   # asteroid.c:599-600
    ppsi2 = np.float32(24.0 / ppsi_hours[0])
    return ModelSpace(variant, rows, types, lo, hi, reopt, ppsi1, ppsi2)


def ls_torque_trend(precision=abi.PREC_FP64, time_step=1e-2, interp=True):
    """LS ellipsoid with torque, de-trended by the phase angle (-DTREND), 12 parameters."""
    return make_variant(True, False, False, True, interp, precision, time_step, trend=True)


def ls_torque2(precision=abi.PREC_FP64, time_step=1e-2, interp=True):
    """LS ellipsoid with a torque that changes at the split time (-DTORQUE2), 15 parameters."""
    return make_variant(True, False, False, True, interp, precision, time_step, torque2=True)


def bc_rotate(precision=abi.PREC_FP64, time_step=1e-2, interp=True, torque=False):
    """BC with the brightness ellipsoid rotated (-DBC -DROTATE), 13 (16 with torque) parameters."""
    return make_variant(torque, True, False, True, interp, precision, time_step, rotate=True)


def rect(precision=abi.PREC_FP64, time_step=1e-2, interp=True, torque=False):
    """Rectangular prism brightness (-DRECT), 10 (13 with torque) parameters."""
    return make_variant(torque, True, False, True, interp, precision, time_step, rect=True)


def ls_torque_nudge(precision=abi.PREC_FP64, time_step=1e-2, interp=True, bc=False, v1s=False):
    """Stage Three (README.md:87-101): LS (or BC) with torque, chi^2 nudged by
    the observed minima (-DNUDGE), 11 (13 with BC) parameters."""
    return make_variant(True, bc, False, True, interp, precision, time_step, nudge=True, v1s=v1s)


def ls_segments(precision=abi.PREC_FP64, time_step=1e-2, interp=True, n_seg=3, one_le=True, torque=False, bc=False):
    """-DSEGMENT -DN_SEG=n_seg [-DONE_LE] [-DTORQUE] [-DBC], LS."""
    return make_variant(torque, bc, False, True, interp, precision, time_step, n_seg=n_seg, one_le=one_le)
