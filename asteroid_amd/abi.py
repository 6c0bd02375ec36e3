"""ctypes mirror of include/asteroid.h and the loader of libasteroid_hip.so.

The structures here are byte-for-byte the C ABI's; the loader fails loudly
when the HIP library has not been built (there is no CPU fallback).
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libasteroid_hip.so")

# Property columns and periodic codes (asteroid.h:115-131).
P_TYPE, P_INDEPENDENT, P_FROZEN, P_ISEG, P_MULTI_SEGMENT, P_PERIODIC = range(6)
N_COLUMNS = 6
PERIODIC, PERIODIC_LAM, SOFT_BOTH, HARD_LEFT, HARD_RIGHT, HARD_BOTH = 1, 0, -1, -2, -3, -4
MAX_PARAMS, MAX_TYPES, MAX_FILTERS, MAX_MINIMA, MAX_SEGMENTS = 40, 16, 8, 10, 4
NOALLOC_MAX_MODELS = 16384

PREC_FP64, PREC_F32SIMPLEX, PREC_MIXED = 0, 1, 2

ERRORS = {
    0: "AST_OK", -1: "AST_E_ARG", -2: "AST_E_STATE", -3: "AST_E_HIP", -4: "AST_E_NOMEM",
    -5: "AST_E_DATA", -6: "AST_E_UNSUPPORTED", -7: "AST_E_NODEVICE", -8: "AST_E_CANCELLED",
    -9: "AST_E_BUSY",
}
E_CANCELLED = -8


class Variant(C.Structure):
    _fields_ = [("torque", C.c_int), ("bc", C.c_int), ("bw_ball", C.c_int), ("p_psi", C.c_int),
                ("interp", C.c_int), ("precision", C.c_int), ("time_step", C.c_double), ("trend", C.c_int),
                ("torque2", C.c_int), ("rotate", C.c_int), ("rect", C.c_int), ("nudge", C.c_int), ("v1s", C.c_int),
                ("n_seg", C.c_int), ("one_le", C.c_int)]


class TypeIds(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "theta_M", "phi_M", "phi_0", "L", "Ti", "Ts", "Tl", "c_tumb", "b_tumb", "Es", "psi_0",
        "c", "b", "theta_R", "phi_R", "kappa", "n_types", "n_params", "A", "T2i", "T2s", "T2l", "Tt", "psi_R")]


class Obs(C.Structure):
    _fields_ = [("mjd", C.c_double), ("V", C.c_float), ("w", C.c_float), ("filter", C.c_int),
                ("pad_", C.c_int), ("E", C.c_double * 3), ("S", C.c_double * 3)]


class Ephem3(C.Structure):
    _fields_ = [("E", (C.c_double * 3) * 3), ("S", (C.c_double * 3) * 3), ("mjd", C.c_double * 3)]


class X2(C.Structure):
    _fields_ = [("reopt", C.c_int), ("ppsi1", C.c_float), ("ppsi2", C.c_float)]


class RunCfg(C.Structure):
    _fields_ = [("n_groups", C.c_int), ("group_size", C.c_int), ("nstages", C.c_int),
                ("n_steps", C.c_int), ("eval_budget", C.c_int64), ("size_min", C.c_double),
                ("ticks_per_launch", C.c_int), ("use_graphs", C.c_int), ("resident_groups", C.c_int)]


class ServeCfg(C.Structure):
    _fields_ = [("max_resident_fits", C.c_int64), ("group_size", C.c_int), ("hbm_budget_bytes", C.c_int64),
                ("work_points_per_launch", C.c_double), ("max_params", C.c_int)]


class Job(C.Structure):
    _fields_ = [("variant", Variant), ("n_params", C.c_int), ("property", C.POINTER(C.c_int)),
                ("n_types", C.c_int), ("types", C.POINTER(C.c_int)), ("limits", C.POINTER(C.c_double)),
                ("x2", X2), ("obs", C.POINTER(Obs)), ("n_data", C.c_int), ("n_filters", C.c_int),
                ("eph", C.POINTER(Ephem3)), ("n_fits", C.c_int), ("nstages", C.c_int), ("n_steps", C.c_int),
                ("eval_budget", C.c_int64), ("size_min", C.c_double), ("seed", C.c_uint64),
                ("params0", C.POINTER(C.c_double)), ("tag", C.c_uint64), ("n_minima", C.c_int),
                ("t_minima", C.POINTER(C.c_float)), ("V_minima", C.POINTER(C.c_float)),
                ("t_segments", C.POINTER(C.c_double)), ("seg_first_obs", C.POINTER(C.c_int))]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("tag", C.c_uint64), ("status", C.c_int), ("n_params", C.c_int),
                ("n_filters", C.c_int), ("n_groups", C.c_int), ("best_group", C.c_int), ("best_f", C.c_double),
                ("best_params", C.c_double * MAX_PARAMS), ("best_dV", C.c_double * MAX_FILTERS),
                ("n_evals", C.c_int64), ("model_points", C.c_int64), ("submit_to_done_ms", C.c_double),
                ("queue_ms", C.c_double), ("run_ms", C.c_double)]


class BatchCfg(C.Structure):
    _fields_ = [("group_size", C.c_int), ("work_points_per_launch", C.c_double), ("use_graphs", C.c_int),
                ("schedule", C.c_int)]


class RmsdCfg(C.Structure):
    _fields_ = [("n_groups", C.c_int), ("group_size", C.c_int), ("n_samples", C.c_int), ("dx", C.c_float),
                ("spherical_k", C.c_int)]


class TuneEntry(C.Structure):
    _fields_ = [("torque", C.c_int), ("bc", C.c_int), ("bw_ball", C.c_int), ("precision", C.c_int),
                ("filter_capacity", C.c_int), ("waves", C.c_int), ("kernel_ms", C.c_double * 3)]


class ServeStats(C.Structure):
    _fields_ = [("pending", C.c_int), ("resident", C.c_int), ("free_fits", C.c_int64),
                ("resident_bytes", C.c_int64), ("launches", C.c_int64), ("evals", C.c_int64),
                ("model_points", C.c_int64), ("kernel_ms", C.c_double)]


assert C.sizeof(Obs) == 72


def type_ids(variant):
    """ast_variant_type_ids of include/asteroid.h (asteroid.h:137-178 order)."""
    torque = bool(variant.torque or variant.torque2)
    bc = bool(variant.bc or variant.rect)
    ids = {}
    k = 0
    for name in ("theta_M", "phi_M", "phi_0", "L"):
        ids[name] = k
        k += 1
    ids["A"] = -1
    if variant.trend:
        ids["A"] = k
        k += 1
    for name in ("Ti", "Ts", "Tl", "T2i", "T2s", "T2l", "Tt"):
        ids[name] = -1
    if torque:
        for name in ("Ti", "Ts", "Tl"):
            ids[name] = k
            k += 1
    if variant.torque2:
        for name in ("T2i", "T2s", "T2l", "Tt"):
            ids[name] = k
            k += 1
    for name in ("c_tumb", "b_tumb", "Es", "psi_0"):
        ids[name] = k
        k += 1
    ids["c"] = ids["b"] = -1
    if bc:
        ids["c"], ids["b"] = k, k + 1
        k += 2
    ids["theta_R"] = ids["phi_R"] = ids["psi_R"] = ids["kappa"] = -1
    if variant.rotate or variant.bw_ball:
        ids["theta_R"], ids["phi_R"] = k, k + 1
        k += 2
    if variant.rotate:
        ids["psi_R"] = k
        k += 1
    if variant.bw_ball:
        ids["kappa"] = k
        k += 1
    ids["n_types"] = k
    ids["n_params"] = (8 + 2 * bc + 3 * bool(variant.rotate) + bool(variant.trend)
                       + (7 if variant.torque2 else 3 * torque) + 3 * bool(variant.bw_ball))
    if variant.n_seg > 1:  # asteroid.h:104-108
        ids["n_params"] += ((6 + (7 if variant.torque2 else 3 * torque) - 2 * bool(variant.one_le))
                            * (variant.n_seg - 1))
    return ids


_lib = None


def load():
    """Load libasteroid_hip.so; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback for the fitting kernels)")
    lib = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    pI, pD, pI64 = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)
    sig = {
        "ast_ctx_create": (I, [C.POINTER(C.c_void_p), C.POINTER(Variant), I]),
        "ast_ctx_destroy": (I, [P]),
        "ast_set_model_space": (I, [P, I, pI, I, pI, pD, C.POINTER(X2)]),
        "ast_upload_curve": (I, [P, C.POINTER(Obs), I, I, C.POINTER(Ephem3)]),
        "ast_set_minima": (I, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), I]),
        "ast_set_segments": (I, [P, C.POINTER(C.c_double), I]),
        "ast_set_segment_starts": (I, [P, pI, C.POINTER(C.c_double), I]),
        "ast_rng_seed": (I, [P, C.c_uint64]),
        "ast_reset_best": (I, [P]),
        "ast_set_start": (I, [P, pD]),
        "ast_run": (I, [P, C.POINTER(RunCfg)]),
        "ast_run_begin": (I, [P, C.POINTER(RunCfg)]),
        "ast_run_continue": (I, [P, I, pI]),
        "ast_run_status": (I, [P, pI64, pI64, pI]),
        "ast_run_cancel": (I, [P]),
        "ast_fetch": (I, [P, pD, pD, pD]),
        "ast_fetch_lanes": (I, [P, pD, pI64, pI64]),
        "ast_chi2_batch": (I, [P, pD, I, pD, pD]),
        "ast_chi2_batch_noalloc": (I, [P, pD, I, pD, pD]),
        "ast_model_curve": (I, [P, pD, I, C.POINTER(Ephem3), I, pD, pD, pD, pD]),
        "ast_periodogram": (I, [P, pD, pD, I, I, pD, pD, pI]),
        "ast_rmsd": (I, [P, C.POINTER(RmsdCfg), C.POINTER(C.c_float), C.POINTER(C.c_float), pI64, pI64, pD, pD]),
        "ast_last_timing": (I, [P, pD, pI]),
        "ast_last_counts": (I, [P, pI64, pI64]),
        "ast_default_model_space": (I, [C.POINTER(Variant), pD, pI, pI, pI, pI, pD, C.POINTER(X2)]),
        "ast_run_jobs": (I, [P, C.POINTER(Job), I, C.POINTER(BatchCfg), C.POINTER(Result)]),
        "ast_autotune": (I, [P, C.POINTER(Job), I, C.POINTER(BatchCfg), I, C.POINTER(TuneEntry), I, pI]),
        "ast_set_waves": (I, [P, C.POINTER(Variant), I, I]),
        "ast_serve_start": (I, [P, C.POINTER(ServeCfg)]),
        "ast_copy_bandwidth": (I, [P, C.c_int64, C.c_int, C.POINTER(C.c_double)]),
        "ast_submit": (I, [P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
        "ast_poll": (I, [P, C.POINTER(Result), I, pI, I]),
        "ast_cancel": (I, [P, C.c_uint64]),
        "ast_serve_stats": (I, [P, C.POINTER(ServeStats)]),
        "ast_serve_stop": (I, [P]),
        "ast_strerror": (C.c_char_p, [I]),
        "ast_last_hip_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


EXPORTED = (
    "ast_ctx_create", "ast_ctx_destroy", "ast_set_model_space", "ast_upload_curve", "ast_rng_seed",
    "ast_reset_best", "ast_set_start", "ast_run", "ast_fetch", "ast_fetch_lanes", "ast_chi2_batch",
    "ast_last_timing", "ast_last_counts", "ast_default_model_space", "ast_strerror",
    "ast_last_hip_error", "ast_serve_start", "ast_submit", "ast_poll", "ast_cancel", "ast_serve_stats",
    "ast_serve_stop", "ast_run_jobs", "ast_autotune", "ast_set_waves", "ast_model_curve", "ast_periodogram",
    "ast_rmsd", "ast_set_minima", "ast_set_segments", "ast_run_begin", "ast_run_continue", "ast_run_status",
    "ast_set_segment_starts", "ast_run_cancel", "ast_chi2_batch_noalloc", "ast_copy_bandwidth",
)
