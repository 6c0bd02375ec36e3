"""ctypes wrapper of the CPU oracle (liboracle.so).  Test infrastructure only:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the product."""
import ctypes as C
import os

import numpy as np

from asteroid_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LITERAL, CANON = 0, 1


class OrcProblem(C.Structure):
    _fields_ = [("v", abi.Variant), ("n_params", C.c_int), ("n_types", C.c_int),
                ("property", C.c_int * (abi.MAX_PARAMS * abi.N_COLUMNS)),
                ("types", C.c_int * abi.MAX_TYPES), ("limits", C.c_double * (2 * abi.MAX_TYPES)),
                ("x2", abi.X2), ("n_data", C.c_int), ("n_filters", C.c_int),
                ("obs", C.POINTER(abi.Obs)), ("eph", abi.Ephem3), ("n_min", C.c_int),
                ("t_min", C.c_float * abi.MAX_MINIMA), ("V_min", C.c_float * abi.MAX_MINIMA),
                ("t_start", C.c_double * abi.MAX_SEGMENTS), ("has_seg_first", C.c_int),
                ("seg_first", C.c_int * abi.MAX_SEGMENTS)]


def _has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " fma " in line + " "
    except OSError:
        pass
    return False


_lib = None


def lib():
    global _lib
    if _lib is None:
        name = "liboracle_fma.so" if _has_fma() else "liboracle.so"
        # AST_ORACLE_LIB selects another build of the same source (the
        # sanitizer test builds one with -fsanitize=address,undefined)
        path = os.environ.get("AST_ORACLE_LIB") or os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        pD, pI, pI64, pU32 = (C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                              C.POINTER(C.c_uint32))
        pP = C.POINTER(OrcProblem)
        L.orc_chi2.argtypes = [pP, C.c_int, pD, C.c_int, pD, pD, pD, pD]
        L.orc_x2params.argtypes = [pP, C.c_int, C.c_int, pD, C.c_int, pD, pI]
        L.orc_params2x.argtypes = [pP, C.c_int, pD, C.c_int, pD]
        L.orc_run.argtypes = [pP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_double,
                              C.c_uint64, C.c_int, pD, pU32, pD, pD, pD, pD, pI64, pI64, C.c_int]
        L.orc_model_curve.argtypes = [pP, C.c_int, pD, C.c_int, C.c_int, pD, pD, pD, pD, C.c_int]
        L.orc_periodogram.argtypes = [pD, pD, C.c_int, C.c_int, pD, pD, pI, pD]
        L.orc_rmsd.argtypes = [pP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_uint64, C.c_int, pD,
                               pU32, C.POINTER(C.c_float), C.POINTER(C.c_float), pI64, pD, C.c_int]
        L.orc_math.argtypes = [C.c_int, pD, pD, pD, C.c_int]
        L.orc_philox.argtypes = [pU32, pU32, pU32]
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_float)]
        L.orc_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(C.c_float)]
        _lib = L
    return _lib


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class Problem:
    """A model space plus a curve, held in the layout the oracle reads."""

    def __init__(self, space, curve, minima=None, t_start=None, seg_first=None):
        """minima = (t_obs, V_obs) for NUDGE variants, t_start = the segment start
        times for SEGMENT variants (both on the curve's clock); seg_first = the
        first observation of each segment, when the observations are split by
        index rather than by t_start."""
        self.space, self.curve, self.minima, self.t_start = space, curve, minima, t_start
        self.seg_first = seg_first
        pb = OrcProblem()
        pb.v = space.variant
        pb.n_params = space.n_params
        pb.n_types = space.n_types
        for k, val in enumerate(space.property.ravel()):
            pb.property[k] = int(val)
        for k, val in enumerate(space.types):
            pb.types[k] = int(val)
        for k, val in enumerate(space.limits):
            pb.limits[k] = float(val)
        pb.x2 = space.x2()
        self._obs = curve.obs_array()
        pb.n_data = curve.n_data
        pb.n_filters = curve.n_filters
        pb.obs = C.cast(self._obs, C.POINTER(abi.Obs))
        pb.eph = curve.ephem3()
        if minima is not None:
            t_obs, V_obs = minima
            pb.n_min = len(t_obs)
            for k in range(len(t_obs)):
                pb.t_min[k] = float(t_obs[k])
                pb.V_min[k] = float(V_obs[k])
        if t_start is not None:
            for k, val in enumerate(t_start):
                pb.t_start[k] = float(val)
        if seg_first is not None:
            pb.has_seg_first = 1
            for k, val in enumerate(seg_first):
                pb.seg_first[k] = int(val)
        self.pb = pb

    def with_space(self, space):
        return Problem(space, self.curve, self.minima, self.t_start, self.seg_first)

    def chi2(self, params, form=CANON, want_vmod=False, want_traj=False):
        params = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        n = params.shape[0]
        chi2 = np.zeros(n)
        dV = np.zeros((n, self.curve.n_filters))
        vm = np.zeros((n, self.curve.n_data)) if want_vmod else None
        tr = np.zeros((n, self.curve.n_data, 6)) if want_traj else None
        rc = lib().orc_chi2(C.byref(self.pb), form, _p(params), n, _p(chi2), _p(dV), _p(vm), _p(tr))
        assert rc == 0, rc
        out = [chi2, dV]
        if want_vmod:
            out.append(vm)
        if want_traj:
            out.append(tr)
        return tuple(out)

    def model_curve(self, params, n_plot=20000, form=CANON, n_threads=1):
        """chi2_plot on the reference's plotting grid: (t, vmod, chi2, dV)."""
        params = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        n = params.shape[0]
        t = np.zeros(n_plot)
        vm = np.zeros((n, n_plot))
        chi2 = np.zeros(n)
        dV = np.zeros((n, self.curve.n_filters))
        rc = lib().orc_model_curve(C.byref(self.pb), form, _p(params), n, n_plot, _p(t), _p(vm), _p(chi2), _p(dV),
                                   n_threads)
        assert rc == 0, rc
        return t, vm, chi2, dV

    def rmsd(self, n_groups, group_size, n_samples, dx, params0, seed=1, reopt=1, spherical_k=False, rng=None,
             form=CANON, n_threads=1):
        """chi2_gpu_rms: per-group float (par_min, par_max), counts (Ntot, Nbad per
        group), f0, f1; rng (per-lane stream counters) is advanced in place."""
        N = self.space.n_params
        rng = np.zeros(n_groups * group_size, np.uint32) if rng is None else rng
        pmin = np.zeros((n_groups, N), np.float32)
        pmax = np.zeros((n_groups, N), np.float32)
        counts = np.zeros((n_groups, 2), np.int64)
        f = np.zeros(2)
        p0 = np.ascontiguousarray(params0, np.float64)
        rc = lib().orc_rmsd(C.byref(self.pb), form, n_groups, group_size, n_samples, float(dx), int(spherical_k),
                            seed, reopt, _p(p0), _p(rng, C.c_uint32), _p(pmin, C.c_float), _p(pmax, C.c_float),
                            _p(counts, C.c_int64), _p(f), n_threads)
        assert rc == 0, rc
        return dict(par_min=pmin, par_max=pmax, counts=counts, f0=f[0], f1=f[1], rng=rng)

    def x2params(self, x, reopt=0, form=CANON):
        x = np.ascontiguousarray(np.atleast_2d(x), np.float64)
        n = x.shape[0]
        params = np.zeros_like(x)
        rcs = np.zeros(n, np.int32)
        assert lib().orc_x2params(C.byref(self.pb), form, reopt, _p(x), n, _p(params), _p(rcs, C.c_int)) == 0
        return params, rcs

    def params2x(self, params, form=CANON):
        params = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        x = np.zeros_like(params)
        assert lib().orc_params2x(C.byref(self.pb), form, _p(params), params.shape[0], _p(x)) == 0
        return x

    def run(self, n_groups, group_size, nstages=1, n_steps=5000, eval_budget=0, size_min=0.0, seed=1,
            reopt=0, params0=None, rng=None, d_f=None, d_params=None, d_dV=None, form=CANON, n_threads=1):
        """One chi2_gpu launch; state arrays (rng counters, per-group bests) are updated in place
        when given, so consecutive calls behave like the reference's host loop."""
        nf, N = self.curve.n_filters, self.space.n_params
        nfit = n_groups * group_size
        rng = np.zeros(nfit, np.uint32) if rng is None else rng
        # setup_kernel's 1e30 (cuda.c:2050) in d_f's CHI_FLOAT (asteroid.c:730):
        # a float under the float-simplex forms, so a group whose every lane
        # failed reports float(1e30)
        sentinel = float(np.float32(1e30)) if self.space.variant.precision != abi.PREC_FP64 else 1e30
        d_f = np.full(n_groups, sentinel) if d_f is None else d_f
        d_params = np.zeros((n_groups, N)) if d_params is None else d_params
        d_dV = np.zeros((n_groups, nf)) if d_dV is None else d_dV
        p0 = np.zeros(N) if params0 is None else np.ascontiguousarray(params0, np.float64)
        lane_f = np.zeros(nfit)
        lane_nev = np.zeros(nfit, np.int64)
        lane_steps = np.zeros(nfit, np.int64)
        rc = lib().orc_run(C.byref(self.pb), form, n_groups, group_size, nstages, n_steps, eval_budget,
                           size_min, seed, reopt, _p(p0), _p(rng, C.c_uint32), _p(d_f), _p(d_params),
                           _p(d_dV), _p(lane_f), _p(lane_nev, C.c_int64), _p(lane_steps, C.c_int64),
                           n_threads)
        assert rc == 0, rc
        return dict(rng=rng, f=d_f, params=d_params, dV=d_dV, lane_f=lane_f, lane_nev=lane_nev,
                    lane_steps=lane_steps)


def periodogram(t, vm, n_clusters=5):
    """minima() (misc.c:65-276): (cl_fr, cl_H, n_minima, H) or None when
    there are more than 10000 minima."""
    t = np.ascontiguousarray(t, np.float64)
    vm = np.ascontiguousarray(vm, np.float64)
    fr = np.zeros(max(n_clusters, 1))
    h = np.zeros(max(n_clusters, 1))
    H = np.zeros(10000)
    nm = C.c_int()
    rc = lib().orc_periodogram(_p(t), _p(vm), len(t), n_clusters, _p(fr), _p(h), C.byref(nm), _p(H))
    if rc != 0:
        return None
    return fr[:n_clusters], h[:n_clusters], nm.value, H


MATH = {"sin": 0, "cos": 1, "tan": 2, "atan2": 3, "log": 4, "log10": 5, "exp": 6, "asin": 7,
        "atan": 8, "expf": 9, "sin_t": 10, "cos_t": 11, "acos": 12, "powf": 13}


def math(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float64)
    out = np.zeros_like(x)
    assert lib().orc_math(MATH[fn], _p(x), _p(y), _p(out), x.size) == 0
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def uniform(seed, fit, k0, n):
    out = np.zeros(n, np.float32)
    lib().orc_uniform(seed, fit, k0, n, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def normal(seed, fit, k0, n):
    out = np.zeros(n, np.float32)
    lib().orc_normal(seed, fit, k0, n, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out
