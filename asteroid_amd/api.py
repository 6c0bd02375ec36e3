"""Python view of the C ABI (include/asteroid.h) over ctypes.

`Context` mirrors the boundary one call per method; errors raise
`AsteroidError` with the ABI's code and message (the reference printed and
exited, cuda_errors.h:4-14).
"""
import ctypes as C

import numpy as np

from . import abi


class AsteroidError(RuntimeError):
    def __init__(self, code, where):
        lib = abi.load()
        msg = lib.ast_strerror(code).decode()
        hip = lib.ast_last_hip_error().decode()
        super().__init__(f"{where}: {abi.ERRORS.get(code, code)} ({msg}){' ' + hip if hip else ''}")
        self.code = code


def _chk(rc, where):
    if rc != 0:
        raise AsteroidError(rc, where)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One fitting context on one GPU (ast_ctx)."""

    def __init__(self, variant, device=0):
        self.lib = abi.load()
        self.variant = variant
        self.ids = abi.type_ids(variant)
        h = C.c_void_p()
        _chk(self.lib.ast_ctx_create(C.byref(h), C.byref(variant), device), "ast_ctx_create")
        self.h = h
        self.space = None
        self.curve = None

    def close(self):
        if self.h:
            self.lib.ast_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model_space(self, space):
        prop = np.ascontiguousarray(space.property.ravel(), np.int32)
        types = np.ascontiguousarray(space.types, np.int32)
        lim = np.ascontiguousarray(space.limits, np.float64)
        x2 = space.x2()
        _chk(self.lib.ast_set_model_space(self.h, space.n_params, prop.ctypes.data_as(C.POINTER(C.c_int)),
                                          space.n_types, types.ctypes.data_as(C.POINTER(C.c_int)), _dp(lim),
                                          C.byref(x2)), "ast_set_model_space")
        self.space = space

    def upload_curve(self, curve):
        obs = curve.obs_array()
        eph = curve.ephem3()
        _chk(self.lib.ast_upload_curve(self.h, obs, curve.n_data, curve.n_filters, C.byref(eph)),
             "ast_upload_curve")
        self.curve = curve

    def set_minima(self, t_obs, V_obs):
        """NUDGE's observed minima (times on the curve's clock, converted magnitudes)."""
        t = np.ascontiguousarray(t_obs, np.float32)
        v = np.ascontiguousarray(V_obs, np.float32)
        _chk(self.lib.ast_set_minima(self.h, t.ctypes.data_as(C.POINTER(C.c_float)),
                                     v.ctypes.data_as(C.POINTER(C.c_float)), len(t)), "ast_set_minima")

    def set_segments(self, t_start):
        """SEGMENT's start times on the curve's clock (before upload_curve)."""
        t = np.ascontiguousarray(t_start, np.float64)
        _chk(self.lib.ast_set_segments(self.h, t.ctypes.data_as(C.POINTER(C.c_double)), len(t)),
             "ast_set_segments")

    def set_segment_starts(self, first_obs, t_start):
        """SEGMENT with the observations' split given as the first observation
        index of each segment, and the plotting grid's as times on the curve's
        clock (before upload_curve)."""
        i = np.ascontiguousarray(first_obs, np.int32)
        t = np.ascontiguousarray(t_start, np.float64)
        _chk(self.lib.ast_set_segment_starts(self.h, i.ctypes.data_as(C.POINTER(C.c_int)),
                                             t.ctypes.data_as(C.POINTER(C.c_double)), len(t)),
             "ast_set_segment_starts")

    def rng_seed(self, seed):
        _chk(self.lib.ast_rng_seed(self.h, int(seed)), "ast_rng_seed")

    def reset_best(self):
        _chk(self.lib.ast_reset_best(self.h), "ast_reset_best")

    def set_start(self, params0):
        p = np.ascontiguousarray(params0, np.float64)
        _chk(self.lib.ast_set_start(self.h, _dp(p)), "ast_set_start")

    def run(self, n_groups, group_size=256, nstages=1, n_steps=5000, eval_budget=0, size_min=0.0,
            ticks_per_launch=0, use_graphs=False, resident_groups=0):
        cfg = abi.RunCfg(n_groups, group_size, nstages, n_steps, eval_budget, size_min, ticks_per_launch,
                         int(use_graphs), resident_groups)
        _chk(self.lib.ast_run(self.h, C.byref(cfg)), "ast_run")
        self._last_groups = n_groups
        self._last_fits = n_groups * group_size

    def run_begin(self, n_groups, group_size=256, nstages=1, n_steps=5000, eval_budget=0, size_min=0.0,
                  ticks_per_launch=0, use_graphs=False, resident_groups=0):
        """Start a run (ast_run_begin); advance it with run_continue."""
        cfg = abi.RunCfg(n_groups, group_size, nstages, n_steps, eval_budget, size_min, ticks_per_launch,
                         int(use_graphs), resident_groups)
        _chk(self.lib.ast_run_begin(self.h, C.byref(cfg)), "ast_run_begin")
        self._last_groups = n_groups
        self._last_fits = n_groups * group_size

    def run_continue(self, max_cycles=0):
        """Up to max_cycles launch cycles (0: to the end); True once every fit has ended."""
        fin = C.c_int(0)
        _chk(self.lib.ast_run_continue(self.h, max_cycles, C.byref(fin)), "ast_run_continue")
        return bool(fin.value)

    def run_cancel(self):
        """End an unfinished run early (ast_run_cancel)."""
        _chk(self.lib.ast_run_cancel(self.h), "ast_run_cancel")

    def run_status(self):
        """(fits started, fits ended, finished) of the current run."""
        st, dn, fin = C.c_int64(0), C.c_int64(0), C.c_int(0)
        _chk(self.lib.ast_run_status(self.h, C.byref(st), C.byref(dn), C.byref(fin)), "ast_run_status")
        return st.value, dn.value, bool(fin.value)

    def fetch(self):
        G, N, NF = self._last_groups, self.space.n_params, self.curve.n_filters
        f = np.zeros(G)
        p = np.zeros((G, N))
        dv = np.zeros((G, NF))
        _chk(self.lib.ast_fetch(self.h, _dp(f), _dp(p), _dp(dv)), "ast_fetch")
        return f, p, dv

    def fetch_lanes(self):
        n = self._last_fits
        f = np.zeros(n)
        ne = np.zeros(n, np.int64)
        ns = np.zeros(n, np.int64)
        _chk(self.lib.ast_fetch_lanes(self.h, _dp(f), ne.ctypes.data_as(C.POINTER(C.c_int64)),
                                      ns.ctypes.data_as(C.POINTER(C.c_int64))), "ast_fetch_lanes")
        return f, ne, ns

    def chi2_batch_noalloc(self, params):
        """chi2_batch through ast_chi2_batch_noalloc (no device allocation)."""
        params = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        n = params.shape[0]
        chi2 = np.zeros(n)
        dv = np.zeros((n, self.curve.n_filters))
        _chk(self.lib.ast_chi2_batch_noalloc(self.h, _dp(params), n, _dp(chi2), _dp(dv)), "ast_chi2_batch_noalloc")
        return chi2, dv

    def chi2_batch(self, params):
        p = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        n = p.shape[0]
        chi2 = np.zeros(n)
        dv = np.zeros((n, self.curve.n_filters))
        _chk(self.lib.ast_chi2_batch(self.h, _dp(p), n, _dp(chi2), _dp(dv)), "ast_chi2_batch")
        return chi2, dv

    def model_curve(self, params, n_plot=20000, curve=None):
        """The -plot evaluation (chi2_plot, cuda.c:1865-1966) of one or more
        models: (t, vmod, chi2, dV) with the grid times t (n_plot), vmod
        (n_models x n_plot), chi2 (n_models) and dV (n_models x n_filters).
        The grid's interior E,S come from the 3-point ephemeris of `curve`
        (default: the uploaded curve)."""
        p = np.ascontiguousarray(np.atleast_2d(params), np.float64)
        n = p.shape[0]
        eph = (curve or self.curve).ephem3()
        t = np.zeros(n_plot)
        vmod = np.zeros((n, n_plot))
        chi2 = np.zeros(n)
        dv = np.zeros((n, self.curve.n_filters))
        _chk(self.lib.ast_model_curve(self.h, _dp(p), n, C.byref(eph), int(n_plot), _dp(t), _dp(vmod), _dp(chi2),
                                      _dp(dv)), "ast_model_curve")
        return t, vmod, chi2, dv

    def periodogram(self, t, vm, n_clusters=5):
        """minima() (misc.c:65-276) on one model curve: (cl_fr, cl_H, n_minima)."""
        t = np.ascontiguousarray(t, np.float64)
        vm = np.ascontiguousarray(vm, np.float64)
        fr = np.zeros(max(n_clusters, 1))
        h = np.zeros(max(n_clusters, 1))
        nm = C.c_int()
        _chk(self.lib.ast_periodogram(self.h, _dp(t), _dp(vm), len(t), int(n_clusters), _dp(fr), _dp(h),
                                      C.byref(nm)), "ast_periodogram")
        return fr[:n_clusters], h[:n_clusters], nm.value

    def rmsd(self, n_groups, group_size, n_samples, dx, spherical_k=False):
        """chi2_gpu_rms (cuda.c:2148-2389) around the start model (set_start):
        per-group float ranges par_min, par_max (n_groups x n_params), this
        call's n_tot and n_bad, and the start model's f0 and threshold f1."""
        N = self.space.n_params
        pmin = np.zeros((n_groups, N), np.float32)
        pmax = np.zeros((n_groups, N), np.float32)
        nt, nb = C.c_int64(), C.c_int64()
        f0, f1 = C.c_double(), C.c_double()
        cfg = abi.RmsdCfg(int(n_groups), int(group_size), int(n_samples), float(dx), int(spherical_k))
        pf = C.POINTER(C.c_float)
        _chk(self.lib.ast_rmsd(self.h, C.byref(cfg), pmin.ctypes.data_as(pf), pmax.ctypes.data_as(pf), C.byref(nt),
                               C.byref(nb), C.byref(f0), C.byref(f1)), "ast_rmsd")
        return dict(par_min=pmin, par_max=pmax, n_tot=nt.value, n_bad=nb.value, f0=f0.value, f1=f1.value)

    def last_timing(self):
        ms = C.c_double()
        n = C.c_int()
        _chk(self.lib.ast_last_timing(self.h, C.byref(ms), C.byref(n)), "ast_last_timing")
        return ms.value, n.value

    def last_counts(self):
        e = C.c_int64()
        p = C.c_int64()
        _chk(self.lib.ast_last_counts(self.h, C.byref(e), C.byref(p)), "ast_last_counts")
        return e.value, p.value


    # ------------------------------------------------------------- batches
    @staticmethod
    def _job(space, curve, n_fits, nstages=1, n_steps=5000, eval_budget=0, size_min=0.0, seed=1, params0=None,
             tag=0, minima=None, t_segments=None, seg_first_obs=None):
        """A ctypes ast_job plus the arrays it points into (kept alive by the caller).
        minima = (t_obs, V_obs) for NUDGE variants, t_segments = the start times
        for SEGMENT variants."""
        keep = dict(prop=np.ascontiguousarray(space.property.ravel(), np.int32),
                    types=np.ascontiguousarray(space.types, np.int32),
                    lim=np.ascontiguousarray(space.limits, np.float64),
                    obs=curve.obs_array(), eph=curve.ephem3(),
                    p0=None if params0 is None else np.ascontiguousarray(params0, np.float64),
                    tmin=None if minima is None else np.ascontiguousarray(minima[0], np.float32),
                    vmin=None if minima is None else np.ascontiguousarray(minima[1], np.float32),
                    tseg=None if t_segments is None else np.ascontiguousarray(t_segments, np.float64),
                    sfirst=None if seg_first_obs is None else np.ascontiguousarray(seg_first_obs, np.int32))
        job = abi.Job(space.variant, space.n_params, keep["prop"].ctypes.data_as(C.POINTER(C.c_int)),
                      space.n_types, keep["types"].ctypes.data_as(C.POINTER(C.c_int)), _dp(keep["lim"]),
                      space.x2(), C.cast(keep["obs"], C.POINTER(abi.Obs)), curve.n_data, curve.n_filters,
                      C.pointer(keep["eph"]), int(n_fits), int(nstages), int(n_steps), int(eval_budget),
                      float(size_min), int(seed), _dp(keep["p0"]) if keep["p0"] is not None else None, int(tag))
        if minima is not None:
            job.n_minima = len(keep["tmin"])
            job.t_minima = keep["tmin"].ctypes.data_as(C.POINTER(C.c_float))
            job.V_minima = keep["vmin"].ctypes.data_as(C.POINTER(C.c_float))
        if t_segments is not None:
            job.t_segments = _dp(keep["tseg"])
        if seg_first_obs is not None:
            job.seg_first_obs = keep["sfirst"].ctypes.data_as(C.POINTER(C.c_int))
        return job, keep

    @staticmethod
    def _result(r):
        return dict(id=r.id, tag=r.tag, status=r.status, best_f=r.best_f, best_group=r.best_group,
                    best_params=np.array(r.best_params[:r.n_params]), best_dV=np.array(r.best_dV[:r.n_filters]),
                    n_evals=r.n_evals, model_points=r.model_points, latency_ms=r.submit_to_done_ms,
                    queue_ms=r.queue_ms, run_ms=r.run_ms)

    def run_jobs(self, jobs, group_size=256, work_points_per_launch=0.0, use_graphs=False, schedule=0):
        """Run a batch of jobs (dicts of Context.submit's arguments) to completion
        in shared launch cycles; returns their results in order."""
        built = [self._job(**j) for j in jobs]
        arr = (abi.Job * len(built))(*[b[0] for b in built])
        out = (abi.Result * len(built))()
        cfg = abi.BatchCfg(int(group_size), float(work_points_per_launch), int(use_graphs), int(schedule))
        _chk(self.lib.ast_run_jobs(self.h, arr, len(built), C.byref(cfg), out), "ast_run_jobs")
        return [self._result(out[k]) for k in range(len(built))]

    def autotune(self, jobs, group_size=256, work_points_per_launch=0.0, cycles=4):
        """Pick each kernel class's register budget by timing `jobs`; returns the report."""
        built = [self._job(**j) for j in jobs]
        arr = (abi.Job * len(built))(*[b[0] for b in built])
        out = (abi.TuneEntry * 64)()
        n = C.c_int()
        cfg = abi.BatchCfg(int(group_size), float(work_points_per_launch), 0, 0)
        _chk(self.lib.ast_autotune(self.h, arr, len(built), C.byref(cfg), int(cycles), out, 64, C.byref(n)),
             "ast_autotune")
        return [dict(torque=e.torque, bc=e.bc, bw_ball=e.bw_ball, precision=e.precision,
                     filter_capacity=e.filter_capacity, waves=e.waves, kernel_ms=list(e.kernel_ms))
                for e in out[:n.value]]

    def set_waves(self, variant, n_filters, waves):
        _chk(self.lib.ast_set_waves(self.h, C.byref(variant), int(n_filters), int(waves)), "ast_set_waves")

    def copy_bandwidth(self, n_bytes, reps=1):
        """Bytes read plus written per second of a plain device copy of n_bytes."""
        out = C.c_double()
        _chk(self.lib.ast_copy_bandwidth(self.h, int(n_bytes), int(reps), C.byref(out)), "ast_copy_bandwidth")
        return out.value

    # ------------------------------------------------------------- serving
    def serve_start(self, max_resident_fits, group_size=256, hbm_budget_bytes=0, work_points_per_launch=0.0,
                    max_params=0):
        cfg = abi.ServeCfg(int(max_resident_fits), int(group_size), int(hbm_budget_bytes),
                           float(work_points_per_launch), int(max_params))
        _chk(self.lib.ast_serve_start(self.h, C.byref(cfg)), "ast_serve_start")
        self._keep = {}

    def submit(self, space, curve, n_fits, nstages=1, n_steps=5000, eval_budget=0, size_min=0.0, seed=1,
               params0=None, tag=0, minima=None, t_segments=None, seg_first_obs=None):
        """Submit a fit job (its own variant, model space and curve); returns its id."""
        job, keep = self._job(space, curve, n_fits, nstages, n_steps, eval_budget, size_min, seed, params0, tag,
                              minima, t_segments, seg_first_obs)
        jid = C.c_uint64()
        _chk(self.lib.ast_submit(self.h, C.byref(job), C.byref(jid)), "ast_submit")
        return jid.value

    def poll(self, max_results=64, timeout_ms=0):
        out = (abi.Result * max_results)()
        n = C.c_int()
        _chk(self.lib.ast_poll(self.h, out, max_results, C.byref(n), int(timeout_ms)), "ast_poll")
        return [self._result(out[k]) for k in range(n.value)]

    def cancel(self, jid):
        _chk(self.lib.ast_cancel(self.h, int(jid)), "ast_cancel")

    def serve_stats(self):
        st = abi.ServeStats()
        _chk(self.lib.ast_serve_stats(self.h, C.byref(st)), "ast_serve_stats")
        return {f: getattr(st, f) for f, _ in abi.ServeStats._fields_}

    def serve_stop(self):
        _chk(self.lib.ast_serve_stop(self.h), "ast_serve_stop")

    def wait_all(self, ids, timeout_s=600.0):
        """Poll until every id in `ids` has a result; returns {id: result}."""
        import time
        want = set(ids)
        got = {}
        t0 = time.time()
        while want - set(got):
            for r in self.poll(256, timeout_ms=100):
                got[r["id"]] = r
            if time.time() - t0 > timeout_s:
                raise TimeoutError(f"{len(want - set(got))} jobs unfinished")
        return got


def default_model_space_c(variant, ppsi_hours=(2.0, 4800.0)):
    """The library's C++ copy of the reference tables (for cross-checking)."""
    lib = abi.load()
    n_params = C.c_int()
    n_types = C.c_int()
    prop = np.zeros(abi.MAX_PARAMS * abi.N_COLUMNS, np.int32)
    types = np.zeros(abi.MAX_TYPES, np.int32)
    lim = np.zeros(2 * abi.MAX_TYPES)
    x2 = abi.X2()
    pp = np.array(ppsi_hours, float)
    _chk(lib.ast_default_model_space(C.byref(variant), _dp(pp), C.byref(n_params),
                                     prop.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n_types),
                                     types.ctypes.data_as(C.POINTER(C.c_int)), _dp(lim), C.byref(x2)),
         "ast_default_model_space")
    n, nt = n_params.value, n_types.value
    return dict(property=prop[:n * 6].reshape(n, 6), types=types[:nt], limits=lim[:2 * nt],
                ppsi1=x2.ppsi1, ppsi2=x2.ppsi2)
