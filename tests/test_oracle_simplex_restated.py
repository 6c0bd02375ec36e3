"""The oracle's Nelder-Mead pinned against a second, independent restatement.

`_chi2_gpu_block` below is a pure-Python reading of one block of the
reference's chi2_gpu kernel (cuda.c:1404-1823): initial point, log-random
vertex offsets with hard-limit sign flips, the argmin sort, centroid and
size test, reflection / expansion / contraction / shrink, the per-stage
reduction over the block and the reseed of the next stage from the block's
best point.  It shares only the elementary pieces with the oracle (the
random stream, x2params / params2x and chi2one, each pinned elsewhere), and
works in numpy scalars of the simplex type so that float and double runs
round where the reference's do.  The compiled oracle (run_group) must agree
with it bit for bit on every lane: final chi^2, evaluation count, step count
and random-stream position, and on the block's stored best parameters.
"""
import collections

import numpy as np
import pytest

import oracle
from asteroid_amd import abi, synth, variants

SMALL = 1e-8  # cuda.c:1455 (a double literal)


def _hard_left(per, lam):  # cuda.c:1495, 1525
    return per in (abi.HARD_BOTH, abi.HARD_LEFT) or (lam == 0 and per == abi.PERIODIC_LAM)


def _hard_right(per, lam):  # cuda.c:1497, 1532
    return per in (abi.HARD_BOTH, abi.HARD_RIGHT) or (lam == 0 and per == abi.PERIODIC_LAM)


class _Block:
    """Per-thread state of one chi2_gpu block, evaluated through the oracle's
    x2params and chi2one."""

    def __init__(self, pb, F, seed, g, G):
        self.pb, self.F, self.seed, self.g, self.G = pb, F, seed, g, G
        self.N = pb.space.n_params
        self.prop = pb.space.property
        self.T = pb.space.ids
        self.types = pb.space.types
        self.seen = collections.Counter()  # simplex moves taken, and how stages ended

    def exp(self, v):
        name = "exp" if self.F is np.float64 else "expf"
        return self.F(oracle.math(name, np.array([float(v)]))[0])

    def eval(self, x, reopt, th):
        """x2params + chi2one (cuda.c:1551-1557); None when x2params refuses x."""
        params, rcs = self.pb.x2params(np.array([float(v) for v in x]), reopt=reopt)
        if rcs[0]:
            return None
        c, dV = self.pb.chi2(params[0])
        th["nev"] += 1
        th["dV"] = dV[0].copy()  # delta_V is the thread's last chi2one output
        return self.F(c[0])

    def stage(self, th, reopt, n_steps, size2_min):
        F, N, prop, T = self.F, self.N, self.prop, self.T
        fit = self.g * self.G + th["t"]

        def rnd():
            r = oracle.uniform(self.seed, fit, th["rng"], 1)[0]
            th["rng"] += 1
            return r

        DX_INI, D2X_INI, DX_RAND = F(0.001), F(-1.2), F(0.0001)  # asteroid.h:205-209
        x = th["x"]
        lam = 0
        for i in range(N):  # cuda.c:1459-1508
            r = rnd()
            t = prop[i][abi.P_TYPE]
            if self.pb.space.variant.bc and not reopt:
                if t == T["c"]:
                    x[0][i] = x[0][self.types[T["c_tumb"]]]
                    continue
                if t == T["b"]:
                    x[0][i] = x[0][self.types[T["b_tumb"]]]
                    continue
            per = prop[i][abi.P_PERIODIC]
            if not reopt or prop[i][abi.P_FROZEN] == -1:
                x[0][i] = F(float(DX_INI) + SMALL + float(r) * (1.0 - 2 * (SMALL + float(DX_INI))))
            else:
                xmin = F(x[0][i] - DX_RAND)
                xmax = F(x[0][i] + DX_RAND)
                if float(xmin) < SMALL and _hard_left(per, lam):
                    xmin = F(SMALL)
                if float(xmax) > 1.0 - SMALL and _hard_right(per, lam):
                    xmax = F(1.0 - SMALL)
                x[0][i] = F(xmin + F(r) * (xmax - xmin))
            if t == T["Es"]:
                lam = int(float(x[0][i]) >= 0.5)
        for j in range(1, N + 1):  # cuda.c:1512-1545
            for i in range(N):
                if i == j - 1:
                    per = prop[i][abi.P_PERIODIC]
                    d2x = rnd()
                    if F is np.float64:
                        dx = DX_INI * self.exp(float(D2X_INI) * float(d2x))
                    else:
                        dx = DX_INI * self.exp(D2X_INI * F(d2x))
                    if rnd() < 0.5:
                        dx = -dx
                        if float(F(x[0][i] + dx)) < SMALL and _hard_left(per, lam):
                            dx = -dx
                    elif float(F(x[0][i] + dx)) > 1.0 - SMALL and _hard_right(per, lam):
                        dx = -dx
                    x[j][i] = F(x[0][i] + dx)
                else:
                    x[j][i] = x[0][i]
        f = th["f"]
        failed = False
        for j in range(N + 1):  # cuda.c:1549-1559
            v = self.eval(x[j], reopt, th)
            if v is None:
                failed = True
                break
            f[j] = v
        ind = th["ind"]
        active = prop[:, abi.P_FROZEN] != 1
        l = 0
        while not failed:  # cuda.c:1568-1735
            l += 1
            taken = [False] * (N + 1)
            for j in range(N + 1):
                fmin, jmin = F(1e30), -1
                for j2 in range(N + 1):
                    if not taken[j2] and f[j2] <= fmin:
                        fmin, jmin = f[j2], j2
                if jmin < 0:
                    ind[0] = 0
                    f[0] = F(1e30)
                    break
                ind[j] = jmin
                taken[jmin] = True
            x0 = np.zeros(N, F)
            for i in range(N):
                s = F(0)
                for j in range(N + 1):
                    s = F(s + x[j][i])
                x0[i] = s / F(N + 1)
            size2 = F(0)
            for j in range(N + 1):
                s = F(0)
                for i in range(N):
                    d = F(x[j][i] - x0[i])
                    s = F(s + d * d)
                size2 = F(size2 + s)
            size2 = size2 / F(N)
            if size2 < size2_min or l > n_steps:
                self.seen["converged" if size2 < size2_min else "step limit"] += 1
                break
            w = ind[N]
            # A frozen coordinate is left unset in the reference's x_r / x_e;
            # here (and in the oracle) it keeps the worst point's value.
            x_r = np.where(active, x0 + F(1.0) * (x0 - x[w]), x[w]).astype(F)
            f_r = self.eval(x_r, reopt, th)
            f_r = F(1e30) if f_r is None else f_r
            if f_r >= f[ind[0]] and f_r < f[ind[N - 1]]:  # reflection
                x[w], f[w] = x_r, f_r
                self.seen["reflect"] += 1
                continue
            if f_r < f[ind[0]]:  # expansion
                x_e = np.where(active, x0 + F(2.0) * (x_r - x0), x_r).astype(F)
                f_e = self.eval(x_e, reopt, th)
                f_e = F(1e30) if f_e is None else f_e
                if f_e < f_r:
                    x[w], f[w] = x_e, f_e
                    self.seen["expand"] += 1
                else:
                    x[w], f[w] = x_r, f_r
                    self.seen["reflect"] += 1
                continue
            x_c = np.where(active, x0 + F(0.5) * (x[w] - x0), x[w]).astype(F)  # contraction
            f_c = self.eval(x_c, reopt, th)
            f_c = F(1e30) if f_c is None else f_c
            if f_c < f[w]:
                x[w], f[w] = x_c, f_c
                self.seen["contract"] += 1
                continue
            self.seen["shrink"] += 1
            bad = False
            b = ind[0]
            for j in range(1, N + 1):  # shrink towards the best point
                v = ind[j]
                x[v] = np.where(active, x[b] + F(0.5) * (x[v] - x[b]), x[v]).astype(F)
                fv = self.eval(x[v], reopt, th)
                if fv is None:
                    bad = True
                else:
                    f[v] = fv
            if bad:
                failed = True
        th["steps"] += l
        th["s_f"] = F(1e30) if failed else f[ind[0]]


def _chi2_gpu_block(pb, F, G, nstages, n_steps, seed, g=0, reopt=0, params0=None):
    """One block (cuda.c:1404-1823) with fresh random streams; returns the
    per-thread records and (d_f, d_params, d_dV) as the block leaves them."""
    blk = _Block(pb, F, seed, g, G)
    N = blk.N
    size_min = F(1e-5) if F is np.float32 else F(1e-10)  # asteroid.h:244-247
    size2_min = F(size_min * size_min)
    threads = []
    for t in range(G):
        th = dict(t=t, rng=0, nev=0, steps=0, s_f=None, dV=None,
                  x=np.zeros((N + 1, N), F), f=np.zeros(N + 1, F), ind=list(range(N + 1)))
        if reopt:  # cuda.c:1409-1416
            th["x"][0] = pb.params2x(np.asarray(params0, np.float64))[0].astype(F)
        threads.append(th)
    s_x0 = None
    thread_min, smin = 0, F(1e30)
    for istage in range(nstages):
        if nstages > 1 and istage > 0:
            for th in threads:
                th["x"][0] = s_x0
        for th in threads:
            blk.stage(th, reopt, n_steps, size2_min)
        thread_min, smin = 0, F(1e30)  # cuda.c:1770-1782 (HUGE = 1e30)
        for j, th in enumerate(threads):
            if th["s_f"] < smin:
                smin, thread_min = th["s_f"], j
        if nstages > 1 and istage < nstages - 1:  # cuda.c:1785-1805
            th = threads[thread_min]
            xb = th["x"][th["ind"][0]]
            if pb.space.variant.p_psi and reopt == 0:
                params, _ = pb.x2params(xb.astype(np.float64), reopt=reopt)
                xb[:] = pb.params2x(params[0])[0].astype(F)
            s_x0 = xb.copy()
            reopt = 1
    th = threads[thread_min]
    best_params, _ = pb.x2params(th["x"][th["ind"][0]].astype(np.float64), reopt=reopt)
    return threads, (smin, best_params[0], th["dV"]), blk.seen


def _problem(v, n=24, span=0.6, seed=21, reopt=0):
    curve, truth = synth.make_curve(seed, v, n, span)
    space = variants.default_model_space(v, reopt=reopt)
    return oracle.Problem(space, curve), truth


# (name, variant, simplex type, curve (n, span, seed), stages, N_STEPS, branches the run must take)
ALL_MOVES = {"reflect", "expand", "contract", "shrink", "step limit"}
CASES = [
    ("ls_torque fp64", lambda: variants.ls_torque(), np.float64, (24, 0.6, 21), 1, 40,
     {"reflect", "expand", "contract", "step limit"}),
    ("ls_torque fp64 two stages", lambda: variants.ls_torque(), np.float64, (24, 0.6, 21), 2, 40,
     {"reflect", "expand", "contract", "step limit"}),
    ("bc_free float simplex", lambda: variants.bc_free(precision=abi.PREC_F32SIMPLEX), np.float32, (24, 0.6, 21),
     2, 40, {"reflect", "expand", "contract", "step limit"}),
    ("ls_torque fp64 long", lambda: variants.ls_torque(), np.float64, (8, 0.3, 3), 1, 1000, ALL_MOVES),
    ("ls_torque float simplex long", lambda: variants.ls_torque(precision=abi.PREC_F32SIMPLEX), np.float32,
     (8, 0.3, 3), 1, 1000, ALL_MOVES | {"converged"}),
]


@pytest.mark.parametrize("name,mk,F,crv,nstages,n_steps,moves", CASES, ids=[c[0] for c in CASES])
def test_oracle_run_equals_restated_chi2_gpu(name, mk, F, crv, nstages, n_steps, moves):
    pb, _ = _problem(mk(), *crv)
    G, seed = 4, 17
    threads, (smin, best, dV), seen = _chi2_gpu_block(pb, F, G, nstages, n_steps, seed)
    r = pb.run(1, G, nstages=nstages, n_steps=n_steps, seed=seed)
    for t, th in enumerate(threads):
        assert r["lane_f"][t] == float(th["s_f"]), (t, r["lane_f"][t], th["s_f"])
        assert r["lane_nev"][t] == th["nev"]
        assert r["lane_steps"][t] == th["steps"]
        assert r["rng"][t] == th["rng"]
    assert r["f"][0] == float(smin)
    assert np.array_equal(r["params"][0], best)
    assert np.array_equal(r["dV"][0], dV)
    # the case must reach the parts of the control flow it is there for
    assert moves <= set(seen), (moves - set(seen), dict(seen))


def test_reoptimisation_block_equals_restated_chi2_gpu():
    v = variants.ls_torque()
    pb, truth = _problem(v, reopt=1)
    threads, (smin, best, dV), _ = _chi2_gpu_block(pb, np.float64, 2, 1, 30, 5, reopt=1, params0=truth)
    r = pb.run(1, 2, nstages=1, n_steps=30, seed=5, reopt=1, params0=truth)
    assert [r["lane_f"][t] for t in range(2)] == [float(th["s_f"]) for th in threads]
    assert [r["lane_nev"][t] for t in range(2)] == [th["nev"] for th in threads]
    assert r["f"][0] == float(smin)
    assert np.array_equal(r["params"][0], best)
