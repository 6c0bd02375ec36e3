"""Seeded synthetic light curves and ephemerides for the benchmark workloads.

The reference ships no data (its light curves and HORIZONS ephemerides are
external files, README.md:11-41), so every workload here is generated:

* a 3-point ephemeris of the asteroid->Earth and asteroid->Sun vectors that
  brackets the observations (the INTERP layout, asteroid.h:328-336), with
  non-degenerate phase angles (the LS formula is singular at zero phase);
* observation times, filters and per-point errors;
* magnitudes of a ground-truth model plus Gaussian noise.

The ground-truth magnitudes come from an independent numpy model of the
same physics (Euler's equations with constant or piecewise-constant torque
integrated by scipy's DOP853 at tight tolerance, body axes from a z-x-z
Euler rotation matrix, the brightness frame turned by a second one, the
Lommel-Seeliger ellipsoid of Muinonen & Lumme 2015, the black-and-white
ball, or the projected area of a rectangular prism, de-trended by the phase
angle), not from the fitting code itself.
"""
import numpy as np

from . import abi

LIGHT_SPEED = 173.144632674  # au/day (asteroid.h:287)


class Curve:
    """One light curve after ingest: the contents of ast_upload_curve."""

    def __init__(self, mjd, V, w, filt, n_filters, E, S, eph=None):
        self.mjd = np.asarray(mjd, np.float64)
        self.V = np.asarray(V, np.float32)
        self.w = np.asarray(w, np.float32)
        self.filter = np.asarray(filt, np.int32)
        self.n_filters = int(n_filters)
        self.E = np.asarray(E, np.float64)   # unit asteroid->Earth per point
        self.S = np.asarray(S, np.float64)   # unit asteroid->Sun per point
        self.eph = eph                       # dict E (3x3 [comp][k]), S, mjd (3,)

    @property
    def n_data(self):
        return len(self.mjd)

    def obs_array(self):
        arr = (abi.Obs * self.n_data)()
        for i in range(self.n_data):
            o = arr[i]
            o.mjd = self.mjd[i]
            o.V = self.V[i]
            o.w = self.w[i]
            o.filter = int(self.filter[i])
            for c in range(3):
                o.E[c] = self.E[i, c]
                o.S[c] = self.S[i, c]
        return arr

    def ephem3(self):
        e = abi.Ephem3()
        if self.eph is None:
            return e
        for c in range(3):
            for k in range(3):
                e.E[c][k] = self.eph["E"][c][k]
                e.S[c][k] = self.eph["S"][c][k]
        for k in range(3):
            e.mjd[k] = self.eph["mjd"][k]
        return e

    def subset(self, n):
        """The first n observations (same ephemeris)."""
        return Curve(self.mjd[:n], self.V[:n], self.w[:n], self.filter[:n], self.n_filters,
                     self.E[:n], self.S[:n], self.eph)


def quad_interp(eph, t):
    """Quadratic Lagrange interpolation of the 3-point ephemeris (misc.c:20-35)."""
    M = np.asarray(eph["mjd"], float)
    t = np.asarray(t, float)
    rr0 = (t - M[1]) * (t - M[2]) / (M[0] - M[1]) / (M[0] - M[2])
    rr1 = (t - M[0]) * (t - M[2]) / (M[1] - M[0]) / (M[1] - M[2])
    rr2 = (t - M[0]) * (t - M[1]) / (M[2] - M[0]) / (M[2] - M[1])
    E = np.stack([eph["E"][c][0] * rr0 + eph["E"][c][1] * rr1 + eph["E"][c][2] * rr2 for c in range(3)], -1)
    S = np.stack([eph["S"][c][0] * rr0 + eph["S"][c][1] * rr1 + eph["S"][c][2] * rr2 for c in range(3)], -1)
    return E, S


def make_ephemeris(rng, t_first, t_last, r_sun=1.3, r_earth=0.3, min_phase_deg=15.0):
    """3 ephemeris points bracketing [t_first, t_last] (days, relative).

    The asteroid sits r_sun au from the Sun and r_earth au from the Earth; the
    Sun and Earth directions seen from it are kept at least min_phase_deg
    apart (an exactly collinear 1.3/0.3/1.0 au triangle has zero phase, where
    the LS brightness is undefined).  Both vectors drift linearly in time.
    """
    for _ in range(1000):
        s_dir = rng.normal(size=3)
        s_dir /= np.linalg.norm(s_dir)
        e_dir = rng.normal(size=3)
        e_dir /= np.linalg.norm(e_dir)
        phase = np.degrees(np.arccos(np.clip(s_dir @ e_dir, -1, 1)))
        if min_phase_deg + 5 <= phase <= 60:
            break
    v_s = rng.normal(scale=0.01, size=3)   # au/day
    v_e = rng.normal(scale=0.02, size=3)
    tk = np.array([t_first - 0.5, 0.5 * (t_first + t_last), t_last + 0.5])
    E = np.array([r_earth * e_dir + v_e * (tt - tk[1]) for tt in tk]).T  # [comp][k]
    S = np.array([r_sun * s_dir + v_s * (tt - tk[1]) for tt in tk]).T
    return {"E": E, "S": S, "mjd": tk}


# ------------------------------------------------------------- truth model
def _euler_rhs(Ii, Is, T):
    mu0, mu1, mu2 = (Is - 1.0) / Ii, (1.0 - Ii) / Is, Ii - Is

    def rhs(_t, y):
        wi, ws, wl, phi, theta, psi = y
        dphi = (wi * np.sin(psi) + ws * np.cos(psi)) / np.sin(theta)
        return [mu0 * ws * wl + T[0], mu1 * wl * wi + T[1], mu2 * wi * ws + T[2],
                dphi, wi * np.cos(psi) - ws * np.sin(psi), wl - dphi * np.cos(theta)]
    return rhs


def _zxz(phi, theta, psi):
    """Rotation matrices Rz(phi) Rx(theta) Rz(psi), shape (n, 3, 3)."""
    cf, sf = np.cos(phi), np.sin(phi)
    ct, st = np.cos(theta), np.sin(theta)
    cp, sp = np.cos(psi), np.sin(psi)
    R = np.empty(phi.shape + (3, 3))
    R[..., 0, 0] = cf * cp - sf * ct * sp
    R[..., 0, 1] = -cf * sp - sf * ct * cp
    R[..., 0, 2] = sf * st
    R[..., 1, 0] = sf * cp + cf * ct * sp
    R[..., 1, 1] = -sf * sp + cf * ct * cp
    R[..., 1, 2] = -cf * st
    R[..., 2, 0] = st * sp
    R[..., 2, 1] = st * cp
    R[..., 2, 2] = ct
    return R


def truth_state(variant, params, ids, t, rtol=1e-12, atol=1e-12):
    """Euler angles and body angular velocity at times t (first = initial state)."""
    from scipy.integrate import solve_ivp
    p = {k: params[v] for k, v in _param_index(ids, variant).items()}
    bt, ct_ = p["b_tumb"], p["c_tumb"]
    Is = (1 + bt * bt) / (bt * bt + ct_ * ct_)
    Ii = (1 + ct_ * ct_) / (bt * bt + ct_ * ct_)
    psi0, Es, L = p["psi_0"], p["Es"], p["L"]
    theta0 = np.arcsin(np.sqrt((Es - 1) / (np.sin(psi0) ** 2 * (1 / Ii - 1 / Is) + 1 / Is - 1)))
    y0 = [L / Ii * np.sin(theta0) * np.sin(psi0), L / Is * np.sin(theta0) * np.cos(psi0),
          L * np.cos(theta0), p["phi_0"], theta0, psi0]
    T = [p["Ti"], p["Ts"], p["Tl"]] if (variant.torque or variant.torque2) else [0.0, 0.0, 0.0]
    if not variant.torque2:
        sol = solve_ivp(_euler_rhs(Ii, Is, T), (t[0], t[-1]), y0, method="DOP853", t_eval=t,
                        rtol=rtol, atol=atol)
        return sol.y  # (6, n)
    # TORQUE2: the second torque from t_split = t0 + Tt (t_last - t0) on
    t_split = t[0] + p["Tt"] * (t[-1] - t[0])
    before = t[t < t_split]
    sol1 = solve_ivp(_euler_rhs(Ii, Is, T), (t[0], t_split), y0, method="DOP853",
                     t_eval=np.append(before, t_split), rtol=rtol, atol=atol)
    T2 = [p["T2i"], p["T2s"], p["T2l"]]
    sol2 = solve_ivp(_euler_rhs(Ii, Is, T2), (t_split, t[-1]), sol1.y[:, -1], method="DOP853",
                     t_eval=t[t >= t_split], rtol=rtol, atol=atol)
    return np.hstack([sol1.y[:, :-1], sol2.y])


def _param_index(ids, variant):
    from .variants import param_positions
    return param_positions(variant)


def truth_vmod(variant, params, t, E, S):
    """Model magnitudes (before the per-filter offset) at times t."""
    ids = abi.type_ids(variant)
    idx = _param_index(ids, variant)
    p = {k: params[v] for k, v in idx.items()}
    y = truth_state(variant, params, ids, np.asarray(t, float))
    phi, theta, psi = y[3], y[4], y[5]
    tM, pM = p["theta_M"], p["phi_M"]
    M = np.array([np.sin(tM) * np.cos(pM), np.sin(tM) * np.sin(pM), np.cos(tM)])
    XM = np.cross([0.0, 1.0, 0.0], M)
    XM /= np.linalg.norm(XM)
    YM = np.cross(M, XM)
    frame = np.stack([XM, YM, M], axis=1)         # M-frame basis as columns
    B = frame[None] @ _zxz(phi, theta, psi)       # columns: b, c, a in inertial axes
    if variant.rotate:                            # the brightness frame, turned in the body frame
        B = B @ _zxz(np.array(p["phi_R"]), np.array(p["theta_R"]), np.array(p["psi_R"]))
    Eb = np.einsum("nij,ni->nj", B, E)            # (Ep_b, Ep_c, Ep_a)
    Sb = np.einsum("nij,ni->nj", B, S)
    trend = 0.0
    if variant.trend:                             # de-trending by the phase angle
        trend = -p["A"] * np.arccos(np.clip(np.sum(E * S, axis=1), -1, 1))
    if variant.bw_ball:
        stR, ctR = np.sin(p["theta_R"]), np.cos(p["theta_R"])
        sfR, cfR = np.sin(p["phi_R"]), np.cos(p["phi_R"])
        a_rot = np.array([stR * sfR, -stR * cfR, ctR])
        cz = Eb @ a_rot
        return -2.5 * np.log10(0.5 * (p["kappa"] * (1 + cz) + 0.5 * (1 - cz))) + trend
    bc = variant.bc or variant.rect
    b = p["b"] if bc else p["b_tumb"]
    c = p["c"] if bc else p["c_tumb"]
    if variant.rect:
        # The reference crosses the inertial axes with Et = (E.b, E.c, E.a) read
        # as an inertial vector; for unit a, b, c = a x b and any Et,
        # |(u x Et) x (v x Et)| = |Et| |Et.(u x v)|, so the three face terms are
        # b |Et.c|, c |Et.b| and b c |Et.a| (|Et| = 1): the projected area of
        # the prism when the frames coincide.
        Et = Eb
        fa = np.abs(np.einsum("ni,ni->n", Et, B[:, :, 2]))
        fb = np.abs(np.einsum("ni,ni->n", Et, B[:, :, 0]))
        fc = np.abs(np.einsum("ni,ni->n", Et, B[:, :, 1]))
        return -2.5 * np.log10(4 * (b * fc + c * fb + b * c * fa)) + trend
    Cm = np.array([1 / b ** 2, 1 / c ** 2, 1.0])
    sS = np.sqrt((Sb * Sb * Cm).sum(1))
    sE = np.sqrt((Eb * Eb * Cm).sum(1))
    ca = (Sb * Eb * Cm).sum(1) / (sS * sE)
    alpha = np.arccos(np.clip(ca, -1, 1))
    sc = np.sqrt(sS ** 2 + sE ** 2 + 2 * sS * sE * ca)
    lam = np.arctan2(sE * np.sin(alpha), sS + sE * ca)
    f = np.cos(lam - alpha) + np.cos(lam) + np.sin(lam) * np.sin(lam - alpha) * np.log(
        1 / np.tan(lam / 2) / np.tan((alpha - lam) / 2))
    return -2.5 * np.log10(b * c * sS * sE / sc * f) + trend


def random_truth(rng, variant, ppsi_hours=(4.0, 10.0), torque_scale=0.2):
    """A plausible tumbling model: SAM or LAM, moderate period, small torque."""
    idx = _param_index(abi.type_ids(variant), variant)
    n = len(idx)
    p = np.zeros(n)
    for _ in range(1000):
        c_t = np.exp(rng.uniform(np.log(0.2), np.log(0.7)))
        b_t = np.exp(np.log(c_t) * rng.uniform(0.2, 0.8))
        Is = (1 + b_t ** 2) / (b_t ** 2 + c_t ** 2)
        Ii = (1 + c_t ** 2) / (b_t ** 2 + c_t ** 2)
        if rng.uniform() < 0.5:   # SAM: 1/Is < Es < 1/Ii
            Es = 1 / Is + rng.uniform(0.2, 0.8) * (1 / Ii - 1 / Is)
            psi_max = np.arctan(np.sqrt(Ii * (Is - 1 / Es) / Is / (1 / Es - Ii)))
            psi0 = rng.uniform(-0.8, 0.8) * psi_max
        else:                     # LAM: 1/Ii < Es < 1
            Es = 1 / Ii + rng.uniform(0.2, 0.8) * (1 - 1 / Ii)
            psi0 = rng.uniform(0, 2 * np.pi)
        arg = (Es - 1) / (np.sin(psi0) ** 2 * (1 / Ii - 1 / Is) + 1 / Is - 1)
        if 0.05 < arg < 0.95:
            break
    period = rng.uniform(*ppsi_hours)
    p[idx["theta_M"]] = rng.uniform(0.3, 2.8)
    p[idx["phi_M"]] = rng.uniform(0, 2 * np.pi)
    p[idx["phi_0"]] = rng.uniform(0, 2 * np.pi)
    p[idx["c_tumb"]] = c_t
    p[idx["b_tumb"]] = b_t
    p[idx["Es"]] = Es
    p[idx["L"]] = 48 * np.pi / period
    p[idx["psi_0"]] = psi0
    if variant.torque or variant.torque2:
        for k in ("Ti", "Ts", "Tl"):
            p[idx[k]] = rng.normal(scale=torque_scale)
    if variant.torque2:
        for k in ("T2i", "T2s", "T2l"):
            p[idx[k]] = rng.normal(scale=torque_scale)
        p[idx["Tt"]] = rng.uniform(0.2, 0.8)
    if variant.trend:
        p[idx["A"]] = rng.uniform(-1.0, 1.0)
    if variant.bc or variant.rect:
        p[idx["c"]] = c_t * np.exp(rng.uniform(-0.2, 0.2))
        p[idx["b"]] = min(1.0, b_t * np.exp(rng.uniform(-0.2, 0.2)))
    if variant.bw_ball or variant.rotate:
        p[idx["theta_R"]] = rng.uniform(0.3, 2.8)
        p[idx["phi_R"]] = rng.uniform(0, 2 * np.pi)
    if variant.rotate:
        p[idx["psi_R"]] = rng.uniform(0, 2 * np.pi)
    if variant.bw_ball:
        p[idx["kappa"]] = np.exp(rng.uniform(np.log(0.01), np.log(0.1)))
    return p


def make_curve(seed, variant, n_obs, span_days, n_filters=1, sigma_range=(0.05, 0.3), truth=None,
               min_phase_deg=15.0):
    """A light curve of n_obs points over span_days with magnitudes of `truth`
    (random if None) plus noise.  Returns (Curve, truth_params)."""
    rng = np.random.default_rng(seed)
    while True:
        t = np.sort(rng.uniform(0.0, span_days, n_obs))
        if np.all(np.diff(t) > 1e-6):
            break
    t = t - t[0]
    eph = make_ephemeris(rng, t[0], t[-1], min_phase_deg=min_phase_deg)
    E, S = quad_interp(eph, t)
    E /= np.linalg.norm(E, axis=1, keepdims=True)
    S /= np.linalg.norm(S, axis=1, keepdims=True)
    if truth is None:
        truth = random_truth(rng, variant)
    vm = truth_vmod(variant, truth, t, E, S)
    filt = rng.integers(0, n_filters, n_obs).astype(np.int32)
    filt[:n_filters] = np.arange(n_filters)  # every filter present
    # per-filter sigma ranges allowed: sigma_range may be a list of ranges
    if np.ndim(sigma_range[0]) == 0:
        ranges = [sigma_range] * n_filters
    else:
        ranges = list(sigma_range)
    sig = np.array([rng.uniform(*ranges[f]) for f in filt])
    offsets = rng.uniform(-1.0, 1.0, n_filters)
    V = vm + offsets[filt] + rng.normal(size=n_obs) * sig
    w = 1.0 / sig ** 2
    return Curve(t, V, w, filt, n_filters, E, S, eph), truth
