// ast_device.h -- device side of the gfx950 fitting core.
//
// One GPU lane owns one model.  The light curve is a read-only table of
// wave-uniform records (every lane of a wave is at the same observation at
// the same time, because the RK4 step count of an interval depends only on
// the observation times, cuda.c:321-323), so it is read through the scalar
// data cache and never staged per lane.  The per-model work follows the
// project's canonical evaluation order (explicit FMA, own elementary
// functions in astmath.h, INTERP ephemeris and RK4 interval constants
// precomputed per observation on the host), which is bitwise the oracle's
// canonical form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/asteroid.h"
#include "astmath.h"
#include "astrng.h"

namespace ast {

constexpr double PI = 3.141592653589793238;  // asteroid.h:16
constexpr double SMALL = 1e-8;               // cuda.c:1455
constexpr float BC_DEV_MAX = 100.0f;         // asteroid.h:248
constexpr int MAXP = AST_MAX_PARAMS;
constexpr int MAXF = AST_MAX_FILTERS;

// One observation as the kernels read it: unit E and S at the observation,
// the RK4 interval from the previous observation (cuda.c:321-323; h, h/2,
// h/6 and the step count), V, w and the filter.  96 bytes.
struct PointRec {
  double E[3];
  double S[3];
  double h, hh, h6;
  float V, w;
  int filter, nsteps;
  int pad[2];
};
static_assert(sizeof(PointRec) == 96, "PointRec layout");

// Parameter positions of a variant (the Property order, asteroid.c:86-125).
template <bool TQ, bool BC, bool BW>
struct Pos {
  static constexpr int o = TQ ? 3 : 0;
  static constexpr int theta_M = 0, phi_M = 1, phi_0 = 2;
  static constexpr int Ti = TQ ? 3 : -1, Ts = TQ ? 4 : -1, Tl = TQ ? 5 : -1;
  static constexpr int c_tumb = 3 + o, b_tumb = 4 + o, Es = 5 + o, L = 6 + o, psi_0 = 7 + o;
  static constexpr int c = BC ? 8 + o : -1, b = BC ? 9 + o : -1;
  static constexpr int ob = o + (BC ? 2 : 0);
  static constexpr int theta_R = BW ? 8 + ob : -1, phi_R = BW ? 9 + ob : -1, kappa = BW ? 10 + ob : -1;
  static constexpr int N = 8 + o + (BC ? 2 : 0) + (BW ? 3 : 0);
};

// Model space by parameter position (the reference indexes limits by type;
// positions and types are in one-to-one correspondence here).
template <class CF>
struct Space {
  int8_t frozen[MAXP];
  int8_t periodic[MAXP];
  int8_t independent[MAXP];
  int32_t p_psi;
  CF lim0[MAXP];
  CF lim1[MAXP];
  float ppsi1, ppsi2;
};

__device__ inline bool hard_left(int per, int LAM) {
  return per == AST_HARD_BOTH || per == AST_HARD_LEFT || (LAM == 0 && per == AST_PERIODIC_LAM);
}
__device__ inline bool hard_right(int per, int LAM) {
  return per == AST_HARD_BOTH || per == AST_HARD_RIGHT || (LAM == 0 && per == AST_PERIODIC_LAM);
}

__device__ inline double cf_exp(double x) { return astm::exp(x); }
__device__ inline float cf_exp(float x) { return astm::expf(x); }
__device__ inline double cf_floor(double x) { return __builtin_floor(x); }
__device__ inline float cf_floor(float x) { return __builtin_floorf(x); }

// ------------------------------------------------------------------ x2params
// cuda.c:1120-1319 (P_PSI branch of the L mapping).  Same operation order
// and types as the oracle.  Returns nonzero on a hard-limit crossing.
template <class CF, bool TQ, bool BC, bool BW>
__device__ int x2params(const CF* x, double* params, const Space<CF>& sp, int reopt) {
  using P = Pos<TQ, BC, BW>;
  constexpr int N = P::N;
  int LAM = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (i == P::Es) LAM = x[i] >= 0.5;
    int per = sp.periodic[i];
    if ((x[i] < 0.0 && hard_left(per, LAM)) || (x[i] > 1.0 && hard_right(per, LAM))) return 1;
  }
  double log_c_tumb = 0, log_b_tumb = 0, Is = 0, Ii = 0, psi_min = 0, psi_max = 0, log_c = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const CF L0 = sp.lim0[i], L1 = sp.lim1[i];
    if (i == P::b_tumb) {
      log_b_tumb = log_c_tumb * (x[i] * (L1 - L0) + L0);
      params[i] = astm::exp(log_b_tumb);
      double b_tumb = params[i];
      double c_tumb = params[P::c_tumb];
      Is = (1.0 + b_tumb * b_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
      Ii = (1.0 + c_tumb * c_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
    } else if (i == P::Es) {
      if (LAM) {
        params[i] = 2.0 * (x[i] - 0.5) * (1.0 - 1.0 / Ii) + 1.0 / Ii;
        psi_min = 0.0;
        psi_max = 2.0 * PI;
      } else {
        params[i] = 2.0 * x[i] * (1.0 / Ii - 1.0 / Is) + 1.0 / Is;
        psi_max = astm::atan(__builtin_sqrt(Ii * (Is - 1.0 / params[i]) / Is / (1.0 / params[i] - Ii)));
        psi_min = -psi_max;
      }
    } else if (i == P::psi_0) {
      params[i] = x[i] * (psi_max - psi_min) + psi_min;
    } else if (BC && i == P::b) {
      double log_b = log_c * (x[i] * (L1 - L0) + L0);
      if (__builtin_fabs(log_b - log_b_tumb) > BC_DEV_MAX) return 1;
      params[i] = astm::exp(log_b);
    } else if (sp.periodic[i] == AST_PERIODIC) {
      params[i] = x[i] * 2.0 * PI;
    } else if (sp.independent[i] == 1) {
      if (!sp.p_psi || i != P::L || reopt) params[i] = x[i] * (L1 - L0) + L0;
      if (i == P::c_tumb) {
        log_c_tumb = params[i];
        params[i] = astm::exp(log_c_tumb);
      } else if (BC && i == P::c) {
        log_c = params[i];
        if (__builtin_fabs(log_c - log_c_tumb) > BC_DEV_MAX) return 1;
        params[i] = astm::exp(log_c);
      } else if (BW && i == P::kappa) {
        params[i] = astm::exp(params[i]);
      } else if (i == P::L && !reopt && sp.p_psi) {
        float dP = sp.ppsi2 - sp.ppsi1;
        params[i] = x[i] * (CF)dP + (CF)sp.ppsi1;
        double EsV = params[P::Es];
        double Einv = 1.0 / EsV;
        double k2;
        if (LAM) k2 = (Is - Ii) * (Einv - 1.0) / ((Ii - 1.0) * (Is - Einv));
        else k2 = (Ii - 1.0) * (Is - Einv) / ((Is - Ii) * (Einv - 1.0));
        double a = 1.0, g = __builtin_sqrt(1.0 - k2), a1, g1;
        for (int ii = 0; ii < 5; ++ii) {  // complete elliptic K by AGM, cuda.c:1272-1279
          a1 = 0.5 * (a + g);
          g1 = __builtin_sqrt(a * g);
          a = a1;
          g = g1;
        }
        if (LAM) params[i] = 4.0 * params[i] * PI / (a + g) * __builtin_sqrt(Ii * Is / (EsV * (Ii - 1.0) * (Is - Einv)));
        else params[i] = 4.0 * params[i] * PI / (a + g) * __builtin_sqrt(Ii * Is / (EsV * (Is - Ii) * (Einv - 1.0)));
      }
    }
  }
  return 0;
}

// ------------------------------------------------------------------ params2x
// cuda.c:1016-1115.
template <class CF, bool TQ, bool BC, bool BW>
__device__ void params2x(CF* x, const double* params, const Space<CF>& sp) {
  using P = Pos<TQ, BC, BW>;
  constexpr int N = P::N;
  int LAM = 0, i_Es = 0;
  double b_tumb = params[P::b_tumb];
  double c_tumb = params[P::c_tumb];
  double Is = (1.0 + b_tumb * b_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
  double Ii = (1.0 + c_tumb * c_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
  for (int i = 0; i < N; ++i) {
    if (sp.frozen[i] != 0) { x[i] = 0; continue; }
    const CF L0 = sp.lim0[i], L1 = sp.lim1[i];
    if (sp.independent[i] == 1) {
      double par = params[i];
      if (sp.periodic[i] == 1) {
        x[i] = par / (2 * PI);
        x[i] = x[i] - cf_floor(x[i]);
      } else {
        if (BC && i == P::c) {
          const CF C0 = sp.lim0[P::c_tumb], C1 = sp.lim1[P::c_tumb];
          x[i] = (astm::log(params[i]) - C0) / (C1 - C0);
          continue;
        }
        if (i == P::c_tumb) par = astm::log(par);
        if (BW && i == P::kappa) par = astm::log(par);
        x[i] = (par - L0) / (L1 - L0);
      }
    } else {
      if (i == P::b_tumb) {
        double par = astm::log(b_tumb) / astm::log(c_tumb);
        x[i] = (par - L0) / (L1 - L0);
      } else if (i == P::Es) {
        i_Es = i;
        LAM = params[i] > 1.0 / Ii;
        if (LAM) x[i] = 0.5 * ((params[i] - 1.0 / Ii) / (1.0 - 1.0 / Ii) + 1.0);
        else x[i] = 0.5 * (params[i] - 1.0 / Is) / (1.0 / Ii - 1.0 / Is);
      } else if (i == P::psi_0) {
        double psi_min, psi_max;
        if (LAM) { psi_min = 0.0; psi_max = 2.0 * PI; }
        else {
          psi_max = astm::atan(__builtin_sqrt(Ii * (Is - 1.0 / params[i_Es]) / Is / (1.0 / params[i_Es] - Ii)));
          psi_min = -psi_max;
        }
        x[i] = (params[i] - psi_min) / (psi_max - psi_min);
      } else if (BC && i == P::b) {
        double par = astm::log(params[i]) / astm::log(params[P::c]);
        x[i] = (par - L0) / (L1 - L0);
      }
    }
  }
}

// ----------------------------------------------------------------- evaluator
__device__ inline double dot3(double ax, double ay, double az, double bx, double by, double bz) {
  return __builtin_fma(az, bz, __builtin_fma(ay, by, ax * bx));
}

// ODE right-hand side.  With FIRST the sin/cos of theta (and psi) come from
// sc[] (those of the state, shared with the body axes), else they are computed.
template <bool TQ, bool FIRST>
__device__ inline void ode(const double* y, double* f, const double* mu, double mu02, const double* sc) {
  if (TQ) {  // cuda.c:46-58
    double s5, c5, s4, c4;
    if (FIRST) { s4 = sc[0]; c4 = sc[1]; s5 = sc[2]; c5 = sc[3]; }
    else { astm::sincos(y[5], &s5, &c5); astm::sincos(y[4], &s4, &c4); }
    f[0] = __builtin_fma(mu[0] * y[1], y[2], mu[3]);
    f[1] = __builtin_fma(mu[1] * y[2], y[0], mu[4]);
    f[2] = __builtin_fma(mu[2] * y[0], y[1], mu[5]);
    f[3] = __builtin_fma(y[1], c5, y[0] * s5) / s4;
    f[4] = __builtin_fma(-y[1], s5, y[0] * c5);
    f[5] = __builtin_fma(-f[3], c4, y[2]);
  } else {   // cuda.c:69-75
    double s2, c2, s1, c1;
    astm::sincos(2.0 * y[2], &s2, &c2);
    if (FIRST) { s1 = sc[0]; c1 = sc[1]; }
    else astm::sincos(y[1], &s1, &c1);
    f[0] = mu[0] * __builtin_fma(-mu[2], c2, mu[1]);
    f[1] = (mu02 * s1) * s2;
    f[2] = c1 * (mu[0] - f[0]);
  }
}

typedef __attribute__((address_space(4))) const PointRec* PointPtr;

// chi^2 of one model (chi2one, cuda.c:81-1011, in-scope branches), summed
// per filter; NF is the compile-time filter capacity (>= n_filters).
template <class CF, bool TQ, bool BC, bool BW, int NF>
__device__ CF chi2_model(const double* params, PointPtr pts, int n_data, int n_filters, int dof,
                         const double* sum_w, CF* dV) {
  using P = Pos<TQ, BC, BW>;
  double sy2[NF], sy[NF];
#pragma unroll
  for (int m = 0; m < NF; ++m) { sy2[m] = 0.0; sy[m] = 0.0; }
  double sTM, cTM, sPM, cPM;
  astm::sincos(params[P::theta_M], &sTM, &cTM);
  astm::sincos(params[P::phi_M], &sPM, &cPM);
  const double M_x = sTM * cPM, M_y = sTM * sPM, M_z = cTM;
  const double XM = __builtin_sqrt(__builtin_fma(M_x, M_x, M_z * M_z));
  const double XM_x = M_z / XM;
  const double XM_z = -M_x / XM;
  const double YM_x = M_y * XM_z;
  const double YM_y = __builtin_fma(-M_x, XM_z, M_z * XM_x);
  const double YM_z = -(M_y * XM_x);
  const double bt = params[P::b_tumb], ct = params[P::c_tumb];
  const double den = __builtin_fma(ct, ct, bt * bt);
  const double Is = __builtin_fma(bt, bt, 1.0) / den;
  const double Ii = __builtin_fma(ct, ct, 1.0) / den;
  const double Is_inv = 1.0 / Is;
  const double Ii_inv = 1.0 / Ii;
  const double Es = params[P::Es], psi0 = params[P::psi_0], L = params[P::L];
  double sP0, cP0;
  astm::sincos(psi0, &sP0, &cP0);
  double theta = astm::asin(__builtin_sqrt((Es - 1.0) / (__builtin_fma(sP0 * sP0, Ii_inv - Is_inv, Is_inv) - 1.0)));
  double phi = params[P::phi_0];
  double psi = psi0;
  constexpr int NODE = TQ ? 6 : 3;
  double mu[6];
  double mu02 = 0.0;
  double y[6];
  // sin/cos of the current theta and psi, shared by the next RK4 step's
  // first stage and the body axes at the observation.
  double sc[4];
  astm::sincos(theta, &sc[0], &sc[1]);
  sc[2] = sP0;
  sc[3] = cP0;
  if (TQ) {
    mu[0] = (Is - 1.0) * Ii_inv;
    mu[1] = (1.0 - Ii) * Is_inv;
    mu[2] = Ii - Is;
    mu[3] = params[P::Ti];
    mu[4] = params[P::Ts];
    mu[5] = params[P::Tl];
    y[0] = ((L * Ii_inv) * sc[0]) * sP0;
    y[1] = ((L * Is_inv) * sc[0]) * cP0;
    y[2] = L * sc[1];
    y[3] = phi;
    y[4] = theta;
    y[5] = psi;
  } else {
    mu[0] = L;
    mu[1] = 0.5 * (Ii_inv + Is_inv);
    mu[2] = 0.5 * (Ii_inv - Is_inv);
    mu02 = mu[0] * mu[2];
    y[0] = phi;
    y[1] = theta;
    y[2] = psi;
  }
  double ib = 0, ic = 0, bcp = 0, kappa = 0, sR = 0, cR = 0, stR = 0, ctR = 0;
  if (BW) {
    astm::sincos(params[P::phi_R], &sR, &cR);
    astm::sincos(params[P::theta_R], &stR, &ctR);
    kappa = params[P::kappa];
  } else {
    const double b = BC ? params[P::b] : bt;
    const double c = BC ? params[P::c] : ct;
    ib = 1.0 / b;
    ic = 1.0 / c;
    bcp = b * c;
  }
  for (int i = 0; i < n_data; ++i) {
    const auto& pr = pts[i];
    const int nsteps = pr.nsteps;  // 0 for the first observation
    const double h = pr.h, hh = pr.hh, h6 = pr.h6;
    for (int l = 0; l < nsteps; ++l) {  // RK4, cuda.c:348-372
      double f[6], K[6], s[6];
      ode<TQ, true>(y, K, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) { s[j] = K[j]; f[j] = __builtin_fma(hh, K[j], y[j]); }
      ode<TQ, false>(f, K, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) { s[j] = __builtin_fma(2.0, K[j], s[j]); f[j] = __builtin_fma(hh, K[j], y[j]); }
      ode<TQ, false>(f, K, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) { s[j] = __builtin_fma(2.0, K[j], s[j]); f[j] = __builtin_fma(h, K[j], y[j]); }
      ode<TQ, false>(f, K, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) y[j] = __builtin_fma(h6, s[j] + K[j], y[j]);
      if (TQ) {
        astm::sincos(y[4], &sc[0], &sc[1]);
        astm::sincos(y[5], &sc[2], &sc[3]);
      } else {
        astm::sincos(y[1], &sc[0], &sc[1]);
      }
    }
    if (TQ) { phi = y[3]; theta = y[4]; psi = y[5]; }
    else { phi = y[0]; theta = y[1]; psi = y[2]; }
    double sf, cf, sp, cp;
    astm::sincos(phi, &sf, &cf);
    const double st = sc[0], ctt = sc[1];
    if (TQ) { sp = sc[2]; cp = sc[3]; }
    else astm::sincos(psi, &sp, &cp);
    const double N_x = __builtin_fma(YM_x, sf, XM_x * cf);
    const double N_y = YM_y * sf;
    const double N_z = __builtin_fma(YM_z, sf, XM_z * cf);
    const double p_x = __builtin_fma(-N_z, M_y, N_y * M_z);
    const double p_y = __builtin_fma(-N_x, M_z, N_z * M_x);
    const double p_z = __builtin_fma(-N_y, M_x, N_x * M_y);
    const double a_x = __builtin_fma(p_x, st, M_x * ctt);
    const double a_y = __builtin_fma(p_y, st, M_y * ctt);
    const double a_z = __builtin_fma(p_z, st, M_z * ctt);
    const double w_x = __builtin_fma(-a_z, N_y, a_y * N_z);
    const double w_y = __builtin_fma(-a_x, N_z, a_z * N_x);
    const double w_z = __builtin_fma(-a_y, N_x, a_x * N_y);
    const double b_x = __builtin_fma(w_x, sp, N_x * cp);
    const double b_y = __builtin_fma(w_y, sp, N_y * cp);
    const double b_z = __builtin_fma(w_z, sp, N_z * cp);
    const double c_x = __builtin_fma(-a_z, b_y, a_y * b_z);
    const double c_y = __builtin_fma(-a_x, b_z, a_z * b_x);
    const double c_z = __builtin_fma(-a_y, b_x, a_x * b_y);
    const double E0 = pr.E[0], E1 = pr.E[1], E2 = pr.E[2];
    double Vmod;
    if (BW) {  // cuda.c:475-497, 603-607
      const double n_x = __builtin_fma(c_x, sR, b_x * cR);
      const double n_y = __builtin_fma(c_y, sR, b_y * cR);
      const double n_z = __builtin_fma(c_z, sR, b_z * cR);
      const double q_x = __builtin_fma(-n_z, a_y, n_y * a_z);
      const double q_y = __builtin_fma(-n_x, a_z, n_z * a_x);
      const double q_z = __builtin_fma(-n_y, a_x, n_x * a_y);
      const double r_x = __builtin_fma(q_x, stR, a_x * ctR);
      const double r_y = __builtin_fma(q_y, stR, a_y * ctR);
      const double r_z = __builtin_fma(q_z, stR, a_z * ctR);
      const double cz = dot3(r_x, r_y, r_z, E0, E1, E2);
      Vmod = -2.5 * astm::log10(0.5 * __builtin_fma(kappa, 1.0 + cz, 0.5 * (1.0 - cz)));
    } else {   // cuda.c:666-689 in the canonical form
      const double S0 = pr.S[0], S1 = pr.S[1], S2 = pr.S[2];
      const double bs_x = b_x * ib, bs_y = b_y * ib, bs_z = b_z * ib;
      const double cs_x = c_x * ic, cs_y = c_y * ic, cs_z = c_z * ic;
      const double uSb = dot3(bs_x, bs_y, bs_z, S0, S1, S2);
      const double uSc = dot3(cs_x, cs_y, cs_z, S0, S1, S2);
      const double uSa = dot3(a_x, a_y, a_z, S0, S1, S2);
      const double uEb = dot3(bs_x, bs_y, bs_z, E0, E1, E2);
      const double uEc = dot3(cs_x, cs_y, cs_z, E0, E1, E2);
      const double uEa = dot3(a_x, a_y, a_z, E0, E1, E2);
      const double sS2 = dot3(uSb, uSc, uSa, uSb, uSc, uSa);
      const double sE2 = dot3(uEb, uEc, uEa, uEb, uEc, uEa);
      const double sS = __builtin_sqrt(sS2);
      const double sE = __builtin_sqrt(sE2);
      const double ca = dot3(uSb, uSc, uSa, uEb, uEc, uEa) / (sS * sE);
      const double sa = __builtin_sqrt(__builtin_fma(-ca, ca, 1.0));
      const double sc = __builtin_sqrt(__builtin_fma((2.0 * sS) * sE, ca, sS2 + sE2));
      const double isc = 1.0 / sc;
      const double cl = __builtin_fma(sE, ca, sS) * isc;
      const double sl = (sE * sa) * isc;
      const double cd = __builtin_fma(cl, ca, sl * sa);
      const double sd = __builtin_fma(-cl, sa, sl * ca);
      const double lg = astm::log(((1.0 + cl) * (1.0 + cd)) / (sl * (-sd)));
      const double inner = __builtin_fma(sl * sd, lg, cd + cl);
      Vmod = -2.5 * astm::log10((((bcp * sS) * sE) * isc) * inner);
    }
    const double w = (double)pr.w;
    const double yv = (double)pr.V - Vmod;
    const int m = pr.filter;
    if (NF == 1) {
      sy2[0] = __builtin_fma(yv * yv, w, sy2[0]);
      sy[0] = __builtin_fma(yv, w, sy[0]);
    } else {
#pragma unroll
      for (int k = 0; k < NF; ++k)
        if (m == k) {
          sy2[k] = __builtin_fma(yv * yv, w, sy2[k]);
          sy[k] = __builtin_fma(yv, w, sy[k]);
        }
    }
  }
  CF chi2a = 0.0;  // cuda.c:884-910
#pragma unroll
  for (int m = 0; m < NF; ++m) {
    if (m < n_filters) {
      CF chi2m = sy2[m] - sy[m] * sy[m] / sum_w[m];
      chi2a = chi2a + chi2m;
      dV[m] = sy[m] / sum_w[m];
    }
  }
  chi2a = chi2a / dof;
  return chi2a;
}

}  // namespace ast
