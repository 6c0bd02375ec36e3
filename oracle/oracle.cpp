// oracle.cpp -- CPU restatement of Asteroid's GPU fitting core.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load this library, and only to check or to
// time beside the GPU path; the product (libasteroid_hip.so and the host
// driver) never calls it.
//
// What it restates (reference = pulsar123/Asteroid):
//   ode_*            ODE_func, cuda.c:13-77
//   chi2_literal     chi2one, cuda.c:81-1011 (in-scope branches: TORQUE or the
//                    3-ODE form, TORQUE2's split interval cuda.c:288-317,
//                    INTERP, ROTATE cuda.c:475-516, LS ellipsoid cuda.c:663-690,
//                    BW_BALL cuda.c:596-607, RECT cuda.c:609-661, BC
//                    cuda.c:568-574, TREND cuda.c:692-697, chi^2 cuda.c:884-910,
//                    NUDGE's model minima and merit cuda.c:716-758, 912-992),
//                    written with the reference's expression order and the C
//                    library's elementary functions, no FMA.
//   chi2_canon       the same model in the project's canonical evaluation
//                    order: explicit FMA where a product feeds an add, the
//                    project's own elementary functions (astmath.h), the
//                    LS angle algebra done with trig identities instead of
//                    atan2/tan/sincos, divisions by b^2, c^2 folded into scaled
//                    axes.  This is the arithmetic the GPU kernels execute, so
//                    the GPU FP64 path must match it bit for bit.
//   x2params         cuda.c:1120-1319 (P_PSI branch; P_PHI/P_BOTH not built)
//   params2x         cuda.c:1016-1115
//   run_group        chi2_gpu, cuda.c:1325-1834: random or reoptimised start,
//                    initial simplex, Nelder-Mead loop with its five chi2one
//                    call sites, block argmin, stage reseed, best-per-block
//                    write.  Random draws come from astrng.h instead of cuRAND.
//   quadratic interpolation of E,S (INTERP)  cuda.c:522-545
//   plot_grid, model_curve, periodogram   the -plot path: read_data.c:304-360,
//                    chi2_plot cuda.c:1865-1966, minima() misc.c:65-276
//   rmsd_group       chi2_gpu_rms, cuda.c:2148-2389
// How each part is pinned (tests/test_oracle_*.py, tests/test_golden.py):
// closed forms (the Lommel-Seeliger sphere, a prism's projected area, an
// observed minimum on a model minimum multiplying chi^2 by P_MIN),
// invariants (torque-free |L| and energy, Sun-Earth reciprocity), an
// independent scipy model of the rotation and brightness (every variant),
// RK4's fourth order, brute-force least squares for delta_V and chi^2, the
// P_psi period, and the literal form against the canonical one; run_group
// against a second, pure-Python restatement of chi2_gpu (bit for bit on every
// lane, every simplex move and both stopping rules reached).  The random
// streams (astrng.h, replacing cuRAND) are pinned by the Philox4x32-10
// known-answer vectors; cuRAND's own streams are not reproduced.
// Behaviour the reference leaves undefined is fixed here and in the kernels
// the same way: coordinates of frozen parameters in reflected, expanded and
// contracted points equal the worst vertex's (cuda.c:1638, 1663, 1696 leave
// them unset); the sorted index array starts as the identity (cuda.c:1346);
// params2x starts with LAM = 0 (cuda.c:1022); the last delta_V of a thread
// starts at 0 (cuda.c:1345).
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/asteroid.h"
#include "../asteroid_amd/csrc/astmath.h"
#include "../asteroid_amd/csrc/astrng.h"

namespace orc {

const double PI = 3.141592653589793238;  // asteroid.h:16
const double SMALL = 1e-8;               // cuda.c:1455
const float BC_DEV_MAX = 100;            // asteroid.h:248

// ----------------------------------------------------------------- problem
struct Problem {
  ast_variant v;
  ast_type_ids T;
  int n_params, n_types;
  int prop[AST_MAX_PARAMS][AST_N_COLUMNS];
  int types[AST_MAX_TYPES];
  double lim[2][AST_MAX_TYPES];
  ast_x2 x2;
  int n_data, n_filters;
  const ast_obs* obs;
  ast_ephem3 eph;
  // NUDGE's observed minima (chi2_struct, asteroid.h:319-337)
  int n_min;
  float t_min[AST_MAX_MINIMA], V_min[AST_MAX_MINIMA];
  // SEGMENT: first observation of each segment (read_data.c:244-250) and the
  // parameter of each type in each segment (Types[N_TYPES][N_SEG],
  // asteroid.c:152-168); n_seg = 0 without SEGMENT.
  int n_seg;
  double seg_t[AST_MAX_SEGMENTS];
  int seg_start[AST_MAX_SEGMENTS];
  int seg_types[AST_MAX_TYPES][AST_MAX_SEGMENTS];
  // TORQUE2 implies TORQUE and RECT implies BC (asteroid.h:32-38)
  bool tq() const { return v.torque || v.torque2; }
  bool bc() const { return v.bc || v.rect; }
};

// ------------------------------------------------------ elementary functions
struct LibmMath {
  static double sin(double x) { return std::sin(x); }
  static double cos(double x) { return std::cos(x); }
  static double tan(double x) { return std::tan(x); }
  static double asin(double x) { return std::asin(x); }
  static double atan(double x) { return std::atan(x); }
  static double atan2(double y, double x) { return std::atan2(y, x); }
  static double acos(double x) { return std::acos(x); }
  static double log(double x) { return std::log(x); }
  static double log10(double x) { return std::log10(x); }
  static double exp(double x) { return std::exp(x); }
  static double sqrt(double x) { return std::sqrt(x); }
  static float expf(float x) { return std::exp(x); }
};
struct AstMath {
  static double sin(double x) { return astm::sin(x); }
  static double cos(double x) { return astm::cos(x); }
  static double tan(double x) { return astm::tan(x); }
  static double asin(double x) { return astm::asin(x); }
  static double atan(double x) { return astm::atan(x); }
  static double atan2(double y, double x) { return astm::atan2(y, x); }
  static double acos(double x) { return astm::acos(x); }
  static double log(double x) { return astm::log(x); }
  static double log10(double x) { return astm::log10(x); }
  static double exp(double x) { return astm::exp(x); }
  static double sqrt(double x) { return astm::p_sqrt(x); }
  static float expf(float x) { return astm::expf(x); }
};
template <class M> inline double cf_exp(double x) { return M::exp(x); }
template <class M> inline float cf_exp(float x) { return M::expf(x); }
inline double cf_floor(double x) { return std::floor(x); }
inline float cf_floor(float x) { return std::floor(x); }

// ------------------------------------------------------ INTERP (cuda.c:528-545)
// Literal form: the reference's expressions.
inline void interp_literal(const ast_ephem3& e, double t, double E[3], double S[3]) {
  const double* M = e.mjd;
  double rr0 = (t - M[1]) * (t - M[2]) / (M[0] - M[1]) / (M[0] - M[2]);
  double rr1 = (t - M[0]) * (t - M[2]) / (M[1] - M[0]) / (M[1] - M[2]);
  double rr2 = (t - M[0]) * (t - M[1]) / (M[2] - M[0]) / (M[2] - M[1]);
  for (int c = 0; c < 3; ++c) {
    E[c] = e.E[c][0] * rr0 + e.E[c][1] * rr1 + e.E[c][2] * rr2;
    S[c] = e.S[c][0] * rr0 + e.S[c][1] * rr1 + e.S[c][2] * rr2;
  }
  double En = std::sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
  double Sn = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
  for (int c = 0; c < 3; ++c) { E[c] = E[c] / En; S[c] = S[c] / Sn; }
}
// Canonical form (FMA chains left to right).
inline void interp_canon(const ast_ephem3& e, double t, double E[3], double S[3]) {
  using astm::p_fma;
  const double* M = e.mjd;
  double d0 = t - M[0], d1 = t - M[1], d2 = t - M[2];
  double rr0 = (d1 * d2) / (M[0] - M[1]) / (M[0] - M[2]);
  double rr1 = (d0 * d2) / (M[1] - M[0]) / (M[1] - M[2]);
  double rr2 = (d0 * d1) / (M[2] - M[0]) / (M[2] - M[1]);
  for (int c = 0; c < 3; ++c) {
    E[c] = p_fma(e.E[c][2], rr2, p_fma(e.E[c][1], rr1, e.E[c][0] * rr0));
    S[c] = p_fma(e.S[c][2], rr2, p_fma(e.S[c][1], rr1, e.S[c][0] * rr0));
  }
  double En = astm::p_sqrt(p_fma(E[2], E[2], p_fma(E[1], E[1], E[0] * E[0])));
  double Sn = astm::p_sqrt(p_fma(S[2], S[2], p_fma(S[1], S[1], S[0] * S[0])));
  for (int c = 0; c < 3; ++c) { E[c] = E[c] / En; S[c] = S[c] / Sn; }
}

// RK4 step count and step of the interval t1 -> t2 (cuda.c:321-323).
inline void rk4_interval(double t1, double t2, double time_step, int* nsteps, double* h) {
  *nsteps = (int)((t2 - t1) / time_step + 1);
  *h = (t2 - t1) / *nsteps;
}

// ============================================================ literal chi2one
// ------------------------------------------------------------------ NUDGE
// The model-minima bookkeeping of chi2one (cuda.c:146-151, 716-758) and the
// merit that uses it (cuda.c:912-992), with the reference's declared types:
// the expressions are written as there, so C's promotions decide which
// parts run in float and which in double.  Constants: asteroid.h:216-230.
namespace nudge {
const int M_MAX = 30;
const float M_MAX2 = 20;
const float DT_MAX = 0.12;
const float DV_MAX = 24;
const float DT_MAX2 = 1.5 * DT_MAX;
const float P_MIN = 0.01;
const float CHI2_0 = 10;
const float CHI2_1 = 30;
const float L_RC = 0.05;
const float L_RC2 = L_RC * L_RC;
const float L_A = 1.0 / (1.0 - L_RC2 / (1.0 + L_RC2));
const float P_MIN2 = 1 / P_MIN - 1;
}  // namespace nudge

struct Nudge {
  int M = 0;
  bool failed = false;  // the reference returns 1e30 at once (cuda.c:735-738)
  float t_mod[nudge::M_MAX], V_mod[nudge::M_MAX];
  float t_old[2], V_old[2];

  // called once per point i (0-based within the curve) with its model magnitude
  void point(const Problem& P, int i, double MJD, double Vmod) {
    using namespace nudge;
    if (failed) return;
    if (i < 2) {
      t_old[i] = MJD;
      V_old[i] = Vmod;
      return;
    }
    if (V_old[1] > V_old[0] && V_old[1] >= Vmod) {
      bool local = 0;
      for (int ii = 0; ii < P.n_min; ii++)
        if (std::fabs(t_old[1] - P.t_min[ii]) < DT_MAX2) local = 1;
      if (local) {
        M++;
        if (M > M_MAX) {
          failed = true;
          return;
        }
        double a = ((Vmod - V_old[1]) / (MJD - t_old[1]) - (V_old[1] - V_old[0]) / (t_old[1] - t_old[0])) /
                   (MJD - t_old[0]);
        double b = (V_old[1] - V_old[0]) / (t_old[1] - t_old[0]) - a * (t_old[1] + t_old[0]);
        double c = V_old[1] - a * t_old[1] * t_old[1] - b * t_old[1];
        t_mod[M - 1] = -b / 2.0 / a;
        V_mod[M - 1] = a * t_mod[M - 1] * t_mod[M - 1] + b * t_mod[M - 1] + c;
      }
    }
    t_old[0] = t_old[1];
    V_old[0] = V_old[1];
    t_old[1] = MJD;
    V_old[1] = Vmod;
  }

  // powf: glibc's in the literal form, the project's in the canonical one
  static float powf_form(int form, float x, float y) { return form == 0 ? ::powf(x, y) : astm::powf(x, y); }

  template <class CF>
  CF merit(const Problem& P, int form, CF chi2a, const CF* delta_V) const {
    using namespace nudge;
    float S_M = 0.0;
    float P_tot = 1.0;
    for (int imod = 0; imod < M; imod++) {
      for (int iobs = 0; iobs < P.n_min; iobs++) {
        float dt = std::fabs(t_mod[imod] - P.t_min[iobs]);
        if (dt < DT_MAX2) {
          if (dt > DT_MAX) {
            float x = (DT_MAX2 - dt) / (DT_MAX2 - DT_MAX);
            S_M = S_M + x * x * (-2.0 * x + 3.0);
          } else {
            S_M = S_M + 1.0;
            float dV = V_mod[imod] + delta_V[0] - P.V_min[iobs];
            if (P.v.v1s && dV > 0.0) dV = 0.0;
            float x = std::sqrt(dt * dt / DT_MAX / DT_MAX + dV * dV / DV_MAX / DV_MAX);
            if (x < 1.0) {
              float P_i = L_A * x * x / (x * x + L_RC2);
              P_tot = P_tot * (P_MIN * (1.0 + P_MIN2 * P_i));
            }
          }
        }
      }
    }
    P_tot = powf_form(form, P_tot, 1.0 / P.n_min);
    if (P_tot < P_MIN) P_tot = P_MIN;
    if (chi2a > CHI2_1) {
      P_tot = 1.0;
    } else if (chi2a > CHI2_0) {
      float x = (CHI2_1 - chi2a) / (CHI2_1 - CHI2_0);
      float beta = x * x * (-2 * x + 3);
      P_tot = powf_form(form, P_tot, beta);
    }
    float P_M;
    if (S_M < M_MAX2) {
      P_M = 1.0;
    } else if (S_M < M_MAX) {
      float x = (S_M - M_MAX2) / (M_MAX - M_MAX2);
      P_M = 1.0 + 3 * x * x * (-2 * x + 3);
    } else {
      P_M = 4.0;
    }
    return chi2a * P_tot * P_M;
  }
};

struct ModelOut {   // optional per-point outputs
  double* vmod;     // n_data
  double* traj;     // n_data*6: phi, theta, psi, Omega_i, Omega_s, Omega_l
  Nudge* nudge;     // NUDGE's model minima (the magnitudes after a failure are NaN)
};

inline void ode_literal(bool torque, const double* y, double* f, const double* mu) {
  if (torque) {  // cuda.c:46-58
    f[0] = mu[0] * y[1] * y[2] + mu[3];
    f[1] = mu[1] * y[2] * y[0] + mu[4];
    f[2] = mu[2] * y[0] * y[1] + mu[5];
    f[3] = (y[0] * std::sin(y[5]) + y[1] * std::cos(y[5])) / std::sin(y[4]);
    f[4] = y[0] * std::cos(y[5]) - y[1] * std::sin(y[5]);
    f[5] = y[2] - f[3] * std::cos(y[4]);
  } else {       // cuda.c:69-75
    f[0] = mu[0] * (mu[1] - mu[2] * std::cos(2.0 * y[2]));
    f[1] = mu[0] * mu[2] * std::sin(y[1]) * std::sin(2.0 * y[2]);
    f[2] = std::cos(y[1]) * (mu[0] - f[0]);
  }
}

// Per-filter sums of one model (cuda.c:195-715).
inline void model_literal(const Problem& P, const double* params, double* sy2, double* sy,
                          double* sw, ModelOut* out, bool accumulate = false) {
  const ast_type_ids& T = P.T;
  const bool torque = P.tq();
  if (!accumulate)
    for (int m = 0; m < P.n_filters; ++m) sy2[m] = sy[m] = sw[m] = 0.0;
  double tM = params[P.types[T.theta_M]], pM = params[P.types[T.phi_M]];
  double M_x = std::sin(tM) * std::cos(pM);
  double M_y = std::sin(tM) * std::sin(pM);
  double M_z = std::cos(tM);
  double XM = std::sqrt(M_z * M_z + M_x * M_x);
  double XM_x = M_z / XM;
  double XM_z = -M_x / XM;
  double YM_x = M_y * XM_z;
  double YM_y = M_z * XM_x - M_x * XM_z;
  double YM_z = -M_y * XM_x;
  double bt = params[P.types[T.b_tumb]], ct = params[P.types[T.c_tumb]];
  double Is = (1.0 + bt * bt) / (bt * bt + ct * ct);
  double Ii = (1.0 + ct * ct) / (bt * bt + ct * ct);
  double Is_inv = 1.0 / Is;
  double Ii_inv = 1.0 / Ii;
  double Es = params[P.types[T.Es]], psi0 = params[P.types[T.psi_0]], L = params[P.types[T.L]];
  double phi = params[P.types[T.phi_0]];
  double theta = std::asin(std::sqrt((Es - 1.0) / (std::sin(psi0) * std::sin(psi0) * (Ii_inv - Is_inv) + Is_inv - 1.0)));
  double psi = psi0;
  double mu[6];
  double Om_i = 0, Om_s = 0, Om_l = 0;
  if (torque) {
    mu[0] = (Is - 1.0) * Ii_inv;
    mu[1] = (1.0 - Ii) * Is_inv;
    mu[2] = Ii - Is;
    mu[3] = params[P.types[T.Ti]];
    mu[4] = params[P.types[T.Ts]];
    mu[5] = params[P.types[T.Tl]];
    Om_i = L * Ii_inv * std::sin(theta) * std::sin(psi);
    Om_s = L * Is_inv * std::sin(theta) * std::cos(psi);
    Om_l = L * std::cos(theta);
  } else {
    double Ip = 0.5 * (Ii_inv + Is_inv);
    double Im = 0.5 * (Ii_inv - Is_inv);
    mu[0] = L;
    mu[1] = Ip;
    mu[2] = Im;
  }
  const int n_ode = torque ? 6 : 3;
  // TORQUE2: the torque changes at this fraction of the span (cuda.c:288-289)
  const double t_split = P.v.torque2 ? params[P.types[T.Tt]] * (P.obs[P.n_data - 1].mjd - P.obs[0].mjd) + P.obs[0].mjd
                                     : 0.0;
  for (int i = 0; i < P.n_data; ++i) {
    if (i > 0) {
     double t1 = P.obs[i - 1].mjd, t2 = P.obs[i].mjd;
     // the interval holding t_split runs as two sub-intervals (cuda.c:290-316)
     const int n_split = (P.v.torque2 && t_split >= t1 && t_split < t2) ? 2 : 1;
     for (int isplit = 0; isplit < n_split; ++isplit) {
      if (n_split == 2) {
        if (isplit == 0) {
          t2 = t_split;
        } else {
          t1 = t_split;
          t2 = P.obs[i].mjd;
          mu[3] = params[P.types[T.T2i]];
          mu[4] = params[P.types[T.T2s]];
          mu[5] = params[P.types[T.T2l]];
        }
      }
      int N_steps;
      double h;
      rk4_interval(t1, t2, P.v.time_step, &N_steps, &h);
      double y[6];
      if (torque) { y[0] = Om_i; y[1] = Om_s; y[2] = Om_l; y[3] = phi; y[4] = theta; y[5] = psi; }
      else { y[0] = phi; y[1] = theta; y[2] = psi; }
      for (int l = 0; l < N_steps; ++l) {  // cuda.c:348-372
        double f[6], K1[6], K2[6], K3[6], K4[6];
        ode_literal(torque, y, K1, mu);
        for (int j = 0; j < n_ode; ++j) f[j] = y[j] + 0.5 * h * K1[j];
        ode_literal(torque, f, K2, mu);
        for (int j = 0; j < n_ode; ++j) f[j] = y[j] + 0.5 * h * K2[j];
        ode_literal(torque, f, K3, mu);
        for (int j = 0; j < n_ode; ++j) f[j] = y[j] + h * K3[j];
        ode_literal(torque, f, K4, mu);
        for (int j = 0; j < n_ode; ++j) y[j] = y[j] + 1 / 6.0 * h * (K1[j] + 2 * K2[j] + 2 * K3[j] + K4[j]);
      }
      if (torque) { Om_i = y[0]; Om_s = y[1]; Om_l = y[2]; phi = y[3]; theta = y[4]; psi = y[5]; }
      else { phi = y[0]; theta = y[1]; psi = y[2]; }
     }
    }
    if (out && out->traj) {
      double* t6 = out->traj + 6 * i;
      t6[0] = phi; t6[1] = theta; t6[2] = psi; t6[3] = Om_i; t6[4] = Om_s; t6[5] = Om_l;
    }
    // Body axes (cuda.c:432-473).
    double cos_phi = std::cos(phi), sin_phi = std::sin(phi);
    double N_x = XM_x * cos_phi + YM_x * sin_phi;
    double N_y = YM_y * sin_phi;
    double N_z = XM_z * cos_phi + YM_z * sin_phi;
    double p_x = N_y * M_z - N_z * M_y;
    double p_y = N_z * M_x - N_x * M_z;
    double p_z = N_x * M_y - N_y * M_x;
    double cos_theta = std::cos(theta), sin_theta = std::sin(theta);
    double a_x = M_x * cos_theta + p_x * sin_theta;
    double a_y = M_y * cos_theta + p_y * sin_theta;
    double a_z = M_z * cos_theta + p_z * sin_theta;
    double w_x = a_y * N_z - a_z * N_y;
    double w_y = a_z * N_x - a_x * N_z;
    double w_z = a_x * N_y - a_y * N_x;
    double sin_psi = std::sin(psi), cos_psi = std::cos(psi);
    double b_x = N_x * cos_psi + w_x * sin_psi;
    double b_y = N_y * cos_psi + w_y * sin_psi;
    double b_z = N_z * cos_psi + w_z * sin_psi;
    double c_x = a_y * b_z - a_z * b_y;
    double c_y = a_z * b_x - a_x * b_z;
    double c_z = a_x * b_y - a_y * b_x;
    if (P.v.bw_ball || P.v.rotate) {  // cuda.c:475-497
      double cp = std::cos(params[P.types[T.phi_R]]), sp = std::sin(params[P.types[T.phi_R]]);
      N_x = b_x * cp + c_x * sp;
      N_y = b_y * cp + c_y * sp;
      N_z = b_z * cp + c_z * sp;
      p_x = N_y * a_z - N_z * a_y;
      p_y = N_z * a_x - N_x * a_z;
      p_z = N_x * a_y - N_y * a_x;
      double ctR = std::cos(params[P.types[T.theta_R]]), stR = std::sin(params[P.types[T.theta_R]]);
      a_x = a_x * ctR + p_x * stR;
      a_y = a_y * ctR + p_y * stR;
      a_z = a_z * ctR + p_z * stR;
    }
    if (P.v.rotate) {  // cuda.c:499-516
      w_x = a_y * N_z - a_z * N_y;
      w_y = a_z * N_x - a_x * N_z;
      w_z = a_x * N_y - a_y * N_x;
      double sR = std::sin(params[P.types[T.psi_R]]), cR = std::cos(params[P.types[T.psi_R]]);
      b_x = N_x * cR + w_x * sR;
      b_y = N_y * cR + w_y * sR;
      b_z = N_z * cR + w_z * sR;
      c_x = a_y * b_z - a_z * b_y;
      c_y = a_z * b_x - a_x * b_z;
      c_z = a_x * b_y - a_y * b_x;
    }
    double E[3], S[3];
    if (P.v.interp) interp_literal(P.eph, P.obs[i].mjd, E, S);
    else for (int c = 0; c < 3; ++c) { E[c] = P.obs[i].E[c]; S[c] = P.obs[i].S[c]; }
    double Ep_b = b_x * E[0] + b_y * E[1] + b_z * E[2];
    double Ep_c = c_x * E[0] + c_y * E[1] + c_z * E[2];
    double Ep_a = a_x * E[0] + a_y * E[1] + a_z * E[2];
    double Sp_b = b_x * S[0] + b_y * S[1] + b_z * S[2];
    double Sp_c = c_x * S[0] + c_y * S[1] + c_z * S[2];
    double Sp_a = a_x * S[0] + a_y * S[1] + a_z * S[2];
    double b = P.bc() ? params[P.types[T.b]] : bt;
    double c = P.bc() ? params[P.types[T.c]] : ct;
    double Vmod;
    if (P.v.bw_ball) {  // cuda.c:603-607
      double cos_zeta = Ep_a;
      double kappa = params[P.types[T.kappa]];
      Vmod = -2.5 * std::log10(0.5 * (kappa * (1 + cos_zeta) + 0.5 * (1 - cos_zeta)));
    } else if (P.v.rect) {  // cuda.c:609-661
      double bx = b * b_x, by = b * b_y, bz = b * b_z;
      double cx = c * c_x, cy = c * c_y, cz = c * c_z;
      double ap_x = a_y * Ep_a - a_z * Ep_c, ap_y = a_z * Ep_b - a_x * Ep_a, ap_z = a_x * Ep_c - a_y * Ep_b;
      double bp_x = by * Ep_a - bz * Ep_c, bp_y = bz * Ep_b - bx * Ep_a, bp_z = bx * Ep_c - by * Ep_b;
      double cp_x = cy * Ep_a - cz * Ep_c, cp_y = cz * Ep_b - cx * Ep_a, cp_z = cx * Ep_c - cy * Ep_b;
      double ab_x = ap_y * bp_z - ap_z * bp_y, ab_y = ap_z * bp_x - ap_x * bp_z, ab_z = ap_x * bp_y - ap_y * bp_x;
      double ac_x = ap_y * cp_z - ap_z * cp_y, ac_y = ap_z * cp_x - ap_x * cp_z, ac_z = ap_x * cp_y - ap_y * cp_x;
      double bc_x = bp_y * cp_z - bp_z * cp_y, bc_y = bp_z * cp_x - bp_x * cp_z, bc_z = bp_x * cp_y - bp_y * cp_x;
      double ab = ab_x * ab_x + ab_y * ab_y + ab_z * ab_z;
      double ac = ac_x * ac_x + ac_y * ac_y + ac_z * ac_z;
      double bcs = bc_x * bc_x + bc_y * bc_y + bc_z * bc_z;
      if (ab < 0.0) ab = 0.0;
      if (bcs < 0.0) bcs = 0.0;
      if (ac < 0.0) ac = 0.0;
      Vmod = -2.5 * std::log10(4 * (std::sqrt(ab) + std::sqrt(bcs) + std::sqrt(ac)));
    } else {            // cuda.c:666-689
      double scalar_Sun = std::sqrt(Sp_b * Sp_b / (b * b) + Sp_c * Sp_c / (c * c) + Sp_a * Sp_a);
      double scalar_Earth = std::sqrt(Ep_b * Ep_b / (b * b) + Ep_c * Ep_c / (c * c) + Ep_a * Ep_a);
      double cos_alpha_p = (Sp_b * Ep_b / (b * b) + Sp_c * Ep_c / (c * c) + Sp_a * Ep_a) / (scalar_Sun * scalar_Earth);
      double sin_alpha_p = std::sqrt(1.0 - cos_alpha_p * cos_alpha_p);
      double alpha_p = std::atan2(sin_alpha_p, cos_alpha_p);
      double scalar = std::sqrt(scalar_Sun * scalar_Sun + scalar_Earth * scalar_Earth + 2 * scalar_Sun * scalar_Earth * cos_alpha_p);
      double cos_lambda_p = (scalar_Sun + scalar_Earth * cos_alpha_p) / scalar;
      double sin_lambda_p = scalar_Earth * sin_alpha_p / scalar;
      double lambda_p = std::atan2(sin_lambda_p, cos_lambda_p);
      Vmod = -2.5 * std::log10(b * c * scalar_Sun * scalar_Earth / scalar *
                               (std::cos(lambda_p - alpha_p) + cos_lambda_p +
                                sin_lambda_p * std::sin(lambda_p - alpha_p) *
                                    std::log(1.0 / std::tan(0.5 * lambda_p) / std::tan(0.5 * (alpha_p - lambda_p)))));
    }
    if (P.v.trend) {  // cuda.c:692-697: de-trending by the solar phase angle
      double alpha = std::acos(Sp_b * Ep_b + Sp_c * Ep_c + Sp_a * Ep_a);
      Vmod = Vmod - params[P.types[T.A]] * alpha;
    }
    if (out && out->vmod) out->vmod[i] = (out->nudge && out->nudge->failed) ? std::nan("") : Vmod;
    if (out && out->nudge) out->nudge->point(P, i, P.obs[i].mjd, Vmod);
    int m = P.obs[i].filter;  // cuda.c:708-714
    double y = P.obs[i].V - Vmod;
    sy2[m] = sy2[m] + y * y * P.obs[i].w;
    sy[m] = sy[m] + y * P.obs[i].w;
    sw[m] = sw[m] + P.obs[i].w;
  }
}

// ========================================================== canonical chi2one
// Templated over the scalar R so an operation-counting type can run it.
// ODE right-hand side.  sc = {sin y_theta, cos y_theta, sin y_psi, cos y_psi}
// when the caller already has them (the first RK4 stage reuses those of the
// state), else nullptr.
template <class R>
inline void ode_canon(bool torque, const R* y, R* f, const R* mu, R mu02, const R* sc = nullptr) {
  using astm::p_fma;
  if (torque) {
    R s5, c5, s4, c4;
    if (sc) { s4 = sc[0]; c4 = sc[1]; s5 = sc[2]; c5 = sc[3]; }
    else { astm::sincos_t(y[5], &s5, &c5); astm::sincos_t(y[4], &s4, &c4); }
    f[0] = p_fma(mu[0] * y[1], y[2], mu[3]);
    f[1] = p_fma(mu[1] * y[2], y[0], mu[4]);
    f[2] = p_fma(mu[2] * y[0], y[1], mu[5]);
    f[3] = p_fma(y[1], c5, y[0] * s5) / s4;
    f[4] = p_fma(-y[1], s5, y[0] * c5);
    f[5] = p_fma(-f[3], c4, y[2]);
  } else {
    R s2, c2, s1, c1;
    astm::sincos_t(R(2.0) * y[2], &s2, &c2);
    if (sc) { s1 = sc[0]; c1 = sc[1]; }
    else astm::sincos_t(y[1], &s1, &c1);
    f[0] = mu[0] * p_fma(-mu[2], c2, mu[1]);
    f[1] = (mu02 * s1) * s2;
    f[2] = c1 * (mu[0] - f[0]);
  }
}

template <class R> inline R dot3(R ax, R ay, R az, R bx, R by, R bz) {
  using astm::p_fma;
  return p_fma(az, bz, p_fma(ay, by, ax * bx));
}

template <class R>
inline void model_canon(const Problem& P, const double* params, R* sy2, R* sy, R* sw, ModelOut* out,
                        bool accumulate = false) {
  using astm::p_fma;
  using astm::p_sqrt;
  using astm::p_val;
  const ast_type_ids& T = P.T;
  const bool torque = P.tq();
  if (!accumulate)
    for (int m = 0; m < P.n_filters; ++m) sy2[m] = sy[m] = sw[m] = R(0.0);
  R sTM, cTM, sPM, cPM;
  astm::sincos_t(R(params[P.types[T.theta_M]]), &sTM, &cTM);
  astm::sincos_t(R(params[P.types[T.phi_M]]), &sPM, &cPM);
  R M_x = sTM * cPM, M_y = sTM * sPM, M_z = cTM;
  R XM = p_sqrt(p_fma(M_x, M_x, M_z * M_z));
  R XM_x = M_z / XM;
  R XM_z = -M_x / XM;
  R YM_x = M_y * XM_z;
  R YM_y = p_fma(-M_x, XM_z, M_z * XM_x);
  R YM_z = -(M_y * XM_x);
  R bt = R(params[P.types[T.b_tumb]]), ct = R(params[P.types[T.c_tumb]]);
  R den = p_fma(ct, ct, bt * bt);
  R Is = p_fma(bt, bt, R(1.0)) / den;
  R Ii = p_fma(ct, ct, R(1.0)) / den;
  R Is_inv = R(1.0) / Is;
  R Ii_inv = R(1.0) / Ii;
  R Es = R(params[P.types[T.Es]]), psi0 = R(params[P.types[T.psi_0]]), L = R(params[P.types[T.L]]);
  R sP0, cP0;
  astm::sincos_t(psi0, &sP0, &cP0);
  R theta = astm::asin(p_sqrt((Es - R(1.0)) / (p_fma(sP0 * sP0, Ii_inv - Is_inv, Is_inv) - R(1.0))));
  R phi = R(params[P.types[T.phi_0]]);
  R psi = psi0;
  R mu[6];
  R mu02 = R(0.0);
  R Om_i = R(0.0), Om_s = R(0.0), Om_l = R(0.0);
  // sin/cos of the current theta and psi, shared by the first RK4 stage of
  // the next step and by the body axes at the observation.
  R sc[4];
  astm::sincos_t(theta, &sc[0], &sc[1]);
  sc[2] = sP0;
  sc[3] = cP0;
  if (torque) {
    mu[0] = (Is - R(1.0)) * Ii_inv;
    mu[1] = (R(1.0) - Ii) * Is_inv;
    mu[2] = Ii - Is;
    mu[3] = R(params[P.types[T.Ti]]);
    mu[4] = R(params[P.types[T.Ts]]);
    mu[5] = R(params[P.types[T.Tl]]);
    Om_i = ((L * Ii_inv) * sc[0]) * sP0;
    Om_s = ((L * Is_inv) * sc[0]) * cP0;
    Om_l = L * sc[1];
  } else {
    mu[0] = L;
    mu[1] = R(0.5) * (Ii_inv + Is_inv);
    mu[2] = R(0.5) * (Ii_inv - Is_inv);
    mu02 = mu[0] * mu[2];
  }
  // Photometric shape (LS) or ball constants (BW_BALL).
  R ib = R(0.0), ic = R(0.0), bcp = R(0.0), kappa = R(0.0);
  R sR = R(0.0), cR = R(0.0), stR = R(0.0), ctR = R(0.0);
  R sPR = R(0.0), cPR = R(0.0);
  if (P.v.rotate) {
    astm::sincos_t(R(params[P.types[T.phi_R]]), &sR, &cR);
    astm::sincos_t(R(params[P.types[T.theta_R]]), &stR, &ctR);
    astm::sincos_t(R(params[P.types[T.psi_R]]), &sPR, &cPR);
  }
  if (P.v.bw_ball) {
    astm::sincos_t(R(params[P.types[T.phi_R]]), &sR, &cR);
    astm::sincos_t(R(params[P.types[T.theta_R]]), &stR, &ctR);
    kappa = R(params[P.types[T.kappa]]);
  } else {
    R b = P.bc() ? R(params[P.types[T.b]]) : bt;
    R c = P.bc() ? R(params[P.types[T.c]]) : ct;
    ib = R(1.0) / b;
    ic = R(1.0) / c;
    bcp = b * c;
  }
  const int n_ode = torque ? 6 : 3;
  // TORQUE2: split time (cuda.c:288-289); the interval holding it runs as two
  // sub-intervals, the second and every later one under the second torque.
  const double t_split = P.v.torque2 ? std::fma(params[P.types[T.Tt]], P.obs[P.n_data - 1].mjd - P.obs[0].mjd,
                                                P.obs[0].mjd)
                                     : 0.0;
  for (int i = 0; i < P.n_data; ++i) {
    if (i > 0) {
     double t1 = P.obs[i - 1].mjd, t2 = P.obs[i].mjd;
     const int n_split = (P.v.torque2 && t_split >= t1 && t_split < t2) ? 2 : 1;
     for (int isplit = 0; isplit < n_split; ++isplit) {
      if (n_split == 2) {
        if (isplit == 0) {
          t2 = t_split;
        } else {
          t1 = t_split;
          t2 = P.obs[i].mjd;
          mu[3] = R(params[P.types[T.T2i]]);
          mu[4] = R(params[P.types[T.T2s]]);
          mu[5] = R(params[P.types[T.T2l]]);
        }
      }
      int N_steps;
      double hd;
      rk4_interval(t1, t2, P.v.time_step, &N_steps, &hd);
      R h = R(hd), hh = R(0.5 * hd), h6 = R(1 / 6.0 * hd);
      R y[6];
      if (torque) { y[0] = Om_i; y[1] = Om_s; y[2] = Om_l; y[3] = phi; y[4] = theta; y[5] = psi; }
      else { y[0] = phi; y[1] = theta; y[2] = psi; }
      for (int l = 0; l < N_steps; ++l) {
        R f[6], K[6], s[6];
        ode_canon(torque, y, K, mu, mu02, sc);
        for (int j = 0; j < n_ode; ++j) { s[j] = K[j]; f[j] = p_fma(hh, K[j], y[j]); }
        ode_canon(torque, f, K, mu, mu02);
        for (int j = 0; j < n_ode; ++j) { s[j] = p_fma(R(2.0), K[j], s[j]); f[j] = p_fma(hh, K[j], y[j]); }
        ode_canon(torque, f, K, mu, mu02);
        for (int j = 0; j < n_ode; ++j) { s[j] = p_fma(R(2.0), K[j], s[j]); f[j] = p_fma(h, K[j], y[j]); }
        ode_canon(torque, f, K, mu, mu02);
        for (int j = 0; j < n_ode; ++j) y[j] = p_fma(h6, s[j] + K[j], y[j]);
        if (torque) {
          astm::sincos_t(y[4], &sc[0], &sc[1]);
          astm::sincos_t(y[5], &sc[2], &sc[3]);
        } else {
          astm::sincos_t(y[1], &sc[0], &sc[1]);
        }
      }
      if (torque) { Om_i = y[0]; Om_s = y[1]; Om_l = y[2]; phi = y[3]; theta = y[4]; psi = y[5]; }
      else { phi = y[0]; theta = y[1]; psi = y[2]; }
     }
    }
    if (out && out->traj) {
      double* t6 = out->traj + 6 * i;
      t6[0] = p_val(phi); t6[1] = p_val(theta); t6[2] = p_val(psi);
      t6[3] = p_val(Om_i); t6[4] = p_val(Om_s); t6[5] = p_val(Om_l);
    }
    R sf, cf, sp, cp;
    astm::sincos_t(phi, &sf, &cf);
    const R st = sc[0], ctt = sc[1];
    if (torque) { sp = sc[2]; cp = sc[3]; }
    else astm::sincos_t(psi, &sp, &cp);
    R N_x = p_fma(YM_x, sf, XM_x * cf);
    R N_y = YM_y * sf;
    R N_z = p_fma(YM_z, sf, XM_z * cf);
    R p_x = p_fma(-N_z, M_y, N_y * M_z);
    R p_y = p_fma(-N_x, M_z, N_z * M_x);
    R p_z = p_fma(-N_y, M_x, N_x * M_y);
    R a_x = p_fma(p_x, st, M_x * ctt);
    R a_y = p_fma(p_y, st, M_y * ctt);
    R a_z = p_fma(p_z, st, M_z * ctt);
    R w_x = p_fma(-a_z, N_y, a_y * N_z);
    R w_y = p_fma(-a_x, N_z, a_z * N_x);
    R w_z = p_fma(-a_y, N_x, a_x * N_y);
    R b_x = p_fma(w_x, sp, N_x * cp);
    R b_y = p_fma(w_y, sp, N_y * cp);
    R b_z = p_fma(w_z, sp, N_z * cp);
    R c_x = p_fma(-a_z, b_y, a_y * b_z);
    R c_y = p_fma(-a_x, b_z, a_z * b_x);
    R c_z = p_fma(-a_y, b_x, a_x * b_y);
    if (P.v.rotate) {  // cuda.c:475-516, the body-axes construction on (b, c, a)
      R n_x = p_fma(c_x, sR, b_x * cR);
      R n_y = p_fma(c_y, sR, b_y * cR);
      R n_z = p_fma(c_z, sR, b_z * cR);
      R q_x = p_fma(-n_z, a_y, n_y * a_z);
      R q_y = p_fma(-n_x, a_z, n_z * a_x);
      R q_z = p_fma(-n_y, a_x, n_x * a_y);
      a_x = p_fma(q_x, stR, a_x * ctR);
      a_y = p_fma(q_y, stR, a_y * ctR);
      a_z = p_fma(q_z, stR, a_z * ctR);
      R v_x = p_fma(-a_z, n_y, a_y * n_z);
      R v_y = p_fma(-a_x, n_z, a_z * n_x);
      R v_z = p_fma(-a_y, n_x, a_x * n_y);
      b_x = p_fma(v_x, sPR, n_x * cPR);
      b_y = p_fma(v_y, sPR, n_y * cPR);
      b_z = p_fma(v_z, sPR, n_z * cPR);
      c_x = p_fma(-a_z, b_y, a_y * b_z);
      c_y = p_fma(-a_x, b_z, a_z * b_x);
      c_z = p_fma(-a_y, b_x, a_x * b_y);
    }
    double Ed[3], Sd[3];
    if (P.v.interp) interp_canon(P.eph, P.obs[i].mjd, Ed, Sd);
    else for (int c = 0; c < 3; ++c) { Ed[c] = P.obs[i].E[c]; Sd[c] = P.obs[i].S[c]; }
    R E0 = R(Ed[0]), E1 = R(Ed[1]), E2 = R(Ed[2]);
    R Vmod;
    if (P.v.bw_ball) {
      R n_x = p_fma(c_x, sR, b_x * cR);
      R n_y = p_fma(c_y, sR, b_y * cR);
      R n_z = p_fma(c_z, sR, b_z * cR);
      R q_x = p_fma(-n_z, a_y, n_y * a_z);
      R q_y = p_fma(-n_x, a_z, n_z * a_x);
      R q_z = p_fma(-n_y, a_x, n_x * a_y);
      R r_x = p_fma(q_x, stR, a_x * ctR);
      R r_y = p_fma(q_y, stR, a_y * ctR);
      R r_z = p_fma(q_z, stR, a_z * ctR);
      R cz = dot3(r_x, r_y, r_z, E0, E1, E2);
      Vmod = R(-2.5) * astm::log10(R(0.5) * p_fma(kappa, R(1.0) + cz, R(0.5) * (R(1.0) - cz)));
    } else if (P.v.rect) {  // cuda.c:609-661
      R bL = R(params[P.types[T.b]]), cL = R(params[P.types[T.c]]);
      R Eb = dot3(b_x, b_y, b_z, E0, E1, E2);
      R Ec = dot3(c_x, c_y, c_z, E0, E1, E2);
      R Ea = dot3(a_x, a_y, a_z, E0, E1, E2);
      R bx = bL * b_x, by = bL * b_y, bz = bL * b_z;
      R cx = cL * c_x, cy = cL * c_y, cz = cL * c_z;
      R ap_x = p_fma(-a_z, Ec, a_y * Ea), ap_y = p_fma(-a_x, Ea, a_z * Eb), ap_z = p_fma(-a_y, Eb, a_x * Ec);
      R bp_x = p_fma(-bz, Ec, by * Ea), bp_y = p_fma(-bx, Ea, bz * Eb), bp_z = p_fma(-by, Eb, bx * Ec);
      R cp_x = p_fma(-cz, Ec, cy * Ea), cp_y = p_fma(-cx, Ea, cz * Eb), cp_z = p_fma(-cy, Eb, cx * Ec);
      R ab_x = p_fma(-ap_z, bp_y, ap_y * bp_z), ab_y = p_fma(-ap_x, bp_z, ap_z * bp_x),
        ab_z = p_fma(-ap_y, bp_x, ap_x * bp_y);
      R ac_x = p_fma(-ap_z, cp_y, ap_y * cp_z), ac_y = p_fma(-ap_x, cp_z, ap_z * cp_x),
        ac_z = p_fma(-ap_y, cp_x, ap_x * cp_y);
      R bc_x = p_fma(-bp_z, cp_y, bp_y * cp_z), bc_y = p_fma(-bp_x, cp_z, bp_z * cp_x),
        bc_z = p_fma(-bp_y, cp_x, bp_x * cp_y);
      R sab = dot3(ab_x, ab_y, ab_z, ab_x, ab_y, ab_z);
      R sac = dot3(ac_x, ac_y, ac_z, ac_x, ac_y, ac_z);
      R sbc = dot3(bc_x, bc_y, bc_z, bc_x, bc_y, bc_z);
      if (p_val(sab) < 0.0) sab = R(0.0);
      if (p_val(sbc) < 0.0) sbc = R(0.0);
      if (p_val(sac) < 0.0) sac = R(0.0);
      Vmod = R(-2.5) * astm::log10(R(4.0) * ((p_sqrt(sab) + p_sqrt(sbc)) + p_sqrt(sac)));
    } else {
      R S0 = R(Sd[0]), S1 = R(Sd[1]), S2 = R(Sd[2]);
      R bs_x = b_x * ib, bs_y = b_y * ib, bs_z = b_z * ib;
      R cs_x = c_x * ic, cs_y = c_y * ic, cs_z = c_z * ic;
      R uSb = dot3(bs_x, bs_y, bs_z, S0, S1, S2);
      R uSc = dot3(cs_x, cs_y, cs_z, S0, S1, S2);
      R uSa = dot3(a_x, a_y, a_z, S0, S1, S2);
      R uEb = dot3(bs_x, bs_y, bs_z, E0, E1, E2);
      R uEc = dot3(cs_x, cs_y, cs_z, E0, E1, E2);
      R uEa = dot3(a_x, a_y, a_z, E0, E1, E2);
      R sS2 = dot3(uSb, uSc, uSa, uSb, uSc, uSa);
      R sE2 = dot3(uEb, uEc, uEa, uEb, uEc, uEa);
      R sS = p_sqrt(sS2);
      R sE = p_sqrt(sE2);
      R ca = dot3(uSb, uSc, uSa, uEb, uEc, uEa) / (sS * sE);
      R sa = p_sqrt(p_fma(-ca, ca, R(1.0)));
      R sc = p_sqrt(p_fma((R(2.0) * sS) * sE, ca, sS2 + sE2));
      R isc = R(1.0) / sc;
      R cl = p_fma(sE, ca, sS) * isc;
      R sl = (sE * sa) * isc;
      R cd = p_fma(cl, ca, sl * sa);     // cos(lambda' - alpha')
      R sd = p_fma(-cl, sa, sl * ca);    // sin(lambda' - alpha')
      R lg = astm::log(((R(1.0) + cl) * (R(1.0) + cd)) / (sl * (-sd)));
      R inner = p_fma(sl * sd, lg, cd + cl);
      Vmod = R(-2.5) * astm::log10((((bcp * sS) * sE) * isc) * inner);
    }
    if (P.v.trend) {  // cuda.c:692-697: acos(S.E) in the inertial frame (a property of the point)
      const double alpha = astm::acos(std::fma(Sd[2], Ed[2], std::fma(Sd[1], Ed[1], Sd[0] * Ed[0])));
      Vmod = p_fma(-R(params[P.types[T.A]]), R(alpha), Vmod);
    }
    if (out && out->vmod) out->vmod[i] = (out->nudge && out->nudge->failed) ? std::nan("") : p_val(Vmod);
    if (out && out->nudge) out->nudge->point(P, i, P.obs[i].mjd, p_val(Vmod));
    int m = P.obs[i].filter;
    R w = R((double)P.obs[i].w);
    R y = R((double)P.obs[i].V) - Vmod;
    sy2[m] = p_fma(y * y, w, sy2[m]);
    sy[m] = p_fma(y, w, sy[m]);
    sw[m] = sw[m] + w;
  }
}

// chi^2 and delta_V from the per-filter sums (cuda.c:884-910).
template <class CF>
inline CF chi2_from_sums(const Problem& P, const double* sy2, const double* sy, const double* sw, CF* dV) {
  CF chi2a = 0.0;
  for (int m = 0; m < P.n_filters; ++m) {
    CF chi2m = sy2[m] - sy[m] * sy[m] / sw[m];
    chi2a = chi2a + chi2m;
    dV[m] = sy[m] / sw[m];
  }
  chi2a = chi2a / (P.n_data - P.n_params - P.n_filters);
  return chi2a;
}

// The model sums in either form.  Under SEGMENT (cuda.c:162-813) each
// segment is an integration of its own from its first observation, with the
// segment's parameters (params[Types[type][iseg]]), and the sums run on over
// all segments in observation order; observations before the first segment
// are not visited.  The segment's sub-problem is the plain model on its
// observations with its parameters in the one-segment layout.
inline void model_sums(const Problem& P, int form, const double* params, double* sy2, double* sy, double* sw,
                       ModelOut* out) {
  if (P.n_seg <= 1) {
    if (form == 0) model_literal(P, params, sy2, sy, sw, out);
    else model_canon<double>(P, params, sy2, sy, sw, out);
    return;
  }
  for (int m = 0; m < P.n_filters; ++m) sy2[m] = sy[m] = sw[m] = 0.0;
  for (int iseg = 0; iseg < P.n_seg; ++iseg) {
    const int i1 = P.seg_start[iseg];
    const int i2 = iseg < P.n_seg - 1 ? P.seg_start[iseg + 1] : P.n_data;  // cuda.c:260-268
    Problem Q = P;
    Q.n_seg = 0;
    Q.obs = P.obs + i1;
    Q.n_data = i2 - i1;
    double ps[AST_MAX_PARAMS] = {0};
    for (int t = 0; t < P.n_types; ++t)
      if (P.types[t] >= 0) ps[P.types[t]] = params[P.seg_types[t][iseg]];
    ModelOut o{nullptr, nullptr, nullptr};
    if (out) {
      o = *out;
      if (o.vmod) o.vmod += i1;
      if (o.traj) o.traj += 6 * (size_t)i1;
    }
    if (form == 0) model_literal(Q, ps, sy2, sy, sw, &o, true);
    else model_canon<double>(Q, ps, sy2, sy, sw, &o, true);
  }
}

// chi^2 of one model; under NUDGE a failed model returns 1e30 with delta_V
// untouched (cuda.c:737), else the nudged chi^2.
template <class CF>
inline CF chi2_eval(const Problem& P, int form, const double* params, CF* dV, ModelOut* out = nullptr) {
  double sy2[AST_MAX_FILTERS], sy[AST_MAX_FILTERS], sw[AST_MAX_FILTERS];
  Nudge nu;
  ModelOut o{nullptr, nullptr, nullptr};
  if (out) o = *out;
  if (P.v.nudge) o.nudge = &nu;
  model_sums(P, form, params, sy2, sy, sw, &o);
  if (P.v.nudge && nu.failed) return CF(1e30);
  CF c = chi2_from_sums<CF>(P, sy2, sy, sw, dV);
  if (P.v.nudge) c = nu.merit<CF>(P, form, c, dV);
  return c;
}

// ================================================================ x2params
inline bool hard_left(int per, int LAM) { return per == AST_HARD_BOTH || per == AST_HARD_LEFT || (LAM == 0 && per == AST_PERIODIC_LAM); }
inline bool hard_right(int per, int LAM) { return per == AST_HARD_BOTH || per == AST_HARD_RIGHT || (LAM == 0 && per == AST_PERIODIC_LAM); }

template <class CF, class M>
int x2params(const Problem& P, const CF* x, double* params, int reopt) {
  const ast_type_ids& T = P.T;
  const int N = P.n_params;
  int LAM = 0;
  for (int i = 0; i < N; ++i) {  // cuda.c:1128-1147
    if (P.prop[i][AST_P_TYPE] == T.Es) LAM = x[i] >= 0.5;
    int per = P.prop[i][AST_P_PERIODIC];
    if ((x[i] < 0.0 && hard_left(per, LAM)) || (x[i] > 1.0 && hard_right(per, LAM))) return 1;
  }
  double log_c_tumb = 0, log_b_tumb = 0, Is = 0, Ii = 0, psi_min = 0, psi_max = 0, log_c = 0;
  for (int i = 0; i < N; ++i) {  // cuda.c:1155-1314
    int t = P.prop[i][AST_P_TYPE];
    CF L0 = (CF)P.lim[0][t], L1 = (CF)P.lim[1][t];
    if (t == T.b_tumb) {
      log_b_tumb = log_c_tumb * (x[i] * (L1 - L0) + L0);
      params[i] = M::exp(log_b_tumb);
      double b_tumb = params[i];
      double c_tumb = params[P.types[T.c_tumb]];
      Is = (1.0 + b_tumb * b_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
      Ii = (1.0 + c_tumb * c_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
    } else if (t == T.Es) {
      if (LAM) {
        params[i] = 2.0 * (x[i] - 0.5) * (1.0 - 1.0 / Ii) + 1.0 / Ii;
        psi_min = 0.0;
        psi_max = 2.0 * PI;
      } else {
        params[i] = 2.0 * x[i] * (1.0 / Ii - 1.0 / Is) + 1.0 / Is;
        psi_max = M::atan(M::sqrt(Ii * (Is - 1.0 / params[i]) / Is / (1.0 / params[i] - Ii)));
        psi_min = -psi_max;
      }
    } else if (t == T.psi_0) {
      params[i] = x[i] * (psi_max - psi_min) + psi_min;
    } else if (P.bc() && t == T.b) {
      double log_b = log_c * (x[i] * (L1 - L0) + L0);
      if (std::fabs(log_b - log_b_tumb) > BC_DEV_MAX) return 1;
      params[i] = M::exp(log_b);
    } else if (P.prop[i][AST_P_PERIODIC] == AST_PERIODIC) {
      params[i] = x[i] * 2.0 * PI;
    } else if (P.prop[i][AST_P_INDEPENDENT] == 1) {
      if (!P.v.p_psi || t != T.L || reopt) params[i] = x[i] * (L1 - L0) + L0;
      if (t == T.c_tumb) {
        log_c_tumb = params[i];
        params[i] = M::exp(log_c_tumb);
      } else if (P.bc() && t == T.c) {
        log_c = params[i];
        if (std::fabs(log_c - log_c_tumb) > BC_DEV_MAX) return 1;
        params[i] = M::exp(log_c);
      } else if (P.v.bw_ball && t == T.kappa) {
        params[i] = M::exp(params[i]);
      } else if (t == T.L && !reopt && P.v.p_psi) {
        float dP = P.x2.ppsi2 - P.x2.ppsi1;
        params[i] = x[i] * (CF)dP + (CF)P.x2.ppsi1;
        // the Es of L's own segment, params[sTypes[T_Es][iseg]] (cuda.c:1263)
        double EsV = params[P.n_seg > 1 ? P.seg_types[T.Es][P.prop[i][AST_P_ISEG]] : P.types[T.Es]];
        double Einv = 1.0 / EsV;
        double k2;
        if (LAM) k2 = (Is - Ii) * (Einv - 1.0) / ((Ii - 1.0) * (Is - Einv));
        else k2 = (Ii - 1.0) * (Is - Einv) / ((Is - Ii) * (Einv - 1.0));
        double a = 1.0, g = M::sqrt(1.0 - k2), a1, g1;
        for (int ii = 0; ii < 5; ++ii) {
          a1 = 0.5 * (a + g);
          g1 = M::sqrt(a * g);
          a = a1;
          g = g1;
        }
        if (LAM) params[i] = 4.0 * params[i] * PI / (a + g) * M::sqrt(Ii * Is / (EsV * (Ii - 1.0) * (Is - Einv)));
        else params[i] = 4.0 * params[i] * PI / (a + g) * M::sqrt(Ii * Is / (EsV * (Is - Ii) * (Einv - 1.0)));
      }
    }
  }
  return 0;
}

// ================================================================ params2x
template <class CF, class M>
void params2x(const Problem& P, CF* x, const double* params) {
  const ast_type_ids& T = P.T;
  const int N = P.n_params;
  int LAM = 0, i_Es = 0;
  double b_tumb = params[P.types[T.b_tumb]];
  double c_tumb = params[P.types[T.c_tumb]];
  double Is = (1.0 + b_tumb * b_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
  double Ii = (1.0 + c_tumb * c_tumb) / (b_tumb * b_tumb + c_tumb * c_tumb);
  for (int i = 0; i < N; ++i) {
    int t = P.prop[i][AST_P_TYPE];
    if (P.prop[i][AST_P_FROZEN] != 0) { x[i] = 0; continue; }
    CF L0 = (CF)P.lim[0][t], L1 = (CF)P.lim[1][t];
    if (P.prop[i][AST_P_INDEPENDENT] == 1) {
      double par = params[i];
      if (P.prop[i][AST_P_PERIODIC] == 1) {
        x[i] = par / (2 * PI);
        x[i] = x[i] - cf_floor(x[i]);
      } else {
        if (P.bc() && t == T.c) {
          CF C0 = (CF)P.lim[0][T.c_tumb], C1 = (CF)P.lim[1][T.c_tumb];
          x[i] = (M::log(params[i]) - C0) / (C1 - C0);
          continue;
        }
        if (t == T.c_tumb) par = M::log(par);
        if (P.v.bw_ball && t == T.kappa) par = M::log(par);
        x[i] = (par - L0) / (L1 - L0);
      }
    } else {
      if (t == T.b_tumb) {
        double par = M::log(b_tumb) / M::log(c_tumb);
        x[i] = (par - L0) / (L1 - L0);
      } else if (t == T.Es) {
        i_Es = i;
        LAM = params[i] > 1.0 / Ii;
        if (LAM) x[i] = 0.5 * ((params[i] - 1.0 / Ii) / (1.0 - 1.0 / Ii) + 1.0);
        else x[i] = 0.5 * (params[i] - 1.0 / Is) / (1.0 / Ii - 1.0 / Is);
      } else if (t == T.psi_0) {
        double psi_min, psi_max;
        if (LAM) { psi_min = 0.0; psi_max = 2.0 * PI; }
        else {
          psi_max = M::atan(M::sqrt(Ii * (Is - 1.0 / params[i_Es]) / Is / (1.0 / params[i_Es] - Ii)));
          psi_min = -psi_max;
        }
        x[i] = (params[i] - psi_min) / (psi_max - psi_min);
      } else if (P.bc() && t == T.b) {
        double par = M::log(params[i]) / M::log(params[P.types[T.c]]);
        x[i] = (par - L0) / (L1 - L0);
      }
    }
  }
}

// ================================================================ simplex
struct RunCfg {
  int n_groups, group_size, nstages, n_steps;
  int64_t eval_budget;
  double size_min;
  uint64_t seed;
};

template <class CF> struct SimplexConst {
  static CF DX_INI() { return (CF)0.001; }   // asteroid.h:205
  static CF D2X_INI() { return (CF)-1.2; }   // asteroid.h:206
  static CF DX_RAND() { return (CF)0.0001; } // asteroid.h:209
};

template <class CF> struct Lane {
  CF x[AST_MAX_PARAMS + 1][AST_MAX_PARAMS];
  CF f[AST_MAX_PARAMS + 1];
  int ind[AST_MAX_PARAMS + 1];
  CF dV[AST_MAX_FILTERS];
  double params[AST_MAX_PARAMS];  // the thread's params[] buffer (cuda.c:1344), last x2params output
  uint32_t rng;      // draws taken
  int64_t nev, steps;
  CF s_f;
};

// One stage of one simplex (cuda.c:1435-1742).  Returns s_f.
template <class CF, class M>
CF run_stage(const Problem& P, int form, const RunCfg& C, int reopt, uint64_t fit, Lane<CF>& S) {
  const ast_type_ids& T = P.T;
  const int N = P.n_params;
  const CF DX_INI = SimplexConst<CF>::DX_INI(), D2X_INI = SimplexConst<CF>::D2X_INI(), DX_RAND = SimplexConst<CF>::DX_RAND();
  const CF ALPHA = 1.0, GAMMA = 2.0, RHO = 0.5, SIGMA = 0.5;  // asteroid.h:239-242
  const CF SIZE2_MIN = (CF)C.size_min * (CF)C.size_min;      // asteroid.h:245
  double* params = S.params;
  const int64_t nev0 = S.nev;
  auto rnd = [&]() { return astrng::uniform01(C.seed, fit, S.rng++); };
  auto eval = [&](const CF* xv, CF* fv) -> bool {  // x2params + chi2one
    if (x2params<CF, M>(P, xv, params, reopt)) return false;
    *fv = chi2_eval<CF>(P, form, params, S.dV);
    S.nev++;
    return true;
  };
  int l = 0;
  int LAM = 0;
  for (int i = 0; i < N; ++i) {  // cuda.c:1459-1508
    float r = rnd();
    int t = P.prop[i][AST_P_TYPE];
    if (P.bc() && !reopt) {
      if (t == T.c) { S.x[0][i] = S.x[0][P.types[T.c_tumb]]; continue; }
      if (t == T.b) { S.x[0][i] = S.x[0][P.types[T.b_tumb]]; continue; }
    }
    int per = P.prop[i][AST_P_PERIODIC];
    if (!reopt || (reopt && P.prop[i][AST_P_FROZEN] == -1)) {
      S.x[0][i] = DX_INI + SMALL + r * (1.0 - 2 * (SMALL + DX_INI));
    } else {
      CF xmin = S.x[0][i] - DX_RAND;
      CF xmax = S.x[0][i] + DX_RAND;
      if (xmin < SMALL && hard_left(per, LAM)) xmin = SMALL;
      if (xmax > 1.0 - SMALL && hard_right(per, LAM)) xmax = 1.0 - SMALL;
      S.x[0][i] = xmin + r * (xmax - xmin);
    }
    if (t == T.Es) LAM = S.x[0][i] >= 0.5;
  }
  for (int j = 1; j < N + 1; ++j) {  // cuda.c:1512-1545
    for (int i = 0; i < N; ++i) {
      if (i == j - 1) {
        int per = P.prop[i][AST_P_PERIODIC];
        float d2x = rnd();
        CF dx = DX_INI * cf_exp<M>((CF)(D2X_INI * d2x));
        if (rnd() < 0.5) {
          dx = -dx;
          if (S.x[0][i] + dx < SMALL && hard_left(per, LAM)) dx = -dx;
        } else {
          if (S.x[0][i] + dx > 1.0 - SMALL && hard_right(per, LAM)) dx = -dx;
        }
        S.x[j][i] = S.x[0][i] + dx;
      } else {
        S.x[j][i] = S.x[0][i];
      }
    }
  }
  bool failed = false;
  for (int j = 0; j < N + 1; ++j) {  // cuda.c:1549-1559
    if (!eval(S.x[j], &S.f[j])) { failed = true; break; }
  }
  while (true) {  // cuda.c:1568-1735
    if (failed) break;
    l++;
    bool ind2[AST_MAX_PARAMS + 1];
    for (int j = 0; j < N + 1; ++j) ind2[j] = false;
    for (int j = 0; j < N + 1; ++j) {
      CF fmin = 1e30;
      int jmin = -1;
      for (int j2 = 0; j2 < N + 1; ++j2)
        if (!ind2[j2] && S.f[j2] <= fmin) { fmin = S.f[j2]; jmin = j2; }
      if (jmin < 0) { S.ind[0] = 0; S.f[S.ind[0]] = 1e30; break; }
      S.ind[j] = jmin;
      ind2[jmin] = true;
    }
    CF x0[AST_MAX_PARAMS];
    for (int i = 0; i < N; ++i) {
      CF sum = 0.0;
      for (int j = 0; j < N + 1; ++j) sum = sum + S.x[j][i];
      x0[i] = sum / (N + 1);
    }
    CF size2 = 0.0;
    for (int j = 0; j < N + 1; ++j) {
      CF sum = 0.0;
      for (int i = 0; i < N; ++i) {
        CF dx = S.x[j][i] - x0[i];
        sum = sum + dx * dx;
      }
      size2 = size2 + sum;
    }
    size2 = size2 / N;
    if (size2 < SIZE2_MIN) break;
    if (l > C.n_steps) break;
    if (C.eval_budget > 0 && S.nev - nev0 >= C.eval_budget) break;
    const int w = S.ind[N];
    CF x_r[AST_MAX_PARAMS];
    for (int i = 0; i < N; ++i)
      x_r[i] = (P.prop[i][AST_P_FROZEN] != 1) ? (CF)(x0[i] + ALPHA * (x0[i] - S.x[w][i])) : S.x[w][i];
    CF f_r;
    if (!eval(x_r, &f_r)) f_r = 1e30;
    if (f_r >= S.f[S.ind[0]] && f_r < S.f[S.ind[N - 1]]) {
      for (int i = 0; i < N; ++i) S.x[w][i] = x_r[i];
      S.f[w] = f_r;
      continue;
    }
    if (f_r < S.f[S.ind[0]]) {
      CF x_e[AST_MAX_PARAMS];
      for (int i = 0; i < N; ++i)
        x_e[i] = (P.prop[i][AST_P_FROZEN] != 1) ? (CF)(x0[i] + GAMMA * (x_r[i] - x0[i])) : x_r[i];
      CF f_e;
      if (!eval(x_e, &f_e)) f_e = 1e30;
      if (f_e < f_r) {
        for (int i = 0; i < N; ++i) S.x[w][i] = x_e[i];
        S.f[w] = f_e;
      } else {
        for (int i = 0; i < N; ++i) S.x[w][i] = x_r[i];
        S.f[w] = f_r;
      }
      continue;
    }
    for (int i = 0; i < N; ++i)
      x_r[i] = (P.prop[i][AST_P_FROZEN] != 1) ? (CF)(x0[i] + RHO * (S.x[w][i] - x0[i])) : S.x[w][i];
    if (!eval(x_r, &f_r)) f_r = 1e30;
    if (f_r < S.f[w]) {
      for (int i = 0; i < N; ++i) S.x[w][i] = x_r[i];
      S.f[w] = f_r;
      continue;
    }
    bool bad = false;
    for (int j = 1; j < N + 1; ++j) {
      const int v = S.ind[j], b = S.ind[0];
      for (int i = 0; i < N; ++i)
        if (P.prop[i][AST_P_FROZEN] != 1) S.x[v][i] = S.x[b][i] + SIGMA * (S.x[v][i] - S.x[b][i]);
      CF fv;
      if (!eval(S.x[v], &fv)) bad = true;
      else S.f[v] = fv;
    }
    if (bad) { failed = true; break; }
  }
  S.steps += l;
  S.s_f = failed ? (CF)1e30 : S.f[S.ind[0]];
  return S.s_f;
}

// One block of chi2_gpu: group g of C.group_size simplexes.
template <class CF, class M>
void run_group(const Problem& P, int form, const RunCfg& C, int reopt_in, const double* params0, int g,
               uint32_t* rng, double* d_f, double* d_params, double* d_dV,
               double* lane_f, int64_t* lane_nev, int64_t* lane_steps) {
  const int G = C.group_size, N = P.n_params;
  std::vector<Lane<CF>> lanes(G);
  int reopt = reopt_in;
  for (int t = 0; t < G; ++t) {
    Lane<CF>& S = lanes[t];
    std::memset(&S, 0, sizeof(S));
    for (int j = 0; j <= N; ++j) S.ind[j] = j;
    S.rng = rng[(size_t)g * G + t];
    if (reopt) {  // cuda.c:1409-1416
      for (int i = 0; i < N; ++i) S.params[i] = params0[i];
      params2x<CF, M>(P, S.x[0], S.params);
    }
  }
  CF s_x0[AST_MAX_PARAMS];
  int thread_min = 0;
  CF smin = (CF)1e30;
  for (int istage = 0; istage < C.nstages; ++istage) {
    if (C.nstages > 1 && istage > 0)
      for (int t = 0; t < G; ++t)
        for (int i = 0; i < N; ++i) lanes[t].x[0][i] = s_x0[i];
    for (int t = 0; t < G; ++t) run_stage<CF, M>(P, form, C, reopt, (uint64_t)g * G + t, lanes[t]);
    thread_min = 0;  // cuda.c:1770-1782
    smin = (CF)1e30;
    for (int j = 0; j < G; ++j)
      if (lanes[j].s_f < smin) { smin = lanes[j].s_f; thread_min = j; }
    if (C.nstages > 1 && istage < C.nstages - 1) {  // cuda.c:1785-1805
      Lane<CF>& S = lanes[thread_min];
      CF* xb = S.x[S.ind[0]];
      if (P.v.p_psi && reopt == 0) {
        x2params<CF, M>(P, xb, S.params, reopt);
        params2x<CF, M>(P, xb, S.params);
      }
      for (int i = 0; i < N; ++i) s_x0[i] = xb[i];
      reopt = 1;
    }
  }
  if (smin < (CF)d_f[g]) {  // cuda.c:1810-1823
    Lane<CF>& S = lanes[thread_min];
    d_f[g] = smin;
    x2params<CF, M>(P, S.x[S.ind[0]], S.params, reopt);
    for (int i = 0; i < N; ++i) d_params[(size_t)g * N + i] = S.params[i];
    for (int m = 0; m < P.n_filters; ++m) d_dV[(size_t)g * P.n_filters + m] = S.dV[m];
  }
  for (int t = 0; t < G; ++t) {
    size_t id = (size_t)g * G + t;
    rng[id] = lanes[t].rng;
    if (lane_f) lane_f[id] = lanes[t].s_f;
    if (lane_nev) lane_nev[id] = lanes[t].nev;
    if (lane_steps) lane_steps[id] = lanes[t].steps;
  }
}

// ==================================================================== RMSD
// The RMSD build's merit function (cuda.c:890-906): the root of the summed
// filter chi^2 over the summed weights, in CHI_FLOAT.
template <class CF>
inline CF rmsd_from_sums(const Problem& P, const double* sy2, const double* sy, const double* sw, CF* dV) {
  CF chi2a = 0.0;
  CF SUM_w = 0.0;
  for (int m = 0; m < P.n_filters; ++m) {
    CF chi2m = sy2[m] - sy[m] * sy[m] / sw[m];
    chi2a = chi2a + chi2m;
    dV[m] = sy[m] / sw[m];
    SUM_w = SUM_w + sw[m];
  }
  return std::sqrt(chi2a / SUM_w);
}

template <class CF>
inline CF rmsd_eval(const Problem& P, int form, const double* params, CF* dV) {
  double sy2[AST_MAX_FILTERS], sy[AST_MAX_FILTERS], sw[AST_MAX_FILTERS];
  Nudge nu;
  ModelOut o{nullptr, nullptr, P.v.nudge ? &nu : nullptr};
  model_sums(P, form, params, sy2, sy, sw, &o);
  if (P.v.nudge && nu.failed) return CF(1e30);
  CF c = rmsd_from_sums<CF>(P, sy2, sy, sw, dV);
  if (P.v.nudge) c = nu.merit<CF>(P, form, c, dV);
  return c;
}

struct RmsdCfg {
  int group_size, n_samples;
  float dx;
  int spherical_k, reopt;
  uint64_t seed;
};

// One block of chi2_gpu_rms (cuda.c:2148-2389): each thread draws
// n_samples models around params0 (a normal shift of dx per scale-free
// coordinate, clamped at hard limits), keeps the range of the parameters of
// those within f1 of the input's RMSD, and the block reduces the ranges
// serially in float.  counts[2g], counts[2g+1]: Ntot, Nbad of the block.
template <class CF, class M>
void rmsd_group(const Problem& P, int form, const RmsdCfg& C, const double* params0, int g, uint32_t* rng,
                float* par_min, float* par_max, int64_t* counts, double* f0f1) {
  const ast_type_ids& T = P.T;
  const int N = P.n_params, G = C.group_size;
  const double SMALL = 1e-8;  // cuda.c:2258
  double prm[AST_MAX_PARAMS];
  for (int i = 0; i < N; ++i) prm[i] = params0[i];
  CF x0[AST_MAX_PARAMS];
  params2x<CF, M>(P, x0, prm);
  CF dV[AST_MAX_FILTERS];
  const CF f0 = rmsd_eval<CF>(P, form, prm, dV);
  const CF f1 = f0 + f0 / std::sqrt((CF)(P.n_data - N - P.n_filters));
  if (f0f1) { f0f1[0] = f0; f0f1[1] = f1; }
  const int iphiM = P.types[T.phi_M], iphi0 = P.types[T.phi_0];
  const double phi_M0 = prm[iphiM], phi_00 = prm[iphi0];
  double phi0 = 0.0;
  const bool sph = C.spherical_k && P.tq();
  auto moments = [&](const double* p, double* Is, double* Ii) {
    const double bt = p[P.types[T.b_tumb]], ct = p[P.types[T.c_tumb]];
    *Is = (1.0 + bt * bt) / (bt * bt + ct * ct);
    *Ii = (1.0 + ct * ct) / (bt * bt + ct * ct);
  };
  if (P.tq()) {
    double Is, Ii;
    moments(prm, &Is, &Ii);
    phi0 = M::atan2(Is * prm[P.types[T.Ts]], Ii * prm[P.types[T.Ti]]);
  }
  std::vector<CF> pmin((size_t)G * N), pmax((size_t)G * N);
  int64_t Ntot_b = 0, Nbad_b = 0;
  for (int t = 0; t < G; ++t) {
    const uint64_t id = (uint64_t)g * G + t;
    uint32_t ctr = rng[id];
    CF* mn = &pmin[(size_t)t * N];
    CF* mx = &pmax[(size_t)t * N];
    for (int i = 0; i < N; ++i) { mn[i] = 1e30; mx[i] = -1e30; }
    int Ntot = 0, Nbad = 0;
    for (int istage = 0; istage < C.n_samples; ++istage) {
      int LAM = 0;
      CF x1[AST_MAX_PARAMS];
      for (int i = 0; i < N; ++i) {
        CF xx = x0[i] + C.dx * astrng::normal01(C.seed, id, ctr);
        ctr += 2;
        const int per = P.prop[i][AST_P_PERIODIC];
        if (xx < SMALL && hard_left(per, LAM)) xx = SMALL;
        else if (xx > 1.0 - SMALL && hard_right(per, LAM)) xx = 1.0 - SMALL;
        x1[i] = xx;
        if (P.prop[i][AST_P_TYPE] == T.Es) LAM = x1[i] >= 0.5;
      }
      if (x2params<CF, M>(P, x1, prm, C.reopt)) continue;
      const CF f = rmsd_eval<CF>(P, form, prm, dV);
      Ntot++;
      if (f < f1) {
        if (prm[iphiM] > phi_M0 + PI) prm[iphiM] = phi_M0 - 2 * PI;
        if (prm[iphiM] < phi_M0 - PI) prm[iphiM] = phi_M0 + 2 * PI;
        if (prm[iphi0] > phi_00 + PI) prm[iphi0] = phi_00 - 2 * PI;
        if (prm[iphi0] < phi_00 - PI) prm[iphi0] = phi_00 + 2 * PI;
        if (sph) {
          double Is, Ii;
          moments(prm, &Is, &Ii);
          const double Ki = Ii * prm[P.types[T.Ti]];
          const double Ks = Is * prm[P.types[T.Ts]];
          const double Kl = prm[P.types[T.Tl]];
          prm[P.types[T.Ti]] = std::sqrt(Ki * Ki + Ks * Ks + Kl * Kl);
          prm[P.types[T.Ts]] = M::acos(Kl / prm[P.types[T.Ti]]);
          double phi = M::atan2(Ks, Ki);
          if (phi > phi0 + PI) phi = phi - 2 * PI;
          if (phi < phi0 - PI) phi = phi + 2 * PI;
          prm[P.types[T.Tl]] = phi;
        }
        for (int i = 0; i < N; ++i) {
          if (prm[i] < mn[i]) mn[i] = prm[i];
          if (prm[i] > mx[i]) mx[i] = prm[i];
        }
      } else {
        Nbad++;
      }
    }
    rng[id] = ctr;
    Ntot_b += Ntot;
    Nbad_b += Nbad;
  }
  for (int i = 0; i < N; ++i) {
    float fmin = 1e30, fmax = -1e30;
    for (int j = 0; j < G; ++j) {
      const float a = pmin[(size_t)j * N + i], b = pmax[(size_t)j * N + i];
      if (a < fmin) fmin = a;
      if (b > fmax) fmax = b;
    }
    par_min[(size_t)g * N + i] = fmin;
    par_max[(size_t)g * N + i] = fmax;
  }
  counts[2 * (size_t)g] = Ntot_b;
  counts[2 * (size_t)g + 1] = Nbad_b;
}

// ================================================================ plotting
// The plotting grid (read_data.c:304-360) as observations: n_plot times
// (t_last - t_first)/(n_plot-1) apart from the first observation (the
// reference's first observation is at time 0 and its grid times are
// iplot*h), both ends copied from the first and last observation, E and S
// included; the interior points' E,S interpolated from the 3-point
// ephemeris and normalised (read_data.c:345-355; literal or canonical
// arithmetic by form).  V = w = 0 and filter 0: only Vmod is read.
inline std::vector<ast_obs> plot_grid(const Problem& P, int form, int n_plot) {
  std::vector<ast_obs> g((size_t)n_plot);
  const ast_obs& first = P.obs[0];
  const ast_obs& last = P.obs[P.n_data - 1];
  const double h = (last.mjd - first.mjd) / (n_plot - 1);
  for (int i = 0; i < n_plot; ++i) {
    ast_obs& o = g[i];
    std::memset(&o, 0, sizeof o);
    if (i == 0 || i == n_plot - 1) {
      const ast_obs& d = i == 0 ? first : last;
      o.mjd = d.mjd;
      for (int c = 0; c < 3; ++c) { o.E[c] = d.E[c]; o.S[c] = d.S[c]; }
    } else {
      o.mjd = first.mjd + i * h;
      if (form == 0) interp_literal(P.eph, o.mjd, o.E, o.S);
      else interp_canon(P.eph, o.mjd, o.E, o.S);
    }
  }
  return g;
}

// chi2_plot (cuda.c:1865-1966, no PROFILES/ANIMATE): chi^2 and delta_V on
// the observations, then Vmod + delta_V[0] (cuda.c:699-703) along the grid,
// integrated from the grid's first time.
template <class CF>
inline CF model_curve(const Problem& P, int form, const double* params, const std::vector<ast_obs>& grid, CF* dV,
                      double* vmod) {
  const CF chi2 = chi2_eval<CF>(P, form, params, dV);
  Problem G = P;
  G.obs = grid.data();
  G.n_data = (int)grid.size();
  G.n_filters = 1;
  double sy2[AST_MAX_FILTERS], sy[AST_MAX_FILTERS], sw[AST_MAX_FILTERS];
  Nudge nu;  // the grid pass finds model minima too, and stops the curve on failure
  ModelOut out{vmod, nullptr, P.v.nudge ? &nu : nullptr};
  if (P.n_seg > 1) {  // the grid's own segment split (read_data.c:357-363, cuda.c:1943-1947)
    int iseg = 0;
    for (int i = 0; i < G.n_data; ++i)
      if (iseg < P.n_seg && grid[i].mjd >= P.seg_t[iseg]) G.seg_start[iseg++] = i;
    if (iseg < P.n_seg) return CF(NAN);
    for (int i = 0; i < G.seg_start[0]; ++i) vmod[i] = std::nan("");  // never computed
    model_sums(G, form, params, sy2, sy, sw, &out);
  } else if (form == 0) {
    model_literal(G, params, sy2, sy, sw, &out);
  } else {
    model_canon<double>(G, params, sy2, sy, sw, &out);
  }
  for (int i = 0; i < G.n_data; ++i) vmod[i] = vmod[i] + dV[0];
  return chi2;
}

// minima() (misc.c:65-276) without PARABOLIC_MAX, MINIMA_SPLINE and
// MINIMA_PRINT: brightness minima (magnitude maxima) of the curve, every
// pairwise interval, a 10000-bin histogram over frequency of the intervals
// within 10% of a whole number of periods, normalised by its mean and
// standard deviation, and the top ncl clusters above 3 sigma.  The left walk
// of a cluster stops at bin 0 (the reference reads H[-1] there,
// misc.c:240-252).  Clusters not found: frequency -1, height 0.  H (10000
// values, may be null) receives the normalised histogram.
inline int periodogram(const double* t, const double* Vm, int Nplot, int ncl, double* cl_fr, double* cl_H,
                       int* n_min, double* H_out) {
  const int NMIN_MAX = 10000;
  std::vector<double> tm;
  for (int i = 1; i <= Nplot - 2; ++i)
    if (Vm[i] > Vm[i - 1] && Vm[i] >= Vm[i + 1]) tm.push_back(t[i]);
  const int N = (int)tm.size();
  if (n_min) *n_min = N;
  if (N > NMIN_MAX) return AST_E_DATA;
  std::vector<double> dt;
  for (int i = 0; i <= N - 2; ++i)
    for (int j = i + 1; j <= N - 1; ++j) dt.push_back(tm[j] - tm[i]);
  std::sort(dt.begin(), dt.end());
  const int NN = 10000;
  const double sgm = 0.1, dt0 = 1.0 / 24.0, dt2 = 80.0 / 24.0, s0 = 3;
  const double fr0 = 1 / dt0, fr2 = 1 / dt2;
  std::vector<double> fr(NN), H(NN);
  std::vector<int> marked(NN, 0);
  double m = 0.0;
  for (int i = 0; i < NN; ++i) {
    fr[i] = fr2 + (double)i / (double)(NN - 1) * (fr0 - fr2);
    H[i] = 0.0;
    const double dt1 = 1 / fr[i];
    for (size_t j = 0; j < dt.size(); ++j) {
      const double dv = std::fabs(dt[j] / dt1 - (int)(dt[j] / dt1 + 0.5));
      if (dv < sgm) H[i] = H[i] + 1.0;
    }
    m = m + H[i];
  }
  m = m / NN;
  double s = 0;
  for (int i = 0; i < NN; ++i) {
    const double dev = H[i] - m;
    s = s + dev * dev;
  }
  s = std::sqrt(s / (NN - 1));
  for (int i = 0; i < NN; ++i) H[i] = (H[i] - m) / s;
  if (H_out) for (int i = 0; i < NN; ++i) H_out[i] = H[i];
  const double cl_min = s0;
  int Ncl = 0;
  for (int j = 0; j < ncl; ++j) cl_H[j] = 0.0;
  for (int j = 0; j < ncl; ++j) {
    double max = -100;
    int imax = -1;
    cl_fr[j] = -1.0;
    for (int i = 0; i < NN; ++i)
      if (marked[i] == 0 && H[i] > max) {
        max = H[i];
        imax = i;
      }
    if (imax > -1 && H[imax] >= cl_min) {
      cl_fr[Ncl] = fr[imax];
      cl_H[Ncl] = H[imax];
      Ncl++;
      int i = imax;
      while (i < NN && H[i] > cl_min) marked[i++] = 1;
      i = imax - 1;
      while (i >= 0 && H[i] > cl_min) marked[i--] = 1;
    }
  }
  return 0;
}

// ============================================================ op counting
// A scalar that counts the floating-point operations performed on it.  The
// canonical evaluator and the elementary functions are templates over the
// scalar type, so running them on Cnt counts exactly the operations the GPU
// executes for one evaluation (the algorithmic FLOP count of the roofline).
namespace cnt {
struct Counts {
  double add, mul, fma, div, sqrt, rnd;
};
thread_local Counts K;
struct R {
  double v;
  R() : v(0) {}
  R(double x) : v(x) {}
};
inline R operator+(R a, R b) { K.add++; return R(a.v + b.v); }
inline R operator-(R a, R b) { K.add++; return R(a.v - b.v); }
inline R operator*(R a, R b) { K.mul++; return R(a.v * b.v); }
inline R operator/(R a, R b) { K.div++; return R(a.v / b.v); }
inline R operator-(R a) { return R(-a.v); }  // sign flip: an operand modifier, not an operation
inline R p_fma(R a, R b, R c) { K.fma++; return R(__builtin_fma(a.v, b.v, c.v)); }
inline R p_sqrt(R a) { K.sqrt++; return R(__builtin_sqrt(a.v)); }
inline R p_div(R a, R b) { K.div++; return R(a.v / b.v); }
inline R p_rint(R a) { K.rnd++; return R(__builtin_rint(a.v)); }
inline R p_floor(R a) { K.rnd++; return R(__builtin_floor(a.v)); }
inline R p_fabs(R a) { return R(__builtin_fabs(a.v)); }
inline R p_copysign(R a, R b) { return R(__builtin_copysign(a.v, b.v)); }
inline R p_flipsign(R a, int flag) { return R(flag ? -a.v : a.v); }
inline R p_xorsign(R a, uint32_t m) { return R(astm::p_xorsign(a.v, m)); }
inline double p_val(R a) { return a.v; }
}  // namespace cnt

}  // namespace orc

// ===================================================================== C API
extern "C" {

// Problem description passed from Python (ctypes mirror in tests/oracle_lib.py).
typedef struct {
  ast_variant v;
  int n_params;
  int n_types;
  int property[AST_MAX_PARAMS * AST_N_COLUMNS];
  int types[AST_MAX_TYPES];
  double limits[2 * AST_MAX_TYPES];  // lower [0,n_types), upper [n_types, 2 n_types)
  ast_x2 x2;
  int n_data;
  int n_filters;
  const ast_obs* obs;
  ast_ephem3 eph;
  int n_min;                      /* NUDGE: observed minima (misc.c:279-323) */
  float t_min[AST_MAX_MINIMA];
  float V_min[AST_MAX_MINIMA];
  double t_start[AST_MAX_SEGMENTS]; /* SEGMENT: segment start times on the curve's clock */
  int has_seg_first;                /* SEGMENT: 1 = the observations split at seg_first[] (the caller's
                                       split on the observer's clock, read_data.c:244-250), 0 = at t_start */
  int seg_first[AST_MAX_SEGMENTS];
} orc_problem;

static int load_problem(const orc_problem* in, orc::Problem* P) {
  if (!in || !in->obs || in->n_params <= 0 || in->n_params > AST_MAX_PARAMS || in->n_types > AST_MAX_TYPES ||
      in->n_filters <= 0 || in->n_filters > AST_MAX_FILTERS)
    return AST_E_ARG;
  P->v = in->v;
  ast_variant_type_ids(&in->v, &P->T);
  P->n_params = in->n_params;
  P->n_types = in->n_types;
  for (int i = 0; i < in->n_params; ++i)
    for (int c = 0; c < AST_N_COLUMNS; ++c) P->prop[i][c] = in->property[i * AST_N_COLUMNS + c];
  for (int t = 0; t < AST_MAX_TYPES; ++t) {
    P->types[t] = t < in->n_types ? in->types[t] : -1;
    P->lim[0][t] = t < in->n_types ? in->limits[t] : 0.0;
    P->lim[1][t] = t < in->n_types ? in->limits[in->n_types + t] : 0.0;
  }
  P->x2 = in->x2;
  P->n_data = in->n_data;
  P->n_filters = in->n_filters;
  P->obs = in->obs;
  P->eph = in->eph;
  if (in->v.nudge && (in->n_min < 1 || in->n_min > AST_MAX_MINIMA)) return AST_E_ARG;
  P->n_min = in->v.nudge ? in->n_min : 0;
  for (int k = 0; k < AST_MAX_MINIMA; ++k) {
    P->t_min[k] = in->t_min[k];
    P->V_min[k] = in->V_min[k];
  }
  P->n_seg = 0;
  if (in->v.n_seg > 1) {
    const int ns = in->v.n_seg;
    if (ns > AST_MAX_SEGMENTS || in->n_params != P->T.n_params) return AST_E_ARG;
    for (int k = 0; k < ns; ++k) P->seg_t[k] = in->t_start[k];
    int iseg = 0;
    if (in->has_seg_first) {
      for (int k = 0; k < ns; ++k) {
        if (in->seg_first[k] < 0 || in->seg_first[k] >= in->n_data || (k > 0 && in->seg_first[k] <= in->seg_first[k - 1]))
          return AST_E_DATA;
        P->seg_start[iseg++] = in->seg_first[k];
      }
    } else {
      for (int i = 0; i < in->n_data; ++i)  // read_data.c:244-250
        if (iseg < ns && in->obs[i].mjd >= in->t_start[iseg]) P->seg_start[iseg++] = i;
    }
    if (iseg < ns) return AST_E_DATA;  // a segment without an opening observation
    for (int t = 0; t < AST_MAX_TYPES; ++t)
      for (int k = 0; k < AST_MAX_SEGMENTS; ++k) P->seg_types[t][k] = -1;
    for (int i = 0; i < in->n_params; ++i) {  // asteroid.c:152-168
      const int t = P->prop[i][AST_P_TYPE];
      if (t < 0 || t >= AST_MAX_TYPES) return AST_E_ARG;
      if (P->prop[i][AST_P_MULTI_SEGMENT] == 0) {
        const int k = P->prop[i][AST_P_ISEG];
        if (k < 0 || k >= ns) return AST_E_ARG;
        P->seg_types[t][k] = i;
      } else {
        for (int k = 0; k < ns; ++k) P->seg_types[t][k] = i;
      }
    }
    for (int t = 0; t < P->n_types; ++t)
      for (int k = 0; k < ns; ++k)
        if (P->types[t] >= 0 && P->seg_types[t][k] < 0) return AST_E_ARG;
    P->n_seg = ns;
  }
  return 0;
}

static bool simplex_float(const ast_variant& v) { return v.precision != AST_PREC_FP64; }

// chi^2 and delta_V of n models; form 0 = literal, 1 = canonical.
int orc_chi2(const orc_problem* pb, int form, const double* params, int n, double* chi2, double* dV,
             double* vmod, double* traj) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) {
    orc::ModelOut out{vmod ? vmod + (size_t)k * P.n_data : nullptr, traj ? traj + (size_t)k * 6 * P.n_data : nullptr};
    const double* p = params + (size_t)k * P.n_params;
    if (simplex_float(P.v)) {
      float d[AST_MAX_FILTERS];
      std::fill(d, d + AST_MAX_FILTERS, std::nanf(""));  // stays NaN for a failed NUDGE model
      chi2[k] = orc::chi2_eval<float>(P, form, p, d, &out);
      for (int m = 0; m < P.n_filters; ++m) dV[(size_t)k * P.n_filters + m] = d[m];
    } else {
      double d[AST_MAX_FILTERS];
      std::fill(d, d + AST_MAX_FILTERS, std::nan(""));
      chi2[k] = orc::chi2_eval<double>(P, form, p, d, &out);
      for (int m = 0; m < P.n_filters; ++m) dV[(size_t)k * P.n_filters + m] = d[m];
    }
  }
  return 0;
}

// The -plot evaluation of n models on an n_plot-point grid: t (n_plot),
// vmod (n*n_plot, model-major), chi2 (n), dV (n*n_filters).
int orc_model_curve(const orc_problem* pb, int form, const double* params, int n, int n_plot, double* t,
                    double* vmod, double* chi2, double* dV, int n_threads) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  if (n_plot < 2) return AST_E_ARG;
  const std::vector<ast_obs> grid = orc::plot_grid(P, form, n_plot);
  if (t) for (int i = 0; i < n_plot; ++i) t[i] = grid[i].mjd;
  auto work = [&](int k) {
    const double* p = params + (size_t)k * P.n_params;
    double* vm = vmod + (size_t)k * n_plot;
    if (simplex_float(P.v)) {
      float d[AST_MAX_FILTERS];
      std::fill(d, d + AST_MAX_FILTERS, std::nanf(""));  // stays NaN for a failed NUDGE model
      chi2[k] = orc::model_curve<float>(P, form, p, grid, d, vm);
      for (int m = 0; m < P.n_filters; ++m) dV[(size_t)k * P.n_filters + m] = d[m];
    } else {
      double d[AST_MAX_FILTERS];
      std::fill(d, d + AST_MAX_FILTERS, std::nan(""));
      chi2[k] = orc::model_curve<double>(P, form, p, grid, d, vm);
      for (int m = 0; m < P.n_filters; ++m) dV[(size_t)k * P.n_filters + m] = d[m];
    }
  };
  if (n_threads <= 1) {
    for (int k = 0; k < n; ++k) work(k);
  } else {
    std::vector<std::thread> th;
    for (int w = 0; w < n_threads; ++w)
      th.emplace_back([&, w]() { for (int k = w; k < n; k += n_threads) work(k); });
    for (auto& x : th) x.join();
  }
  return 0;
}

// chi2_gpu_rms over n_groups blocks (threaded over blocks); rng holds the
// per-lane stream counters and is advanced in place.
int orc_rmsd(const orc_problem* pb, int form, int n_groups, int group_size, int n_samples, float dx, int spherical_k,
             uint64_t seed, int reopt, const double* params0, uint32_t* rng, float* par_min, float* par_max,
             int64_t* counts, double* f0f1, int n_threads) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  if (n_groups <= 0 || group_size <= 0 || n_samples < 0) return AST_E_ARG;
  orc::RmsdCfg C{group_size, n_samples, dx, spherical_k, reopt, seed};
  auto work = [&](int g) {
    double* ff = g == 0 ? f0f1 : nullptr;
    if (simplex_float(P.v)) {
      if (form) orc::rmsd_group<float, orc::AstMath>(P, form, C, params0, g, rng, par_min, par_max, counts, ff);
      else orc::rmsd_group<float, orc::LibmMath>(P, form, C, params0, g, rng, par_min, par_max, counts, ff);
    } else {
      if (form) orc::rmsd_group<double, orc::AstMath>(P, form, C, params0, g, rng, par_min, par_max, counts, ff);
      else orc::rmsd_group<double, orc::LibmMath>(P, form, C, params0, g, rng, par_min, par_max, counts, ff);
    }
  };
  if (n_threads <= 1) {
    for (int g = 0; g < n_groups; ++g) work(g);
  } else {
    std::vector<std::thread> th;
    for (int w = 0; w < n_threads; ++w)
      th.emplace_back([&, w]() { for (int g = w; g < n_groups; g += n_threads) work(g); });
    for (auto& t : th) t.join();
  }
  return 0;
}

int orc_periodogram(const double* t, const double* vm, int n, int ncl, double* cl_fr, double* cl_H, int* n_min,
                    double* H) {
  if (!t || !vm || n < 0 || ncl < 0) return AST_E_ARG;
  return orc::periodogram(t, vm, n, ncl, cl_fr, cl_H, n_min, H);
}

// x2params on n points (x in double, converted to the simplex float type).
int orc_x2params(const orc_problem* pb, int form, int reopt, const double* x, int n, double* params, int* rcs) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) {
    const double* xk = x + (size_t)k * P.n_params;
    double* pk = params + (size_t)k * P.n_params;
    if (simplex_float(P.v)) {
      float xf[AST_MAX_PARAMS];
      for (int i = 0; i < P.n_params; ++i) xf[i] = (float)xk[i];
      rcs[k] = form ? orc::x2params<float, orc::AstMath>(P, xf, pk, reopt) : orc::x2params<float, orc::LibmMath>(P, xf, pk, reopt);
    } else {
      rcs[k] = form ? orc::x2params<double, orc::AstMath>(P, xk, pk, reopt) : orc::x2params<double, orc::LibmMath>(P, xk, pk, reopt);
    }
  }
  return 0;
}

int orc_params2x(const orc_problem* pb, int form, const double* params, int n, double* x) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  for (int k = 0; k < n; ++k) {
    const double* pk = params + (size_t)k * P.n_params;
    double* xk = x + (size_t)k * P.n_params;
    if (simplex_float(P.v)) {
      float xf[AST_MAX_PARAMS];
      if (form) orc::params2x<float, orc::AstMath>(P, xf, pk);
      else orc::params2x<float, orc::LibmMath>(P, xf, pk);
      for (int i = 0; i < P.n_params; ++i) xk[i] = xf[i];
    } else {
      if (form) orc::params2x<double, orc::AstMath>(P, xk, pk);
      else orc::params2x<double, orc::LibmMath>(P, xk, pk);
    }
  }
  return 0;
}

// One chi2_gpu launch over all groups, threaded over groups.
int orc_run(const orc_problem* pb, int form, int n_groups, int group_size, int nstages, int n_steps,
            int64_t eval_budget, double size_min, uint64_t seed, int reopt, const double* params0,
            uint32_t* rng, double* d_f, double* d_params, double* d_dV, double* lane_f, int64_t* lane_nev,
            int64_t* lane_steps, int n_threads) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  if (n_groups <= 0 || group_size <= 0 || nstages <= 0) return AST_E_ARG;
  orc::RunCfg C{n_groups, group_size, nstages, n_steps, eval_budget, size_min, seed};
  if (C.size_min <= 0) C.size_min = simplex_float(P.v) ? 1e-5 : 1e-10;
  auto work = [&](int g) {
    if (simplex_float(P.v)) {
      if (form) orc::run_group<float, orc::AstMath>(P, form, C, reopt, params0, g, rng, d_f, d_params, d_dV, lane_f, lane_nev, lane_steps);
      else orc::run_group<float, orc::LibmMath>(P, form, C, reopt, params0, g, rng, d_f, d_params, d_dV, lane_f, lane_nev, lane_steps);
    } else {
      if (form) orc::run_group<double, orc::AstMath>(P, form, C, reopt, params0, g, rng, d_f, d_params, d_dV, lane_f, lane_nev, lane_steps);
      else orc::run_group<double, orc::LibmMath>(P, form, C, reopt, params0, g, rng, d_f, d_params, d_dV, lane_f, lane_nev, lane_steps);
    }
  };
  if (n_threads <= 1) {
    for (int g = 0; g < n_groups; ++g) work(g);
  } else {
    std::vector<std::thread> th;
    for (int w = 0; w < n_threads; ++w)
      th.emplace_back([&, w]() { for (int g = w; g < n_groups; g += n_threads) work(g); });
    for (auto& t : th) t.join();
  }
  return 0;
}

// Elementary functions over arrays (fn: 0 sin, 1 cos, 2 tan, 3 atan2(x,y as y,x), 4 log,
// 5 log10, 6 exp, 7 asin, 8 atan, 9 expf, 10/11 sin/cos of sincos_t, 12 acos, 13 powf).  For atan2, x holds y and y holds x.
int orc_math(int fn, const double* x, const double* y, double* out, int n) {
  for (int k = 0; k < n; ++k) {
    double a = x[k];
    switch (fn) {
      case 0: out[k] = astm::sin(a); break;
      case 1: out[k] = astm::cos(a); break;
      case 2: out[k] = astm::tan(a); break;
      case 3: out[k] = astm::atan2(a, y[k]); break;
      case 4: out[k] = astm::log(a); break;
      case 5: out[k] = astm::log10(a); break;
      case 6: out[k] = astm::exp(a); break;
      case 7: out[k] = astm::asin(a); break;
      case 8: out[k] = astm::atan(a); break;
      case 9: out[k] = astm::expf((float)a); break;
      case 10: { double sn, cs; astm::sincos_t(a, &sn, &cs); out[k] = sn; break; }
      case 11: { double sn, cs; astm::sincos_t(a, &sn, &cs); out[k] = cs; break; }
      case 12: out[k] = astm::acos(a); break;
      case 13: out[k] = astm::powf((float)a, (float)y[k]); break;
      default: return AST_E_ARG;
    }
  }
  return 0;
}

// Operation counts of one canonical evaluation of `params` on the problem's
// curve: out[0..5] = add/sub, mul, fma, div, sqrt, rint/floor.
int orc_count_ops(const orc_problem* pb, const double* params, double* out) {
  orc::Problem P;
  int rc = load_problem(pb, &P);
  if (rc) return rc;
  if (P.n_seg > 1) return AST_E_UNSUPPORTED;  // counted per segment by the caller
  orc::cnt::K = orc::cnt::Counts{0, 0, 0, 0, 0, 0};
  orc::cnt::R sy2[AST_MAX_FILTERS], sy[AST_MAX_FILTERS], sw[AST_MAX_FILTERS];
  orc::model_canon<orc::cnt::R>(P, params, sy2, sy, sw, nullptr);
  const orc::cnt::Counts& k = orc::cnt::K;
  out[0] = k.add; out[1] = k.mul; out[2] = k.fma; out[3] = k.div; out[4] = k.sqrt; out[5] = k.rnd;
  return 0;
}

int orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  astrng::philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = c[i];
  return 0;
}

int orc_uniform(uint64_t seed, uint64_t fit, uint32_t k0, int n, float* out) {
  for (int k = 0; k < n; ++k) out[k] = astrng::uniform01(seed, fit, k0 + (uint32_t)k);
  return 0;
}

// n normal deviates of one stream from draw k0 on (two draws each).
int orc_normal(uint64_t seed, uint64_t fit, uint32_t k0, int n, float* out) {
  for (int k = 0; k < n; ++k) out[k] = astrng::normal01(seed, fit, k0 + 2u * (uint32_t)k);
  return 0;
}

}  // extern "C"
