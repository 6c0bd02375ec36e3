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
// h/6 and the step count), V, w, the filter and one per-point value (aux):
// the phase angle acos(S.E) under TREND, else the time.  96 bytes.  A table
// of n_data records carries one more record after them (TailRec): TIME_STEP
// and, for NUDGE, the observed minima.
struct PointRec {
  double E[3];
  double S[3];
  double h, hh, h6;
  float V, w;
  int filter, nsteps;
  double aux;    // TREND: phase angle acos(S.E); otherwise the time (TORQUE2's split, NUDGE's minima)
};
static_assert(sizeof(PointRec) == 96, "PointRec layout");

// The record after a table's points: the variant's TIME_STEP (for TORQUE2's
// split interval) and NUDGE's observed brightness minima (asteroid.h:216-230,
// 319-337; misc.c:279-323): times relative to the curve's first observation
// and magnitudes in the converted system, as floats; v1s is the V1S macro
// (one-sided magnitude distance, cuda.c:940-945).
constexpr int NOBS_MAX = 10;  // asteroid.h:219
struct TailRec {
  int n_obs;
  int v1s;
  float t_obs[NOBS_MAX];
  float V_obs[NOBS_MAX];
  double ts;
};
static_assert(sizeof(TailRec) == sizeof(PointRec), "the tail record takes one PointRec slot");

// Model variant as one compile-time code: the reference's macros
// (makefile:3-36) as bit flags.  TORQUE2 implies TORQUE and RECT implies BC
// (asteroid.h:32-38).
// VF_SEG is SEGMENT (cuda.c:162-813) with three segments and ONE_LE; VF_SEG4
// makes it four segments, VF_SEGLE gives every segment its own L and Es (no
// ONE_LE).
enum : int { VF_TQ = 1, VF_BC = 2, VF_BW = 4, VF_TREND = 8, VF_TQ2 = 16, VF_ROT = 32, VF_RECT = 64, VF_NUDGE = 128,
             VF_SEG = 256, VF_SEG4 = 512, VF_SEGLE = 1024 };
template <int VC>
struct VarF {
  static constexpr bool TQ = (VC & (VF_TQ | VF_TQ2)) != 0;  // constant torque
  static constexpr bool T2 = (VC & VF_TQ2) != 0;            // a second torque after a split time
  static constexpr bool BC = (VC & (VF_BC | VF_RECT)) != 0; // photometric b, c
  static constexpr bool BW = (VC & VF_BW) != 0;             // black-and-white ball
  static constexpr bool TR = (VC & VF_TREND) != 0;          // de-trending by the phase angle
  static constexpr bool RO = (VC & VF_ROT) != 0;            // brightness frame rotated
  static constexpr bool RC = (VC & VF_RECT) != 0;           // rectangular prism brightness
  static constexpr bool NU = (VC & VF_NUDGE) != 0;          // model minima nudged to observed ones
  static constexpr bool SG = (VC & VF_SEG) != 0;            // SEGMENT: the curve split into segments
  static constexpr int NSEG = SG ? ((VC & VF_SEG4) ? 4 : 3) : 1;
  static constexpr bool ONE_LE = SG && !(VC & VF_SEGLE);    // L and Es shared by the segments
};
#define AST_VFLAGS(VC)                                                                           \
  constexpr bool TQ = VarF<VC>::TQ, BC = VarF<VC>::BC, BW = VarF<VC>::BW, T2 = VarF<VC>::T2,   \
                 TR = VarF<VC>::TR, RO = VarF<VC>::RO, RC = VarF<VC>::RC, NU = VarF<VC>::NU,   \
                 SG = VarF<VC>::SG

// Parameter positions of a variant (the Property order, asteroid.c:86-125).
template <int VC>
struct Pos {
  using F = VarF<VC>;
  static constexpr int t = F::TR ? 1 : 0;
  static constexpr int o = t + (F::TQ ? 3 : 0) + (F::T2 ? 4 : 0);
  static constexpr int theta_M = 0, phi_M = 1, phi_0 = 2;
  static constexpr int A = F::TR ? 3 : -1;
  static constexpr int Ti = F::TQ ? 3 + t : -1, Ts = F::TQ ? 4 + t : -1, Tl = F::TQ ? 5 + t : -1;
  static constexpr int T2i = F::T2 ? 6 + t : -1, T2s = F::T2 ? 7 + t : -1, T2l = F::T2 ? 8 + t : -1,
                       Tt = F::T2 ? 9 + t : -1;
  static constexpr int c_tumb = 3 + o, b_tumb = 4 + o, Es = 5 + o, L = 6 + o, psi_0 = 7 + o;
  static constexpr int c = F::BC ? 8 + o : -1, b = F::BC ? 9 + o : -1;
  static constexpr int ob = o + (F::BC ? 2 : 0);
  static constexpr int theta_R = (F::BW || F::RO) ? 8 + ob : -1, phi_R = (F::BW || F::RO) ? 9 + ob : -1;
  static constexpr int psi_R = F::RO ? 10 + ob : -1, kappa = F::BW ? 10 + ob : -1;
  static constexpr int N0 = 8 + o + (F::BC ? 2 : 0) + ((F::BW || F::RO) ? 3 : 0);
  // SEGMENT (asteroid.c:128-150, asteroid.h:104-108): after the first
  // segment's parameters, each further segment repeats, in the same order,
  // those that are not multi-segment: theta_M, phi_M, phi_0, the torque
  // components, Es and L (unless ONE_LE) and psi_0; c_tumb, b_tumb and the
  // photometric c, b are shared.  rep(r) is the first-segment position of the r-th repeated
  // parameter, base(i) the first-segment position of parameter i (its role)
  // and segof(i) its segment.
  static_assert(!F::SG || (VC & ~(VF_SEG | VF_SEG4 | VF_SEGLE | VF_TQ | VF_BC)) == 0,
                "SEGMENT is built for the LS ellipsoid with or without torque and photometric b, c");
  static constexpr int NSEG = F::NSEG;
  static constexpr int R = F::SG ? 4 + (F::TQ ? 3 : 0) + (F::ONE_LE ? 0 : 2) : 0;
  static constexpr int rep(int r) {
    return r < 3 ? r : (F::TQ && r < 6) ? r : (!F::ONE_LE && r == R - 3) ? Es : (!F::ONE_LE && r == R - 2) ? L : psi_0;
  }
  static constexpr int rank(int bp) {
    for (int r = 0; r < R; ++r)
      if (rep(r) == bp) return r;
    return -1;
  }
  static constexpr int base(int i) { return i < N0 ? i : rep((i - N0) % R); }
  static constexpr int segof(int i) { return i < N0 ? 0 : (i - N0) / R + 1; }
  static constexpr int seg(int k, int bp) { return (k == 0 || rank(bp) < 0) ? bp : N0 + (k - 1) * R + rank(bp); }
  static constexpr bool is_Es(int i) { return base(i) == Es; }
  static constexpr bool is_L(int i) { return base(i) == L; }
  static constexpr bool is_psi0(int i) { return base(i) == psi_0; }
  static constexpr int N = N0 + R * (NSEG - 1);
  static_assert(N <= MAXP && N < 64, "parameter count beyond the slot layout (and the 6-bit vertex index)");
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
template <class CF, int VC>
__device__ int x2params(const CF* x, double* params, const Space<CF>& sp, int reopt) {
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  int LAM = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (P::is_Es(i)) LAM = x[i] >= 0.5;
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
    } else if (P::is_Es(i)) {
      if (LAM) {
        params[i] = 2.0 * (x[i] - 0.5) * (1.0 - 1.0 / Ii) + 1.0 / Ii;
        psi_min = 0.0;
        psi_max = 2.0 * PI;
      } else {
        params[i] = 2.0 * x[i] * (1.0 / Ii - 1.0 / Is) + 1.0 / Is;
        psi_max = astm::atan(__builtin_sqrt(Ii * (Is - 1.0 / params[i]) / Is / (1.0 / params[i] - Ii)));
        psi_min = -psi_max;
      }
    } else if (P::is_psi0(i)) {
      params[i] = x[i] * (psi_max - psi_min) + psi_min;
    } else if (BC && i == P::b) {
      double log_b = log_c * (x[i] * (L1 - L0) + L0);
      if (__builtin_fabs(log_b - log_b_tumb) > BC_DEV_MAX) return 1;
      params[i] = astm::exp(log_b);
    } else if (sp.periodic[i] == AST_PERIODIC) {
      params[i] = x[i] * 2.0 * PI;
    } else if (sp.independent[i] == 1) {
      if (!sp.p_psi || !P::is_L(i) || reopt) params[i] = x[i] * (L1 - L0) + L0;
      if (i == P::c_tumb) {
        log_c_tumb = params[i];
        params[i] = astm::exp(log_c_tumb);
      } else if (BC && i == P::c) {
        log_c = params[i];
        if (__builtin_fabs(log_c - log_c_tumb) > BC_DEV_MAX) return 1;
        params[i] = astm::exp(log_c);
      } else if (BW && i == P::kappa) {
        params[i] = astm::exp(params[i]);
      } else if (P::is_L(i) && !reopt && sp.p_psi) {
        float dP = sp.ppsi2 - sp.ppsi1;
        params[i] = x[i] * (CF)dP + (CF)sp.ppsi1;
        double EsV = params[P::seg(P::segof(i), P::Es)];  // the segment's Es (cuda.c:1263)
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

// Segment builds call the map through one out-of-line copy.  With it inlined
// at every simplex proposal site, the builds with torque and per-segment L
// and Es (29 and 38 parameters) gave fit trajectories that left the CPU
// restatement's at some lanes (hipcc of ROCm 7.2 for gfx950, at -O1, -O2 and
// -O3 alike), while the same map evaluated on its own matched it bitwise; the
// out-of-line copy matches bit for bit.  Other builds keep it inlined.
template <class CF, int VC>
__device__ __attribute__((noinline)) int x2params_ool(const CF* x, double* params, const Space<CF>& sp, int reopt) {
  return x2params<CF, VC>(x, params, sp, reopt);
}
template <class CF, int VC>
__device__ inline int x2params_at(const CF* x, double* params, const Space<CF>& sp, int reopt) {
  if constexpr (VarF<VC>::SG) return x2params_ool<CF, VC>(x, params, sp, reopt);
  else return x2params<CF, VC>(x, params, sp, reopt);
}

// ------------------------------------------------------------------ params2x
// cuda.c:1016-1115.
template <class CF, int VC>
__device__ void params2x(CF* x, const double* params, const Space<CF>& sp) {
  using P = Pos<VC>;
  AST_VFLAGS(VC);
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
      } else if (P::is_Es(i)) {
        i_Es = i;
        LAM = params[i] > 1.0 / Ii;
        if (LAM) x[i] = 0.5 * ((params[i] - 1.0 / Ii) / (1.0 - 1.0 / Ii) + 1.0);
        else x[i] = 0.5 * (params[i] - 1.0 / Is) / (1.0 / Ii - 1.0 / Is);
      } else if (P::is_psi0(i)) {
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
__device__ inline double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <class T>
__device__ inline T dot3(T ax, T ay, T az, T bx, T by, T bz) {
  return fmx(az, bz, fmx(ay, by, ax * bx));
}

// Spin frame of the model (cuda.c:190-200): the angular-momentum axis M and
// the two axes XM, YM perpendicular to it (XM_y = 0).
template <class T>
struct SpinFrame {
  T M_x, M_y, M_z, XM_x, XM_z, YM_x, YM_y, YM_z;
};

// Body axes a, b, c at one observation from the Euler angles' sines and
// cosines (cuda.c:432-473): N in the XM-YM plane at phi, p = M x N, a at
// theta from M towards p, w = a x N, b at psi from N towards w, c = a x b.
template <class T>
__device__ inline void body_axes(const SpinFrame<T>& F, T sf, T cf, T st, T ctt, T sp, T cp, T* a, T* b, T* c) {
  const T N_x = fmx(F.YM_x, sf, F.XM_x * cf);
  const T N_y = F.YM_y * sf;
  const T N_z = fmx(F.YM_z, sf, F.XM_z * cf);
  const T p_x = fmx(-N_z, F.M_y, N_y * F.M_z);
  const T p_y = fmx(-N_x, F.M_z, N_z * F.M_x);
  const T p_z = fmx(-N_y, F.M_x, N_x * F.M_y);
  a[0] = fmx(p_x, st, F.M_x * ctt);
  a[1] = fmx(p_y, st, F.M_y * ctt);
  a[2] = fmx(p_z, st, F.M_z * ctt);
  const T w_x = fmx(-a[2], N_y, a[1] * N_z);
  const T w_y = fmx(-a[0], N_z, a[2] * N_x);
  const T w_z = fmx(-a[1], N_x, a[0] * N_y);
  b[0] = fmx(w_x, sp, N_x * cp);
  b[1] = fmx(w_y, sp, N_y * cp);
  b[2] = fmx(w_z, sp, N_z * cp);
  c[0] = fmx(-a[2], b[1], a[1] * b[2]);
  c[1] = fmx(-a[0], b[2], a[2] * b[0]);
  c[2] = fmx(-a[1], b[0], a[0] * b[1]);
}

// ODE right-hand side.  With FIRST the sin/cos of theta (and psi) come from
// sc[] (those of the state, shared with the body axes), else they are computed.
template <bool TQ, bool FIRST>
__device__ inline void ode(const double* y, double* f, const double* mu, double mu02, const double* sc) {
  if (TQ) {  // cuda.c:46-58
    double s5, c5, s4, c4;
    if (FIRST) { s4 = sc[0]; c4 = sc[1]; s5 = sc[2]; c5 = sc[3]; }
    else { astm::sincos_t(y[5], &s5, &c5); astm::sincos_t(y[4], &s4, &c4); }
    f[0] = __builtin_fma(mu[0] * y[1], y[2], mu[3]);
    f[1] = __builtin_fma(mu[1] * y[2], y[0], mu[4]);
    f[2] = __builtin_fma(mu[2] * y[0], y[1], mu[5]);
    f[3] = astm::div_rn(__builtin_fma(y[1], c5, y[0] * s5), s4);
    f[4] = __builtin_fma(-y[1], s5, y[0] * c5);
    f[5] = __builtin_fma(-f[3], c4, y[2]);
  } else {   // cuda.c:69-75
    double s2, c2, s1, c1;
    astm::sincos_t(2.0 * y[2], &s2, &c2);
    if (FIRST) { s1 = sc[0]; c1 = sc[1]; }
    else astm::sincos_t(y[1], &s1, &c1);
    f[0] = mu[0] * __builtin_fma(-mu[2], c2, mu[1]);
    f[1] = (mu02 * s1) * s2;
    f[2] = c1 * (mu[0] - f[0]);
  }
}

typedef __attribute__((address_space(4))) const PointRec* PointPtr;

// An empty asm that takes a wave-uniform value in scalar registers and
// "returns" it: the value must be computed (loaded) before this point and
// stays in registers after it.  No instruction is emitted.
template <class T> __device__ inline void pin_sgpr(T& v) { asm volatile("" : "+s"(v)); }
typedef __attribute__((address_space(4))) const TailRec* TailPtr;

// ------------------------------------------------------------------ NUDGE
// Constants of asteroid.h:216-230, evaluated with the reference's types
// (float constants, derived ones through double where it writes double).
constexpr int NU_M_MAX = 30;
constexpr float NU_M_MAX2 = 20.0f;
constexpr float NU_DT_MAX = 0.12f;
constexpr float NU_DV_MAX = 24.0f;
constexpr float NU_DT_MAX2 = (float)(1.5 * (double)NU_DT_MAX);
constexpr float NU_P_MIN = 0.01f;
constexpr float NU_CHI2_0 = 10.0f;
constexpr float NU_CHI2_1 = 30.0f;
constexpr float NU_L_RC = 0.05f;
constexpr float NU_L_RC2 = NU_L_RC * NU_L_RC;
constexpr float NU_L_A = (float)(1.0 / (1.0 - (double)NU_L_RC2 / (1.0 + (double)NU_L_RC2)));
constexpr float NU_P_MIN2 = 1 / NU_P_MIN - 1;

// Model minima of one evaluation (cuda.c:716-758): a point whose magnitude
// is above both neighbours' (a brightness minimum) and lies within DT_MAX2
// of an observed minimum is located by the parabola through the three
// points and kept; the 31st such minimum fails the model (1e30).  Times and
// magnitudes are kept as floats, the parabola in double, as there.  The
// located minima live in arrays of the caller's (t_mod, V_mod): apart from
// the struct, so its scalars stay in registers.
struct NudgeState {
  float t0, t1, V0, V1;  // the last two points (t_old[], V_old[] of the reference)
  float* t_mod;
  float* V_mod;
  int M = 0;
  bool failed = false;

  __device__ void point(int i, double mjd, double Vmod, TailPtr mr) {
    if (failed) return;
    if (i == 0) {
      t0 = (float)mjd;
      V0 = (float)Vmod;
      return;
    }
    if (i == 1) {
      t1 = (float)mjd;
      V1 = (float)Vmod;
      return;
    }
    if (V1 > V0 && (double)V1 >= Vmod) {
      bool local = false;
      for (int ii = 0; ii < mr->n_obs; ++ii)
        if (__builtin_fabsf(t1 - mr->t_obs[ii]) < NU_DT_MAX2) local = true;
      if (local) {
        ++M;
        if (M > NU_M_MAX) {
          failed = true;
          return;
        }
        const float q = (V1 - V0) / (t1 - t0);
        const double a = ((Vmod - (double)V1) / (mjd - (double)t1) - (double)q) / (mjd - (double)t0);
        const double b = (double)q - a * (double)(t1 + t0);
        const double c = (double)V1 - a * (double)t1 * (double)t1 - b * (double)t1;
        const float tm = (float)(-b / 2.0 / a);
        t_mod[M - 1] = tm;
        V_mod[M - 1] = (float)(a * (double)tm * (double)tm + b * (double)tm + c);
      }
    }
    t0 = t1;
    V0 = V1;
    t1 = (float)mjd;
    V1 = (float)Vmod;
  }

  // The merit (cuda.c:912-992): chi2 times the reward for model minima
  // close to observed ones in (t, V) and the penalty for too many of them.
  template <class CF>
  __device__ CF merit(CF chi2a, CF dv0, TailPtr mr) const {
    float S_M = 0.0f, P_tot = 1.0f;
    for (int imod = 0; imod < M; ++imod) {
      for (int iobs = 0; iobs < mr->n_obs; ++iobs) {
        const float dt = __builtin_fabsf(t_mod[imod] - mr->t_obs[iobs]);
        if (dt < NU_DT_MAX2) {
          if (dt > NU_DT_MAX) {
            const float x = (NU_DT_MAX2 - dt) / (NU_DT_MAX2 - NU_DT_MAX);
            S_M = (float)((double)S_M + (double)(x * x) * (-2.0 * (double)x + 3.0));
          } else {
            S_M = (float)((double)S_M + 1.0);
            float dV = (float)(((CF)V_mod[imod] + dv0) - (CF)mr->V_obs[iobs]);
            if (mr->v1s && dV > 0.0f) dV = 0.0f;
            const float x = __builtin_sqrtf(dt * dt / NU_DT_MAX / NU_DT_MAX + dV * dV / NU_DV_MAX / NU_DV_MAX);
            if (x < 1.0f) {
              const float P_i = NU_L_A * x * x / (x * x + NU_L_RC2);
              P_tot = (float)((double)P_tot * ((double)NU_P_MIN * (1.0 + (double)(NU_P_MIN2 * P_i))));
            }
          }
        }
      }
    }
    P_tot = astm::powf(P_tot, (float)(1.0 / mr->n_obs));
    if (P_tot < NU_P_MIN) P_tot = NU_P_MIN;
    if (chi2a > (CF)NU_CHI2_1) {
      P_tot = 1.0f;
    } else if (chi2a > (CF)NU_CHI2_0) {
      const float x = (float)(((CF)NU_CHI2_1 - chi2a) / (CF)(NU_CHI2_1 - NU_CHI2_0));
      const float beta = x * x * (-2 * x + 3);
      P_tot = astm::powf(P_tot, beta);
    }
    float P_M;
    if (S_M < NU_M_MAX2) {
      P_M = 1.0f;
    } else if (S_M < NU_M_MAX) {
      const float x = (S_M - NU_M_MAX2) / (NU_M_MAX - NU_M_MAX2);
      P_M = (float)(1.0 + (double)(3 * x * x * (-2 * x + 3)));
    } else {
      P_M = 4.0f;
    }
    return chi2a * (CF)P_tot * (CF)P_M;
  }
};

// Per-model constants of the brightness.
struct ModelConsts {
  double ib, ic, bcp, kappa, sR, cR, stR, ctR;
  double sPR, cPR;  // ROTATE: sin, cos of psi_R
};

#define AST_AI __device__ __attribute__((always_inline)) inline

// Body axes at one observation (FP64, both forms) from the spin frame, phi
// and the sines and cosines of theta (sc[0..1]) and, with torque, of psi
// (sc[2..3]).
template <bool TQ>
AST_AI void axes_at(const SpinFrame<double>& Fd, double phi, double psi, const double* sc, double* a, double* b,
                    double* c) {
  double sf, cf, sp, cp;
  astm::sincos_t(phi, &sf, &cf);
  const double st = sc[0], ctt = sc[1];
  if (TQ) { sp = sc[2]; cp = sc[3]; }
  else astm::sincos_t(psi, &sp, &cp);
  body_axes<double>(Fd, sf, cf, st, ctt, sp, cp, a, b, c);
}

// ROTATE (cuda.c:475-516): the brightness ellipsoid's axes turned by phi_R,
// theta_R, psi_R about the kinematic ones, by the same construction as the
// body axes with (b, c, a) in place of the spin frame; the first two steps
// are also those of the BW_BALL axis (bw_arg_f64).
AST_AI void rotate_axes(double* a, double* b, double* c, const ModelConsts& K) {
  const double N_x = __builtin_fma(c[0], K.sR, b[0] * K.cR);
  const double N_y = __builtin_fma(c[1], K.sR, b[1] * K.cR);
  const double N_z = __builtin_fma(c[2], K.sR, b[2] * K.cR);
  const double p_x = __builtin_fma(-N_z, a[1], N_y * a[2]);
  const double p_y = __builtin_fma(-N_x, a[2], N_z * a[0]);
  const double p_z = __builtin_fma(-N_y, a[0], N_x * a[1]);
  a[0] = __builtin_fma(p_x, K.stR, a[0] * K.ctR);
  a[1] = __builtin_fma(p_y, K.stR, a[1] * K.ctR);
  a[2] = __builtin_fma(p_z, K.stR, a[2] * K.ctR);
  const double w_x = __builtin_fma(-a[2], N_y, a[1] * N_z);
  const double w_y = __builtin_fma(-a[0], N_z, a[2] * N_x);
  const double w_z = __builtin_fma(-a[1], N_x, a[0] * N_y);
  b[0] = __builtin_fma(w_x, K.sPR, N_x * K.cPR);
  b[1] = __builtin_fma(w_y, K.sPR, N_y * K.cPR);
  b[2] = __builtin_fma(w_z, K.sPR, N_z * K.cPR);
  c[0] = __builtin_fma(-a[2], b[1], a[1] * b[2]);
  c[1] = __builtin_fma(-a[0], b[2], a[2] * b[0]);
  c[2] = __builtin_fma(-a[1], b[0], a[0] * b[1]);
}

// RECT (cuda.c:609-661): magnitude of a rectangular prism of half-lengths
// 1, bL, cL along a, b, c at phase 0, as the reference writes it: each axis
// (b and c scaled by their lengths) is crossed with the vector whose x, y, z
// components are E's projections on b, c, a, and the three face terms are
// the norms of the pairwise cross products of those.
AST_AI double rect_mag(const double* a, const double* b, const double* c, const double* E, double bL, double cL) {
  const double Eb = dot3(b[0], b[1], b[2], E[0], E[1], E[2]);
  const double Ec = dot3(c[0], c[1], c[2], E[0], E[1], E[2]);
  const double Ea = dot3(a[0], a[1], a[2], E[0], E[1], E[2]);
  const double bx = bL * b[0], by = bL * b[1], bz = bL * b[2];
  const double cx = cL * c[0], cy = cL * c[1], cz = cL * c[2];
  const double ap_x = __builtin_fma(-a[2], Ec, a[1] * Ea), ap_y = __builtin_fma(-a[0], Ea, a[2] * Eb),
               ap_z = __builtin_fma(-a[1], Eb, a[0] * Ec);
  const double bp_x = __builtin_fma(-bz, Ec, by * Ea), bp_y = __builtin_fma(-bx, Ea, bz * Eb),
               bp_z = __builtin_fma(-by, Eb, bx * Ec);
  const double cp_x = __builtin_fma(-cz, Ec, cy * Ea), cp_y = __builtin_fma(-cx, Ea, cz * Eb),
               cp_z = __builtin_fma(-cy, Eb, cx * Ec);
  const double ab_x = __builtin_fma(-ap_z, bp_y, ap_y * bp_z), ab_y = __builtin_fma(-ap_x, bp_z, ap_z * bp_x),
               ab_z = __builtin_fma(-ap_y, bp_x, ap_x * bp_y);
  const double ac_x = __builtin_fma(-ap_z, cp_y, ap_y * cp_z), ac_y = __builtin_fma(-ap_x, cp_z, ap_z * cp_x),
               ac_z = __builtin_fma(-ap_y, cp_x, ap_x * cp_y);
  const double bc_x = __builtin_fma(-bp_z, cp_y, bp_y * cp_z), bc_y = __builtin_fma(-bp_x, cp_z, bp_z * cp_x),
               bc_z = __builtin_fma(-bp_y, cp_x, bp_x * cp_y);
  double sab = dot3(ab_x, ab_y, ab_z, ab_x, ab_y, ab_z);
  double sac = dot3(ac_x, ac_y, ac_z, ac_x, ac_y, ac_z);
  double sbc = dot3(bc_x, bc_y, bc_z, bc_x, bc_y, bc_z);
  if (sab < 0.0) sab = 0.0;
  if (sbc < 0.0) sbc = 0.0;
  if (sac < 0.0) sac = 0.0;
  return -2.5 * astm::log10(4.0 * ((astm::sqrt_rn(sab) + astm::sqrt_rn(sbc)) + astm::sqrt_rn(sac)));
}

// Sun and Earth directions in the body frame scaled by the ellipsoid's axes
// (cuda.c:666-671 in the canonical form): u_S = (b.S/b, c.S/c, a.S) and u_E.
struct ScaledDirs {
  double uSb, uSc, uSa, uEb, uEc, uEa;
};
AST_AI ScaledDirs scaled_dirs(const double* a, const double* b, const double* c, const double* E,
                              const double* S, const ModelConsts& K) {
  const double bs_x = b[0] * K.ib, bs_y = b[1] * K.ib, bs_z = b[2] * K.ib;
  const double cs_x = c[0] * K.ic, cs_y = c[1] * K.ic, cs_z = c[2] * K.ic;
  ScaledDirs u;
  u.uSb = dot3(bs_x, bs_y, bs_z, S[0], S[1], S[2]);
  u.uSc = dot3(cs_x, cs_y, cs_z, S[0], S[1], S[2]);
  u.uSa = dot3(a[0], a[1], a[2], S[0], S[1], S[2]);
  u.uEb = dot3(bs_x, bs_y, bs_z, E[0], E[1], E[2]);
  u.uEc = dot3(cs_x, cs_y, cs_z, E[0], E[1], E[2]);
  u.uEa = dot3(a[0], a[1], a[2], E[0], E[1], E[2]);
  return u;
}

// Lommel-Seeliger magnitude from the scaled directions (cuda.c:672-689), FP64
// canonical form.
AST_AI double ls_tail_f64(const ScaledDirs& u, double bcp) {
  const double sS2 = dot3(u.uSb, u.uSc, u.uSa, u.uSb, u.uSc, u.uSa);
  const double sE2 = dot3(u.uEb, u.uEc, u.uEa, u.uEb, u.uEc, u.uEa);
  const double sS = astm::sqrt_rn(sS2);
  const double sE = astm::sqrt_rn(sE2);
  const double ca = astm::div_rn(dot3(u.uSb, u.uSc, u.uSa, u.uEb, u.uEc, u.uEa), sS * sE);
  const double sa = astm::sqrt_rn(__builtin_fma(-ca, ca, 1.0));
  const double scn = astm::sqrt_rn(__builtin_fma((2.0 * sS) * sE, ca, sS2 + sE2));
  const double isc = astm::div_rn(1.0, scn);
  const double cl = __builtin_fma(sE, ca, sS) * isc;
  const double sl = (sE * sa) * isc;
  const double cd = __builtin_fma(cl, ca, sl * sa);
  const double sd = __builtin_fma(-cl, sa, sl * ca);
  const double lg = astm::log(astm::div_rn((1.0 + cl) * (1.0 + cd), sl * (-sd)));
  const double inner = __builtin_fma(sl * sd, lg, cd + cl);
  return -2.5 * astm::log10((((bcp * sS) * sE) * isc) * inner);
}

// Lambert + ball magnitude (cuda.c:475-497, 603-607): the reflecting a axis
// turned by theta_R, phi_R, and cos zeta = r.E.  Returns the argument of the
// final log10.
AST_AI double bw_arg_f64(const double* a, const double* b, const double* c, const double* E, const ModelConsts& K) {
  const double n_x = __builtin_fma(c[0], K.sR, b[0] * K.cR);
  const double n_y = __builtin_fma(c[1], K.sR, b[1] * K.cR);
  const double n_z = __builtin_fma(c[2], K.sR, b[2] * K.cR);
  const double q_x = __builtin_fma(-n_z, a[1], n_y * a[2]);
  const double q_y = __builtin_fma(-n_x, a[2], n_z * a[0]);
  const double q_z = __builtin_fma(-n_y, a[0], n_x * a[1]);
  const double r_x = __builtin_fma(q_x, K.stR, a[0] * K.ctR);
  const double r_y = __builtin_fma(q_y, K.stR, a[1] * K.ctR);
  const double r_z = __builtin_fma(q_z, K.stR, a[2] * K.ctR);
  const double cz = dot3(r_x, r_y, r_z, E[0], E[1], E[2]);
  return 0.5 * __builtin_fma(K.kappa, 1.0 + cz, 0.5 * (1.0 - cz));
}

// ------------------------------------------------- mixed-precision brightness
// The mixed form keeps the Euler-angle integration, the body axes, the
// scaled Sun and Earth directions and the chi^2 sums in FP64, and evaluates
// the brightness functions (square roots, divisions, logarithms) in FP32
// with the hardware reciprocal square root, reciprocal and log2 (about 1 ulp
// each).  The axes stay FP64 because the c-axis projections are scaled by
// 1/c (up to ~100 for flat shapes), which would magnify FP32 rounding of the
// axes to ~1e-5 mag.

// log2 of a float with an absolute error of about 2^-24: the exponent is
// taken out exactly and the hardware log2 sees the mantissa in [0.5, 1).
__device__ inline float log2_f32(float x) {
  return (float)__builtin_amdgcn_frexp_expf(x) + __builtin_amdgcn_logf(__builtin_amdgcn_frexp_mantf(x));
}
// The same, summed in FP64 (for the magnitude itself).
__device__ inline double log2_f32x(float x) {
  return (double)__builtin_amdgcn_frexp_expf(x) + (double)__builtin_amdgcn_logf(__builtin_amdgcn_frexp_mantf(x));
}

constexpr float LN2_F = 0.693147180559945309f;
constexpr double M2P5_LOG10_2 = -0.75257498915995295;  // -2.5 log10(2): magnitude from log2

// 1/sqrt(x) in FP64 from the hardware estimate and one Newton step.
__device__ inline double rsq_nr(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-(x * r), r, 1.0);
  return __builtin_fma(0.5 * r, e, r);
}

// Lommel-Seeliger magnitude of the mixed form, without cancellation at any
// phase.  With s^, e^ the unit scaled Sun and Earth directions, t = u_S + u_E
// and t^ its unit vector, h = s^ + e^: the phase angle alpha' is 2 s' with
// cos s' = |h|/2, and lambda' (the angle from s^ to t^) gives
// d' = lambda' - alpha'/2 with cos d' = t^.h/|h|.  With x = cos s'/cos d',
//   cos(alpha'-lambda') + cos lambda' + sin lambda' sin(lambda'-alpha')
//     ln(cot(lambda'/2) cot((alpha'-lambda')/2)) = 2 cos^2 d' g(x),
//   g(x) = x - (1 - x^2) artanh x,
// which vanishes like x^3 towards backscatter (x -> 0).  Below X0 g comes from
// its series sum_k 2 x^(2k+1)/((2k-1)(2k+1)); above, from the closed form with
// y = 1 - x = |s^ - t^| |e^ - t^| / (2 cos d') (that is 2 sin(lambda'/2)
// sin((alpha'-lambda')/2) / cos d'), so nothing cancels as x -> 1 either.  The
// unit vectors, their sum and differences are formed in FP64, the rest in
// FP32; the magnitude error stays below ~1.2e-6 for every geometry.
constexpr float MX_SERIES_X0 = 0.6f;
constexpr int MX_SERIES_TERMS = 14;  // 0.36^14 < 1e-6 of the leading term at X0

__device__ inline double ls_tail_mx(const ScaledDirs& ud, float bcp) {
  const double sS2 = dot3(ud.uSb, ud.uSc, ud.uSa, ud.uSb, ud.uSc, ud.uSa);
  const double sE2 = dot3(ud.uEb, ud.uEc, ud.uEa, ud.uEb, ud.uEc, ud.uEa);
  const double tb = ud.uSb + ud.uEb, tc = ud.uSc + ud.uEc, ta = ud.uSa + ud.uEa;
  const double tt2 = dot3(tb, tc, ta, tb, tc, ta);
  const double rS = rsq_nr(sS2), rE = rsq_nr(sE2), rT = rsq_nr(tt2);
  const double sb = ud.uSb * rS, sc = ud.uSc * rS, sa = ud.uSa * rS;  // s^
  const double eb = ud.uEb * rE, ec = ud.uEc * rE, ea = ud.uEa * rE;  // e^
  const double qb = tb * rT, qc = tc * rT, qa = ta * rT;              // t^
  const float hb = (float)(sb + eb), hc = (float)(sc + ec), ha = (float)(sa + ea);
  const float pb = (float)(sb - qb), pc = (float)(sc - qc), pa = (float)(sa - qa);
  const float wb = (float)(eb - qb), wc = (float)(ec - qc), wa = (float)(ea - qa);
  const float sS = (float)(sS2 * rS), sE = (float)(sE2 * rE), isc = (float)rT;
  const float hh = dot3(hb, hc, ha, hb, hc, ha);
  const float rh = __builtin_amdgcn_rsqf(hh);
  const float cdp = dot3((float)qb, (float)qc, (float)qa, hb, hc, ha) * rh;  // cos d'
  const float icd = __builtin_amdgcn_rcpf(cdp);
  const float x = (0.5f * (hh * rh)) * icd;
  const float pp = dot3(pb, pc, pa, pb, pc, pa), ww = dot3(wb, wc, wa, wb, wc, wa);
  const float y = (0.5f * (__builtin_amdgcn_sqrtf(pp) * __builtin_amdgcn_sqrtf(ww))) * icd;  // 1 - x
  // series
  const float xs = __builtin_fminf(x, MX_SERIES_X0);
  const float x2 = xs * xs;
  float acc = 2.0f / ((2.0f * MX_SERIES_TERMS - 1.0f) * (2.0f * MX_SERIES_TERMS + 1.0f));
#pragma unroll
  for (int k = MX_SERIES_TERMS - 1; k >= 1; --k) acc = fmx(acc, x2, 2.0f / ((2.0f * k - 1.0f) * (2.0f * k + 1.0f)));
  const float gs = (acc * x2) * xs;
  // closed form (y = 0 only at zero phase, where the limit is g = x)
  const float yc = __builtin_fmaxf(y, 1e-30f);
  const float art = (0.5f * LN2_F) * log2_f32((2.0f - y) * __builtin_amdgcn_rcpf(yc));  // artanh x
  const float gc = x - (y * (2.0f - y)) * art;
  const float g = x < MX_SERIES_X0 ? gs : gc;
  const float inner = ((2.0f * cdp) * cdp) * g;
  return M2P5_LOG10_2 * log2_f32x((((bcp * sS) * sE) * isc) * inner);
}

// Parameter BP (a first-segment position) of segment k: its own copy if it
// is one that each segment repeats, the shared one otherwise.
template <int VC, int BP>
AST_AI double segval(const double* params, int k) {
  using P = Pos<VC>;
  if constexpr (P::NSEG == 1 || P::rank(BP) < 0) {
    return params[BP];
  } else {
    double v = params[P::seg(0, BP)];
    if (k == 1) v = params[P::seg(1, BP)];
    if (k == 2) v = params[P::seg(2, BP)];
    if constexpr (P::NSEG > 3)
      if (k == 3) v = params[P::seg(3, BP)];
    return v;
  }
}

// The rotation at the start of segment k (k = 0: the start of the model;
// cuda.c:190-252): the spin frame M, XM, YM (XM_y = 0) from theta_M and
// phi_M, theta from Es and psi_0, the initial Euler angles (and, with
// torque, the angular velocity from L) and the ODE constants.
template <int VC>
AST_AI void begin_rotation(const double* params, int k, double Is_inv, double Ii_inv, double Is, double Ii,
                           SpinFrame<double>& Fd, double* y, double* mu, double& mu02, double* sc) {
  using P = Pos<VC>;
  constexpr bool TQ = VarF<VC>::TQ;
  double sTM, cTM, sPM, cPM;
  astm::sincos_t(segval<VC, P::theta_M>(params, k), &sTM, &cTM);
  astm::sincos_t(segval<VC, P::phi_M>(params, k), &sPM, &cPM);
  Fd.M_x = sTM * cPM;
  Fd.M_y = sTM * sPM;
  Fd.M_z = cTM;
  const double XM = __builtin_sqrt(__builtin_fma(Fd.M_x, Fd.M_x, Fd.M_z * Fd.M_z));
  Fd.XM_x = Fd.M_z / XM;
  Fd.XM_z = -Fd.M_x / XM;
  Fd.YM_x = Fd.M_y * Fd.XM_z;
  Fd.YM_y = __builtin_fma(-Fd.M_x, Fd.XM_z, Fd.M_z * Fd.XM_x);
  Fd.YM_z = -(Fd.M_y * Fd.XM_x);
  const double Es = segval<VC, P::Es>(params, k), psi0 = segval<VC, P::psi_0>(params, k);
  const double L = segval<VC, P::L>(params, k);
  double sP0, cP0;
  astm::sincos_t(psi0, &sP0, &cP0);
  const double theta = astm::asin(__builtin_sqrt((Es - 1.0) / (__builtin_fma(sP0 * sP0, Ii_inv - Is_inv, Is_inv) - 1.0)));
  const double phi = segval<VC, P::phi_0>(params, k);
  astm::sincos_t(theta, &sc[0], &sc[1]);
  sc[2] = sP0;
  sc[3] = cP0;
  if constexpr (TQ) {
    mu[0] = (Is - 1.0) * Ii_inv;
    mu[1] = (1.0 - Ii) * Is_inv;
    mu[2] = Ii - Is;
    mu[3] = segval<VC, P::Ti>(params, k);
    mu[4] = segval<VC, P::Ts>(params, k);
    mu[5] = segval<VC, P::Tl>(params, k);
    y[0] = ((L * Ii_inv) * sc[0]) * sP0;
    y[1] = ((L * Is_inv) * sc[0]) * cP0;
    y[2] = L * sc[1];
    y[3] = phi;
    y[4] = theta;
    y[5] = psi0;
  } else {
    mu[0] = L;
    mu[1] = 0.5 * (Ii_inv + Is_inv);
    mu[2] = 0.5 * (Ii_inv - Is_inv);
    mu02 = mu[0] * mu[2];
    y[0] = phi;
    y[1] = theta;
    y[2] = psi0;
  }
}

// chi^2 of one model (chi2one, cuda.c:81-1011, in-scope branches), summed
// per filter; NF is the compile-time filter capacity (>= n_filters).  With MX
// (the mixed-precision form) the Euler angles and angular velocity are
// integrated exactly as in the FP64 form, and each observation's geometry
// and brightness run in FP32; the chi^2 sums stay FP64.
// FORM selects what the model is scored by: FORM_CHI2, chi^2 per degree of
// freedom (cuda.c:907-908); FORM_PLOT (chi2one called with Nplot > 0,
// cuda.c:699-703, 884-885), where the points are a plotting grid, each
// point's Vmod + dv0 goes to vplot[i * vstride], nothing is summed and the
// result is 0; FORM_RMSD (the RMSD build, cuda.c:890-906), the root of the
// summed filter chi^2 over the summed weights.
enum { FORM_CHI2 = 0, FORM_PLOT = 1, FORM_RMSD = 2 };
template <class CF, int VC, int NF, bool MX = false, int FORM = FORM_CHI2>
__device__ CF chi2_model(const double* params, PointPtr pts, int n_data, int n_filters, int dof,
                         const double* sum_w, CF* dV, double* vplot = nullptr, size_t vstride = 0,
                         double dv0 = 0.0) {
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  static_assert(!(TR && (T2 || NU)), "PointRec.aux is the phase angle under TREND and the time otherwise");
  double sy2[NF], sy[NF];
#pragma unroll
  for (int m = 0; m < NF; ++m) { sy2[m] = 0.0; sy[m] = 0.0; }
  const double bt = params[P::b_tumb], ct = params[P::c_tumb];
  const double den = __builtin_fma(ct, ct, bt * bt);
  const double Is = __builtin_fma(bt, bt, 1.0) / den;
  const double Ii = __builtin_fma(ct, ct, 1.0) / den;
  const double Is_inv = 1.0 / Is;
  const double Ii_inv = 1.0 / Ii;
  constexpr int NODE = TQ ? 6 : 3;
  SpinFrame<double> Fd;
  double mu[6];
  double mu02 = 0.0;
  double y[6];
  // sin/cos of the current theta and psi, shared by the next RK4 step's
  // first stage and the body axes at the observation.
  double sc[4];
  begin_rotation<VC>(params, 0, Is_inv, Ii_inv, Is, Ii, Fd, y, mu, mu02, sc);
  double phi, theta, psi;
  ModelConsts K{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if constexpr (RO) {
    astm::sincos_t(params[P::phi_R], &K.sR, &K.cR);
    astm::sincos_t(params[P::theta_R], &K.stR, &K.ctR);
    astm::sincos_t(params[P::psi_R], &K.sPR, &K.cPR);
  }
  if (BW) {
    astm::sincos_t(params[P::phi_R], &K.sR, &K.cR);
    astm::sincos_t(params[P::theta_R], &K.stR, &K.ctR);
    K.kappa = params[P::kappa];
  } else {
    const double b = BC ? params[P::b] : bt;
    const double c = BC ? params[P::c] : ct;
    K.ib = 1.0 / b;
    K.ic = 1.0 / c;
    K.bcp = b * c;
  }
  const float bcpf = (float)K.bcp;  // mixed form
  float nu_t[NU ? NU_M_MAX : 1], nu_V[NU ? NU_M_MAX : 1];
  NudgeState nu;
  nu.t_mod = nu_t;
  nu.V_mod = nu_V;
  TailPtr mr = (TailPtr)(pts + n_data);
  double trendA = 0.0, t_split = 0.0;
  if constexpr (TR) trendA = params[P::A];
  bool switched = false;
  if constexpr (T2)  // cuda.c:291: the torque changes at this fraction of the span
    t_split = __builtin_fma(params[P::Tt], pts[n_data - 1].aux - pts[0].aux, pts[0].aux);
  // The next interval's RK4 constants are read one point ahead, so their
  // scalar loads complete behind this point's work instead of stalling the
  // next integration (pts[n_data], the tail record, keeps the look-ahead
  // inside the table; its values are never used).
  int nx_steps = pts[0].nsteps;  // 0 for the first observation
  double nx_h = pts[0].h, nx_hh = pts[0].hh, nx_h6 = pts[0].h6;
  // The observation's own fields are read one point ahead too, and all of
  // the next record is read right after the body axes: the scalar loads and
  // the LDS table reads share one wait counter, so loads issued anywhere
  // else would be waited for at the next sin/cos, a few instructions on,
  // while after the axes the point's brightness and chi^2 (no table read)
  // cover their latency.
  double nx_E[3] = {pts[0].E[0], pts[0].E[1], pts[0].E[2]};
  double nx_S[3] = {pts[0].S[0], pts[0].S[1], pts[0].S[2]};
  float nx_V = pts[0].V, nx_w = pts[0].w;
  int nx_filter = pts[0].filter;
  for (int i = 0; i < n_data; ++i) {
    const auto& pr = pts[i];
    int nsteps = nx_steps;
    double h = nx_h, hh = nx_hh, h6 = nx_h6;
    const double E[3] = {nx_E[0], nx_E[1], nx_E[2]};
    const double S[3] = {nx_S[0], nx_S[1], nx_S[2]};
    const float pV = nx_V, pw = nx_w;
    const int pfilter = nx_filter;
    if constexpr (SG) {
      // The first point of segment k (its record's step count is -1 - k;
      // a table that opens with segment 0 has no marker there): the
      // rotation starts again from the segment's own parameters (cuda.c:163-252
      // inside the segment loop); the shape, and so the brightness
      // constants, are shared.
      if (nsteps < 0) begin_rotation<VC>(params, -1 - nsteps, Is_inv, Ii_inv, Is, Ii, Fd, y, mu, mu02, sc);
    }
    int nsub = 1;
    double t2_end = 0.0;
    if constexpr (T2) {
      // The interval holding t_split runs as two sub-intervals, the second
      // under the second torque (cuda.c:288-317); the torque then stays.
      if (i > 0 && !switched && t_split >= pts[i - 1].aux && t_split < pr.aux) {
        const double t1 = pts[i - 1].aux;
        nsub = 2;
        t2_end = pr.aux;
        nsteps = (int)((t_split - t1) / mr->ts + 1);
        h = (t_split - t1) / nsteps;
        hh = 0.5 * h;
        h6 = 1 / 6.0 * h;
      }
    }
    for (int isub = 0; isub < nsub; ++isub) {
    if constexpr (T2) {
      if (isub == 1) {
        nsteps = (int)((t2_end - t_split) / mr->ts + 1);
        h = (t2_end - t_split) / nsteps;
        hh = 0.5 * h;
        h6 = 1 / 6.0 * h;
        mu[3] = params[P::T2i];
        mu[4] = params[P::T2s];
        mu[5] = params[P::T2l];
        switched = true;
      }
    }
    for (int l = 0; l < nsteps; ++l) {  // RK4, cuda.c:348-372
      double f[6], K4[6], s[6];
      ode<TQ, true>(y, K4, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) { s[j] = K4[j]; f[j] = __builtin_fma(hh, K4[j], y[j]); }
      ode<TQ, false>(f, K4, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) { s[j] = __builtin_fma(2.0, K4[j], s[j]); f[j] = __builtin_fma(hh, K4[j], y[j]); }
      ode<TQ, false>(f, K4, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) { s[j] = __builtin_fma(2.0, K4[j], s[j]); f[j] = __builtin_fma(h, K4[j], y[j]); }
      ode<TQ, false>(f, K4, mu, mu02, sc);
#pragma unroll
      for (int j = 0; j < NODE; ++j) y[j] = __builtin_fma(h6, s[j] + K4[j], y[j]);
      if (TQ) {
        astm::sincos_t(y[4], &sc[0], &sc[1]);
        astm::sincos_t(y[5], &sc[2], &sc[3]);
      } else {
        astm::sincos_t(y[1], &sc[0], &sc[1]);
      }
    }
    }  // sub-intervals
    if (TQ) { phi = y[3]; theta = y[4]; psi = y[5]; }
    else { phi = y[0]; theta = y[1]; psi = y[2]; }
    double a[3], b[3], c[3];
    axes_at<TQ>(Fd, phi, psi, sc, a, b, c);
    __builtin_amdgcn_sched_barrier(0);  // keep the next record's loads below the axes' table read
    nx_steps = pts[i + 1].nsteps;
    nx_h = pts[i + 1].h;
    nx_hh = pts[i + 1].hh;
    nx_h6 = pts[i + 1].h6;
#pragma unroll
    for (int k = 0; k < 3; ++k) { nx_E[k] = pts[i + 1].E[k]; nx_S[k] = pts[i + 1].S[k]; }
    nx_V = pts[i + 1].V;
    nx_w = pts[i + 1].w;
    nx_filter = pts[i + 1].filter;
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (RO) rotate_axes(a, b, c, K);
    double Vmod;
    if constexpr (RC) {
      Vmod = rect_mag(a, b, c, E, params[P::b], params[P::c]);
    } else if (BW) {
      const double arg = bw_arg_f64(a, b, c, E, K);
      Vmod = MX ? M2P5_LOG10_2 * log2_f32x((float)arg) : -2.5 * astm::log10(arg);
    } else {
      const ScaledDirs u = scaled_dirs(a, b, c, E, S, K);
      if (MX) {
        Vmod = ls_tail_mx(u, bcpf);
      } else {
        Vmod = ls_tail_f64(u, K.bcp);
      }
    }
    if constexpr (TR) Vmod = __builtin_fma(-trendA, pr.aux, Vmod);  // cuda.c:692-697
    // The next record's values are pinned in SGPRs here, after the
    // brightness (an empty asm that "changes" them): the compiler can then
    // neither sink their loads to this point nor drop them and read the
    // record again next to a table read.
    pin_sgpr(nx_steps);
    pin_sgpr(nx_h);
    pin_sgpr(nx_hh);
    pin_sgpr(nx_h6);
#pragma unroll
    for (int k = 0; k < 3; ++k) { pin_sgpr(nx_E[k]); pin_sgpr(nx_S[k]); }
    pin_sgpr(nx_V);
    pin_sgpr(nx_w);
    pin_sgpr(nx_filter);
    if constexpr (FORM == FORM_PLOT) {
      // Under NUDGE the reference stops the curve at a failed model
      // (cuda.c:735-738); the rest of the curve is NaN here.
      vplot[(size_t)i * vstride] = (NU && nu.failed) ? __builtin_nan("") : Vmod + dv0;
      if constexpr (NU) nu.point(i, pr.aux, Vmod, mr);
      continue;
    }
    const double w = (double)pw;
    const double yv = (double)pV - Vmod;
    const int m = pfilter;
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
    if constexpr (NU) nu.point(i, pr.aux, Vmod, mr);
  }
  if constexpr (FORM == FORM_PLOT) return CF(0.0);
  // A failed NUDGE model returns before delta_V is written (cuda.c:737).
  if constexpr (NU)
    if (nu.failed) return CF(1e30);
  CF chi2a = 0.0;  // cuda.c:884-910
  CF sum_all = 0.0;
#pragma unroll
  for (int m = 0; m < NF; ++m) {
    if (m < n_filters) {
      CF chi2m = sy2[m] - sy[m] * sy[m] / sum_w[m];
      chi2a = chi2a + chi2m;
      dV[m] = sy[m] / sum_w[m];
      if constexpr (FORM == FORM_RMSD) sum_all = sum_all + sum_w[m];
    }
  }
  if constexpr (FORM == FORM_RMSD) {
    // the root is taken in double: rounded to float it is the correctly
    // rounded float root
    chi2a = (CF)__builtin_sqrt((double)(chi2a / sum_all));
  } else {
    chi2a = chi2a / dof;
  }
  if constexpr (NU) chi2a = nu.merit<CF>(chi2a, dV[0], mr);
  return chi2a;
}

}  // namespace ast
