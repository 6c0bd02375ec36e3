// ast_capi.hip -- host side of libasteroid_hip.so (the C ABI of include/asteroid.h).
//
// Pieces:
//  * validation of model spaces and light curves, and the per-observation
//    table the kernels read (INTERP ephemeris evaluated once per observation
//    and the RK4 interval constants, in the canonical FMA order);
//  * a slot pool: the device state of every resident simplex, lane-fastest;
//  * device jobs: one curve + model space + random streams + per-group bests
//    + one descriptor per group;
//  * the cycle: one launch of the fitting kernel per kernel class (variant,
//    simplex precision, filter capacity) over every running group, each class
//    on its own stream, timed with HIP events;
//  * ast_run (the reference's chi2_gpu launch) = one job run to completion;
//  * the serving scheduler (ast_serve_*, ast_submit/poll/cancel): a thread
//    that admits jobs into free slots while their HBM footprint fits, runs
//    cycles over everything resident (continuous batching) and retires jobs
//    as they finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/asteroid.h"
#include "ast_kernels.h"

namespace ast {
template <int VC> hipError_t launch_variant(const LaunchArgs& a);
extern template hipError_t launch_variant<0>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BC>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BW>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BW>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_TREND>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ2>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BC | VF_ROT>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC | VF_ROT>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_RECT>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_RECT>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_NUDGE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC | VF_NUDGE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_SEG>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_SEG | VF_SEG4>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_SEG | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_SEG | VF_SEG4 | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_SEG>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_SEG | VF_SEG4>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_SEG | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_SEG | VF_SEG4 | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BC | VF_SEG>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BC | VF_SEG | VF_SEG4>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BC | VF_SEG | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC | VF_SEG>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC | VF_SEG | VF_SEG4>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC | VF_SEG | VF_SEGLE>(const LaunchArgs&);
extern template hipError_t launch_variant<VF_TQ | VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE>(const LaunchArgs&);
}  // namespace ast

using namespace ast;

static thread_local char g_hip_err[256] = "";

// Device memory of ast_chi2_batch_noalloc: the models, chi^2 and delta_V of
// up to AST_NOALLOC_MAX_MODELS models (6.4 MB), one copy per GPU, held until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// the device array: its address on each device comes from
// hipGetSymbolAddress (scratch_setup, called by ast_ctx_create with the
// device current) and only that address is passed to kernels and copies.
constexpr size_t NOALLOC_PARAMS = (size_t)AST_NOALLOC_MAX_MODELS * AST_MAX_PARAMS;
constexpr size_t NOALLOC_CHI2 = (size_t)AST_NOALLOC_MAX_MODELS;
constexpr size_t NOALLOC_DV = (size_t)AST_NOALLOC_MAX_MODELS * AST_MAX_FILTERS;
static __device__ __align__(256) double scratch[NOALLOC_PARAMS + NOALLOC_CHI2 + NOALLOC_DV];
constexpr int SCRATCH_MAX_DEVICES = 64;
static double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {nullptr};
static std::mutex g_scratch_mu;

static void set_err(const char* msg) { snprintf(g_hip_err, sizeof g_hip_err, "%s", msg); }

#define HIPCHK(call)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      snprintf(g_hip_err, sizeof g_hip_err, "%s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
      return e_ == hipErrorOutOfMemory ? AST_E_NOMEM : AST_E_HIP;                                      \
    }                                                                                                  \
  } while (0)

#define RETIF(x)            \
  do {                      \
    int r_ = (x);           \
    if (r_) return r_;      \
  } while (0)

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

template <class T>
int dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)p, count * sizeof(T)));
  return 0;
}
template <class T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// ------------------------------------------------------------ validation
// The variant codes built (one kernel family each), indexed by variant_index.
constexpr int VARIANT_CODES[] = {0, VF_TQ, VF_BC, VF_TQ | VF_BC, VF_BW, VF_TQ | VF_BW,
                                 VF_TQ | VF_TREND, VF_TQ2, VF_BC | VF_ROT, VF_TQ | VF_BC | VF_ROT,
                                 VF_RECT, VF_TQ | VF_RECT, VF_TQ | VF_NUDGE, VF_TQ | VF_BC | VF_NUDGE,
                                 VF_SEG, VF_SEG | VF_SEG4, VF_SEG | VF_SEGLE, VF_SEG | VF_SEG4 | VF_SEGLE,
                                 VF_TQ | VF_SEG, VF_TQ | VF_SEG | VF_SEG4, VF_TQ | VF_SEG | VF_SEGLE,
                                 VF_TQ | VF_SEG | VF_SEG4 | VF_SEGLE,
                                 VF_BC | VF_SEG, VF_BC | VF_SEG | VF_SEG4, VF_BC | VF_SEG | VF_SEGLE,
                                 VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE, VF_TQ | VF_BC | VF_SEG,
                                 VF_TQ | VF_BC | VF_SEG | VF_SEG4, VF_TQ | VF_BC | VF_SEG | VF_SEGLE,
                                 VF_TQ | VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE};
constexpr int N_VARIANTS = (int)(sizeof VARIANT_CODES / sizeof VARIANT_CODES[0]);

// Variant code of the macros (TORQUE2 implies TORQUE, RECT implies BC).
int variant_code(const ast_variant& v) {
  const bool t2 = v.torque2, tq = v.torque || t2, rc = v.rect, bc = v.bc || rc;
  return (tq && !t2 ? VF_TQ : 0) | (t2 ? VF_TQ2 : 0) | (bc && !rc ? VF_BC : 0) | (v.bw_ball ? VF_BW : 0) |
         (v.trend ? VF_TREND : 0) | (v.rotate ? VF_ROT : 0) | (rc ? VF_RECT : 0) | (v.nudge ? VF_NUDGE : 0) |
         (v.n_seg > 1 ? VF_SEG | (v.n_seg == 4 ? VF_SEG4 : 0) | (v.one_le ? 0 : VF_SEGLE) : 0);
}

// Index of the variant's kernel family, or -1 if it is not built.
int variant_index(const ast_variant& v) {
  const int code = variant_code(v);
  for (int k = 0; k < N_VARIANTS; ++k)
    if (VARIANT_CODES[k] == code) return k;
  return -1;
}

bool variant_supported(const ast_variant& v) {
  if (variant_index(v) < 0) return false;
  // NUDGE's merit jumps where a model minimum appears or vanishes, so the
  // mixed form's 1e-6 chi^2 tolerance cannot hold for it; the reference's
  // Stage Three runs -DACC (README.md:89-93).
  if (v.nudge && v.precision == AST_PREC_MIXED) return false;
  // SEGMENT is built for three or four segments (the reference's T_START
  // tables, asteroid.h:40-46), ONE_LE or not, on the LS ellipsoid with or
  // without constant torque and photometric b, c
  if (v.n_seg < 0 || v.n_seg > AST_MAX_SEGMENTS || v.n_seg == 2) return false;
  if (v.n_seg > 1 && (v.bw_ball || v.trend || v.torque2 || v.rotate || v.rect || v.nudge)) return false;
  if (v.precision != AST_PREC_FP64 && v.precision != AST_PREC_F32SIMPLEX && v.precision != AST_PREC_MIXED)
    return false;
  if (!(v.time_step > 0)) return false;
  return true;
}

struct SpaceH {  // a validated model space
  ast_variant v{};
  ast_type_ids T{};
  int n_params = 0, n_types = 0;
  int prop[AST_MAX_PARAMS][AST_N_COLUMNS]{};
  double lim[2][AST_MAX_TYPES]{};
  ast_x2 x2{};
};

int make_space(const ast_variant& v, int n_params, const int* property, int n_types, const int* types,
               const double* limits, const ast_x2* x2, SpaceH* out) {
  if (!property || !types || !limits || !x2) return AST_E_ARG;
  if (!variant_supported(v)) return AST_E_UNSUPPORTED;
  SpaceH S;
  S.v = v;
  ast_variant_type_ids(&v, &S.T);
  if (n_params != S.T.n_params || n_types != S.T.n_types) return AST_E_ARG;
  // The kernels fix each parameter's role by its position: the Property
  // order of the reference (asteroid.c:86-125).
  const int expect[] = {S.T.theta_M, S.T.phi_M, S.T.phi_0, S.T.A, S.T.Ti, S.T.Ts, S.T.Tl, S.T.T2i,
                        S.T.T2s, S.T.T2l, S.T.Tt, S.T.c_tumb, S.T.b_tumb, S.T.Es, S.T.L, S.T.psi_0,
                        S.T.c, S.T.b, S.T.theta_R, S.T.phi_R, S.T.psi_R, S.T.kappa};
  int pos = 0;
  for (int k = 0; k < (int)(sizeof expect / sizeof expect[0]); ++k) {
    if (expect[k] < 0) continue;
    if (property[pos * AST_N_COLUMNS + AST_P_TYPE] != expect[k]) return AST_E_UNSUPPORTED;
    if (types[expect[k]] != pos) return AST_E_UNSUPPORTED;
    int fr = property[pos * AST_N_COLUMNS + AST_P_FROZEN];
    if (fr < -1 || fr > 1) return AST_E_ARG;
    pos++;
  }
  // SEGMENT (asteroid.c:128-150): the first segment's shape parameters
  // (c_tumb, b_tumb, and the photometric c, b under BC) and under ONE_LE its
  // Es and L are multi-segment (asteroid.c:88-120), the rest repeat for each
  // further segment in the same order
  if (v.n_seg > 1) {
    std::vector<int> rep;
    for (int i = 0; i < pos; ++i) {
      const int t = property[i * AST_N_COLUMNS + AST_P_TYPE];
      const bool shared = t == S.T.c_tumb || t == S.T.b_tumb || (S.T.c >= 0 && t == S.T.c) ||
                          (S.T.b >= 0 && t == S.T.b) || (v.one_le && (t == S.T.Es || t == S.T.L));
      if (property[i * AST_N_COLUMNS + AST_P_MULTI_SEGMENT] != (shared ? 1 : 0)) return AST_E_UNSUPPORTED;
      if (property[i * AST_N_COLUMNS + AST_P_ISEG] != 0) return AST_E_UNSUPPORTED;
      if (!shared) rep.push_back(t);
    }
    for (int k = 1; k < v.n_seg; ++k)
      for (int t : rep) {
        if (pos >= n_params) return AST_E_UNSUPPORTED;
        const int* r = property + pos * AST_N_COLUMNS;
        if (r[AST_P_TYPE] != t || r[AST_P_ISEG] != k || r[AST_P_MULTI_SEGMENT] != 0) return AST_E_UNSUPPORTED;
        if (r[AST_P_FROZEN] < -1 || r[AST_P_FROZEN] > 1) return AST_E_ARG;
        pos++;
      }
  }
  if (pos != n_params) return AST_E_UNSUPPORTED;
  for (int i = 0; i < n_params; ++i)
    for (int k = 0; k < AST_N_COLUMNS; ++k) S.prop[i][k] = property[i * AST_N_COLUMNS + k];
  for (int t = 0; t < n_types; ++t) {
    S.lim[0][t] = limits[t];
    S.lim[1][t] = limits[n_types + t];
  }
  S.n_params = n_params;
  S.n_types = n_types;
  S.x2 = *x2;
  *out = S;
  return AST_OK;
}

template <class CF>
Space<CF> space_table(const SpaceH& S) {
  Space<CF> sp;
  std::memset(&sp, 0, sizeof sp);
  for (int i = 0; i < S.n_params; ++i) {
    int t = S.prop[i][AST_P_TYPE];
    sp.frozen[i] = (int8_t)S.prop[i][AST_P_FROZEN];
    sp.periodic[i] = (int8_t)S.prop[i][AST_P_PERIODIC];
    sp.independent[i] = (int8_t)S.prop[i][AST_P_INDEPENDENT];
    sp.lim0[i] = (CF)S.lim[0][t];
    sp.lim1[i] = (CF)S.lim[1][t];
  }
  sp.p_psi = S.v.p_psi;
  sp.ppsi1 = S.x2.ppsi1;
  sp.ppsi2 = S.x2.ppsi2;
  return sp;
}

// Canonical INTERP evaluation at time t (the oracle's interp_canon).
void interp_canon(const ast_ephem3& e, double t, double E[3], double S[3]) {
  const double* M = e.mjd;
  double d0 = t - M[0], d1 = t - M[1], d2 = t - M[2];
  double rr0 = (d1 * d2) / (M[0] - M[1]) / (M[0] - M[2]);
  double rr1 = (d0 * d2) / (M[1] - M[0]) / (M[1] - M[2]);
  double rr2 = (d0 * d1) / (M[2] - M[0]) / (M[2] - M[1]);
  for (int k = 0; k < 3; ++k) {
    E[k] = std::fma(e.E[k][2], rr2, std::fma(e.E[k][1], rr1, e.E[k][0] * rr0));
    S[k] = std::fma(e.S[k][2], rr2, std::fma(e.S[k][1], rr1, e.S[k][0] * rr0));
  }
  double En = std::sqrt(std::fma(E[2], E[2], std::fma(E[1], E[1], E[0] * E[0])));
  double Sn = std::sqrt(std::fma(S[2], S[2], std::fma(S[1], S[1], S[0] * S[0])));
  for (int k = 0; k < 3; ++k) { E[k] = E[k] / En; S[k] = S[k] / Sn; }
}

struct CurveH {  // a validated light curve as the kernels read it
  std::vector<PointRec> pts;  // n_data observations, then the minima record (zero unless NUDGE)
  double sum_w[AST_MAX_FILTERS] = {0};
  int n_data = 0, n_filters = 0;
  int n_all = 0;          // observations given (SEGMENT leaves those before the first segment out of the table)
  int64_t rk4_steps = 0;  // RK4 steps of one evaluation
  double t_first = 0, t_last = 0;  // times of the first and last observation
  PointRec first{};                // the first observation (the plotting grid starts there)
};

// Phase angle of unit Sun and Earth directions, acos(S.E) in the canonical
// form (TREND, cuda.c:694): a property of the point, computed once.
double phase_angle(const double E[3], const double S[3]) {
  return astm::acos(std::fma(S[2], E[2], std::fma(S[1], E[1], S[0] * E[0])));
}

// SEGMENT's split (read_data.c:244-250, and 357-363 for the plotting grid):
// point i opens segment s when its time is at or after seg_t[s].  Returns
// the number of segments opened.
int segment_starts(int n_seg, int n, const double* t, const double* seg_t, int* start) {
  int k = 0;
  for (int i = 0; i < n; ++i)
    if (k < n_seg && t[i] >= seg_t[k]) start[k++] = i;
  return k;
}

// The observation table keeps the observations from the first segment on;
// the first record of segment k > 0 carries the step count -1 - k, where the
// evaluator starts that segment's rotation, and the first record none.
// sum_w covers the table; the degrees of freedom keep every observation
// (n_all, cuda.c:907).
// With seg_first (the caller's split, e.g. on the observer's clock as
// read_data.c:246 does, before the light-time correction) the segments open
// at those observation indices, which must increase and lie in the table.
int split_segments(int n_seg, const ast_obs* obs, const double* seg_t, const int* seg_first, CurveH* C) {
  int start[AST_MAX_SEGMENTS];
  if (seg_first) {
    for (int k = 0; k < n_seg; ++k)
      if (seg_first[k] < 0 || seg_first[k] >= C->n_data || (k > 0 && seg_first[k] <= seg_first[k - 1]))
        return AST_E_DATA;
    for (int k = 0; k < n_seg; ++k) start[k] = seg_first[k];
  } else {
    if (!seg_t) return AST_E_STATE;
    std::vector<double> t(C->n_data);
    for (int i = 0; i < C->n_data; ++i) t[i] = obs[i].mjd;
    if (segment_starts(n_seg, C->n_data, t.data(), seg_t, start) < n_seg) return AST_E_DATA;
  }
  const int i0 = start[0], n = C->n_data - i0;
  std::vector<PointRec> pts(C->pts.begin() + i0, C->pts.end());  // the tail record comes along
  for (int m = 0; m < C->n_filters; ++m) C->sum_w[m] = 0.0;
  C->rk4_steps = 0;
  for (int i = 0; i < n; ++i) {
    PointRec& p = pts[i];
    for (int s = 0; s < n_seg; ++s)
      if (start[s] == i + i0) {
        p.nsteps = s == 0 ? 0 : -1 - s;
        p.h = p.hh = p.h6 = 0.0;
      }
    if (p.nsteps > 0) C->rk4_steps += p.nsteps;
    C->sum_w[p.filter] = C->sum_w[p.filter] + (double)p.w;
  }
  for (int m = 0; m < C->n_filters; ++m)
    if (!(C->sum_w[m] > 0)) return AST_E_DATA;  // a filter with no observation in the segments
  C->pts = std::move(pts);
  C->n_data = n;
  return AST_OK;
}

int make_curve(const ast_variant& v, const ast_obs* obs, int n_data, int n_filters, const ast_ephem3* eph,
               CurveH* out, const double* seg_t = nullptr, const int* seg_first = nullptr) {
  if (!obs || n_data <= 0 || n_filters <= 0) return AST_E_ARG;
  if (n_filters > AST_MAX_FILTERS) return AST_E_DATA;
  if (v.interp && !eph) return AST_E_ARG;
  CurveH C;
  C.pts.assign((size_t)n_data + 1, PointRec{});
  std::vector<int> seen(n_filters, 0);
  for (int i = 0; i < n_data; ++i) {
    const ast_obs& o = obs[i];
    if (o.filter < 0 || o.filter >= n_filters) return AST_E_DATA;
    if (i > 0 && !(o.mjd > obs[i - 1].mjd)) return AST_E_DATA;  // read_data.c:81-85
    if (!std::isfinite(o.mjd) || !std::isfinite(o.V) || !(o.w > 0)) return AST_E_DATA;
    PointRec& p = C.pts[i];
    std::memset(&p, 0, sizeof p);
    if (v.interp) interp_canon(*eph, o.mjd, p.E, p.S);
    else for (int k = 0; k < 3; ++k) { p.E[k] = o.E[k]; p.S[k] = o.S[k]; }
    if (i > 0) {  // cuda.c:321-323
      double t1 = obs[i - 1].mjd, t2 = o.mjd;
      int nsteps = (int)((t2 - t1) / v.time_step + 1);
      double h = (t2 - t1) / nsteps;
      p.nsteps = nsteps;
      C.rk4_steps += nsteps;
      p.h = h;
      p.hh = 0.5 * h;
      p.h6 = 1 / 6.0 * h;
    }
    p.V = o.V;
    p.w = o.w;
    p.filter = o.filter;
    p.aux = v.trend ? phase_angle(p.E, p.S) : o.mjd;
    C.sum_w[o.filter] = C.sum_w[o.filter] + (double)o.w;
    seen[o.filter] = 1;
  }
  for (int m = 0; m < n_filters; ++m)
    if (!seen[m]) return AST_E_DATA;
  TailRec tail{};
  tail.ts = v.time_step;
  std::memcpy(&C.pts[n_data], &tail, sizeof tail);
  C.n_data = n_data;
  C.n_all = n_data;
  C.n_filters = n_filters;
  C.first = C.pts[0];
  C.t_first = obs[0].mjd;
  C.t_last = obs[n_data - 1].mjd;
  if (v.n_seg > 1) RETIF(split_segments(v.n_seg, obs, seg_t, seg_first, &C));
  *out = std::move(C);
  return AST_OK;
}

// The plotting grid (read_data.c:304-360): n_plot times evenly spaced from
// the first to the last observation.  The two ends are the first and last
// observations themselves, E and S included; the interior points take E,S
// from the 3-point ephemeris (read_data.c:345, the INTERP quadratic, then
// normalised).  RK4 intervals between grid points as between observations
// (cuda.c:321-323).  V, w and the filter are unused on a grid.
int make_grid(const ast_variant& v, const CurveH& C, const ast_ephem3& eph, int n_plot, std::vector<PointRec>* out,
              std::vector<double>* t_out, const double* seg_t = nullptr, int* g0 = nullptr) {
  if (n_plot < 2 || C.n_data < 1) return AST_E_ARG;
  std::vector<PointRec> g((size_t)n_plot + 1);
  g[n_plot] = C.pts[C.n_data];  // the minima record
  std::vector<double> t((size_t)n_plot);
  const double h = (C.t_last - C.t_first) / (n_plot - 1);
  for (int i = 0; i < n_plot; ++i) {
    PointRec& p = g[i];
    std::memset(&p, 0, sizeof p);
    if (i == 0 || i == n_plot - 1) {
      const PointRec& d = i == 0 ? C.first : C.pts[C.n_data - 1];
      t[i] = i == 0 ? C.t_first : C.t_last;
      for (int k = 0; k < 3; ++k) { p.E[k] = d.E[k]; p.S[k] = d.S[k]; }
    } else {
      t[i] = C.t_first + (double)i * h;
      interp_canon(eph, t[i], p.E, p.S);
    }
    p.aux = v.trend ? phase_angle(p.E, p.S) : t[i];
    if (i > 0) {
      const double t1 = t[i - 1], t2 = t[i];
      const int nsteps = (int)((t2 - t1) / v.time_step + 1);
      const double hs = (t2 - t1) / nsteps;
      p.nsteps = nsteps;
      p.h = hs;
      p.hh = 0.5 * hs;
      p.h6 = 1 / 6.0 * hs;
    }
  }
  if (g0) *g0 = 0;
  if (v.n_seg > 1) {  // SEGMENT: each segment's first grid point restarts the rotation
    int start[AST_MAX_SEGMENTS];
    if (!seg_t || !g0) return AST_E_STATE;
    if (segment_starts(v.n_seg, n_plot, t.data(), seg_t, start) < v.n_seg) return AST_E_DATA;
    for (int s = 0; s < v.n_seg; ++s)
      if (start[s] > 0) {
        PointRec& p = g[start[s]];
        p.nsteps = -1 - s;
        p.h = p.hh = p.h6 = 0.0;
      }
    *g0 = start[0];  // grid points before it are not computed (their magnitudes are NaN)
  }
  *out = std::move(g);
  if (t_out) *t_out = std::move(t);
  return AST_OK;
}

// NUDGE's observed minima as a tail record (misc.c:279-323 fills the same
// table): 1..AST_MAX_MINIMA finite times and magnitudes.
int make_minima(const ast_variant& v, const float* t_obs, const float* V_obs, int n_obs, TailRec* out) {
  if (!t_obs || !V_obs || n_obs < 1 || n_obs > AST_MAX_MINIMA) return AST_E_ARG;
  TailRec r{};
  r.n_obs = n_obs;
  r.v1s = v.v1s ? 1 : 0;
  for (int k = 0; k < n_obs; ++k) {
    if (!std::isfinite(t_obs[k]) || !std::isfinite(V_obs[k])) return AST_E_ARG;
    r.t_obs[k] = t_obs[k];
    r.V_obs[k] = V_obs[k];
  }
  *out = r;
  return AST_OK;
}

// The curve's tail record with a job's minima (TIME_STEP kept).
int set_tail_minima(const ast_variant& v, const float* t_obs, const float* V_obs, int n_obs, CurveH* C) {
  TailRec r{};
  RETIF(make_minima(v, t_obs, V_obs, n_obs, &r));
  r.ts = v.time_step;
  std::memcpy(&C->pts[C->n_data], &r, sizeof r);
  return AST_OK;
}

int nf_class(int n_filters) { return n_filters <= 1 ? 1 : (n_filters <= 4 ? 4 : 8); }

hipError_t dispatch(const ast_variant& v, LaunchArgs& a) {
  a.precision = v.precision;
  switch (variant_code(v)) {
    case 0: return launch_variant<0>(a);
    case VF_BC: return launch_variant<VF_BC>(a);
    case VF_BW: return launch_variant<VF_BW>(a);
    case VF_TQ: return launch_variant<VF_TQ>(a);
    case VF_TQ | VF_BC: return launch_variant<VF_TQ | VF_BC>(a);
    case VF_TQ | VF_BW: return launch_variant<VF_TQ | VF_BW>(a);
    case VF_TQ | VF_TREND: return launch_variant<VF_TQ | VF_TREND>(a);
    case VF_TQ2: return launch_variant<VF_TQ2>(a);
    case VF_BC | VF_ROT: return launch_variant<VF_BC | VF_ROT>(a);
    case VF_TQ | VF_BC | VF_ROT: return launch_variant<VF_TQ | VF_BC | VF_ROT>(a);
    case VF_RECT: return launch_variant<VF_RECT>(a);
    case VF_TQ | VF_RECT: return launch_variant<VF_TQ | VF_RECT>(a);
    case VF_TQ | VF_NUDGE: return launch_variant<VF_TQ | VF_NUDGE>(a);
    case VF_TQ | VF_BC | VF_NUDGE: return launch_variant<VF_TQ | VF_BC | VF_NUDGE>(a);
    case VF_SEG: return launch_variant<VF_SEG>(a);
    case VF_SEG | VF_SEG4: return launch_variant<VF_SEG | VF_SEG4>(a);
    case VF_SEG | VF_SEGLE: return launch_variant<VF_SEG | VF_SEGLE>(a);
    case VF_SEG | VF_SEG4 | VF_SEGLE: return launch_variant<VF_SEG | VF_SEG4 | VF_SEGLE>(a);
    case VF_TQ | VF_SEG: return launch_variant<VF_TQ | VF_SEG>(a);
    case VF_TQ | VF_SEG | VF_SEG4: return launch_variant<VF_TQ | VF_SEG | VF_SEG4>(a);
    case VF_TQ | VF_SEG | VF_SEGLE: return launch_variant<VF_TQ | VF_SEG | VF_SEGLE>(a);
    case VF_TQ | VF_SEG | VF_SEG4 | VF_SEGLE: return launch_variant<VF_TQ | VF_SEG | VF_SEG4 | VF_SEGLE>(a);
    case VF_BC | VF_SEG: return launch_variant<VF_BC | VF_SEG>(a);
    case VF_BC | VF_SEG | VF_SEG4: return launch_variant<VF_BC | VF_SEG | VF_SEG4>(a);
    case VF_BC | VF_SEG | VF_SEGLE: return launch_variant<VF_BC | VF_SEG | VF_SEGLE>(a);
    case VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE: return launch_variant<VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE>(a);
    case VF_TQ | VF_BC | VF_SEG: return launch_variant<VF_TQ | VF_BC | VF_SEG>(a);
    case VF_TQ | VF_BC | VF_SEG | VF_SEG4: return launch_variant<VF_TQ | VF_BC | VF_SEG | VF_SEG4>(a);
    case VF_TQ | VF_BC | VF_SEG | VF_SEGLE: return launch_variant<VF_TQ | VF_BC | VF_SEG | VF_SEGLE>(a);
    case VF_TQ | VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE: return launch_variant<VF_TQ | VF_BC | VF_SEG | VF_SEG4 | VF_SEGLE>(a);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------------------------- slot pool
// Device state of up to `nf` resident simplexes.  Element storage is 8 bytes
// wide so that float and double simplexes share the pool; arrays have room
// for P parameters, the largest parameter count of what the pool was sized
// for (a kernel class packs its rows by its own count, so a class of N
// parameters touches (N + 1) N simplex rows whatever P is).
constexpr int MAX_CLASSES = N_VARIANTS * 3 * 3;  // variants x 3 filter capacities x 3 precision forms
constexpr int NSTREAMS = 4;      // classes share streams round-robin
constexpr size_t slot_bytes(size_t P) {
  return 8 * ((P + 1) * P + (P + 1) + 4 * P + AST_MAX_FILTERS + 2) + (P + 1) + 4 * 7;
}

struct Pool {
  size_t nf = 0;        // fits
  int G = 0;            // group size the slots are cut for
  int n_slots = 0;      // nf / G
  int P = 0;            // parameter capacity of a slot
  void *X = nullptr, *F = nullptr, *X0 = nullptr, *XR = nullptr, *XP = nullptr, *DV = nullptr;
  void *fr = nullptr, *sf = nullptr;
  uint8_t* IND = nullptr;
  double* PRM = nullptr;
  int32_t *ctl = nullptr, *lstep = nullptr, *nev = nullptr, *nev_tot = nullptr, *steps_tot = nullptr;
  int32_t *fitid = nullptr, *work = nullptr;
  int32_t *g_reopt = nullptr, *g_stage = nullptr, *g_done = nullptr;
  int32_t* job_done = nullptr;     // [n_slots]: a job's finished-group counter sits at its first slot
  std::vector<int32_t> job_done_h; // host copy, read back once per cycle
  std::vector<int> free_slots;
};

void pool_free(Pool& p) {
  dev_free(p.X); dev_free(p.F); dev_free(p.X0); dev_free(p.XR); dev_free(p.XP); dev_free(p.DV);
  dev_free(p.fr); dev_free(p.sf); dev_free(p.IND); dev_free(p.PRM); dev_free(p.ctl); dev_free(p.lstep);
  dev_free(p.nev); dev_free(p.nev_tot); dev_free(p.steps_tot); dev_free(p.fitid); dev_free(p.work);
  dev_free(p.g_reopt); dev_free(p.g_stage);
  dev_free(p.g_done); dev_free(p.job_done);
  p.nf = 0;
  p.n_slots = 0;
  p.P = 0;
  p.free_slots.clear();
}

int pool_alloc(Pool& p, size_t nf, int G, int n_params) {
  pool_free(p);
  if (n_params <= 0 || n_params > AST_MAX_PARAMS) return AST_E_ARG;
  const size_t P = (size_t)n_params;
  int rc = 0;
#define ALLOC(ptr, type, count) if (!rc) rc = dev_alloc((type**)&(ptr), (count))
  ALLOC(p.X, double, (P + 1) * P * nf);
  ALLOC(p.F, double, (P + 1) * nf);
  ALLOC(p.X0, double, P * nf);
  ALLOC(p.XR, double, P * nf);
  ALLOC(p.XP, double, P * nf);
  ALLOC(p.DV, double, AST_MAX_FILTERS * nf);
  ALLOC(p.fr, double, nf);
  ALLOC(p.sf, double, nf);
  ALLOC(p.IND, uint8_t, (P + 1) * nf);
  ALLOC(p.PRM, double, P * nf);
  ALLOC(p.ctl, int32_t, nf);
  ALLOC(p.lstep, int32_t, nf);
  ALLOC(p.nev, int32_t, nf);
  ALLOC(p.nev_tot, int32_t, nf);
  ALLOC(p.steps_tot, int32_t, nf);
  ALLOC(p.fitid, int32_t, nf);
  ALLOC(p.work, int32_t, nf);
  ALLOC(p.g_reopt, int32_t, nf / G);
  ALLOC(p.g_stage, int32_t, nf / G);
  ALLOC(p.g_done, int32_t, nf / G);
  ALLOC(p.job_done, int32_t, nf / G);
#undef ALLOC
  if (rc) { pool_free(p); return rc; }
  p.nf = nf;
  p.G = G;
  p.P = n_params;
  p.n_slots = (int)(nf / G);
  p.job_done_h.assign((size_t)p.n_slots, 0);
  p.free_slots.clear();
  for (int s = p.n_slots - 1; s >= 0; --s) p.free_slots.push_back(s);
  return 0;
}

template <class CF>
SlotBufs<CF> pool_bufs(const Pool& p, int cls) {
  SlotBufs<CF> B;
  B.X = (CF*)p.X; B.F = (CF*)p.F; B.IND = p.IND; B.X0 = (CF*)p.X0; B.XR = (CF*)p.XR; B.XP = (CF*)p.XP;
  B.PRM = p.PRM; B.DV = (CF*)p.DV; B.ctl = p.ctl; B.lstep = p.lstep; B.nev = p.nev; B.nev_tot = p.nev_tot;
  B.steps_tot = p.steps_tot; B.fr = (CF*)p.fr; B.sf = (CF*)p.sf; B.fitid = p.fitid; B.work = p.work;
  B.g_reopt = p.g_reopt; B.g_stage = p.g_stage;
  B.g_done = p.g_done; B.nf = p.nf;
  return B;
}

// --------------------------------------------------------------- jobs
struct RunParams {
  int n_groups = 0, group_size = 0, nstages = 1, n_steps = 0;
  int64_t eval_budget = 0;
  double size_min = 0;
  uint64_t seed = 0;
};

struct DevJob {
  uint64_t id = 0, tag = 0;
  uint64_t uid = 0;            // unique within the context (launch-batch membership)
  SpaceH space;
  RunParams rp;
  int n_data = 0, n_filters = 0;
  int n_all = 0;               // observations in the degrees of freedom
  int cls = 0;                 // kernel class index (for streams and counters)
  double cost = 1;             // one evaluation, in observation-equivalents (eval_cost)
  int ticks = 1;               // evaluation rounds per launch
  int cf = 8;                  // simplex float bytes
  // Workgroups the job runs on (its slots): every group (n_workers =
  // rp.n_groups), or fewer for a single-stage job, whose lanes then take the
  // job's fits from a device queue as their fits end (refill).
  int n_workers = 0;
  bool refill = false;
  bool reduced = false;        // refill: the groups' results have been formed
  // device memory (owned unless borrowed)
  PointRec* pts = nullptr;
  double* sum_w = nullptr;
  void* sp = nullptr;          // Space<CF>
  double* params0 = nullptr;
  uint32_t* rng = nullptr;
  void* best_f = nullptr;
  double* best_params = nullptr;
  double* best_dV = nullptr;
  int32_t* done = nullptr;     // finished-group counter
  void* fit_f = nullptr;       // per fit: stage-final best chi^2 (CF), evaluations, steps
  int32_t* fit_nev = nullptr;
  int32_t* fit_steps = nullptr;
  double* fit_params = nullptr;  // refill: per fit, its best point and last delta_V
  double* fit_dV = nullptr;
  int32_t* queue = nullptr;    // refill: next fit to start, fits finished
  void* descs = nullptr;       // GroupDesc<CF>[n_workers] (device, for the init kernel)
  std::vector<unsigned char> desc_host;  // the same descriptors on the host
  bool borrowed = false;       // rng and best arrays belong to the context
  std::vector<int> slots;
  size_t bytes = 0;
  int groups_done = 0;
  int status = AST_OK;
  bool cancelled = false;
  Clock::time_point t_submit, t_admit, t_done;
  double run_ms = 0;
};

void job_free(DevJob& j) {
  dev_free(j.pts); dev_free(j.sum_w); dev_free(j.sp); dev_free(j.params0);
  dev_free(j.fit_f); dev_free(j.fit_nev); dev_free(j.fit_steps); dev_free(j.fit_params); dev_free(j.fit_dV);
  dev_free(j.queue);
  j.done = nullptr;  // the pool's
  dev_free(j.descs);
  if (!j.borrowed) { dev_free(j.rng); dev_free(j.best_f); dev_free(j.best_params); dev_free(j.best_dV); }
}

int ticks_for(double cost, double work) {
  double t = work / (cost > 1 ? cost : 1);
  return t < 1 ? 1 : (t > 4096 ? 4096 : (int)t);
}

// Cost of one evaluation in observation-equivalents: the observations plus
// the RK4 steps weighted by a step's FLOPs over an observation's (the
// per_rk4 / per_point rows of the FLOP table: about 1.9 with torque, 1.4
// without, 2.8 for the cheaper Lambert + ball brightness).  Launch sizes
// follow it, so sparse curves (many RK4 steps per observation) and dense
// ones do similar work per launch.
double eval_cost(const ast_variant& v, const CurveH& C) {
  const double w = v.torque ? (v.bw_ball ? 2.8 : 1.9) : 1.4;
  return (double)C.n_data + w * (double)C.rk4_steps;
}

// Device bytes of a job: its slots (simplex state), random-stream counters,
// per-fit results, per-group bests and its curve table.
size_t job_bytes(int n_groups, int group_size, int n_workers, int n_params, int n_filters, size_t n_records) {
  const size_t fits = (size_t)n_groups * group_size;
  const bool refill = n_workers < n_groups;
  return (size_t)n_workers * group_size * slot_bytes((size_t)n_params) + fits * (4 + 8 + 8 + (refill ? 8 * (n_params + n_filters) : 0)) +
         (size_t)n_groups * 256 + sizeof(PointRec) * n_records;
}

// Device memory of one job (its own streams and bests unless borrowed).  The
// clears run on `s`, the stream the job's kernels run on, so they are ordered
// before its first launch.
int job_upload(DevJob& j, const Pool& pool, const CurveH& C, const double* params0, hipStream_t s) {
  if (j.slots.empty() || j.slots[0] < 0 || j.slots[0] >= pool.n_slots) return AST_E_STATE;
  if (j.n_workers <= 0) j.n_workers = j.rp.n_groups;
  if ((int)j.slots.size() != j.n_workers || j.n_workers > j.rp.n_groups) return AST_E_STATE;
  j.refill = j.n_workers < j.rp.n_groups;
  if (j.refill && j.rp.nstages != 1) return AST_E_ARG;  // a stage reseed needs the whole group
  const SpaceH& S = j.space;
  j.cf = S.v.precision == AST_PREC_FP64 ? 8 : 4;
  j.n_data = C.n_data;
  j.n_all = C.n_all;
  j.n_filters = C.n_filters;
  const int N = S.n_params, G = j.rp.n_groups;
  const size_t fits = (size_t)G * j.rp.group_size;
  RETIF(dev_alloc(&j.pts, C.pts.size()));
  RETIF(dev_alloc(&j.sum_w, (size_t)AST_MAX_FILTERS));
  RETIF(dev_alloc((char**)&j.sp, sizeof(Space<double>)));
  RETIF(dev_alloc(&j.params0, (size_t)AST_MAX_PARAMS));
  HIPCHK(hipMemcpy(j.pts, C.pts.data(), sizeof(PointRec) * C.pts.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(j.sum_w, C.sum_w, sizeof(double) * AST_MAX_FILTERS, hipMemcpyHostToDevice));
  if (j.cf == 8) {
    Space<double> sp = space_table<double>(S);
    HIPCHK(hipMemcpy(j.sp, &sp, sizeof sp, hipMemcpyHostToDevice));
  } else {
    Space<float> sp = space_table<float>(S);
    HIPCHK(hipMemcpy(j.sp, &sp, sizeof sp, hipMemcpyHostToDevice));
  }
  double p0[AST_MAX_PARAMS] = {0};
  if (params0)
    for (int i = 0; i < N; ++i) p0[i] = params0[i];
  HIPCHK(hipMemcpy(j.params0, p0, sizeof p0, hipMemcpyHostToDevice));
  j.done = pool.job_done + j.slots[0];
  HIPCHK(hipMemsetAsync(j.done, 0, sizeof(int32_t), s));
  if (!j.borrowed) {
    RETIF(dev_alloc(&j.rng, fits));
    RETIF(dev_alloc((char**)&j.best_f, (size_t)j.cf * G));
    RETIF(dev_alloc(&j.best_params, (size_t)G * AST_MAX_PARAMS));
    RETIF(dev_alloc(&j.best_dV, (size_t)G * AST_MAX_FILTERS));
    std::vector<char> hf((size_t)j.cf * G);
    for (int g = 0; g < G; ++g) {
      if (j.cf == 8) { double v = 1e30; std::memcpy(&hf[(size_t)g * 8], &v, 8); }
      else { float v = 1e30f; std::memcpy(&hf[(size_t)g * 4], &v, 4); }
    }
    HIPCHK(hipMemcpy(j.best_f, hf.data(), hf.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(j.rng, 0, fits * sizeof(uint32_t), s));
    HIPCHK(hipMemsetAsync(j.best_params, 0, sizeof(double) * G * AST_MAX_PARAMS, s));
    HIPCHK(hipMemsetAsync(j.best_dV, 0, sizeof(double) * G * AST_MAX_FILTERS, s));
  }
  RETIF(dev_alloc((char**)&j.fit_f, (size_t)j.cf * fits));
  RETIF(dev_alloc(&j.fit_nev, fits));
  RETIF(dev_alloc(&j.fit_steps, fits));
  HIPCHK(hipMemsetAsync(j.fit_nev, 0, fits * sizeof(int32_t), s));
  HIPCHK(hipMemsetAsync(j.fit_steps, 0, fits * sizeof(int32_t), s));
  if (j.refill) {
    RETIF(dev_alloc(&j.fit_params, fits * (size_t)N));
    RETIF(dev_alloc(&j.fit_dV, fits * (size_t)C.n_filters));
    RETIF(dev_alloc(&j.queue, (size_t)2));
    const int32_t q[2] = {(int32_t)((size_t)j.n_workers * j.rp.group_size), 0};
    HIPCHK(hipMemcpy(j.queue, q, sizeof q, hipMemcpyHostToDevice));
  }
  j.reduced = false;
  j.bytes = job_bytes(j.rp.n_groups, j.rp.group_size, j.n_workers, N, C.n_filters, C.pts.size());
  return 0;
}

template <class CF>
int job_descs_typed(DevJob& j, int ticks) {
  const int G = j.n_workers;
  std::vector<GroupDesc<CF>> d(G);
  double sm = j.rp.size_min > 0 ? j.rp.size_min : (j.cf == 8 ? 1e-10 : 1e-5);  // asteroid.h:233-237
  CF smc = (CF)sm;
  for (int g = 0; g < G; ++g) {
    GroupDesc<CF>& D = d[g];
    std::memset(&D, 0, sizeof D);
    D.pts = j.pts;
    D.sum_w = j.sum_w;
    D.sp = (const Space<CF>*)j.sp;
    D.params0 = j.params0;
    D.rng = j.rng;
    D.best_f = (CF*)j.best_f;
    D.best_params = j.best_params;
    D.best_dV = j.best_dV;
    D.job_done = j.done;
    D.seed = j.rp.seed;
    D.size2_min = smc * smc;  // asteroid.h:245
    D.n_data = j.n_data;
    D.n_filters = j.n_filters;
    D.dof = j.n_all - j.space.n_params - j.n_filters;
    D.nstages = j.rp.nstages;
    D.n_steps = j.rp.n_steps;
    D.eval_budget = (int)(j.rp.eval_budget > 0x7fffffff ? 0x7fffffff : j.rp.eval_budget);
    D.reopt = j.space.x2.reopt;
    D.slot = j.slots[g];
    D.group_in_job = g;
    D.ticks = ticks;
    D.fit_f = (CF*)j.fit_f;
    D.fit_nev = j.fit_nev;
    D.fit_steps = j.fit_steps;
    D.refill = j.refill ? 1 : 0;
    D.n_fits = j.rp.n_groups * j.rp.group_size;
    D.queue = j.queue;
    D.fit_params = j.fit_params;
    D.fit_dV = j.fit_dV;
  }
  if (!j.descs) RETIF(dev_alloc((char**)&j.descs, sizeof(GroupDesc<CF>) * G));
  HIPCHK(hipMemcpy(j.descs, d.data(), sizeof(GroupDesc<CF>) * G, hipMemcpyHostToDevice));
  j.desc_host.resize(sizeof(GroupDesc<CF>) * G);
  std::memcpy(j.desc_host.data(), d.data(), j.desc_host.size());
  return 0;
}

int job_descs(DevJob& j, int ticks) {
  return j.cf == 8 ? job_descs_typed<double>(j, ticks) : job_descs_typed<float>(j, ticks);
}

// The shapes the kernels assume: one workgroup of pool.G threads per group,
// every group on its own slot of the pool, the job's arrays sized for its
// groups.  Checked on the host before every launch.
bool job_fits_pool(const DevJob& j, const Pool& pool) {
  if (j.rp.group_size != pool.G || j.rp.n_groups <= 0) return false;
  if (j.n_workers <= 0 || j.n_workers > j.rp.n_groups) return false;
  if ((int)j.slots.size() != j.n_workers || !j.descs || !j.fit_f || !j.fit_nev || !j.fit_steps) return false;
  if (j.refill && (j.rp.nstages != 1 || !j.queue || !j.fit_params || !j.fit_dV)) return false;
  if (j.space.n_params > pool.P || j.n_filters > AST_MAX_FILTERS || j.n_filters <= 0) return false;
  for (int s : j.slots)
    if (s < 0 || s >= pool.n_slots) return false;
  return (size_t)pool.n_slots * (size_t)pool.G <= pool.nf;
}

// Launch the init or fit kernel over all groups of one job.
hipError_t job_launch(DevJob& j, const Pool& pool, int op, hipStream_t s) {
  if (!job_fits_pool(j, pool)) {
    set_err("job does not fit the slot pool");
    return hipErrorInvalidValue;
  }
  LaunchArgs a{};
  a.op = op;
  a.stream = s;
  a.descs = j.descs;
  SlotBufs<double> bd;
  SlotBufs<float> bf;
  if (j.cf == 8) { bd = pool_bufs<double>(pool, j.cls); a.bufs = &bd; }
  else { bf = pool_bufs<float>(pool, j.cls); a.bufs = &bf; }
  a.n_groups = j.n_workers;
  a.group_size = j.rp.group_size;
  a.n_filters = j.n_filters;
  return dispatch(j.space.v, a);
}

// A job whose workgroups have all finished: with refill, form each group's
// result from its fits' results (group_best_kernel) and wait for it.
int job_reduce(DevJob& j, hipStream_t s) {
  if (!j.refill || j.reduced) return 0;
  const int G = j.rp.group_size, N = j.space.n_params;
  if (j.cf == 8)
    hipLaunchKernelGGL((group_best_kernel<double>), dim3(j.rp.n_groups), dim3(64), 0, s, (const double*)j.fit_f,
                       j.fit_params, j.fit_dV, G, N, j.n_filters, j.rp.n_groups, (double*)j.best_f, j.best_params,
                       j.best_dV);
  else
    hipLaunchKernelGGL((group_best_kernel<float>), dim3(j.rp.n_groups), dim3(64), 0, s, (const float*)j.fit_f,
                       j.fit_params, j.fit_dV, G, N, j.n_filters, j.rp.n_groups, (float*)j.best_f, j.best_params,
                       j.best_dV);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  j.reduced = true;
  return 0;
}

// All fits of the job have ended.
bool job_complete(const DevJob& j) { return j.groups_done >= j.n_workers; }

// Kernel class of a job: one fit-kernel instantiation (variant, filter
// capacity, precision form).
int class_of(const ast_variant& v, int n_filters) {
  int k = variant_index(v);
  int f = n_filters <= 1 ? 0 : (n_filters <= 4 ? 1 : 2);
  int p = v.precision == AST_PREC_FP64 ? 0 : (v.precision == AST_PREC_F32SIMPLEX ? 1 : 2);
  return (k * 3 + f) * 3 + p;
}
int stream_of(int cls) { return cls % NSTREAMS; }

// One launch per kernel class and cycle: the descriptors of every running
// group of the class's jobs, concatenated in device memory.  Rebuilt when
// the set of running jobs of the class changes (between cycles, with no
// kernel in flight).
struct ClassBatch {
  std::vector<uint64_t> members;  // job uids, in descriptor order
  void* d_descs = nullptr;
  size_t cap = 0;                 // bytes
  int n_groups = 0;
};

// --------------------------------------------------------------- serving
struct Serve {
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  bool stop = false;
  std::deque<std::unique_ptr<DevJob>> pending;  // submitted, not admitted
  std::vector<std::unique_ptr<DevJob>> resident;
  std::deque<ast_result> done;
  std::vector<uint64_t> cancels;
  std::map<uint64_t, std::pair<CurveH, std::vector<double>>> inputs;  // curve + params0 until admission
  uint64_t next_id = 1;
  ast_serve_cfg cfg{};
  int64_t resident_bytes = 0;
  int64_t launches = 0;
  double kernel_ms = 0;
  int64_t evals = 0, points = 0;
  int error = 0;
};

}  // namespace

// ------------------------------------------------------------------- RMSD
struct RmsdOut {
  float* par_min;
  float* par_max;
  int64_t n_tot, n_bad;
  double f0, f1;
  float ms;
};


// ------------------------------------------------------------- periodogram
// The periodogram of the minima of a model light curve (minima(),
// misc.c:65-276, the default branch: minima at grid times, step-function
// histogram).  Constants of misc.c:117-121 and NMIN_MAX (misc.c:70).
constexpr int PGRAM_BINS = 10000;        // NN
constexpr double PGRAM_SGM = 0.1;        // sgm
constexpr double PGRAM_DT0 = 1.0 / 24.0; // dt0, days
constexpr double PGRAM_DT2 = 80.0 / 24.0;// dt2, days
constexpr double PGRAM_S0 = 3;           // s0
constexpr int PGRAM_MAX_MINIMA = 10000;  // NMIN_MAX

// Histogram counts: for each frequency bin i (period p = 1/fr[i]), the
// number of pairs of minima a < b whose interval is within a fraction sgm of
// a whole number of periods (misc.c:128-146).  Each thread owns one bin;
// blockIdx.y takes every n_chunks-th first minimum a, and the per-chunk
// counts go to partial[chunk][bin] (summed on the host, exactly).  The
// interval and the ratio are computed as the reference does (a subtraction
// and an IEEE division), so the counts equal the reference's whatever the
// order of the pairs.
__global__ __launch_bounds__(256) void minima_hist_kernel(const double* __restrict__ tm, int n,
                                                         const double* __restrict__ period, int n_bins,
                                                         double sgm, int n_chunks, int32_t* __restrict__ partial) {
  const int bin = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (bin >= n_bins) return;
  const double p = period[bin];
  int32_t cnt = 0;
  for (int a = chunk; a < n - 1; a += n_chunks) {
    const double ta = tm[a];
    for (int b = a + 1; b < n; ++b) {
      const double q = (tm[b] - ta) / p;
      const double dv = fabs(q - (double)(int)(q + 0.5));
      cnt += dv < sgm ? 1 : 0;
    }
  }
  partial[(size_t)chunk * n_bins + bin] = cnt;
}

// Frequency of bin i (misc.c:128).
inline double pgram_fr(int i) {
  const double fr0 = 1 / PGRAM_DT0, fr2 = 1 / PGRAM_DT2;
  return fr2 + (double)i / (double)(PGRAM_BINS - 1) * (fr0 - fr2);
}

int pgram_counts(hipStream_t s, const std::vector<double>& tm, std::vector<int32_t>* counts, double* ms_out) {
  counts->assign(PGRAM_BINS, 0);
  *ms_out = 0;
  const int n = (int)tm.size();
  if (n < 2) return AST_OK;
  std::vector<double> period(PGRAM_BINS);
  for (int i = 0; i < PGRAM_BINS; ++i) period[i] = 1 / pgram_fr(i);
  const int bx = (PGRAM_BINS + 255) / 256;
  const int n_chunks = std::max(1, std::min(n - 1, (2048 + bx - 1) / bx));
  double *td = nullptr, *pd = nullptr;
  int32_t* part = nullptr;
  int rc = dev_alloc(&td, (size_t)n);
  if (!rc) rc = dev_alloc(&pd, (size_t)PGRAM_BINS);
  if (!rc) rc = dev_alloc(&part, (size_t)n_chunks * PGRAM_BINS);
  if (rc) { dev_free(td); dev_free(pd); dev_free(part); return rc; }
  std::vector<int32_t> host((size_t)n_chunks * PGRAM_BINS);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he == hipSuccess) he = hipMemcpy(td, tm.data(), sizeof(double) * n, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(pd, period.data(), sizeof(double) * PGRAM_BINS, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipEventRecord(e0, s);
  if (he == hipSuccess) {
    hipLaunchKernelGGL(minima_hist_kernel, dim3(bx, n_chunks), dim3(256), 0, s, td, n, pd, PGRAM_BINS, PGRAM_SGM,
                       n_chunks, part);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipEventRecord(e1, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
  if (he == hipSuccess) he = hipMemcpy(host.data(), part, sizeof(int32_t) * host.size(), hipMemcpyDeviceToHost);
  if (he != hipSuccess) {
    (void)hipDeviceSynchronize();
    snprintf(g_hip_err, sizeof g_hip_err, "%s (periodogram)", hipGetErrorString(he));
    rc = AST_E_HIP;
  } else {
    for (int ch = 0; ch < n_chunks; ++ch)
      for (int i = 0; i < PGRAM_BINS; ++i) (*counts)[i] += host[(size_t)ch * PGRAM_BINS + i];
    *ms_out = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dev_free(td);
  dev_free(pd);
  dev_free(part);
  return rc;
}

// Normalisation and clusters (misc.c:148-266), on the host in the
// reference's order.  A cluster's left walk stops at bin 0 (the reference
// reads H[-1] there, misc.c:240-252).  Clusters not found get frequency -1
// and height 0.
void pgram_clusters(const std::vector<int32_t>& counts, int ncl, double* cl_fr, double* cl_H) {
  const int NN = PGRAM_BINS;
  std::vector<double> H(NN), fr(NN);
  std::vector<int> marked(NN, 0);
  double m = 0.0;
  for (int i = 0; i < NN; ++i) {
    fr[i] = pgram_fr(i);
    H[i] = (double)counts[i];
    m = m + H[i];
  }
  m = m / NN;
  double sd = 0;
  for (int i = 0; i < NN; ++i) {
    const double dev = H[i] - m;
    sd = sd + dev * dev;
  }
  sd = std::sqrt(sd / (NN - 1));
  for (int i = 0; i < NN; ++i) H[i] = (H[i] - m) / sd;
  const double cl_min = PGRAM_S0;
  for (int j = 0; j < ncl; ++j) { cl_fr[j] = -1.0; cl_H[j] = 0.0; }
  int Ncl = 0;
  for (int j = 0; j < ncl; ++j) {
    double mx = -100;
    int imax = -1;
    cl_fr[j] = -1.0;
    for (int i = 0; i < NN; ++i)
      if (marked[i] == 0 && H[i] > mx) { mx = H[i]; imax = i; }
    if (imax > -1 && H[imax] >= cl_min) {
      cl_fr[Ncl] = fr[imax];
      cl_H[Ncl] = H[imax];
      Ncl++;
      for (int i = imax; i < NN && H[i] > cl_min; ++i) marked[i] = 1;
      for (int i = imax - 1; i >= 0 && H[i] > cl_min; --i) marked[i] = 1;
    }
  }
}

struct ast_ctx {
  int device = 0;
  hipStream_t streams[NSTREAMS] = {nullptr};
  hipEvent_t ev0[NSTREAMS] = {nullptr}, ev1[NSTREAMS] = {nullptr};
  ast_variant v{};
  // reference path state
  bool space_set = false;
  SpaceH space;
  bool curve_set = false;
  CurveH curve;
  PointRec* pts = nullptr;   // for ast_chi2_batch
  double* sum_w = nullptr;
  TailRec minima{};           // NUDGE's observed minima (in the curve table's tail record)
  bool minima_set = false;
  double seg_t[AST_MAX_SEGMENTS]{};  // SEGMENT's start times
  int seg_first[AST_MAX_SEGMENTS]{};  // and the first observation of each segment, when given
  bool seg_set = false, seg_first_set = false;
  double params0[AST_MAX_PARAMS]{};
  uint64_t seed = 0;
  uint32_t* rng = nullptr;
  size_t rng_cap = 0;
  void* best_f = nullptr;
  double* best_params = nullptr;
  double* best_dV = nullptr;
  size_t best_cap = 0;
  int best_cf = 8;
  Pool pool;
  DevJob* last_job = nullptr;  // lanes of the last ast_run (slot rows 0..fits-1)
  std::unique_ptr<DevJob> ref_job;
  int last_groups = 0, last_group_size = 0;
  bool run_active = false;     // an ast_run_begin whose job has not finished
  bool run_graphs = false;
  long run_cycles = 0, run_max_cycles = 0;
  double last_ms = 0;
  int last_launches = 0;
  int64_t last_evals = 0, last_points = 0;
  std::unique_ptr<Serve> serve;
  ClassBatch batch[MAX_CLASSES];
  std::atomic<uint64_t> next_uid{1};
  int waves[MAX_CLASSES];  // fit-kernel register budget per class (4, 3 or 2 waves per SIMD)
  hipEvent_t ev_fork = nullptr, ev_join[NSTREAMS] = {nullptr};  // graph capture fork/join
  hipEvent_t ev_pre[NSTREAMS] = {nullptr};  // work queued on the class streams before a graph launch
  hipGraphExec_t cg_exec = nullptr;  // the captured cycle
  std::vector<uint64_t> cg_key;
  ast_ctx() { for (int k = 0; k < MAX_CLASSES; ++k) waves[k] = 4; }
};

namespace {

int set_device(ast_ctx* c) {
  HIPCHK(hipSetDevice(c->device));
  return 0;
}

// One cycle: one fit-kernel launch per kernel class over every running
// group of every job of that class, every stream drained at the end; adds
// the cycle's kernel time (the longest stream).
int run_cycle_once(ast_ctx* c, std::vector<DevJob*>& jobs, double* kernel_ms, int* launches, bool graphs);

// A failed cycle may leave kernels in flight: drain the device before the
// caller frees anything they use.
int run_cycle(ast_ctx* c, std::vector<DevJob*>& jobs, double* kernel_ms, int* launches, bool graphs = false) {
  int rc = run_cycle_once(c, jobs, kernel_ms, launches, graphs);
  if (rc) (void)hipDeviceSynchronize();
  return rc;
}

// Bring the class's concatenated descriptors up to date with `members`.
int class_batch_update(ast_ctx* c, int k, const std::vector<DevJob*>& members) {
  ClassBatch& cb = c->batch[k];
  std::vector<uint64_t> ids;
  for (DevJob* j : members) ids.push_back(j->uid);
  if (ids == cb.members && cb.d_descs) return 0;
  size_t bytes = 0;
  int groups = 0;
  for (DevJob* j : members) {
    bytes += j->desc_host.size();
    groups += j->n_workers;
  }
  if (bytes > cb.cap) {
    dev_free(cb.d_descs);
    cb.cap = 0;
    RETIF(dev_alloc((char**)&cb.d_descs, bytes));
    cb.cap = bytes;
  }
  std::vector<unsigned char> h(bytes);
  size_t off = 0;
  for (DevJob* j : members) {
    std::memcpy(h.data() + off, j->desc_host.data(), j->desc_host.size());
    off += j->desc_host.size();
  }
  HIPCHK(hipMemcpy(cb.d_descs, h.data(), bytes, hipMemcpyHostToDevice));
  cb.members = ids;
  cb.n_groups = groups;
  return 0;
}

// One fit-kernel launch of class k over the class batch, on its stream.
hipError_t launch_class(ast_ctx* c, int k, const DevJob& j0) {
  LaunchArgs a{};
  a.op = 1;
  a.stream = c->streams[stream_of(k)];
  a.descs = c->batch[k].d_descs;
  SlotBufs<double> bd;
  SlotBufs<float> bf;
  if (j0.cf == 8) { bd = pool_bufs<double>(c->pool, k); a.bufs = &bd; }
  else { bf = pool_bufs<float>(c->pool, k); a.bufs = &bf; }
  a.n_groups = c->batch[k].n_groups;
  a.group_size = c->pool.G;
  a.n_filters = j0.n_filters;
  a.waves = c->waves[k];
  return dispatch(j0.space.v, a);
}

// The cycle as a HIP graph: every class launch, forked from stream 0 onto
// the classes' streams and joined back.  Valid while the class batches, the
// pool and the register budgets are unchanged (the key).
int cycle_graph(ast_ctx* c, std::vector<DevJob*>* members, const bool* used_s) {
  std::vector<uint64_t> key{(uint64_t)(uintptr_t)c->pool.X, (uint64_t)c->pool.nf, (uint64_t)c->pool.G};
  for (int k = 0; k < MAX_CLASSES; ++k) {
    if (members[k].empty()) continue;
    key.push_back((uint64_t)k);
    key.push_back((uint64_t)c->waves[k]);
    key.push_back((uint64_t)(uintptr_t)c->batch[k].d_descs);
    key.push_back((uint64_t)c->batch[k].n_groups);
    for (uint64_t u : c->batch[k].members) key.push_back(u);
  }
  if (c->cg_exec && key == c->cg_key) return 0;
  if (c->cg_exec) (void)hipGraphExecDestroy(c->cg_exec);
  c->cg_exec = nullptr;
  c->cg_key.clear();
  const hipStream_t s0 = c->streams[0];
  HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipEventRecord(c->ev_fork, s0);
  for (int k = 1; k < NSTREAMS && e == hipSuccess; ++k)
    if (used_s[k]) e = hipStreamWaitEvent(c->streams[k], c->ev_fork, 0);
  for (int k = 0; k < MAX_CLASSES && e == hipSuccess; ++k)
    if (!members[k].empty()) e = launch_class(c, k, *members[k][0]);
  for (int k = 1; k < NSTREAMS && e == hipSuccess; ++k)
    if (used_s[k]) {
      e = hipEventRecord(c->ev_join[k], c->streams[k]);
      if (e == hipSuccess) e = hipStreamWaitEvent(s0, c->ev_join[k], 0);
    }
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s0, &g);  // always end the capture
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&c->cg_exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    c->cg_exec = nullptr;
    snprintf(g_hip_err, sizeof g_hip_err, "%s (cycle graph)", hipGetErrorString(e));
    return AST_E_HIP;
  }
  c->cg_key = key;
  return 0;
}

int run_cycle_once(ast_ctx* c, std::vector<DevJob*>& jobs, double* kernel_ms, int* launches, bool graphs) {
  std::vector<DevJob*> members[MAX_CLASSES];
  bool used_s[NSTREAMS] = {false};
  for (DevJob* j : jobs) {
    if (job_complete(*j)) continue;
    if (!job_fits_pool(*j, c->pool)) {
      set_err("job does not fit the slot pool");
      return AST_E_STATE;
    }
    members[j->cls].push_back(j);
    used_s[stream_of(j->cls)] = true;
  }
  int n_cls = 0;
  for (int k = 0; k < MAX_CLASSES; ++k)
    if (!members[k].empty()) {
      RETIF(class_batch_update(c, k, members[k]));
      n_cls++;
    }
  double ms_max = 0;
  std::vector<int32_t> done(jobs.size());
  if (graphs && n_cls > 0) {
    // one graph launch on stream 0; every class stream joins back into it
    RETIF(cycle_graph(c, members, used_s));
    const hipStream_t s0 = c->streams[0];
    // The graph runs on stream 0, but a job's init kernel and clears were
    // queued on its class's stream: stream 0 waits for them, or the first
    // fit launch could read slot state before it is written.
    for (int k = 1; k < NSTREAMS; ++k)
      if (used_s[k]) {
        HIPCHK(hipEventRecord(c->ev_pre[k], c->streams[k]));
        HIPCHK(hipStreamWaitEvent(s0, c->ev_pre[k], 0));
      }
    HIPCHK(hipEventRecord(c->ev0[0], s0));
    HIPCHK(hipGraphLaunch(c->cg_exec, s0));
    HIPCHK(hipEventRecord(c->ev1[0], s0));
    *launches += n_cls;
    HIPCHK(hipStreamSynchronize(s0));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0[0], c->ev1[0]));
    ms_max = ms;
  } else {
    for (int k = 0; k < NSTREAMS; ++k)
      if (used_s[k]) HIPCHK(hipEventRecord(c->ev0[k], c->streams[k]));
    for (int k = 0; k < MAX_CLASSES; ++k) {
      if (members[k].empty()) continue;
      HIPCHK(launch_class(c, k, *members[k][0]));
      (*launches)++;
    }
    for (int k = 0; k < NSTREAMS; ++k)
      if (used_s[k]) HIPCHK(hipEventRecord(c->ev1[k], c->streams[k]));
    for (int k = 0; k < NSTREAMS; ++k) {
      bool any = used_s[k];
      for (size_t i = 0; i < jobs.size() && !any; ++i) any = stream_of(jobs[i]->cls) == k;
      if (!any) continue;
      HIPCHK(hipStreamSynchronize(c->streams[k]));
      if (used_s[k]) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0[k], c->ev1[k]));
        if (ms > ms_max) ms_max = ms;
      }
    }
  }
  // Every job's finished-group counter in one read-back (they sit at the
  // jobs' first slots; the streams are drained, so the counters are final).
  Pool& P = c->pool;
  if (!jobs.empty())
    HIPCHK(hipMemcpy(P.job_done_h.data(), P.job_done, sizeof(int32_t) * P.n_slots, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < jobs.size(); ++i) done[i] = P.job_done_h[(size_t)jobs[i]->slots[0]];
  *kernel_ms += ms_max;
  for (size_t i = 0; i < jobs.size(); ++i) {
    DevJob& j = *jobs[i];
    const bool was = job_complete(j);
    j.groups_done = done[i];
    if (!was && job_complete(j)) {
      RETIF(job_reduce(j, c->streams[stream_of(j.cls)]));
      j.t_done = Clock::now();
    }
  }
  return 0;
}

// Grow the reference path's random-stream counters and per-group bests.
int ensure_streams(ast_ctx* c, size_t fits, size_t groups) {
  if (c->rng_cap < fits) {
    uint32_t* nr = nullptr;
    RETIF(dev_alloc(&nr, fits));
    HIPCHK(hipMemset(nr, 0, fits * sizeof(uint32_t)));
    if (c->rng) HIPCHK(hipMemcpy(nr, c->rng, c->rng_cap * sizeof(uint32_t), hipMemcpyDeviceToDevice));
    dev_free(c->rng);
    c->rng = nr;
    c->rng_cap = fits;
  }
  const int cf = c->v.precision == AST_PREC_FP64 ? 8 : 4;
  if (c->best_cap < groups || !c->best_f) {
    void* nf = nullptr;
    double *np = nullptr, *nd = nullptr;
    int rc = dev_alloc((char**)&nf, (size_t)cf * groups);
    if (!rc) rc = dev_alloc(&np, groups * AST_MAX_PARAMS);
    if (!rc) rc = dev_alloc(&nd, groups * AST_MAX_FILTERS);
    if (rc) { dev_free(nf); dev_free(np); dev_free(nd); return rc; }
    std::vector<char> hf((size_t)cf * groups);
    for (size_t g = 0; g < groups; ++g) {
      if (cf == 8) { double v = 1e30; std::memcpy(&hf[g * 8], &v, 8); }
      else { float v = 1e30f; std::memcpy(&hf[g * 4], &v, 4); }
    }
    HIPCHK(hipMemcpy(nf, hf.data(), hf.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(np, 0, groups * AST_MAX_PARAMS * sizeof(double)));
    HIPCHK(hipMemset(nd, 0, groups * AST_MAX_FILTERS * sizeof(double)));
    if (c->best_f) {
      HIPCHK(hipMemcpy(nf, c->best_f, (size_t)cf * c->best_cap, hipMemcpyDeviceToDevice));
      HIPCHK(hipMemcpy(np, c->best_params, c->best_cap * AST_MAX_PARAMS * sizeof(double), hipMemcpyDeviceToDevice));
      HIPCHK(hipMemcpy(nd, c->best_dV, c->best_cap * AST_MAX_FILTERS * sizeof(double), hipMemcpyDeviceToDevice));
    }
    dev_free(c->best_f);
    dev_free(c->best_params);
    dev_free(c->best_dV);
    c->best_f = nf;
    c->best_params = np;
    c->best_dV = nd;
    c->best_cap = groups;
    c->best_cf = cf;
  }
  // the clears and copies above ran on the null stream, which the context's
  // non-blocking streams do not wait for
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int reset_best_f(ast_ctx* c) {
  if (!c->best_f) return 0;
  const size_t cf = (size_t)c->best_cf, n = c->best_cap;
  std::vector<char> hf(cf * n);
  for (size_t g = 0; g < n; ++g) {
    if (cf == 8) { double v = 1e30; std::memcpy(&hf[g * 8], &v, 8); }
    else { float v = 1e30f; std::memcpy(&hf[g * 4], &v, 4); }
  }
  HIPCHK(hipMemcpy(c->best_f, hf.data(), cf * n, hipMemcpyHostToDevice));
  return 0;
}

// Finished job -> result record (host).
int job_result(DevJob& j, ast_result* r) {
  std::memset(r, 0, sizeof *r);
  r->id = j.id;
  r->tag = j.tag;
  r->status = j.cancelled ? AST_E_CANCELLED : j.status;
  r->n_params = j.space.n_params;
  r->n_filters = j.n_filters;
  r->n_groups = j.rp.n_groups;
  const int G = j.rp.n_groups, N = j.space.n_params, NF = j.n_filters;
  std::vector<double> f(G), p((size_t)G * N), dv((size_t)G * NF);
  if (!j.cancelled) {
    if (j.cf == 8) {
      HIPCHK(hipMemcpy(f.data(), j.best_f, sizeof(double) * G, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(G);
      HIPCHK(hipMemcpy(tmp.data(), j.best_f, sizeof(float) * G, hipMemcpyDeviceToHost));
      for (int g = 0; g < G; ++g) f[g] = tmp[g];
    }
    HIPCHK(hipMemcpy(p.data(), j.best_params, sizeof(double) * G * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dv.data(), j.best_dV, sizeof(double) * G * NF, hipMemcpyDeviceToHost));
  }
  // Best of the job's groups (strict <, first wins: asteroid.c:741-749).
  int gb = 0;
  double fb = 1e32;
  for (int g = 0; g < G; ++g)
    if (f[g] < fb) { fb = f[g]; gb = g; }
  r->best_f = j.cancelled ? 1e30 : fb;
  r->best_group = gb;
  for (int i = 0; i < N; ++i) r->best_params[i] = p.empty() ? 0 : p[(size_t)gb * N + i];
  for (int m = 0; m < NF; ++m) r->best_dV[m] = dv.empty() ? 0 : dv[(size_t)gb * NF + m];
  return 0;
}

// Evaluations a job has done so far: the lanes' work counters over its slots
// (consecutive slots are read in one copy).
int job_evals(const ast_ctx* c, const DevJob& j, int64_t* evals) {
  const size_t G = (size_t)j.rp.group_size;
  std::vector<int32_t> tmp;
  int64_t s = 0;
  for (size_t a = 0; a < j.slots.size();) {
    size_t b = a + 1;
    while (b < j.slots.size() && j.slots[b] == j.slots[b - 1] + 1) ++b;
    tmp.resize((b - a) * G);
    HIPCHK(hipMemcpy(tmp.data(), c->pool.work + (size_t)j.slots[a] * G, sizeof(int32_t) * tmp.size(),
                     hipMemcpyDeviceToHost));
    for (int32_t w : tmp) s += w;
    a = b;
  }
  *evals = s;
  return 0;
}

// A job from its C description: validated model space and light curve,
// run parameters and kernel class; the curve and start model are returned
// for the upload.
int make_job(ast_ctx* c, const ast_job* job, int G, std::unique_ptr<DevJob>* out, CurveH* curve,
             std::vector<double>* p0) {
  if (job->n_fits <= 0 || job->n_fits % G || job->nstages <= 0 || job->n_steps < 0 || job->eval_budget < 0)
    return AST_E_ARG;
  if (job->variant.nudge &&
      (job->n_minima < 1 || job->n_minima > AST_MAX_MINIMA || !job->t_minima || !job->V_minima))
    return AST_E_ARG;
  if (job->variant.n_seg > 1 && !job->t_segments) return AST_E_ARG;
  if (job->variant.n_seg > 1)
    for (int k = 0; k < job->variant.n_seg; ++k)
      if (!std::isfinite(job->t_segments[k])) return AST_E_ARG;
  std::unique_ptr<DevJob> j(new DevJob());
  j->uid = c->next_uid++;
  RETIF(make_space(job->variant, job->n_params, job->property, job->n_types, job->types, job->limits, &job->x2,
                   &j->space));
  CurveH C;
  RETIF(make_curve(job->variant, job->obs, job->n_data, job->n_filters, job->eph, &C,
                   job->variant.n_seg > 1 ? job->t_segments : nullptr,
                   job->variant.n_seg > 1 ? job->seg_first_obs : nullptr));
  if (C.n_all - j->space.n_params - C.n_filters <= 0) return AST_E_DATA;  // dof over every observation
  if (job->variant.nudge) RETIF(set_tail_minima(job->variant, job->t_minima, job->V_minima, job->n_minima, &C));
  j->rp.n_groups = job->n_fits / G;
  j->rp.group_size = G;
  j->rp.nstages = job->nstages;
  j->rp.n_steps = job->n_steps;
  j->rp.eval_budget = job->eval_budget;
  j->rp.size_min = job->size_min;
  j->rp.seed = job->seed;
  j->tag = job->tag;
  j->cls = class_of(job->variant, job->n_filters);
  j->n_data = C.n_data;
  j->n_filters = C.n_filters;
  j->cost = eval_cost(job->variant, C);
  j->t_submit = Clock::now();
  p0->clear();
  if (job->params0) p0->assign(job->params0, job->params0 + j->space.n_params);
  *curve = std::move(C);
  *out = std::move(j);
  return AST_OK;
}

// Upload `jobs` into slots 0.., run them in shared cycles and collect their
// results (results may be null).  cycle_limit > 0 stops after that many
// cycles (the autotuner's fixed work); 0 runs every job to completion.
int batch_run(ast_ctx* c, const std::vector<const ast_job*>& jobs, int G, double work, long cycle_limit,
              ast_result* results, double* kernel_ms, int* launches, bool graphs, int schedule = 0) {
  const int n_jobs = (int)jobs.size();
  std::vector<std::unique_ptr<DevJob>> J(n_jobs);
  std::vector<CurveH> curves(n_jobs);
  std::vector<std::vector<double>> p0(n_jobs);
  size_t groups = 0;
  int n_params = 1;
  for (int i = 0; i < n_jobs; ++i) {
    RETIF(make_job(c, jobs[i], G, &J[i], &curves[i], &p0[i]));
    J[i]->id = (uint64_t)i;
    groups += (size_t)J[i]->rp.n_groups;
    n_params = std::max(n_params, J[i]->space.n_params);
  }
  RETIF(set_device(c));
  if ((size_t)c->pool.n_slots < groups || c->pool.G != G || c->pool.P < n_params)
    RETIF(pool_alloc(c->pool, groups * G, G, n_params));
  c->last_groups = 0;  // the pool's rows no longer hold an ast_run's lanes
  // Work of one simplex of each job: an upper bound of its evaluations (the
  // budget, or the step limit) times the cost of one.
  std::vector<double> evals(n_jobs), wj(n_jobs);
  double w_max = 0;
  for (int i = 0; i < n_jobs; ++i) {
    const DevJob& j = *J[i];
    double e = (double)(j.space.n_params + 1) + (double)j.rp.n_steps * (j.space.n_params + 2);
    if (j.rp.eval_budget > 0 && (double)j.rp.eval_budget < e) e = (double)j.rp.eval_budget;
    evals[i] = e * j.rp.nstages;
    wj[i] = evals[i] * j.cost;
    if (wj[i] > w_max) w_max = wj[i];
  }
  // Launch order (slots, descriptor order): the heaviest jobs first so they
  // are always dispatched first (schedules 0 and 2), or the jobs' own order.
  std::vector<int> order(n_jobs);
  for (int i = 0; i < n_jobs; ++i) order[i] = i;
  if (schedule != 1)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return wj[a] > wj[b]; });
  std::vector<DevJob*> run;
  int slot = 0;
  const Clock::time_point t0 = Clock::now();
  int rc = 0;
  for (int oi = 0; oi < n_jobs && !rc; ++oi) {
    const int i = order[oi];
    DevJob& j = *J[i];
    for (int g = 0; g < j.rp.n_groups; ++g) j.slots.push_back(slot++);
    const hipStream_t s = c->streams[stream_of(j.cls)];
    // Ticks per launch: a share of the job's own total, so that every job
    // ends near the cycle the heaviest one needs and the GPU stays full to
    // the end (schedule 0), or equal work per launch for every job (1, 2).
    int ticks = ticks_for(j.cost, work);
    if (schedule == 0) {
      const double t = std::ceil(evals[i] * work / w_max);
      ticks = t < 1 ? 1 : (t > 4096 ? 4096 : (int)t);
    }
    rc = job_upload(j, c->pool, curves[i], p0[i].empty() ? nullptr : p0[i].data(), s);
    if (!rc) rc = job_descs(j, ticks);
    if (!rc) {
      hipError_t e = job_launch(j, c->pool, 0, s);
      if (e != hipSuccess) {
        snprintf(g_hip_err, sizeof g_hip_err, "%s (init kernel)", hipGetErrorString(e));
        rc = AST_E_HIP;
      }
    }
    j.t_submit = j.t_admit = t0;
    j.ticks = ticks;
    run.push_back(&j);
  }
  // Every tick completes one evaluation of each running simplex; a stage
  // needs at most (N+1) + n_steps*(N+2) of them.
  long max_cycles = 0;
  for (DevJob* j : run) {
    const double per_stage = (double)(j->space.n_params + 1) + (double)j->rp.n_steps * (j->space.n_params + 2);
    const long mc = (long)((per_stage * j->rp.nstages + 16.0) / j->ticks) + 4;
    if (mc > max_cycles) max_cycles = mc;
  }
  for (long L = 0; !rc; ++L) {
    bool all = true;
    for (DevJob* j : run) all = all && job_complete(*j);
    if (all || (cycle_limit > 0 && L >= cycle_limit)) break;
    if (L > max_cycles) {
      set_err("fit did not terminate");
      rc = AST_E_STATE;
      break;
    }
    rc = run_cycle(c, run, kernel_ms, launches, graphs);
  }
  if (rc) (void)hipDeviceSynchronize();
  int64_t ev_tot = 0, pts_tot = 0;
  for (int i = 0; i < n_jobs; ++i) {
    DevJob& j = *J[i];
    if (results) {
      ast_result& r = results[i];
      std::memset(&r, 0, sizeof r);
      if (!rc && !j.slots.empty()) {
        int e = job_result(j, &r);
        int64_t ev = 0;
        if (!e) e = job_evals(c, j, &ev);
        if (e && !rc) rc = e;
        r.n_evals = ev;
        r.model_points = ev * j.n_data;
        r.submit_to_done_ms = std::chrono::duration<double, std::milli>(j.t_done - t0).count();
        r.run_ms = r.submit_to_done_ms;
        ev_tot += ev;
        pts_tot += ev * j.n_data;
      } else {
        r.status = rc ? rc : AST_E_STATE;
        r.best_f = 1e30;
      }
      r.id = (uint64_t)i;
      r.tag = j.tag;
    }
    job_free(j);
  }
  if (results) {
    c->last_evals = ev_tot;
    c->last_points = pts_tot;
  }
  return rc;
}

// ------------------------------------------------------- serving thread
// Upload an admitted job and start its groups.  Its slots were reserved
// under the scheduler lock when it was admitted.
int serve_admit(ast_ctx* c, Serve& S, DevJob& j, const CurveH& curve, const std::vector<double>& p0) {
  const hipStream_t s = c->streams[stream_of(j.cls)];
  RETIF(job_upload(j, c->pool, curve, p0.empty() ? nullptr : p0.data(), s));
  RETIF(job_descs(j, ticks_for(j.cost, S.cfg.work_points_per_launch)));
  HIPCHK(job_launch(j, c->pool, 0, s));
  return 0;
}

void serve_loop(ast_ctx* c) {
  Serve& S = *c->serve;
  (void)hipSetDevice(c->device);
  for (;;) {
    std::vector<std::unique_ptr<DevJob>> admitted;
    std::vector<std::pair<CurveH, std::vector<double>>> admitted_in;
    {
      std::unique_lock<std::mutex> lk(S.mu);
      if (S.resident.empty() && S.pending.empty())
        S.cv_submit.wait_for(lk, std::chrono::milliseconds(20));
      if (S.stop && S.pending.empty() && S.resident.empty()) break;
      if (S.stop) {  // cancel everything left
        for (auto& j : S.pending) S.cancels.push_back(j->id);
        for (auto& j : S.resident) S.cancels.push_back(j->id);
      }
      // Cancellations of pending jobs complete at once.
      for (uint64_t id : S.cancels) {
        for (auto it = S.pending.begin(); it != S.pending.end(); ++it)
          if ((*it)->id == id) {
            ast_result r;
            std::memset(&r, 0, sizeof r);
            r.id = id;
            r.tag = (*it)->tag;
            r.status = AST_E_CANCELLED;
            r.best_f = 1e30;
            S.done.push_back(r);
            S.inputs.erase(id);
            S.pending.erase(it);
            S.cv_done.notify_all();
            break;
          }
        for (auto& j : S.resident)
          if (j->id == id) j->cancelled = true;
      }
      S.cancels.clear();
      // Admission: FIFO, while the job fits the free slots and its footprint
      // fits the HBM budget.  A multi-stage job needs a slot per group (its
      // stage reseeds need whole groups).  A single-stage job starts on as
      // many of its groups' worth of slots as are free once at least half of
      // its groups' worth is (or the whole pool, for a job larger than twice
      // the pool), and its lanes take the rest of its fits as their fits end
      // (refill): free slots do not wait behind a job that does not fit
      // whole, and no job runs on so few slots that it trails the others by
      // more than twice its time on all of them.  The slots are taken off the
      // free list here, so later jobs of the same pass see what is left.
      while (!S.pending.empty()) {
        DevJob& j = *S.pending.front();
        auto in = S.inputs.find(j.id);
        const int64_t nfree = (int64_t)c->pool.free_slots.size();
        const int64_t G_job = (int64_t)j.rp.n_groups;
        const int64_t least = std::min<int64_t>((G_job + 1) / 2, (int64_t)c->pool.n_slots);
        const int64_t need = j.rp.nstages == 1 ? (nfree >= least ? std::min<int64_t>(G_job, nfree) : least) : G_job;
        const int64_t bytes = (int64_t)job_bytes(j.rp.n_groups, j.rp.group_size, (int)need, j.space.n_params,
                                                 j.n_filters, in->second.first.pts.size());
        if (nfree < need) break;
        if (S.cfg.hbm_budget_bytes > 0 && S.resident_bytes + bytes > S.cfg.hbm_budget_bytes &&
            !(S.resident.empty() && admitted.empty()))
          break;
        j.slots.clear();
        for (int64_t g = 0; g < need; ++g) {
          j.slots.push_back(c->pool.free_slots.back());
          c->pool.free_slots.pop_back();
        }
        j.n_workers = (int)need;
        j.bytes = (size_t)bytes;
        S.resident_bytes += bytes;
        admitted_in.push_back(std::move(in->second));
        S.inputs.erase(in);
        admitted.push_back(std::move(S.pending.front()));
        S.pending.pop_front();
      }
    }
    for (size_t k = 0; k < admitted.size(); ++k) {
      DevJob& j = *admitted[k];
      int rc = serve_admit(c, S, j, admitted_in[k].first, admitted_in[k].second);
      if (rc) j.status = rc;
      j.t_admit = Clock::now();
      std::lock_guard<std::mutex> lk(S.mu);
      S.resident.push_back(std::move(admitted[k]));
    }
    std::vector<DevJob*> run;
    {
      std::lock_guard<std::mutex> lk(S.mu);
      for (auto& j : S.resident)
        if (!j->cancelled && j->status == AST_OK) run.push_back(j.get());
    }
    if (!run.empty()) {
      double ms = 0;
      int nl = 0;
      auto t0 = Clock::now();
      int rc = run_cycle(c, run, &ms, &nl);
      double wall = ms_since(t0);
      std::lock_guard<std::mutex> lk(S.mu);
      if (rc) {
        S.error = rc;
        for (DevJob* j : run) j->status = rc;
      }
      S.kernel_ms += ms;
      S.launches += nl;
      for (DevJob* j : run) j->run_ms += wall;
    }
    // Retire finished, failed and cancelled jobs.
    std::vector<std::unique_ptr<DevJob>> finished;
    {
      std::lock_guard<std::mutex> lk(S.mu);
      for (auto it = S.resident.begin(); it != S.resident.end();) {
        DevJob& j = **it;
        if (j.cancelled || j.status != AST_OK || job_complete(j)) {
          finished.push_back(std::move(*it));
          it = S.resident.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (auto& jp : finished) {
      DevJob& j = *jp;
      ast_result r;
      int rc = j.status == AST_OK ? job_result(j, &r) : 0;
      if (j.status != AST_OK || rc) {
        std::memset(&r, 0, sizeof r);
        r.id = j.id;
        r.tag = j.tag;
        r.status = rc ? rc : j.status;
        r.best_f = 1e30;
      }
      int64_t ev = 0;
      if (!j.cancelled && j.status == AST_OK) job_evals(c, j, &ev);
      r.n_evals = ev;
      r.model_points = ev * j.n_data;
      r.submit_to_done_ms = ms_since(j.t_submit);
      r.queue_ms = std::chrono::duration<double, std::milli>(j.t_admit - j.t_submit).count();
      r.run_ms = j.run_ms;
      std::lock_guard<std::mutex> lk(S.mu);
      for (int s : j.slots) c->pool.free_slots.push_back(s);
      S.resident_bytes -= (int64_t)j.bytes;
      S.evals += ev;
      S.points += ev * j.n_data;
      job_free(j);
      S.done.push_back(r);
      S.cv_done.notify_all();
    }
  }
}

}  // namespace

// One RMSD launch pair: the input model, then the shifted models.
template <class CF>
int rmsd_run(ast_ctx* c, const ast_rmsd_cfg& cfg, RmsdOut* out) {
  const int N = c->space.n_params, G = cfg.group_size, ng = cfg.n_groups;
  const size_t lanes = (size_t)ng * G;
  RETIF(ensure_streams(c, lanes, (size_t)ng));
  const Space<CF> sp_h = space_table<CF>(c->space);
  Space<CF>* sp = nullptr;
  RmsdRef<CF>* ref = nullptr;
  double* p0 = nullptr;
  float *pmin = nullptr, *pmax = nullptr;
  int32_t* cnt = nullptr;
  int rc = dev_alloc(&sp, 1);
  if (!rc) rc = dev_alloc(&ref, 1);
  if (!rc) rc = dev_alloc(&p0, (size_t)N);
  if (!rc) rc = dev_alloc(&pmin, (size_t)ng * N);
  if (!rc) rc = dev_alloc(&pmax, (size_t)ng * N);
  if (!rc) rc = dev_alloc(&cnt, (size_t)ng * 2);
  auto release = [&]() { dev_free(sp); dev_free(ref); dev_free(p0); dev_free(pmin); dev_free(pmax); dev_free(cnt); };
  if (rc) { release(); return rc; }
  RmsdArgs<CF> A{};
  A.pts = c->pts;
  A.sum_w = c->sum_w;
  A.n_data = c->curve.n_data;
  A.n_filters = c->curve.n_filters;
  A.dof = c->curve.n_all - N - c->curve.n_filters;
  A.sp = sp;
  A.reopt = c->space.x2.reopt;
  A.params0 = p0;
  A.ref = ref;
  A.rng = c->rng;
  A.seed = c->seed;
  A.n_samples = cfg.n_samples;
  A.dx = cfg.dx;
  A.spherical_k = cfg.spherical_k;
  A.par_min = pmin;
  A.par_max = pmax;
  A.counts = cnt;
  hipStream_t s = c->streams[0];
  hipError_t he = hipMemcpy(sp, &sp_h, sizeof sp_h, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(p0, c->params0, sizeof(double) * N, hipMemcpyHostToDevice);
  LaunchArgs a{};
  a.stream = s;
  a.n_filters = c->curve.n_filters;
  a.n_groups = ng;
  a.group_size = G;
  a.rmsd = &A;
  a.op = 4;
  if (he == hipSuccess) he = dispatch(c->v, a);
  if (he == hipSuccess) he = hipEventRecord(c->ev0[0], s);
  a.op = 5;
  if (he == hipSuccess) he = dispatch(c->v, a);
  if (he == hipSuccess) he = hipEventRecord(c->ev1[0], s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, c->ev0[0], c->ev1[0]);
  std::vector<int32_t> ch((size_t)ng * 2);
  RmsdRef<CF> rh;
  if (he == hipSuccess) he = hipMemcpy(out->par_min, pmin, sizeof(float) * ng * N, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(out->par_max, pmax, sizeof(float) * ng * N, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(ch.data(), cnt, sizeof(int32_t) * ch.size(), hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(&rh, ref, sizeof rh, hipMemcpyDeviceToHost);
  if (he != hipSuccess) {
    (void)hipDeviceSynchronize();
    snprintf(g_hip_err, sizeof g_hip_err, "%s (rmsd)", hipGetErrorString(he));
    release();
    return AST_E_HIP;
  }
  out->n_tot = out->n_bad = 0;
  for (int g = 0; g < ng; ++g) {
    out->n_tot += ch[(size_t)g * 2];
    out->n_bad += ch[(size_t)g * 2 + 1];
  }
  out->f0 = rh.f0;
  out->f1 = rh.f1;
  out->ms = ms;
  release();
  return AST_OK;
}

extern "C" {

const char* ast_strerror(int code) {
  switch (code) {
    case AST_OK: return "ok";
    case AST_E_ARG: return "bad argument";
    case AST_E_STATE: return "call out of order or job did not terminate";
    case AST_E_HIP: return "HIP runtime error";
    case AST_E_NOMEM: return "out of memory";
    case AST_E_DATA: return "input data rejected";
    case AST_E_UNSUPPORTED: return "variant or model space not supported";
    case AST_E_NODEVICE: return "no HIP device";
    case AST_E_CANCELLED: return "job cancelled";
    case AST_E_BUSY: return "context is serving";
    default: return "unknown error";
  }
}

const char* ast_last_hip_error(void) { return g_hip_err; }

// The device address of `scratch` on `device` (current), looked up once per
// device for the process.
static hipError_t scratch_setup(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  g_scratch_ptr[device] = (double*)ptr;
  return hipSuccess;
}

int ast_ctx_create(ast_ctx** out, const ast_variant* v, int device) {
  if (!out || !v) return AST_E_ARG;
  *out = nullptr;
  if (!variant_supported(*v)) return AST_E_UNSUPPORTED;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return AST_E_NODEVICE;
  if (device < 0 || device >= n) return AST_E_ARG;
  ast_ctx* c = new (std::nothrow) ast_ctx();
  if (!c) return AST_E_NOMEM;
  c->device = device;
  c->v = *v;
  hipError_t e = hipSetDevice(device);
  for (int k = 0; k < NSTREAMS && e == hipSuccess; ++k) {
    e = hipStreamCreateWithFlags(&c->streams[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0[k]);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1[k]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_pre[k], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = scratch_setup(device);
  if (e != hipSuccess) {
    set_err(hipGetErrorString(e));
    ast_ctx_destroy(c);
    return AST_E_HIP;
  }
  *out = c;
  return AST_OK;
}

int ast_ctx_destroy(ast_ctx* c) {
  if (!c) return AST_E_ARG;
  if (c->serve) ast_serve_stop(c);
  (void)hipSetDevice(c->device);
  if (c->ref_job) job_free(*c->ref_job);
  pool_free(c->pool);
  for (int k = 0; k < MAX_CLASSES; ++k) dev_free(c->batch[k].d_descs);
  dev_free(c->pts);
  dev_free(c->sum_w);
  dev_free(c->rng);
  dev_free(c->best_f);
  dev_free(c->best_params);
  dev_free(c->best_dV);
  if (c->cg_exec) (void)hipGraphExecDestroy(c->cg_exec);
  for (int k = 0; k < NSTREAMS; ++k) {
    if (c->streams[k]) (void)hipStreamDestroy(c->streams[k]);
    if (c->ev0[k]) (void)hipEventDestroy(c->ev0[k]);
    if (c->ev1[k]) (void)hipEventDestroy(c->ev1[k]);
    if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
    if (c->ev_pre[k]) (void)hipEventDestroy(c->ev_pre[k]);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  delete c;
  return AST_OK;
}

int ast_set_model_space(ast_ctx* c, int n_params, const int* property, int n_types, const int* types,
                        const double* limits, const ast_x2* x2) {
  if (!c) return AST_E_ARG;
  SpaceH S;
  RETIF(make_space(c->v, n_params, property, n_types, types, limits, x2, &S));
  c->space = S;
  c->space_set = true;
  return AST_OK;
}

int ast_upload_curve(ast_ctx* c, const ast_obs* obs, int n_data, int n_filters, const ast_ephem3* eph) {
  if (!c) return AST_E_ARG;
  if (c->v.n_seg > 1 && !c->seg_set) return AST_E_STATE;
  CurveH C;
  RETIF(make_curve(c->v, obs, n_data, n_filters, eph, &C, c->seg_set ? c->seg_t : nullptr,
                   c->seg_first_set ? c->seg_first : nullptr));
  {  // the tail record: the context's minima with this curve's TIME_STEP
    TailRec tail = c->minima;
    tail.ts = c->v.time_step;
    std::memcpy(&C.pts[C.n_data], &tail, sizeof tail);  // after the table (SEGMENT may drop leading observations)
  }
  RETIF(set_device(c));
  dev_free(c->pts);
  dev_free(c->sum_w);
  RETIF(dev_alloc(&c->pts, C.pts.size()));
  RETIF(dev_alloc(&c->sum_w, (size_t)AST_MAX_FILTERS));
  HIPCHK(hipMemcpy(c->pts, C.pts.data(), sizeof(PointRec) * C.pts.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->sum_w, C.sum_w, sizeof(double) * AST_MAX_FILTERS, hipMemcpyHostToDevice));
  c->curve = std::move(C);
  c->curve_set = true;
  return AST_OK;
}

int ast_set_minima(ast_ctx* c, const float* t_obs, const float* V_obs, int n_obs) {
  if (!c) return AST_E_ARG;
  if (!c->v.nudge) return AST_E_STATE;
  TailRec r{};
  RETIF(make_minima(c->v, t_obs, V_obs, n_obs, &r));
  c->minima = r;
  c->minima_set = true;
  if (c->curve_set) {
    const int n = c->curve.n_data;
    r.ts = c->v.time_step;
    std::memcpy(&c->curve.pts[n], &r, sizeof r);
    RETIF(set_device(c));
    HIPCHK(hipMemcpy(c->pts + n, &c->curve.pts[n], sizeof(PointRec), hipMemcpyHostToDevice));
  }
  return AST_OK;
}

int ast_set_segments(ast_ctx* c, const double* t_start, int n_seg) {
  if (!c) return AST_E_ARG;
  if (c->v.n_seg <= 1) return AST_E_STATE;
  if (!t_start || n_seg != c->v.n_seg) return AST_E_ARG;
  for (int k = 0; k < n_seg; ++k)
    if (!std::isfinite(t_start[k])) return AST_E_ARG;
  for (int k = 0; k < n_seg; ++k) c->seg_t[k] = t_start[k];
  c->seg_set = true;
  c->seg_first_set = false;
  return AST_OK;
}

int ast_set_segment_starts(ast_ctx* c, const int* first_obs, const double* t_start, int n_seg) {
  if (!c) return AST_E_ARG;
  if (c->v.n_seg <= 1) return AST_E_STATE;
  if (!first_obs || !t_start || n_seg != c->v.n_seg) return AST_E_ARG;
  for (int k = 0; k < n_seg; ++k)
    if (!std::isfinite(t_start[k]) || first_obs[k] < 0 || (k > 0 && first_obs[k] <= first_obs[k - 1]))
      return AST_E_ARG;
  for (int k = 0; k < n_seg; ++k) {
    c->seg_t[k] = t_start[k];
    c->seg_first[k] = first_obs[k];
  }
  c->seg_set = true;
  c->seg_first_set = true;
  return AST_OK;
}

int ast_rng_seed(ast_ctx* c, uint64_t seed) {
  if (!c) return AST_E_ARG;
  if (c->run_active) return AST_E_STATE;  // the running job counts on its streams
  RETIF(set_device(c));
  c->seed = seed;
  if (c->rng) HIPCHK(hipMemset(c->rng, 0, c->rng_cap * sizeof(uint32_t)));
  RETIF(reset_best_f(c));
  HIPCHK(hipDeviceSynchronize());  // before any kernel on the non-blocking streams reads them
  return AST_OK;
}

int ast_reset_best(ast_ctx* c) {
  if (!c) return AST_E_ARG;
  if (c->run_active) return AST_E_STATE;
  RETIF(set_device(c));
  return reset_best_f(c);
}

int ast_set_start(ast_ctx* c, const double* params0) {
  if (!c || !params0) return AST_E_ARG;
  if (!c->space_set) return AST_E_STATE;
  for (int i = 0; i < c->space.n_params; ++i) c->params0[i] = params0[i];
  return AST_OK;
}

int ast_run_begin(ast_ctx* c, const ast_run_cfg* rc) {
  if (!c || !rc) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (!c->space_set || !c->curve_set || (c->v.nudge && !c->minima_set)) return AST_E_STATE;
  const int G = rc->group_size;
  if (rc->n_groups <= 0 || G < 64 || G > 1024 || (G & 63) || rc->nstages <= 0 || rc->n_steps < 0 ||
      rc->eval_budget < 0 || rc->resident_groups < 0)
    return AST_E_ARG;
  const int workers = rc->resident_groups > 0 && rc->resident_groups < rc->n_groups ? rc->resident_groups
                                                                                     : rc->n_groups;
  if (workers < rc->n_groups && rc->nstages != 1) return AST_E_ARG;  // stage reseeds need whole groups
  if (c->curve.n_all - c->space.n_params - c->curve.n_filters <= 0) return AST_E_DATA;
  RETIF(set_device(c));
  c->run_active = false;
  c->last_groups = 0;
  const size_t fits = (size_t)rc->n_groups * G, lanes = (size_t)workers * G;
  RETIF(ensure_streams(c, fits, rc->n_groups));
  if (c->pool.nf < lanes || c->pool.G != G || c->pool.P < c->space.n_params)
    RETIF(pool_alloc(c->pool, lanes, G, c->space.n_params));
  // The reference path is one job that borrows the context's streams and bests.
  if (c->ref_job) job_free(*c->ref_job);
  c->ref_job.reset(new DevJob());
  DevJob& j = *c->ref_job;
  j.uid = c->next_uid++;
  j.space = c->space;
  j.rp.n_groups = rc->n_groups;
  j.rp.group_size = G;
  j.rp.nstages = rc->nstages;
  j.rp.n_steps = rc->n_steps;
  j.rp.eval_budget = rc->eval_budget;
  j.rp.size_min = rc->size_min;
  j.rp.seed = c->seed;
  j.n_workers = workers;
  j.borrowed = true;
  j.rng = c->rng;
  j.best_f = c->best_f;
  j.best_params = c->best_params;
  j.best_dV = c->best_dV;
  j.cls = class_of(c->v, c->curve.n_filters);
  const hipStream_t s = c->streams[stream_of(j.cls)];
  for (int g = 0; g < workers; ++g) j.slots.push_back(g);
  RETIF(job_upload(j, c->pool, c->curve, c->params0, s));
  // The context's streams and bests were cleared by synchronous-API calls on
  // the null stream, which the job's non-blocking stream does not wait for.
  HIPCHK(hipDeviceSynchronize());
  int ticks;
  if (rc->ticks_per_launch > 0) {
    ticks = rc->ticks_per_launch > 4096 ? 4096 : rc->ticks_per_launch;
  } else {  // about 100 ms of kernel per launch when the GPU is full
    double tot = (double)lanes * c->curve.n_data;
    double t = 3.0e9 / (tot > 1 ? tot : 1);
    ticks = (int)(t < 16 ? 16 : (t > 4096 ? 4096 : t));
  }
  j.ticks = ticks;
  RETIF(job_descs(j, ticks));
  HIPCHK(job_launch(j, c->pool, 0, s));
  c->last_ms = 0;
  c->last_launches = 0;
  c->last_evals = 0;
  c->last_points = 0;
  // Every tick completes one evaluation of each running simplex; a stage
  // needs at most (N+1) + n_steps*(N+2) of them, and with refill a lane runs
  // its share of the fits one after another.
  const double per_stage = (double)(c->space.n_params + 1) + (double)rc->n_steps * (c->space.n_params + 2);
  const double rounds = std::ceil((double)fits / (double)lanes) + 1.0;
  c->run_max_cycles = (long)((per_stage * rc->nstages * rounds + 16.0) / ticks) + 4;
  c->run_cycles = 0;
  c->run_graphs = rc->use_graphs != 0;
  c->run_active = true;
  return AST_OK;
}

int ast_run_continue(ast_ctx* c, int max_cycles, int* finished) {
  if (!c) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (!c->run_active || !c->ref_job) return AST_E_STATE;
  RETIF(set_device(c));
  DevJob& j = *c->ref_job;
  std::vector<DevJob*> jobs{&j};
  for (long L = 0; !job_complete(j) && (max_cycles <= 0 || L < max_cycles); ++L) {
    if (c->run_cycles > c->run_max_cycles) {
      c->run_active = false;
      set_err("fit did not terminate");
      return AST_E_STATE;
    }
    int e = run_cycle(c, jobs, &c->last_ms, &c->last_launches, c->run_graphs);
    if (e) {
      c->run_active = false;
      return e;
    }
    c->run_cycles++;
  }
  int64_t ev = 0;
  RETIF(job_evals(c, j, &ev));
  c->last_evals = ev;
  c->last_points = ev * c->curve.n_data;
  const bool fin = job_complete(j);
  if (fin) {
    c->run_active = false;
    c->last_groups = j.rp.n_groups;
    c->last_group_size = j.rp.group_size;
  }
  if (finished) *finished = fin ? 1 : 0;
  return AST_OK;
}

int ast_run_status(ast_ctx* c, int64_t* fits_started, int64_t* fits_done, int* finished) {
  if (!c) return AST_E_ARG;
  if (!c->ref_job) return AST_E_STATE;
  RETIF(set_device(c));
  const DevJob& j = *c->ref_job;
  const int64_t all = (int64_t)j.rp.n_groups * j.rp.group_size;
  int64_t st = 0, dn = 0;
  if (j.refill) {  // the queue's counters (a finished lane takes one more start than there are fits)
    int32_t q[2] = {0, 0};
    HIPCHK(hipMemcpy(q, j.queue, sizeof q, hipMemcpyDeviceToHost));
    st = std::min<int64_t>(q[0], all);
    dn = std::min<int64_t>(q[1], all);
  } else {  // whole groups end together
    st = all;
    dn = (int64_t)j.groups_done * j.rp.group_size;
  }
  if (fits_started) *fits_started = st;
  if (fits_done) *fits_done = dn;
  if (finished) *finished = job_complete(j) ? 1 : 0;
  return AST_OK;
}

int ast_run_cancel(ast_ctx* c) {
  if (!c) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (!c->run_active) return AST_OK;
  RETIF(set_device(c));
  HIPCHK(hipDeviceSynchronize());  // nothing of the run is in flight after a cycle; drain anyway
  c->run_active = false;
  return AST_OK;
}

int ast_run(ast_ctx* c, const ast_run_cfg* rc) {
  RETIF(ast_run_begin(c, rc));
  int fin = 0;
  RETIF(ast_run_continue(c, 0, &fin));
  return fin ? AST_OK : AST_E_STATE;
}

int ast_fetch(ast_ctx* c, double* f, double* params, double* dV) {
  if (!c) return AST_E_ARG;
  if (!c->last_groups) return AST_E_STATE;
  RETIF(set_device(c));
  const int G = c->last_groups, N = c->space.n_params, NF = c->curve.n_filters;
  if (f) {
    if (c->best_cf == 8) {
      HIPCHK(hipMemcpy(f, c->best_f, sizeof(double) * G, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(G);
      HIPCHK(hipMemcpy(tmp.data(), c->best_f, sizeof(float) * G, hipMemcpyDeviceToHost));
      for (int g = 0; g < G; ++g) f[g] = tmp[g];
    }
  }
  if (params) HIPCHK(hipMemcpy(params, c->best_params, sizeof(double) * G * N, hipMemcpyDeviceToHost));
  if (dV) HIPCHK(hipMemcpy(dV, c->best_dV, sizeof(double) * G * NF, hipMemcpyDeviceToHost));
  return AST_OK;
}

int ast_fetch_lanes(ast_ctx* c, double* f_last, int64_t* n_evals, int64_t* n_steps) {
  if (!c) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (!c->last_groups || !c->ref_job) return AST_E_STATE;
  RETIF(set_device(c));
  const DevJob& j = *c->ref_job;
  const size_t fits = (size_t)j.rp.n_groups * j.rp.group_size;
  if (f_last) {
    if (j.cf == 8) {
      HIPCHK(hipMemcpy(f_last, j.fit_f, sizeof(double) * fits, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(fits);
      HIPCHK(hipMemcpy(tmp.data(), j.fit_f, sizeof(float) * fits, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < fits; ++k) f_last[k] = tmp[k];
    }
  }
  std::vector<int32_t> tmp(fits);
  if (n_evals) {
    HIPCHK(hipMemcpy(tmp.data(), j.fit_nev, sizeof(int32_t) * fits, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < fits; ++k) n_evals[k] = tmp[k];
  }
  if (n_steps) {
    HIPCHK(hipMemcpy(tmp.data(), j.fit_steps, sizeof(int32_t) * fits, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < fits; ++k) n_steps[k] = tmp[k];
  }
  return AST_OK;
}

int ast_chi2_batch(ast_ctx* c, const double* params, int n, double* chi2, double* dV) {
  if (!c || !params || n <= 0 || !chi2 || !dV) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;  // its streams and events belong to the scheduler
  if (!c->space_set || !c->curve_set || (c->v.nudge && !c->minima_set)) return AST_E_STATE;
  if (c->curve.n_all - c->space.n_params - c->curve.n_filters <= 0) return AST_E_DATA;
  RETIF(set_device(c));
  double *pd = nullptr, *cd = nullptr, *dd = nullptr;
  int rc = dev_alloc(&pd, (size_t)n * c->space.n_params);
  if (!rc) rc = dev_alloc(&cd, (size_t)n);
  if (!rc) rc = dev_alloc(&dd, (size_t)n * c->curve.n_filters);
  if (rc) { dev_free(pd); dev_free(cd); dev_free(dd); return rc; }
  int result = AST_OK;
  hipStream_t s = c->streams[0];
  hipError_t he = hipMemcpy(pd, params, sizeof(double) * n * c->space.n_params, hipMemcpyHostToDevice);
  LaunchArgs a{};
  a.op = 2;
  a.stream = s;
  a.pts = c->pts;
  a.sum_w = c->sum_w;
  a.n_data = c->curve.n_data;
  a.n_filters = c->curve.n_filters;
  a.dof = c->curve.n_all - c->space.n_params - c->curve.n_filters;
  a.params = pd;
  a.n = n;
  a.chi2 = cd;
  a.dV = dd;
  if (he == hipSuccess) he = hipEventRecord(c->ev0[0], s);
  if (he == hipSuccess) he = dispatch(c->v, a);
  if (he == hipSuccess) he = hipEventRecord(c->ev1[0], s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, c->ev0[0], c->ev1[0]);
  if (he == hipSuccess) he = hipMemcpy(chi2, cd, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(dV, dd, sizeof(double) * n * c->curve.n_filters, hipMemcpyDeviceToHost);
  if (he != hipSuccess) {
    snprintf(g_hip_err, sizeof g_hip_err, "%s (chi2 batch)", hipGetErrorString(he));
    result = AST_E_HIP;
  } else {
    c->last_ms = ms;
    c->last_launches = 1;
    c->last_evals = n;
    c->last_points = (int64_t)n * c->curve.n_data;
  }
  dev_free(pd);
  dev_free(cd);
  dev_free(dd);
  return result;
}

// A plain streaming copy of n 16-byte words (one load and one store per
// word, grid-stride) and a fill, for ast_copy_bandwidth.
__global__ void __launch_bounds__(256) copy_words_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
__global__ void __launch_bounds__(256) fill_words_kernel(uint4* __restrict__ dst, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = make_uint4(v, v ^ 1u, v ^ 2u, v ^ 3u);
}

int ast_copy_bandwidth(ast_ctx* c, int64_t bytes, int reps, double* bytes_per_s) {
  if (!c || bytes < 16 || reps <= 0 || !bytes_per_s) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;  // its streams and events belong to the scheduler
  RETIF(set_device(c));
  const size_t n = (size_t)bytes / 16;
  uint4 *src = nullptr, *dst = nullptr;
  int rc = dev_alloc(&src, n);
  if (!rc) rc = dev_alloc(&dst, n);
  if (rc) { dev_free(src); dev_free(dst); return rc; }
  hipStream_t s = c->streams[0];
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_words_kernel, dim3(blocks), dim3(256), 0, s, src, n, 0x5a5a5a5au);
  hipError_t he = hipGetLastError();
  if (he == hipSuccess) he = hipEventRecord(c->ev0[0], s);
  for (int r = 0; r < reps && he == hipSuccess; ++r) {
    hipLaunchKernelGGL(copy_words_kernel, dim3(blocks), dim3(256), 0, s, src, dst, n);
    he = hipGetLastError();
  }
  if (he == hipSuccess) he = hipEventRecord(c->ev1[0], s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, c->ev0[0], c->ev1[0]);
  dev_free(src);
  dev_free(dst);
  if (he != hipSuccess) {
    snprintf(g_hip_err, sizeof g_hip_err, "%s (copy bandwidth)", hipGetErrorString(he));
    return AST_E_HIP;
  }
  c->last_ms = ms;
  c->last_launches = reps;
  *bytes_per_s = ms > 0 ? 2.0 * 16.0 * (double)n * reps / (ms * 1e-3) : 0.0;
  return AST_OK;
}

int ast_chi2_batch_noalloc(ast_ctx* c, const double* params, int n, double* chi2, double* dV) {
  if (!c || !params || n <= 0 || !chi2 || !dV) return AST_E_ARG;
  if (n > AST_NOALLOC_MAX_MODELS) {
    snprintf(g_hip_err, sizeof g_hip_err, "ast_chi2_batch_noalloc takes at most %d models per call (got %d)",
             (int)AST_NOALLOC_MAX_MODELS, n);
    return AST_E_ARG;
  }
  if (c->serve) return AST_E_BUSY;
  if (!c->space_set || !c->curve_set || (c->v.nudge && !c->minima_set)) return AST_E_STATE;
  if (c->curve.n_all - c->space.n_params - c->curve.n_filters <= 0) return AST_E_DATA;
  double* const base = c->device >= 0 && c->device < SCRATCH_MAX_DEVICES ? g_scratch_ptr[c->device] : nullptr;
  if (!base) return AST_E_STATE;
  RETIF(set_device(c));
  double* pd = base;                     // models, n * n_params
  double* cd = base + NOALLOC_PARAMS;    // chi^2, n
  double* dd = cd + NOALLOC_CHI2;        // delta_V, n * n_filters
  const hipStream_t s = c->streams[0];
  const int N = c->space.n_params, NF = c->curve.n_filters;
  LaunchArgs a{};
  a.op = 2;
  a.stream = s;
  a.pts = c->pts;
  a.sum_w = c->sum_w;
  a.n_data = c->curve.n_data;
  a.n_filters = NF;
  a.dof = c->curve.n_all - N - NF;
  a.params = pd;
  a.n = n;
  a.chi2 = cd;
  a.dV = dd;
  hipError_t he = hipMemcpyAsync(pd, params, sizeof(double) * (size_t)n * N, hipMemcpyHostToDevice, s);
  if (he == hipSuccess) he = hipEventRecord(c->ev0[0], s);
  if (he == hipSuccess) he = dispatch(c->v, a);
  if (he == hipSuccess) he = hipEventRecord(c->ev1[0], s);
  if (he == hipSuccess) he = hipMemcpyAsync(chi2, cd, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipMemcpyAsync(dV, dd, sizeof(double) * (size_t)n * NF, hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, c->ev0[0], c->ev1[0]);
  if (he != hipSuccess) {
    snprintf(g_hip_err, sizeof g_hip_err, "%s (chi2 batch, no allocation)", hipGetErrorString(he));
    return AST_E_HIP;
  }
  c->last_ms = ms;
  c->last_launches = 1;
  c->last_evals = n;
  c->last_points = (int64_t)n * c->curve.n_data;
  return AST_OK;
}

int ast_rmsd(ast_ctx* c, const ast_rmsd_cfg* cfg, float* par_min, float* par_max, int64_t* n_tot, int64_t* n_bad,
             double* f0, double* f1) {
  if (!c || !cfg || !par_min || !par_max) return AST_E_ARG;
  if (cfg->n_groups <= 0 || cfg->group_size <= 0 || cfg->group_size > 1024 || cfg->n_samples < 0 ||
      !(cfg->dx >= 0))
    return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (c->run_active) return AST_E_STATE;  // it shares the context's random streams
  if (!c->space_set || !c->curve_set || (c->v.nudge && !c->minima_set)) return AST_E_STATE;
  if (c->v.n_seg > 1) return AST_E_UNSUPPORTED;
  if (c->curve.n_all - c->space.n_params - c->curve.n_filters <= 0) return AST_E_DATA;
  RETIF(set_device(c));
  RmsdOut o{par_min, par_max, 0, 0, 0, 0, 0};
  const int rc = c->v.precision == AST_PREC_FP64 ? rmsd_run<double>(c, *cfg, &o) : rmsd_run<float>(c, *cfg, &o);
  if (rc) return rc;
  if (n_tot) *n_tot = o.n_tot;
  if (n_bad) *n_bad = o.n_bad;
  if (f0) *f0 = o.f0;
  if (f1) *f1 = o.f1;
  c->last_ms = o.ms;
  c->last_launches = 1;
  c->last_evals = o.n_tot;
  c->last_points = o.n_tot * c->curve.n_data;
  return AST_OK;
}

int ast_model_curve(ast_ctx* c, const double* params, int n, const ast_ephem3* eph, int n_plot, double* t,
                    double* vmod, double* chi2, double* dV) {
  if (!c || !params || n <= 0 || !eph || n_plot < 2 || !vmod || !chi2 || !dV) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (!c->space_set || !c->curve_set || (c->v.nudge && !c->minima_set)) return AST_E_STATE;
  if (c->curve.n_all - c->space.n_params - c->curve.n_filters <= 0) return AST_E_DATA;
  std::vector<PointRec> grid;
  std::vector<double> tg;
  int g0 = 0;
  RETIF(make_grid(c->v, c->curve, *eph, n_plot, &grid, &tg, c->seg_t, &g0));
  RETIF(set_device(c));
  const int np = c->space.n_params, nf = c->curve.n_filters;
  PointRec* gd = nullptr;
  double *pd = nullptr, *cd = nullptr, *dd = nullptr, *vd = nullptr;
  int rc = dev_alloc(&gd, grid.size());
  if (!rc) rc = dev_alloc(&pd, (size_t)n * np);
  if (!rc) rc = dev_alloc(&cd, (size_t)n);
  if (!rc) rc = dev_alloc(&dd, (size_t)n * nf);
  if (!rc) rc = dev_alloc(&vd, (size_t)n * n_plot);
  auto release = [&]() { dev_free(gd); dev_free(pd); dev_free(cd); dev_free(dd); dev_free(vd); };
  if (rc) { release(); return rc; }
  hipStream_t s = c->streams[0];
  hipError_t he = hipMemcpy(gd, grid.data(), sizeof(PointRec) * grid.size(), hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(pd, params, sizeof(double) * n * np, hipMemcpyHostToDevice);
  LaunchArgs a{};
  a.op = 3;
  a.stream = s;
  a.pts = c->pts;
  a.sum_w = c->sum_w;
  a.n_data = c->curve.n_data;
  a.n_filters = nf;
  a.dof = c->curve.n_all - np - nf;
  a.params = pd;
  a.n = n;
  a.chi2 = cd;
  a.dV = dd;
  a.grid = gd;
  a.n_plot = n_plot;
  a.vmod = vd;
  if (he == hipSuccess) he = hipEventRecord(c->ev0[0], s);
  if (he == hipSuccess) he = dispatch(c->v, a);
  if (he == hipSuccess) he = hipEventRecord(c->ev1[0], s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, c->ev0[0], c->ev1[0]);
  if (he == hipSuccess) he = hipMemcpy(chi2, cd, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(dV, dd, sizeof(double) * n * nf, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(vmod, vd, sizeof(double) * n * (size_t)n_plot, hipMemcpyDeviceToHost);
  if (he == hipSuccess)
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < g0; ++i) vmod[(size_t)k * n_plot + i] = std::nan("");
  int result = AST_OK;
  if (he != hipSuccess) {
    (void)hipDeviceSynchronize();
    snprintf(g_hip_err, sizeof g_hip_err, "%s (model curve)", hipGetErrorString(he));
    result = AST_E_HIP;
  } else {
    if (t) std::memcpy(t, tg.data(), sizeof(double) * n_plot);
    c->last_ms = ms;
    c->last_launches = 1;
    c->last_evals = n;
    c->last_points = (int64_t)n * (c->curve.n_data + n_plot);
  }
  release();
  return result;
}

int ast_periodogram(ast_ctx* c, const double* t, const double* vm, int n, int n_clusters, double* cl_fr,
                    double* cl_H, int* n_minima) {
  if (!c || !t || !vm || n < 0 || n_clusters < 0 || (n_clusters > 0 && (!cl_fr || !cl_H))) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  // Brightness minima are maxima of the magnitude (misc.c:74-92).
  std::vector<double> tm;
  for (int i = 1; i <= n - 2; ++i)
    if (vm[i] > vm[i - 1] && vm[i] >= vm[i + 1]) tm.push_back(t[i]);
  if (n_minima) *n_minima = (int)tm.size();
  if (tm.size() > (size_t)PGRAM_MAX_MINIMA) {
    set_err("more than 10000 minima (misc.c:80-84)");
    return AST_E_DATA;
  }
  std::vector<int32_t> counts;
  RETIF(set_device(c));
  RETIF(pgram_counts(c->streams[0], tm, &counts, &c->last_ms));
  c->last_launches = tm.size() >= 2 ? 1 : 0;
  pgram_clusters(counts, n_clusters, cl_fr, cl_H);
  return AST_OK;
}

int ast_last_timing(ast_ctx* c, double* kernel_ms, int* launches) {
  if (!c) return AST_E_ARG;
  if (kernel_ms) *kernel_ms = c->last_ms;
  if (launches) *launches = c->last_launches;
  return AST_OK;
}

int ast_last_counts(ast_ctx* c, int64_t* evals, int64_t* model_points) {
  if (!c) return AST_E_ARG;
  if (evals) *evals = c->last_evals;
  if (model_points) *model_points = c->last_points;
  return AST_OK;
}

// ------------------------------------------------------------------ serving
int ast_serve_start(ast_ctx* c, const ast_serve_cfg* cfg) {
  if (!c || !cfg) return AST_E_ARG;
  if (c->serve || c->run_active) return AST_E_STATE;
  const int G = cfg->group_size ? cfg->group_size : 256;
  const int P = cfg->max_params ? cfg->max_params : AST_MAX_PARAMS;
  if (G < 64 || G > 1024 || (G & 63) || cfg->max_resident_fits < G || P < 0 || P > AST_MAX_PARAMS)
    return AST_E_ARG;
  RETIF(set_device(c));
  size_t fits = (size_t)(cfg->max_resident_fits / G) * G;
  RETIF(pool_alloc(c->pool, fits, G, P));
  c->serve.reset(new Serve());
  c->serve->cfg = *cfg;
  c->serve->cfg.group_size = G;
  if (c->serve->cfg.work_points_per_launch <= 0) c->serve->cfg.work_points_per_launch = 16000;
  c->serve->th = std::thread(serve_loop, c);
  return AST_OK;
}

int ast_submit(ast_ctx* c, const ast_job* job, uint64_t* id) {
  if (!c || !job || !id) return AST_E_ARG;
  if (!c->serve) return AST_E_STATE;
  Serve& S = *c->serve;
  std::unique_ptr<DevJob> j;
  CurveH C;
  std::vector<double> p0;
  RETIF(make_job(c, job, S.cfg.group_size, &j, &C, &p0));
  // a multi-stage job needs a slot per group; a single-stage one can run on
  // fewer (its lanes take the rest of its fits from its queue)
  if (j->rp.nstages != 1 && j->rp.n_groups > c->pool.n_slots) {
    set_err("a multi-stage job has more groups than the pool has slots");
    return AST_E_NOMEM;
  }
  if (j->space.n_params > c->pool.P) {
    set_err("the job has more parameters than the serving slots hold (ast_serve_cfg.max_params)");
    return AST_E_ARG;
  }
  std::lock_guard<std::mutex> lk(S.mu);
  j->id = S.next_id++;
  *id = j->id;
  S.inputs[j->id] = std::make_pair(std::move(C), std::move(p0));
  S.pending.push_back(std::move(j));
  S.cv_submit.notify_all();
  return AST_OK;
}

int ast_run_jobs(ast_ctx* c, const ast_job* jobs, int n_jobs, const ast_batch_cfg* cfg, ast_result* results) {
  if (!c || !jobs || n_jobs <= 0 || !results) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (c->run_active) return AST_E_STATE;  // the slot pool holds a running ast_run_begin job
  const int G = (cfg && cfg->group_size) ? cfg->group_size : 256;
  if (G < 64 || G > 1024 || (G & 63)) return AST_E_ARG;
  const double work = (cfg && cfg->work_points_per_launch > 0) ? cfg->work_points_per_launch : 16000.0;
  std::vector<const ast_job*> js;
  for (int i = 0; i < n_jobs; ++i) js.push_back(&jobs[i]);
  c->last_ms = 0;
  c->last_launches = 0;
  return batch_run(c, js, G, work, 0, results, &c->last_ms, &c->last_launches, cfg && cfg->use_graphs,
                   cfg ? cfg->schedule : 0);
}

int ast_set_waves(ast_ctx* c, const ast_variant* v, int n_filters, int waves) {
  if (!c || !v || n_filters <= 0 || n_filters > AST_MAX_FILTERS || (waves != 2 && waves != 3 && waves != 4))
    return AST_E_ARG;
  if (!variant_supported(*v)) return AST_E_UNSUPPORTED;
  c->waves[class_of(*v, n_filters)] = waves;
  return AST_OK;
}

int ast_autotune(ast_ctx* c, const ast_job* jobs, int n_jobs, const ast_batch_cfg* cfg, int cycles,
                 ast_tune_entry* out, int max_out, int* n_out) {
  if (!c || !jobs || n_jobs <= 0 || (max_out > 0 && !out) || !n_out) return AST_E_ARG;
  if (c->serve) return AST_E_BUSY;
  if (c->run_active) return AST_E_STATE;
  *n_out = 0;
  const int G = (cfg && cfg->group_size) ? cfg->group_size : 256;
  if (G < 64 || G > 1024 || (G & 63)) return AST_E_ARG;
  const double work = (cfg && cfg->work_points_per_launch > 0) ? cfg->work_points_per_launch : 16000.0;
  if (cycles <= 0) cycles = 4;
  // the jobs of each kernel class
  std::map<int, std::vector<const ast_job*>> by_class;
  for (int i = 0; i < n_jobs; ++i) {
    const ast_job& j = jobs[i];
    if (j.n_filters <= 0 || j.n_filters > AST_MAX_FILTERS) return AST_E_DATA;
    if (!variant_supported(j.variant)) return AST_E_UNSUPPORTED;
    by_class[class_of(j.variant, j.n_filters)].push_back(&j);
  }
  const int cand[3] = {4, 3, 2};
  for (auto& kv : by_class) {
    const int k = kv.first;
    double ms[3] = {0, 0, 0};
    int best = 4;
    double best_ms = 0;
    // Two interleaved passes over the candidates, each kept at its faster
    // pass: the builds often run within a few percent of each other, so one
    // timing each would let clock drift between them pick the build.
    for (int pass = 0; pass < 2; ++pass) {
      for (int w = 0; w < 3; ++w) {
        if (cand[w] < 4 && G > 256) continue;
        c->waves[k] = cand[w];
        int launches = 0;
        double t = 0;
        // the same jobs from the same start each time: identical work, so the
        // kernel times compare the register budgets alone
        RETIF(batch_run(c, kv.second, G, work, cycles, nullptr, &t, &launches, false));
        if (pass == 0 || t < ms[w]) ms[w] = t;
      }
    }
    for (int w = 0; w < 3; ++w)
      if (ms[w] > 0 && (best_ms == 0 || ms[w] < best_ms)) { best_ms = ms[w]; best = cand[w]; }
    c->waves[k] = best;
    if (*n_out < max_out) {
      ast_tune_entry& e = out[(*n_out)++];
      std::memset(&e, 0, sizeof e);
      const ast_job& j = *kv.second[0];
      e.torque = j.variant.torque;
      e.bc = j.variant.bc;
      e.bw_ball = j.variant.bw_ball;
      e.precision = j.variant.precision;
      e.filter_capacity = j.n_filters <= 1 ? 1 : (j.n_filters <= 4 ? 4 : AST_MAX_FILTERS);
      e.waves = best;
      for (int w = 0; w < 3; ++w) e.kernel_ms[w] = ms[w];
    }
  }
  return AST_OK;
}

int ast_poll(ast_ctx* c, ast_result* out, int max, int* n_out, int timeout_ms) {
  if (!c || !out || max <= 0 || !n_out) return AST_E_ARG;
  if (!c->serve) return AST_E_STATE;
  Serve& S = *c->serve;
  std::unique_lock<std::mutex> lk(S.mu);
  if (S.done.empty() && timeout_ms > 0)
    S.cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !S.done.empty(); });
  int n = 0;
  while (n < max && !S.done.empty()) {
    out[n++] = S.done.front();
    S.done.pop_front();
  }
  *n_out = n;
  return S.error ? S.error : AST_OK;
}

int ast_cancel(ast_ctx* c, uint64_t id) {
  if (!c) return AST_E_ARG;
  if (!c->serve) return AST_E_STATE;
  std::lock_guard<std::mutex> lk(c->serve->mu);
  c->serve->cancels.push_back(id);
  c->serve->cv_submit.notify_all();
  return AST_OK;
}

int ast_serve_stats(ast_ctx* c, ast_serve_stats_t* st) {
  if (!c || !st) return AST_E_ARG;
  if (!c->serve) return AST_E_STATE;
  Serve& S = *c->serve;
  std::lock_guard<std::mutex> lk(S.mu);
  st->pending = (int)S.pending.size();
  st->resident = (int)S.resident.size();
  st->free_fits = (int64_t)c->pool.free_slots.size() * S.cfg.group_size;
  st->resident_bytes = S.resident_bytes;
  st->launches = S.launches;
  st->kernel_ms = S.kernel_ms;
  st->evals = S.evals;
  st->model_points = S.points;
  return AST_OK;
}

int ast_serve_stop(ast_ctx* c) {
  if (!c) return AST_E_ARG;
  if (!c->serve) return AST_E_STATE;
  {
    std::lock_guard<std::mutex> lk(c->serve->mu);
    c->serve->stop = true;
    c->serve->cv_submit.notify_all();
  }
  if (c->serve->th.joinable()) c->serve->th.join();
  c->serve.reset();
  return AST_OK;
}

}  // extern "C"
