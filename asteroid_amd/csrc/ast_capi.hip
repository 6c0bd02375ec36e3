// ast_capi.hip -- host side of libasteroid_hip.so (the C ABI of include/asteroid.h).
//
// Owns the device buffers of one context, builds the per-observation table
// the kernels read (INTERP ephemeris evaluated once per observation and the
// RK4 interval constants, in the canonical FMA order), drives the chunked
// fitting launches until every group has finished, and times every launch
// with HIP events on the context's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/asteroid.h"
#include "ast_kernels.h"

namespace ast {
template <bool TQ, bool BC, bool BW> hipError_t launch_variant(const LaunchArgs& a);
extern template hipError_t launch_variant<false, false, false>(const LaunchArgs&);
extern template hipError_t launch_variant<false, true, false>(const LaunchArgs&);
extern template hipError_t launch_variant<false, false, true>(const LaunchArgs&);
extern template hipError_t launch_variant<true, false, false>(const LaunchArgs&);
extern template hipError_t launch_variant<true, true, false>(const LaunchArgs&);
extern template hipError_t launch_variant<true, false, true>(const LaunchArgs&);
}  // namespace ast

using namespace ast;

static thread_local char g_hip_err[256] = "";

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      snprintf(g_hip_err, sizeof g_hip_err, "%s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
      return e_ == hipErrorOutOfMemory ? AST_E_NOMEM : AST_E_HIP;                      \
    }                                                                                  \
  } while (0)

namespace {

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

struct Job {  // device buffers of the simplex state, typed by the simplex float
  size_t fits_cap = 0, groups_cap = 0;
  int N = 0;
  void* X = nullptr; void* F = nullptr; uint8_t* IND = nullptr; void* X0 = nullptr; void* XR = nullptr;
  void* XP = nullptr; double* PRM = nullptr; void* DV = nullptr;
  int32_t *ctl = nullptr, *lstep = nullptr, *nev = nullptr, *nev_tot = nullptr, *steps_tot = nullptr;
  void* fr = nullptr; void* sf = nullptr;
  int32_t *g_reopt = nullptr, *g_stage = nullptr, *g_done = nullptr, *active = nullptr;
};

}  // namespace

struct ast_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  ast_variant v{};
  ast_type_ids T{};
  int cf_size = 8;  // bytes of the simplex float
  // model space
  bool space_set = false;
  int n_params = 0, n_types = 0;
  int prop[AST_MAX_PARAMS][AST_N_COLUMNS]{};
  int types[AST_MAX_TYPES]{};
  double lim[2][AST_MAX_TYPES]{};
  ast_x2 x2{};
  Space<double> sp_d{};
  Space<float> sp_f{};
  // curve
  bool curve_set = false;
  int n_data = 0, n_filters = 0;
  PointRec* pts = nullptr;
  double* sum_w = nullptr;
  // start model
  double params0[AST_MAX_PARAMS]{};
  double* params0_d = nullptr;
  // random streams and per-group bests
  uint64_t seed = 0;
  uint32_t* rng = nullptr;
  size_t rng_cap = 0;
  void* best_f = nullptr;       // CF[groups]
  double* best_params = nullptr;
  double* best_dV = nullptr;
  size_t best_cap = 0;
  Job job;
  int last_groups = 0, last_group_size = 0;
  // timing / counts of the last run
  double last_ms = 0;
  int last_launches = 0;
  int64_t last_evals = 0, last_points = 0;
  std::vector<hipEvent_t> ev;
};

namespace {

bool variant_supported(const ast_variant& v) {
  if (v.bc && v.bw_ball) return false;
  if (v.precision != AST_PREC_FP64 && v.precision != AST_PREC_F32SIMPLEX) return false;
  if (!(v.time_step > 0)) return false;
  return true;
}

hipError_t dispatch(ast_ctx* c, LaunchArgs& a) {
  a.precision = c->v.precision;
  a.stream = c->stream;
  a.sp_d = &c->sp_d;
  a.sp_f = &c->sp_f;
  a.pts = c->pts;
  a.sum_w = c->sum_w;
  a.n_data = c->n_data;
  a.n_filters = c->n_filters;
  a.dof = c->n_data - c->n_params - c->n_filters;
  const bool tq = c->v.torque, bc = c->v.bc, bw = c->v.bw_ball;
  if (!tq && !bc && !bw) return launch_variant<false, false, false>(a);
  if (!tq && bc) return launch_variant<false, true, false>(a);
  if (!tq && bw) return launch_variant<false, false, true>(a);
  if (tq && !bc && !bw) return launch_variant<true, false, false>(a);
  if (tq && bc) return launch_variant<true, true, false>(a);
  return launch_variant<true, false, true>(a);
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

template <class CF>
void build_space(const ast_ctx* c, Space<CF>* sp) {
  std::memset(sp, 0, sizeof(*sp));
  for (int i = 0; i < c->n_params; ++i) {
    int t = c->prop[i][AST_P_TYPE];
    sp->frozen[i] = (int8_t)c->prop[i][AST_P_FROZEN];
    sp->periodic[i] = (int8_t)c->prop[i][AST_P_PERIODIC];
    sp->independent[i] = (int8_t)c->prop[i][AST_P_INDEPENDENT];
    sp->lim0[i] = (CF)c->lim[0][t];
    sp->lim1[i] = (CF)c->lim[1][t];
  }
  sp->p_psi = c->v.p_psi;
  sp->ppsi1 = c->x2.ppsi1;
  sp->ppsi2 = c->x2.ppsi2;
}

void free_job(Job& j) {
  dev_free(j.X); dev_free(j.F); dev_free(j.IND); dev_free(j.X0); dev_free(j.XR); dev_free(j.XP);
  dev_free(j.PRM); dev_free(j.DV); dev_free(j.ctl); dev_free(j.lstep); dev_free(j.nev); dev_free(j.nev_tot);
  dev_free(j.steps_tot); dev_free(j.fr); dev_free(j.sf); dev_free(j.g_reopt); dev_free(j.g_stage);
  dev_free(j.g_done); dev_free(j.active);
  j.fits_cap = j.groups_cap = 0;
}

int ensure_job(ast_ctx* c, size_t fits, size_t groups) {
  Job& j = c->job;
  const int N = c->n_params;
  if (j.fits_cap >= fits && j.groups_cap >= groups && j.N == N) return 0;
  free_job(j);
  const size_t cf = (size_t)c->cf_size;
  int rc = 0;
#define ALLOC(ptr, type, count) if (!rc) rc = dev_alloc((type**)&(ptr), (count))
  ALLOC(j.X, char, cf * (N + 1) * N * fits);
  ALLOC(j.F, char, cf * (N + 1) * fits);
  ALLOC(j.IND, uint8_t, (size_t)(N + 1) * fits);
  ALLOC(j.X0, char, cf * N * fits);
  ALLOC(j.XR, char, cf * N * fits);
  ALLOC(j.XP, char, cf * N * fits);
  ALLOC(j.PRM, double, (size_t)N * fits);
  ALLOC(j.DV, char, cf * AST_MAX_FILTERS * fits);
  ALLOC(j.ctl, int32_t, fits);
  ALLOC(j.lstep, int32_t, fits);
  ALLOC(j.nev, int32_t, fits);
  ALLOC(j.nev_tot, int32_t, fits);
  ALLOC(j.steps_tot, int32_t, fits);
  ALLOC(j.fr, char, cf * fits);
  ALLOC(j.sf, char, cf * fits);
  ALLOC(j.g_reopt, int32_t, groups);
  ALLOC(j.g_stage, int32_t, groups);
  ALLOC(j.g_done, int32_t, groups);
  ALLOC(j.active, int32_t, 1);
#undef ALLOC
  if (rc) { free_job(j); return rc; }
  j.fits_cap = fits;
  j.groups_cap = groups;
  j.N = N;
  return 0;
}

// Grow the random-stream counters and per-group bests, keeping old entries.
int ensure_streams(ast_ctx* c, size_t fits, size_t groups) {
  if (c->rng_cap < fits) {
    uint32_t* nr = nullptr;
    int rc = dev_alloc(&nr, fits);
    if (rc) return rc;
    HIPCHK(hipMemsetAsync(nr, 0, fits * sizeof(uint32_t), c->stream));
    if (c->rng) HIPCHK(hipMemcpyAsync(nr, c->rng, c->rng_cap * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_free(c->rng);
    c->rng = nr;
    c->rng_cap = fits;
  }
  if (c->best_cap < groups || !c->best_f) {
    const size_t cf = (size_t)c->cf_size;
    void* nf = nullptr;
    double *np = nullptr, *nd = nullptr;
    int rc = dev_alloc((char**)&nf, cf * groups);
    if (!rc) rc = dev_alloc(&np, groups * AST_MAX_PARAMS);
    if (!rc) rc = dev_alloc(&nd, groups * AST_MAX_FILTERS);
    if (rc) { dev_free(nf); dev_free(np); dev_free(nd); return rc; }
    std::vector<char> hf(cf * groups);
    for (size_t g = 0; g < groups; ++g) {
      if (cf == 8) { double v = 1e30; std::memcpy(&hf[g * 8], &v, 8); }
      else { float v = 1e30f; std::memcpy(&hf[g * 4], &v, 4); }
    }
    HIPCHK(hipMemcpy(nf, hf.data(), cf * groups, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(np, 0, groups * AST_MAX_PARAMS * sizeof(double)));
    HIPCHK(hipMemset(nd, 0, groups * AST_MAX_FILTERS * sizeof(double)));
    if (c->best_f) {
      HIPCHK(hipMemcpy(nf, c->best_f, cf * c->best_cap, hipMemcpyDeviceToDevice));
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
  }
  return 0;
}

int reset_best_f(ast_ctx* c) {
  if (!c->best_f) return 0;
  const size_t cf = (size_t)c->cf_size, n = c->best_cap;
  std::vector<char> hf(cf * n);
  for (size_t g = 0; g < n; ++g) {
    if (cf == 8) { double v = 1e30; std::memcpy(&hf[g * 8], &v, 8); }
    else { float v = 1e30f; std::memcpy(&hf[g * 4], &v, 4); }
  }
  HIPCHK(hipMemcpy(c->best_f, hf.data(), cf * n, hipMemcpyHostToDevice));
  return 0;
}

template <class CF>
void fill_bufs(ast_ctx* c, FitBufs<CF>* B) {
  Job& j = c->job;
  B->X = (CF*)j.X; B->F = (CF*)j.F; B->IND = j.IND; B->X0 = (CF*)j.X0; B->XR = (CF*)j.XR; B->XP = (CF*)j.XP;
  B->PRM = j.PRM; B->DV = (CF*)j.DV; B->ctl = j.ctl; B->lstep = j.lstep; B->nev = j.nev; B->nev_tot = j.nev_tot;
  B->steps_tot = j.steps_tot; B->rng = c->rng; B->fr = (CF*)j.fr; B->sf = (CF*)j.sf; B->g_reopt = j.g_reopt;
  B->g_stage = j.g_stage; B->g_done = j.g_done; B->g_best_f = (CF*)c->best_f; B->g_best_params = c->best_params;
  B->g_best_dV = c->best_dV; B->active_groups = j.active;
}

template <class CF>
int run_typed(ast_ctx* c, const ast_run_cfg* rc) {
  const size_t fits = (size_t)rc->n_groups * rc->group_size;
  int e = ensure_streams(c, fits, rc->n_groups);
  if (e) return e;
  e = ensure_job(c, fits, rc->n_groups);
  if (e) return e;
  FitBufs<CF> B;
  fill_bufs<CF>(c, &B);
  FitCfg<CF> C;
  C.n_fits = (int)fits;
  C.nstages = rc->nstages;
  C.n_steps = rc->n_steps;
  C.eval_budget = (int)(rc->eval_budget > 0x7fffffff ? 0x7fffffff : rc->eval_budget);
  double smin = rc->size_min > 0 ? rc->size_min : (c->v.precision == AST_PREC_FP64 ? 1e-10 : 1e-5);
  CF sm = (CF)smin;
  C.size2_min = sm * sm;
  C.seed = c->seed;
  C.n_data = c->n_data;
  C.n_filters = c->n_filters;
  C.dof = c->n_data - c->n_params - c->n_filters;
  C.reopt = c->x2.reopt;
  int ticks = rc->ticks_per_launch;
  if (ticks <= 0) {
    double work = (double)fits * (double)c->n_data;
    double tt = 3.0e9 / (work > 1 ? work : 1);
    ticks = tt < 16 ? 16 : (tt > 4096 ? 4096 : (int)tt);
  }
  C.ticks = ticks;
  HIPCHK(hipMemcpy(c->params0_d, c->params0, sizeof(c->params0), hipMemcpyHostToDevice));
  LaunchArgs a{};
  a.op = 0;
  a.bufs = &B;
  a.cfg = &C;
  a.params0 = c->params0_d;
  a.n_groups = rc->n_groups;
  a.group_size = rc->group_size;
  HIPCHK(dispatch(c, a));
  a.op = 1;
  c->last_ms = 0;
  c->last_launches = 0;
  // Upper bound on launches: every tick completes one evaluation of each
  // running simplex; a stage needs at most (N+1) + n_steps*(N+2) of them.
  const double per_stage = (double)(c->n_params + 1) + (double)rc->n_steps * (c->n_params + 2);
  const double max_ticks = per_stage * rc->nstages + 16.0;
  const long max_launches = (long)(max_ticks / ticks) + 4;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int result = 0;
  for (long L = 0;; ++L) {
    if (L > max_launches) { result = AST_E_STATE; snprintf(g_hip_err, sizeof g_hip_err, "fit did not terminate"); break; }
    hipError_t he = hipMemsetAsync(c->job.active, 0, sizeof(int32_t), c->stream);
    if (he == hipSuccess) he = hipEventRecord(e0, c->stream);
    if (he == hipSuccess) he = dispatch(c, a);
    if (he == hipSuccess) he = hipEventRecord(e1, c->stream);
    int32_t active = 0;
    if (he == hipSuccess) he = hipMemcpyAsync(&active, c->job.active, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
    float ms = 0;
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
    if (he != hipSuccess) {
      snprintf(g_hip_err, sizeof g_hip_err, "%s (fit launch %ld)", hipGetErrorString(he), L);
      result = AST_E_HIP;
      break;
    }
    c->last_ms += ms;
    c->last_launches++;
    if (active == 0) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (result) return result;
  std::vector<int32_t> nev(fits);
  HIPCHK(hipMemcpy(nev.data(), c->job.nev_tot, fits * sizeof(int32_t), hipMemcpyDeviceToHost));
  int64_t s = 0;
  for (size_t k = 0; k < fits; ++k) s += nev[k];
  c->last_evals = s;
  c->last_points = s * c->n_data;
  c->last_groups = rc->n_groups;
  c->last_group_size = rc->group_size;
  return 0;
}

}  // namespace

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
    default: return "unknown error";
  }
}

const char* ast_last_hip_error(void) { return g_hip_err; }

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
  ast_variant_type_ids(v, &c->T);
  c->cf_size = v->precision == AST_PREC_FP64 ? 8 : 4;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&c->params0_d, sizeof(c->params0));
  if (e != hipSuccess) {
    snprintf(g_hip_err, sizeof g_hip_err, "%s", hipGetErrorString(e));
    delete c;
    return AST_E_HIP;
  }
  *out = c;
  return AST_OK;
}

int ast_ctx_destroy(ast_ctx* c) {
  if (!c) return AST_E_ARG;
  (void)hipSetDevice(c->device);
  free_job(c->job);
  dev_free(c->pts);
  dev_free(c->sum_w);
  dev_free(c->params0_d);
  dev_free(c->rng);
  dev_free(c->best_f);
  dev_free(c->best_params);
  dev_free(c->best_dV);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return AST_OK;
}

int ast_set_model_space(ast_ctx* c, int n_params, const int* property, int n_types, const int* types,
                        const double* limits, const ast_x2* x2) {
  if (!c || !property || !types || !limits || !x2) return AST_E_ARG;
  if (n_params != c->T.n_params || n_types != c->T.n_types) return AST_E_ARG;
  // The kernels fix each parameter's role by its position: the Property
  // order of the reference (asteroid.c:86-125).
  const int expect[] = {c->T.theta_M, c->T.phi_M, c->T.phi_0, c->T.Ti, c->T.Ts, c->T.Tl, c->T.c_tumb, c->T.b_tumb,
                        c->T.Es, c->T.L, c->T.psi_0, c->T.c, c->T.b, c->T.theta_R, c->T.phi_R, c->T.kappa};
  int pos = 0;
  for (int k = 0; k < 16; ++k) {
    if (expect[k] < 0) continue;
    if (property[pos * AST_N_COLUMNS + AST_P_TYPE] != expect[k]) return AST_E_UNSUPPORTED;
    if (types[expect[k]] != pos) return AST_E_UNSUPPORTED;
    int fr = property[pos * AST_N_COLUMNS + AST_P_FROZEN];
    if (fr < -1 || fr > 1) return AST_E_ARG;
    pos++;
  }
  for (int i = 0; i < n_params; ++i)
    for (int k = 0; k < AST_N_COLUMNS; ++k) c->prop[i][k] = property[i * AST_N_COLUMNS + k];
  for (int t = 0; t < n_types; ++t) {
    c->types[t] = types[t];
    c->lim[0][t] = limits[t];
    c->lim[1][t] = limits[n_types + t];
  }
  c->n_params = n_params;
  c->n_types = n_types;
  c->x2 = *x2;
  build_space<double>(c, &c->sp_d);
  build_space<float>(c, &c->sp_f);
  c->space_set = true;
  return AST_OK;
}

int ast_upload_curve(ast_ctx* c, const ast_obs* obs, int n_data, int n_filters, const ast_ephem3* eph) {
  if (!c || !obs || n_data <= 0 || n_filters <= 0) return AST_E_ARG;
  if (n_filters > AST_MAX_FILTERS) return AST_E_DATA;
  if (c->v.interp && !eph) return AST_E_ARG;
  std::vector<PointRec> h(n_data);
  std::vector<double> sw(AST_MAX_FILTERS, 0.0);
  std::vector<int> seen(n_filters, 0);
  for (int i = 0; i < n_data; ++i) {
    const ast_obs& o = obs[i];
    if (o.filter < 0 || o.filter >= n_filters) return AST_E_DATA;
    if (i > 0 && !(o.mjd > obs[i - 1].mjd)) return AST_E_DATA;  // read_data.c:81-85
    if (!std::isfinite(o.mjd) || !std::isfinite(o.V) || !(o.w > 0)) return AST_E_DATA;
    PointRec& p = h[i];
    std::memset(&p, 0, sizeof p);
    if (c->v.interp) interp_canon(*eph, o.mjd, p.E, p.S);
    else for (int k = 0; k < 3; ++k) { p.E[k] = o.E[k]; p.S[k] = o.S[k]; }
    if (i > 0) {  // cuda.c:321-323
      double t1 = obs[i - 1].mjd, t2 = o.mjd;
      int nsteps = (int)((t2 - t1) / c->v.time_step + 1);
      double hh = (t2 - t1) / nsteps;
      p.nsteps = nsteps;
      p.h = hh;
      p.hh = 0.5 * hh;
      p.h6 = 1 / 6.0 * hh;
    }
    p.V = o.V;
    p.w = o.w;
    p.filter = o.filter;
    sw[o.filter] = sw[o.filter] + (double)o.w;
    seen[o.filter] = 1;
  }
  for (int m = 0; m < n_filters; ++m)
    if (!seen[m]) return AST_E_DATA;
  (void)hipSetDevice(c->device);
  dev_free(c->pts);
  dev_free(c->sum_w);
  int rc = dev_alloc(&c->pts, (size_t)n_data);
  if (!rc) rc = dev_alloc(&c->sum_w, (size_t)AST_MAX_FILTERS);
  if (rc) return rc;
  HIPCHK(hipMemcpy(c->pts, h.data(), sizeof(PointRec) * n_data, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->sum_w, sw.data(), sizeof(double) * AST_MAX_FILTERS, hipMemcpyHostToDevice));
  c->n_data = n_data;
  c->n_filters = n_filters;
  c->curve_set = true;
  return AST_OK;
}

int ast_rng_seed(ast_ctx* c, uint64_t seed) {
  if (!c) return AST_E_ARG;
  (void)hipSetDevice(c->device);
  c->seed = seed;
  if (c->rng) HIPCHK(hipMemset(c->rng, 0, c->rng_cap * sizeof(uint32_t)));
  return reset_best_f(c);
}

int ast_reset_best(ast_ctx* c) {
  if (!c) return AST_E_ARG;
  (void)hipSetDevice(c->device);
  return reset_best_f(c);
}

int ast_set_start(ast_ctx* c, const double* params0) {
  if (!c || !params0 || !c->space_set) return c && !c->space_set ? AST_E_STATE : AST_E_ARG;
  for (int i = 0; i < c->n_params; ++i) c->params0[i] = params0[i];
  return AST_OK;
}

int ast_run(ast_ctx* c, const ast_run_cfg* rc) {
  if (!c || !rc) return AST_E_ARG;
  if (!c->space_set || !c->curve_set) return AST_E_STATE;
  const int G = rc->group_size;
  if (rc->n_groups <= 0 || G < 64 || G > 1024 || (G & 63) || rc->nstages <= 0 || rc->n_steps < 0 ||
      rc->eval_budget < 0)
    return AST_E_ARG;
  if (c->n_data - c->n_params - c->n_filters <= 0) return AST_E_DATA;
  (void)hipSetDevice(c->device);
  return c->v.precision == AST_PREC_FP64 ? run_typed<double>(c, rc) : run_typed<float>(c, rc);
}

int ast_fetch(ast_ctx* c, double* f, double* params, double* dV) {
  if (!c) return AST_E_ARG;
  if (!c->last_groups) return AST_E_STATE;
  (void)hipSetDevice(c->device);
  const int G = c->last_groups, N = c->n_params, NF = c->n_filters;
  if (f) {
    if (c->cf_size == 8) {
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
  if (!c->last_groups) return AST_E_STATE;
  (void)hipSetDevice(c->device);
  const size_t fits = (size_t)c->last_groups * c->last_group_size;
  if (f_last) {
    if (c->cf_size == 8) {
      HIPCHK(hipMemcpy(f_last, c->job.sf, sizeof(double) * fits, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp(fits);
      HIPCHK(hipMemcpy(tmp.data(), c->job.sf, sizeof(float) * fits, hipMemcpyDeviceToHost));
      for (size_t k = 0; k < fits; ++k) f_last[k] = tmp[k];
    }
  }
  std::vector<int32_t> tmp(fits);
  if (n_evals) {
    HIPCHK(hipMemcpy(tmp.data(), c->job.nev_tot, sizeof(int32_t) * fits, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < fits; ++k) n_evals[k] = tmp[k];
  }
  if (n_steps) {
    HIPCHK(hipMemcpy(tmp.data(), c->job.steps_tot, sizeof(int32_t) * fits, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < fits; ++k) n_steps[k] = tmp[k];
  }
  return AST_OK;
}

int ast_chi2_batch(ast_ctx* c, const double* params, int n, double* chi2, double* dV) {
  if (!c || !params || n <= 0 || !chi2 || !dV) return AST_E_ARG;
  if (!c->space_set || !c->curve_set) return AST_E_STATE;
  if (c->n_data - c->n_params - c->n_filters <= 0) return AST_E_DATA;
  (void)hipSetDevice(c->device);
  double *pd = nullptr, *cd = nullptr, *dd = nullptr;
  int rc = dev_alloc(&pd, (size_t)n * c->n_params);
  if (!rc) rc = dev_alloc(&cd, (size_t)n);
  if (!rc) rc = dev_alloc(&dd, (size_t)n * c->n_filters);
  if (rc) { dev_free(pd); dev_free(cd); dev_free(dd); return rc; }
  int result = AST_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipMemcpy(pd, params, sizeof(double) * n * c->n_params, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  LaunchArgs a{};
  a.op = 2;
  a.params = pd;
  a.n = n;
  a.chi2 = cd;
  a.dV = dd;
  if (he == hipSuccess) he = hipEventRecord(e0, c->stream);
  if (he == hipSuccess) he = dispatch(c, a);
  if (he == hipSuccess) he = hipEventRecord(e1, c->stream);
  if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  float ms = 0;
  if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
  if (he == hipSuccess) he = hipMemcpy(chi2, cd, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (he == hipSuccess) he = hipMemcpy(dV, dd, sizeof(double) * n * c->n_filters, hipMemcpyDeviceToHost);
  if (he != hipSuccess) {
    snprintf(g_hip_err, sizeof g_hip_err, "%s (chi2 batch)", hipGetErrorString(he));
    result = AST_E_HIP;
  } else {
    c->last_ms = ms;
    c->last_launches = 1;
    c->last_evals = n;
    c->last_points = (int64_t)n * c->n_data;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dev_free(pd);
  dev_free(cd);
  dev_free(dd);
  return result;
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

}  // extern "C"
