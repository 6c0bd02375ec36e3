// ast_kernels.h -- the fitting kernels (template definitions; each variant
// is instantiated in its own translation unit, ast_k_*.hip).
//
// The simplex of chi2_gpu (cuda.c:1424-1807) calls chi2one from five places
// (initial vertices, reflection, expansion, contraction, shrink), so the
// lanes of a wavefront diverge across the most expensive call in the
// program.  Here the Nelder-Mead control flow is a per-lane state machine:
// every "tick" each lane holds at most one pending point, all lanes
// evaluate their points through ONE call of the evaluator (divergence-free,
// since every model walks the same observation loop), and each lane then
// advances its own state machine with the result until it needs the next
// point.  The sequence of points a lane evaluates, the comparisons it makes
// and every stored value are those of the reference control flow, so a
// lane's trajectory is the oracle's trajectory.
//
// Simplex state lives in HBM in lane-fastest (structure-of-arrays) layout,
// so every access of a wavefront is one contiguous, coalesced segment.  A
// workgroup is one reseed group (the reference's block): after every tick
// it votes on whether all of its simplexes finished the stage; if so the
// block argmin, stage reseed or best-per-group write of cuda.c:1739-1823 run
// in place.  Kernel launches are chunks of ticks; all state persists in
// HBM between launches.
#pragma once
#include "ast_device.h"

namespace ast {

enum Phase : int {
  PH_START = 0,     // begin a stage: draw the start point and the simplex
  PH_INIT = 1,      // pending/complete: chi^2 of initial vertex j
  PH_BEGIN = 2,     // internal: top of the simplex loop (sort, centroid, tests)
  PH_REFLECT = 3,   // pending/complete: reflected point
  PH_EXPAND = 4,    // pending/complete: expanded point
  PH_CONTRACT = 5,  // pending/complete: contracted point
  PH_SHRINK = 6,    // pending/complete: shrunk vertex ind[j]
  PH_SHRINK_NEXT = 7,  // internal: shrink the next vertex
  PH_STAGE_DONE = 8,
  PH_FIT_DONE = 9
};

__device__ inline bool phase_pending(int ph) {
  return ph == PH_INIT || ph == PH_REFLECT || ph == PH_EXPAND || ph == PH_CONTRACT || ph == PH_SHRINK;
}

// Simplex state of the resident simplexes ("slots"), lane-fastest.  A group
// of G simplexes owns slot fits [slot*G, slot*G + G) while it runs; nf is the
// slot capacity.  Element (row, fit) of an array of fixed type (IND, PRM and
// the int32 counters) lives at row * nf + fit.  The arrays of simplex floats
// (CF) are made of 8-byte cells so that float and double groups can share the
// pool side by side: with KC = 8 / sizeof(CF), element (row, fit = slot*G + t)
// lives at CF index row * (KC*nf) + KC*slot*G + t, i.e. a float group packs
// its G values into the first half of its slot's cells of the row.  Every
// access of a wavefront stays one contiguous segment.
template <class CF>
struct SlotBufs {
  CF* X;        // [(N+1)*N]  simplex vertices
  CF* F;        // [N+1]      chi^2 at the vertices
  uint8_t* IND; // [N+1]      sorted vertex indices
  CF* X0;       // [N]        centroid
  CF* XR;       // [N]        reflected point
  CF* XP;       // [N]        expanded / contracted point
  double* PRM;  // [N]        the thread's params[] buffer (cuda.c:1344)
  CF* DV;       // [MAXF]     delta_V of the last evaluation
  int32_t* ctl;       // phase | j<<4 | failed<<10 | bad<<11 (j <= N <= 63)
  int32_t* lstep;     // simplex steps in this stage (l, cuda.c:1436)
  int32_t* nev;       // chi^2 evaluations in this stage
  int32_t* nev_tot;   // evaluations, all stages of this job
  int32_t* steps_tot; // steps, all stages of this job
  CF* fr;             // f_r kept for the expansion test
  CF* sf;             // stage-final best chi^2 (s_f, cuda.c:1739-1742)
  int32_t* fitid;     // the job's index of the fit the lane runs (its random stream)
  int32_t* work;      // evaluations the lane did since its group started (all fits)
  int32_t* g_reopt;   // per group slot: s_x2_params.reopt
  int32_t* g_stage;   // per group slot: current stage
  int32_t* g_done;    // per group slot: all stages finished
  size_t nf;          // slot capacity (fits)
};

// One group of one job, as a launch sees it (read once per workgroup,
// wave-uniform).  The job's curve, model space, start model, random streams
// and per-group best arrays are the job's own, so where a group runs (its
// slot, its neighbours in the launch) does not change its results.
template <class CF>
struct GroupDesc {
  const PointRec* pts;
  const double* sum_w;
  const Space<CF>* sp;
  const double* params0;
  uint32_t* rng;          // the job's stream counters, by job fit index
  CF* best_f;             // d_f of the job, by group
  double* best_params;    // d_params, [group][N]
  double* best_dV;        // d_dV, [group][n_filters]
  int32_t* job_done;      // groups of the job that finished
  uint64_t seed;
  CF size2_min;
  int n_data, n_filters, dof;
  int nstages, n_steps, eval_budget, reopt;
  int slot;               // group slot
  int group_in_job;       // the group (with refill: the worker) index within the job
  int ticks;              // evaluation rounds in this launch
  // Per-fit results of the job, written by each fit when it ends: its
  // stage-final best chi^2 (1e30 if it failed), evaluations and simplex steps
  // over all its stages.
  CF* fit_f;
  int32_t* fit_nev;
  int32_t* fit_steps;
  // Refill (single-stage jobs run by fewer workgroups than the job has
  // groups): a lane whose fit ended takes the next unstarted fit of the job
  // from queue[0] (queue[1] counts finished fits) and keeps the ended fit's
  // best point (x2params of its best vertex) and its last delta_V, from which
  // group_best_kernel forms each group's result once every fit has ended.
  int refill;
  int n_fits;
  int32_t* queue;
  double* fit_params;     // [n_fits][N]
  double* fit_dV;         // [n_fits][n_filters]
};

template <class CF>
struct LaneRegs {
  int phase, j, failed, bad, l, nev, nev_tot, steps_tot, work;
  uint32_t rng;
  CF fr, sf;
};

// Per-stage settings a lane needs while advancing (copied out of the descriptor).
template <class CF>
struct StageCfg {
  uint64_t seed;
  CF size2_min;
  int n_steps, eval_budget;
};

template <class CF, int VC>
struct Machine {
  using P = Pos<VC>;
  static constexpr int N = P::N;
  static constexpr bool BC = VarF<VC>::BC;
  const Space<CF>& sp;
  const SlotBufs<CF>& B;
  const StageCfg<CF>& C;
  size_t fit, nf;   // slot fit index and slot capacity (fixed-type arrays)
  size_t fo, rs;    // the same element in the CF arrays: offset and row stride
  uint64_t jfit;    // fit index within the job (random stream)
  int reopt;

  __device__ CF& X(int v, int i) const { return B.X[((size_t)v * N + i) * rs + fo]; }
  __device__ CF& F(int v) const { return B.F[(size_t)v * rs + fo]; }
  __device__ uint8_t& IND(int v) const { return B.IND[(size_t)v * nf + fit]; }
  __device__ CF& X0(int i) const { return B.X0[(size_t)i * rs + fo]; }
  __device__ CF& XR(int i) const { return B.XR[(size_t)i * rs + fo]; }
  __device__ CF& XP(int i) const { return B.XP[(size_t)i * rs + fo]; }
  __device__ double& PRM(int i) const { return B.PRM[(size_t)i * nf + fit]; }

  // x2params on a point, through the persistent params[] buffer.
  __device__ bool propose(const CF* xv) const {
    double prm[N];
#pragma unroll
    for (int i = 0; i < N; ++i) prm[i] = PRM(i);
    int rc = x2params_at<CF, VC>(xv, prm, sp, reopt);
#pragma unroll
    for (int i = 0; i < N; ++i) PRM(i) = prm[i];
    return rc == 0;
  }
  __device__ bool propose_row(int v) const {
    CF xv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) xv[i] = X(v, i);
    return propose(xv);
  }
  __device__ bool propose_xr() const {
    CF xv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) xv[i] = XR(i);
    return propose(xv);
  }
  __device__ bool propose_xp() const {
    CF xv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) xv[i] = XP(i);
    return propose(xv);
  }

  __device__ float rnd(LaneRegs<CF>& R) const { return astrng::uniform01(C.seed, jfit, R.rng++); }

  // Stage start (cuda.c:1456-1546): x[0] random or around the reoptimised
  // point, then the N displaced vertices.
  __device__ void start(LaneRegs<CF>& R) const {
    const CF DX_INI = (CF)0.001, D2X_INI = (CF)-1.2, DX_RAND = (CF)0.0001;  // asteroid.h:205-209
    int LAM = 0;
    CF x0v[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x0v[i] = X(0, i);
    for (int i = 0; i < N; ++i) {
      float r = rnd(R);
      if (BC && !reopt) {
        if (i == P::c) { x0v[i] = x0v[P::c_tumb]; continue; }
        if (i == P::b) { x0v[i] = x0v[P::b_tumb]; continue; }
      }
      int per = sp.periodic[i];
      CF xi;
      if (!reopt || (reopt && sp.frozen[i] == -1)) {
        xi = DX_INI + SMALL + r * (1.0 - 2 * (SMALL + DX_INI));
      } else {
        CF xmin = x0v[i] - DX_RAND;
        CF xmax = x0v[i] + DX_RAND;
        if (xmin < SMALL && hard_left(per, LAM)) xmin = SMALL;
        if (xmax > 1.0 - SMALL && hard_right(per, LAM)) xmax = 1.0 - SMALL;
        xi = xmin + r * (xmax - xmin);
      }
      x0v[i] = xi;
      if (P::is_Es(i)) LAM = xi >= 0.5;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) X(0, i) = x0v[i];
    for (int jj = 1; jj < N + 1; ++jj) {
      for (int i = 0; i < N; ++i) {
        if (i == jj - 1) {
          int per = sp.periodic[i];
          float d2x = rnd(R);
          CF dx = DX_INI * cf_exp((CF)(D2X_INI * d2x));
          if (rnd(R) < 0.5) {
            dx = -dx;
            if (x0v[i] + dx < SMALL && hard_left(per, LAM)) dx = -dx;
          } else {
            if (x0v[i] + dx > 1.0 - SMALL && hard_right(per, LAM)) dx = -dx;
          }
          X(jj, i) = x0v[i] + dx;
        } else {
          X(jj, i) = x0v[i];
        }
      }
    }
  }

  __device__ void finish(LaneRegs<CF>& R) const {
    R.sf = R.failed ? (CF)1e30 : F(IND(0));
    R.steps_tot += R.l;
    R.phase = PH_STAGE_DONE;
  }

  // A fit's own results, when it ends (refill): stage-final best chi^2,
  // counts, the physical parameters of its best vertex (x2params through the
  // lane's params[] buffer, as the group's writer does at cuda.c:1810-1823)
  // and the delta_V of its last evaluation (cuda.c:1816-1822).
  __device__ void end_fit(const LaneRegs<CF>& R, const GroupDesc<CF>& D) const {
    D.fit_f[jfit] = R.sf;
    D.fit_nev[jfit] = R.nev_tot;
    D.fit_steps[jfit] = R.steps_tot;
    D.rng[jfit] = R.rng;
    if (!R.failed) {
      CF xb[N];
      double prm[N];
      const int b = IND(0);
#pragma unroll
      for (int i = 0; i < N; ++i) { xb[i] = X(b, i); prm[i] = PRM(i); }
      x2params_at<CF, VC>(xb, prm, sp, reopt);
#pragma unroll
      for (int i = 0; i < N; ++i) D.fit_params[jfit * N + i] = prm[i];
    }
    for (int m = 0; m < D.n_filters; ++m) D.fit_dV[jfit * D.n_filters + m] = B.DV[(size_t)m * rs + fo];
  }

  // The lane's state for a new fit (as fit_init_kernel leaves a group's
  // lanes), then its stage start.
  __device__ void begin_fit(LaneRegs<CF>& R, const GroupDesc<CF>& D) const {
#pragma unroll
    for (int v = 0; v <= N; ++v) IND(v) = (uint8_t)v;
    double prm[N];
#pragma unroll
    for (int i = 0; i < N; ++i) prm[i] = reopt ? D.params0[i] : 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) PRM(i) = prm[i];
    CF x[N];
    if (reopt) params2x<CF, VC>(x, prm, sp);
    else
      for (int i = 0; i < N; ++i) x[i] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) X(0, i) = x[i];
    for (int m = 0; m < MAXF; ++m) B.DV[(size_t)m * rs + fo] = 0;
    R.phase = PH_START;
    R.nev_tot = 0;
    R.steps_tot = 0;
    R.rng = D.rng[jfit];
    R.fr = 0;
    R.sf = (CF)1e30;
    advance(R, (CF)0, false);
  }

  __device__ void replace_worst_with_xr(CF fv) const {
    const int w = IND(N);
#pragma unroll
    for (int i = 0; i < N; ++i) X(w, i) = XR(i);
    F(w) = fv;
  }
  __device__ void replace_worst_with_xp(CF fv) const {
    const int w = IND(N);
#pragma unroll
    for (int i = 0; i < N; ++i) X(w, i) = XP(i);
    F(w) = fv;
  }

  // Advance the lane until it has a pending point or its stage is over.
  // `val` is the chi^2 of the pending point when has_result.
  __device__ __attribute__((noinline)) void advance(LaneRegs<CF>& R, CF val, bool has_result) const {
    const CF ALPHA = 1.0, GAMMA = 2.0, RHO = 0.5, SIGMA = 0.5;  // asteroid.h:239-242
    if (!has_result && R.phase != PH_START) return;
    for (;;) {
      switch (R.phase) {
        case PH_START: {
          start(R);
          R.l = 0;
          R.nev = 0;
          R.failed = 0;
          R.bad = 0;
          R.j = 0;
          R.phase = PH_INIT;
          if (propose_row(0)) return;
          R.failed = 1;
          R.phase = PH_BEGIN;
          break;
        }
        case PH_INIT: {  // cuda.c:1550-1559
          F(R.j) = val;
          R.j++;
          if (R.j <= N) {
            if (propose_row(R.j)) return;
            R.failed = 1;
          }
          R.phase = PH_BEGIN;
          break;
        }
        case PH_BEGIN: {  // cuda.c:1568-1645
          if (R.failed) { finish(R); return; }
          R.l++;
          CF f[N + 1];
          int ind[N + 1];
#pragma unroll
          for (int v = 0; v <= N; ++v) { f[v] = F(v); ind[v] = IND(v); }
          bool taken[N + 1];
#pragma unroll
          for (int v = 0; v <= N; ++v) taken[v] = false;
          for (int v = 0; v <= N; ++v) {
            CF fmin = 1e30;
            int jmin = -1;
#pragma unroll
            for (int v2 = 0; v2 <= N; ++v2)
              if (!taken[v2] && f[v2] <= fmin) { fmin = f[v2]; jmin = v2; }
            if (jmin < 0) {  // NaN values left (cuda.c:1592-1598)
              ind[0] = 0;
              F(0) = 1e30;
              break;
            }
            ind[v] = jmin;
#pragma unroll
            for (int v2 = 0; v2 <= N; ++v2)
              if (v2 == jmin) taken[v2] = true;
          }
#pragma unroll
          for (int v = 0; v <= N; ++v) IND(v) = (uint8_t)ind[v];
          CF x0[N];
          for (int i = 0; i < N; ++i) {
            CF sum = 0.0;
            for (int v = 0; v < N + 1; ++v) sum = sum + X(v, i);
            x0[i] = sum / (N + 1);
          }
          CF size2 = 0.0;
          for (int v = 0; v < N + 1; ++v) {
            CF sum = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
              CF dx = X(v, i) - x0[i];
              sum = sum + dx * dx;
            }
            size2 = size2 + sum;
          }
          size2 = size2 / N;
          if (size2 < C.size2_min || R.l > C.n_steps || (C.eval_budget > 0 && R.nev >= C.eval_budget)) {
            finish(R);
            return;
          }
          const int w = ind[N];
#pragma unroll
          for (int i = 0; i < N; ++i) {
            X0(i) = x0[i];
            CF xw = X(w, i);
            XR(i) = (sp.frozen[i] != 1) ? (CF)(x0[i] + ALPHA * (x0[i] - xw)) : xw;
          }
          R.phase = PH_REFLECT;
          if (propose_xr()) return;
          val = 1e30;
          break;
        }
        case PH_REFLECT: {  // cuda.c:1646-1712 (decisions after the reflection)
          R.fr = val;
          const CF fb = F(IND(0));
          if (val >= fb && val < F(IND(N - 1))) {
            replace_worst_with_xr(val);
            R.phase = PH_BEGIN;
            break;
          }
          if (val < fb) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
              CF xr = XR(i);
              XP(i) = (sp.frozen[i] != 1) ? (CF)(X0(i) + GAMMA * (xr - X0(i))) : xr;
            }
            R.phase = PH_EXPAND;
            if (propose_xp()) return;
            val = 1e30;
            break;
          }
          const int w = IND(N);
#pragma unroll
          for (int i = 0; i < N; ++i) {
            CF xw = X(w, i);
            XP(i) = (sp.frozen[i] != 1) ? (CF)(X0(i) + RHO * (xw - X0(i))) : xw;
          }
          R.phase = PH_CONTRACT;
          if (propose_xp()) return;
          val = 1e30;
          break;
        }
        case PH_EXPAND: {  // cuda.c:1671-1689
          if (val < R.fr) replace_worst_with_xp(val);
          else replace_worst_with_xr(R.fr);
          R.phase = PH_BEGIN;
          break;
        }
        case PH_CONTRACT: {  // cuda.c:1703-1712
          if (val < F(IND(N))) {
            replace_worst_with_xp(val);
            R.phase = PH_BEGIN;
            break;
          }
          R.bad = 0;
          R.j = 1;
          R.phase = PH_SHRINK_NEXT;
          break;
        }
        case PH_SHRINK: {  // cuda.c:1726
          F(IND(R.j)) = val;
          R.j++;
          R.phase = PH_SHRINK_NEXT;
          break;
        }
        case PH_SHRINK_NEXT: {  // cuda.c:1716-1733
          while (R.j <= N) {
            const int v = IND(R.j), b = IND(0);
            CF xv[N];
#pragma unroll
            for (int i = 0; i < N; ++i) {
              CF xi = X(v, i);
              if (sp.frozen[i] != 1) xi = X(b, i) + SIGMA * (xi - X(b, i));
              xv[i] = xi;
              X(v, i) = xi;
            }
            if (propose(xv)) { R.phase = PH_SHRINK; return; }
            R.bad = 1;
            R.j++;
          }
          if (R.bad) R.failed = 1;
          R.phase = PH_BEGIN;
          break;
        }
        default:
          return;
      }
    }
  }
};

// --------------------------------------------------------------- init kernel
// A group's start (one workgroup of G threads per new group): identity sort
// order, stage-0 start point (params2x of the reoptimisation model,
// cuda.c:1409-1416), counters, group flags.
template <class CF, int VC>
__global__ void fit_init_kernel(const GroupDesc<CF>* descs, SlotBufs<CF> B) {
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  constexpr int KC = 8 / sizeof(CF);
  const GroupDesc<CF>& D = descs[blockIdx.x];
  const size_t nf = B.nf, rs = KC * nf;
  const size_t fit = (size_t)D.slot * blockDim.x + threadIdx.x;
  const size_t fo = (size_t)D.slot * blockDim.x * KC + threadIdx.x;
  for (int v = 0; v <= N; ++v) B.IND[v * nf + fit] = (uint8_t)v;
  double prm[N];
  for (int i = 0; i < N; ++i) prm[i] = D.reopt ? D.params0[i] : 0.0;
  for (int i = 0; i < N; ++i) B.PRM[i * nf + fit] = prm[i];
  CF x[N];
  if (D.reopt) params2x<CF, VC>(x, prm, *D.sp);
  else
    for (int i = 0; i < N; ++i) x[i] = 0;
  for (int i = 0; i < N; ++i) B.X[(size_t)i * rs + fo] = x[i];
  for (int m = 0; m < MAXF; ++m) B.DV[m * rs + fo] = 0;
  B.ctl[fit] = PH_START;
  B.lstep[fit] = 0;
  B.nev[fit] = 0;
  B.nev_tot[fit] = 0;
  B.steps_tot[fit] = 0;
  B.fitid[fit] = D.group_in_job * (int)blockDim.x + (int)threadIdx.x;
  B.work[fit] = 0;
  B.fr[fo] = 0;
  B.sf[fo] = (CF)1e30;
  if (threadIdx.x == 0) {
    B.g_reopt[D.slot] = D.reopt;
    B.g_stage[D.slot] = 0;
    B.g_done[D.slot] = 0;
  }
}

// One lane's evaluation inside the fitting kernel: parameters from the
// lane's params[] buffer, delta_V to its last-dV slot.  Kept out of line so
// the evaluator's register allocation does not compete with the simplex
// bookkeeping live across it.  W (the calling kernel's waves per SIMD) makes
// one copy per register budget: a function shared by kernels of different
// budgets is compiled for the smallest, 128 VGPRs, where the FP64 evaluator
// spills.  The mixed form keeps the one 128-VGPR copy, which measured faster
// than copies given more registers.
template <class CF, int VC, int NF, bool MX, int W>
__device__ __attribute__((noinline)) CF eval_lane(const double* PRM, CF* DV, size_t nf, size_t fit, size_t rs,
                                                   size_t fo, const PointRec* pts_g, int n_data, int n_filters, int dof,
                                                   const double* sum_w) {
  constexpr int N = Pos<VC>::N;
  // Arguments of a device function arrive in vector registers; the table
  // pointer and the sizes are wave-uniform, so move them to scalar registers
  // to keep the observation loop on scalar loads and scalar branches.
  const uint64_t pu = (uint64_t)(uintptr_t)pts_g;
  const uint64_t plo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pu);
  const uint64_t phi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pu >> 32));
  PointPtr pts = (PointPtr)(uintptr_t)((phi << 32) | plo);
  n_data = __builtin_amdgcn_readfirstlane(n_data);
  n_filters = __builtin_amdgcn_readfirstlane(n_filters);
  dof = __builtin_amdgcn_readfirstlane(dof);
  double prm[N];
#pragma unroll
  for (int i = 0; i < N; ++i) prm[i] = PRM[(size_t)i * nf + fit];
  CF dv[NF];
  if constexpr (VarF<VC>::NU) {
    // a failed NUDGE model leaves delta_V as the lane's last evaluation left it
#pragma unroll
    for (int m = 0; m < NF; ++m) dv[m] = m < n_filters ? DV[(size_t)m * rs + fo] : CF(0);
  }
  CF fv = chi2_model<CF, VC, NF, MX>(prm, pts, n_data, n_filters, dof, sum_w, dv);
#pragma unroll
  for (int m = 0; m < NF; ++m)
    if (m < n_filters) DV[(size_t)m * rs + fo] = dv[m];
  return fv;
}

// The next `1` from a counter shared by many lanes: one atomic per
// wavefront for all of its lanes that ask in this pass, each getting its own
// value in lane order.  queue[1] counts the requests (each follows a fit's
// end), queue[0] hands out the fit indices.
__device__ inline int take_fit(int32_t* queue) {
  const uint64_t mask = __ballot(1);
  const int lane = (int)__lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  const int rank = __popcll(mask & ((1ull << lane) - 1ull));
  int base = 0;
  if (lane == leader) {
    const int cnt = __popcll(mask);
    base = atomicAdd(&queue[0], cnt);
    atomicAdd(&queue[1], cnt);
  }
  base = __shfl(base, leader);
  return base + rank;
}

// Each group's result from its fits' own results, once every fit of the job
// has ended (refill): the block argmin and the best-per-group write of
// cuda.c:1739-1742, 1810-1823 -- the lowest chi^2 of the group's fits, the
// first one on ties, replaces the stored best only when strictly lower.
template <class CF>
__global__ __launch_bounds__(64) void group_best_kernel(const CF* fit_f, const double* fit_params,
                                                        const double* fit_dV, int G, int N, int n_filters,
                                                        int n_groups, CF* best_f, double* best_params,
                                                        double* best_dV) {
  const int g = blockIdx.x;
  if (threadIdx.x != 0 || g >= n_groups) return;
  int tm = 0;
  CF smin = (CF)1e30;
  for (int k = 0; k < G; ++k) {
    const CF f = fit_f[(size_t)g * G + k];
    if (f < smin) { smin = f; tm = k; }
  }
  if (smin < best_f[g]) {
    const size_t j = (size_t)g * G + tm;
    best_f[g] = smin;
    for (int i = 0; i < N; ++i) best_params[(size_t)g * N + i] = fit_params[j * N + i];
    for (int m = 0; m < n_filters; ++m) best_dV[(size_t)g * n_filters + m] = fit_dV[j * n_filters + m];
  }
}

// ---------------------------------------------------------------- fit kernel
// One workgroup = one group (the reference's block); it runs D.ticks rounds
// of "evaluate the pending points, advance every lane", then saves its state.
// W is the register budget as waves per SIMD: 4 (128 VGPRs, groups of up to
// 1024 simplexes), 3 (168 VGPRs) or 2 (256 VGPRs), both for groups of up to
// 256.  The arithmetic is the same for every W; only speed differs, and the
// autotuner picks W per kernel class by measurement.
template <class CF, int VC, int NF, bool MX, int W>
__global__ __launch_bounds__(W == 4 ? 1024 : 256) __attribute__((amdgpu_waves_per_eu(W, W)))
void fit_kernel(const GroupDesc<CF>* descs, SlotBufs<CF> B) {
  astm::sincos_table_load();  // the evaluator's sin/cos table into LDS (whole workgroup)
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  __shared__ CF s_f[1024];
  __shared__ CF s_x0[MAXP];
  __shared__ int s_tm;
  __shared__ CF s_smin;
  constexpr int KC = 8 / sizeof(CF);
  const GroupDesc<CF>& D = descs[blockIdx.x];
  const int t = threadIdx.x;
  const int G = blockDim.x;
  const int slot = D.slot;
  const size_t nf = B.nf, rs = KC * nf;
  const size_t fit = (size_t)slot * G + t;
  const size_t fo = (size_t)slot * G * KC + t;
  if (B.g_done[slot]) return;  // same value for every thread of the group
  uint64_t jfit = (uint64_t)B.fitid[fit];  // group_in_job * G + t unless the lane has refilled
  const Space<CF>& sp = *D.sp;
  const StageCfg<CF> SC{D.seed, D.size2_min, D.n_steps, D.eval_budget};
  const int nstages = D.nstages, ticks = D.ticks;
  int reopt = B.g_reopt[slot];
  int stage = B.g_stage[slot];
  LaneRegs<CF> R;
  const int ctl = B.ctl[fit];
  R.phase = ctl & 15;
  R.j = (ctl >> 4) & 63;
  R.failed = (ctl >> 10) & 1;
  R.bad = (ctl >> 11) & 1;
  R.l = B.lstep[fit];
  R.nev = B.nev[fit];
  R.nev_tot = B.nev_tot[fit];
  R.steps_tot = B.steps_tot[fit];
  R.work = B.work[fit];
  R.rng = D.rng[jfit];
  R.fr = B.fr[fo];
  R.sf = B.sf[fo];
  const bool refill = D.refill != 0;
  bool done = false;
  {
    Machine<CF, VC> M{sp, B, SC, fit, nf, fo, rs, jfit, reopt};
    M.advance(R, (CF)0, false);  // a lane at a stage start draws its simplex first
  }
  for (int tick = 0; tick < ticks && !done; ++tick) {
    Machine<CF, VC> M{sp, B, SC, fit, nf, fo, rs, jfit, reopt};
    if (phase_pending(R.phase)) {
      CF fv = eval_lane<CF, VC, NF, MX, MX ? 4 : W>(B.PRM, B.DV, nf, fit, rs, fo, D.pts, D.n_data, D.n_filters, D.dof,
                                            D.sum_w);
      R.nev++;
      R.nev_tot++;
      R.work++;
      M.advance(R, fv, true);
    }
    if (refill) {
      // A lane whose fit ended keeps its results and starts the job's next
      // unstarted fit (one that fails at its start ends at once and the lane
      // takes another); with the queue empty it idles.  The group is done
      // when every lane idles.
      while (R.phase == PH_STAGE_DONE) {
        Machine<CF, VC> Me{sp, B, SC, fit, nf, fo, rs, jfit, reopt};
        Me.end_fit(R, D);
        const int nj = take_fit(D.queue);
        if (nj >= D.n_fits) {
          R.phase = PH_FIT_DONE;
          break;
        }
        jfit = (uint64_t)nj;
        Machine<CF, VC> Mb{sp, B, SC, fit, nf, fo, rs, jfit, reopt};
        Mb.begin_fit(R, D);
      }
      if (__syncthreads_count(R.phase != PH_FIT_DONE) == 0) done = true;
      continue;
    }
    const int running = __syncthreads_count(phase_pending(R.phase));
    if (running) continue;
    // Every simplex of the group finished the stage (cuda.c:1739-1823).
    s_f[t] = R.sf;
    __syncthreads();
    if (t == 0) {
      int tm = 0;
      CF smin = (CF)1e30;
      for (int k = 0; k < G; ++k)
        if (s_f[k] < smin) { smin = s_f[k]; tm = k; }
      s_tm = tm;
      s_smin = smin;
    }
    __syncthreads();
    const int tm = s_tm;
    const CF smin = s_smin;
    if (nstages > 1 && stage < nstages - 1) {
      if (t == tm) {
        CF xb[N];
        const int b = M.IND(0);
#pragma unroll
        for (int i = 0; i < N; ++i) xb[i] = M.X(b, i);
        if (sp.p_psi && reopt == 0) {  // cuda.c:1791-1798
          double prm[N];
#pragma unroll
          for (int i = 0; i < N; ++i) prm[i] = M.PRM(i);
          x2params_at<CF, VC>(xb, prm, sp, reopt);
          params2x<CF, VC>(xb, prm, sp);
#pragma unroll
          for (int i = 0; i < N; ++i) { M.PRM(i) = prm[i]; M.X(b, i) = xb[i]; }
        }
#pragma unroll
        for (int i = 0; i < N; ++i) s_x0[i] = xb[i];
      }
      __syncthreads();
      reopt = 1;
      stage++;
#pragma unroll
      for (int i = 0; i < N; ++i) M.X(0, i) = s_x0[i];
      R.phase = PH_START;
      Machine<CF, VC> M2{sp, B, SC, fit, nf, fo, rs, jfit, reopt};
      M2.advance(R, (CF)0, false);
    } else {
      const int g = D.group_in_job;
      if (t == tm && smin < D.best_f[g]) {  // cuda.c:1810-1823
        D.best_f[g] = smin;
        CF xb[N];
        double prm[N];
        const int b = M.IND(0);
#pragma unroll
        for (int i = 0; i < N; ++i) { xb[i] = M.X(b, i); prm[i] = M.PRM(i); }
        x2params_at<CF, VC>(xb, prm, sp, reopt);
#pragma unroll
        for (int i = 0; i < N; ++i) {
          M.PRM(i) = prm[i];
          D.best_params[(size_t)g * N + i] = prm[i];
        }
        for (int m = 0; m < D.n_filters; ++m)
          D.best_dV[(size_t)g * D.n_filters + m] = B.DV[(size_t)m * rs + fo];
      }
      D.fit_f[jfit] = R.sf;
      D.fit_nev[jfit] = R.nev_tot;
      D.fit_steps[jfit] = R.steps_tot;
      stage++;
      R.phase = PH_FIT_DONE;
      done = true;
    }
  }
  B.ctl[fit] = R.phase | (R.j << 4) | (R.failed << 10) | (R.bad << 11);
  B.lstep[fit] = R.l;
  B.nev[fit] = R.nev;
  B.nev_tot[fit] = R.nev_tot;
  B.steps_tot[fit] = R.steps_tot;
  B.work[fit] = R.work;
  B.fitid[fit] = (int32_t)jfit;
  D.rng[jfit] = R.rng;
  B.fr[fo] = R.fr;
  B.sf[fo] = R.sf;
  if (t == 0) {
    B.g_reopt[slot] = reopt;
    B.g_stage[slot] = stage;
    B.g_done[slot] = done ? 1 : 0;
    if (done) atomicAdd(D.job_done, 1);
  }
}

// ------------------------------------------------------------- chi2 batch
template <class CF, int VC, int NF, bool MX>
__global__ __launch_bounds__(256) void chi2_batch_kernel(const PointRec* pts_g, const double* sum_w, int n_data,
                                                        int n_filters, int dof, const double* params, int n,
                                                        double* chi2, double* dV) {
  astm::sincos_table_load();  // the evaluator's sin/cos table into LDS (whole workgroup)
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  PointPtr pts = (PointPtr)(uintptr_t)pts_g;
  double prm[N];
#pragma unroll
  for (int i = 0; i < N; ++i) prm[i] = params[(size_t)k * N + i];
  CF dv[NF];
#pragma unroll
  for (int m = 0; m < NF; ++m) dv[m] = NU ? CF(__builtin_nan("")) : CF(0);  // NaN: a failed NUDGE model
  CF c = chi2_model<CF, VC, NF, MX>(prm, pts, n_data, n_filters, dof, sum_w, dv);
  chi2[k] = c;
  for (int m = 0; m < n_filters; ++m) dV[(size_t)k * n_filters + m] = dv[m];
}

// ------------------------------------------------------------ model curve
// The -plot evaluation (chi2_plot, cuda.c:1865-1966, without PROFILES and
// ANIMATE), one model per lane: chi^2 and delta_V of the model against the
// observations (step one, cuda.c:1939-1942), then the model magnitude plus
// delta_V[0] at each point of the plotting grid (step two, cuda.c:1949-1950),
// integrated along the grid from its first time.  vmod is model-major: lane
// k writes vmod[k * n_plot + i].
template <class CF, int VC, int NF, bool MX>
__global__ __launch_bounds__(256) void model_curve_kernel(const PointRec* pts_g, const double* sum_w, int n_data,
                                                         int n_filters, int dof, const PointRec* grid_g, int n_plot,
                                                         const double* params, int n, double* chi2, double* dV,
                                                         double* vmod) {
  astm::sincos_table_load();  // the evaluator's sin/cos table into LDS (whole workgroup)
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double prm[N];
#pragma unroll
  for (int i = 0; i < N; ++i) prm[i] = params[(size_t)k * N + i];
  CF dv[NF];
#pragma unroll
  for (int m = 0; m < NF; ++m) dv[m] = NU ? CF(__builtin_nan("")) : CF(0);  // NaN: a failed NUDGE model
  CF c = chi2_model<CF, VC, NF, MX>(prm, (PointPtr)(uintptr_t)pts_g, n_data, n_filters, dof, sum_w, dv);
  chi2[k] = c;
  for (int m = 0; m < n_filters; ++m) dV[(size_t)k * n_filters + m] = dv[m];
  chi2_model<CF, VC, NF, MX, FORM_PLOT>(prm, (PointPtr)(uintptr_t)grid_g, n_plot, n_filters, dof, sum_w,
                                               dv, vmod + (size_t)k * n_plot, 1, (double)dv[0]);
}

// ------------------------------------------------------------------- RMSD
// Confidence intervals by random shifts around an input model
// (chi2_gpu_rms, cuda.c:2148-2389, the RMSD build).  RmsdRef holds what
// every thread of the reference computes first (cuda.c:2217-2247): the
// input model in scale-free coordinates, its RMSD f0 and the acceptance
// threshold f1 = f0 + f0/sqrt(N_data - N_PARAMS - N_filters).
template <class CF>
struct RmsdRef {
  CF x0[MAXP];
  CF f0, f1;
  double phi_M0, phi_00, phi_K0;  // the input model's phi_M, phi_0 and torque azimuth
};

template <class CF>
struct RmsdArgs {
  const PointRec* pts;
  const double* sum_w;
  int n_data, n_filters, dof;
  const Space<CF>* sp;
  int reopt;
  const double* params0;  // d_params0
  RmsdRef<CF>* ref;
  uint32_t* rng;          // per-lane stream counters (the curandState array)
  uint64_t seed;
  int n_samples;          // Nstages: random points per lane
  float dx;               // dx_rand
  int spherical_k;        // SPHERICAL_K
  float* par_min;         // [group][N_PARAMS] (dpar_min)
  float* par_max;
  int32_t* counts;        // [group][2]: Ntot, Nbad
};

template <class CF, int VC, int NF, bool MX>
__global__ __launch_bounds__(64) void rmsd_ref_kernel(RmsdArgs<CF> A) {
  astm::sincos_table_load();  // the evaluator's sin/cos table into LDS (whole workgroup)
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  if (threadIdx.x != 0) return;
  double prm[N];
  for (int i = 0; i < N; ++i) prm[i] = A.params0[i];
  CF x[N];
  params2x<CF, VC>(x, prm, *A.sp);
  CF dv[NF];
  const CF f0 = chi2_model<CF, VC, NF, MX, FORM_RMSD>(prm, (PointPtr)(uintptr_t)A.pts, A.n_data,
                                                              A.n_filters, A.dof, A.sum_w, dv);
  RmsdRef<CF>& R = *A.ref;
  for (int i = 0; i < N; ++i) R.x0[i] = x[i];
  R.f0 = f0;
  R.f1 = f0 + f0 / (CF)__builtin_sqrt((double)(CF)A.dof);
  R.phi_M0 = prm[P::phi_M];
  R.phi_00 = prm[P::phi_0];
  R.phi_K0 = 0.0;
  if (TQ) {
    const double bt = prm[P::b_tumb], ct = prm[P::c_tumb];
    const double Is = (1.0 + bt * bt) / (bt * bt + ct * ct);
    const double Ii = (1.0 + ct * ct) / (bt * bt + ct * ct);
    R.phi_K0 = astm::atan2(Is * prm[P::Ts], Ii * prm[P::Ti]);
  }
}

// One lane per thread of the reference: n_samples shifted models each,
// the range of the accepted ones per parameter, reduced per group in lane
// order by the group's first thread (cuda.c:2356-2381).
template <class CF, int VC, int NF, bool MX, int W>
__global__ __launch_bounds__(W == 4 ? 1024 : 256) __attribute__((amdgpu_waves_per_eu(W, W))) void rmsd_kernel(
    RmsdArgs<CF> A) {
  astm::sincos_table_load();  // the evaluator's sin/cos table into LDS (whole workgroup)
  using P = Pos<VC>;
  AST_VFLAGS(VC);
  constexpr int N = P::N;
  constexpr double PI = 3.141592653589793238;
  constexpr double SMALL = 1e-8;  // cuda.c:2258
  __shared__ float s_min[1024], s_max[1024];
  __shared__ int32_t s_cnt[2][1024];
  const int t = threadIdx.x;
  const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + t;
  const Space<CF>& sp = *A.sp;
  const RmsdRef<CF>& R = *A.ref;
  CF pmin[N], pmax[N];
  for (int i = 0; i < N; ++i) { pmin[i] = (CF)1e30; pmax[i] = (CF)-1e30; }
  int32_t ntot = 0, nbad = 0;
  uint32_t ctr = A.rng[id];
  double prm[N];
  for (int s = 0; s < A.n_samples; ++s) {
    int LAM = 0;
    CF x1[N];
    for (int i = 0; i < N; ++i) {  // cuda.c:2276-2287
      const float z = astrng::normal01(A.seed, id, ctr);
      ctr += 2;
      CF xx = R.x0[i] + (CF)(A.dx * z);
      const int per = sp.periodic[i];
      if (xx < SMALL && hard_left(per, LAM)) xx = (CF)SMALL;
      else if (xx > 1.0 - SMALL && hard_right(per, LAM)) xx = (CF)(1.0 - SMALL);
      x1[i] = xx;
      if (P::is_Es(i)) LAM = x1[i] >= (CF)0.5;
    }
    if (x2params_at<CF, VC>(x1, prm, sp, A.reopt)) continue;
    CF dv[NF];
    const CF f = chi2_model<CF, VC, NF, MX, FORM_RMSD>(prm, (PointPtr)(uintptr_t)A.pts, A.n_data,
                                                               A.n_filters, A.dof, A.sum_w, dv);
    ntot++;
    if (f < R.f1) {
      // cuda.c:2304-2311, as written: an angle beyond +-pi of the input's is
      // set to the input's value -2 pi, which the second test then moves to
      // the input's value +2 pi
      if (prm[P::phi_M] > R.phi_M0 + PI) prm[P::phi_M] = R.phi_M0 - 2 * PI;
      if (prm[P::phi_M] < R.phi_M0 - PI) prm[P::phi_M] = R.phi_M0 + 2 * PI;
      if (prm[P::phi_0] > R.phi_00 + PI) prm[P::phi_0] = R.phi_00 - 2 * PI;
      if (prm[P::phi_0] < R.phi_00 - PI) prm[P::phi_0] = R.phi_00 + 2 * PI;
      if (TQ && A.spherical_k) {  // cuda.c:2312-2333
        const double bt = prm[P::b_tumb], ct = prm[P::c_tumb];
        const double Is = (1.0 + bt * bt) / (bt * bt + ct * ct);
        const double Ii = (1.0 + ct * ct) / (bt * bt + ct * ct);
        const double Ki = Ii * prm[P::Ti], Ks = Is * prm[P::Ts], Kl = prm[P::Tl];
        const double K = __builtin_sqrt(Ki * Ki + Ks * Ks + Kl * Kl);
        prm[P::Ti] = K;
        prm[P::Ts] = astm::acos(Kl / K);
        double phi = astm::atan2(Ks, Ki);
        if (phi > R.phi_K0 + PI) phi = phi - 2 * PI;
        if (phi < R.phi_K0 - PI) phi = phi + 2 * PI;
        prm[P::Tl] = phi;
      }
      for (int i = 0; i < N; ++i) {
        if (prm[i] < pmin[i]) pmin[i] = (CF)prm[i];
        if (prm[i] > pmax[i]) pmax[i] = (CF)prm[i];
      }
    } else {
      nbad++;
    }
  }
  A.rng[id] = ctr;
  s_cnt[0][t] = ntot;
  s_cnt[1][t] = nbad;
  for (int i = 0; i < N; ++i) {
    s_min[t] = (float)pmin[i];
    s_max[t] = (float)pmax[i];
    __syncthreads();
    if (t == 0) {
      float fmin = 1e30f, fmax = -1e30f;
      for (int j = 0; j < (int)blockDim.x; ++j) {
        if (s_min[j] < fmin) fmin = s_min[j];
        if (s_max[j] > fmax) fmax = s_max[j];
      }
      A.par_min[(size_t)blockIdx.x * N + i] = fmin;
      A.par_max[(size_t)blockIdx.x * N + i] = fmax;
    }
    __syncthreads();
  }
  if (t == 0) {
    int32_t a = 0, b = 0;
    for (int j = 0; j < (int)blockDim.x; ++j) { a += s_cnt[0][j]; b += s_cnt[1][j]; }
    A.counts[(size_t)blockIdx.x * 2] = a;
    A.counts[(size_t)blockIdx.x * 2 + 1] = b;
  }
}

// --------------------------------------------------------- launch interface
struct LaunchArgs {
  int op;  // 0 init groups, 1 fit ticks, 2 chi2 batch, 3 model curve
  int precision;
  hipStream_t stream;
  // fit: descriptors of the groups of this launch (device), their count, G
  const void* descs;   // GroupDesc<CF>*
  const void* bufs;    // SlotBufs<CF>*
  int n_groups, group_size, n_filters;
  int waves;           // fit kernel register budget: 4, 3 or 2 waves per SIMD
  // batch
  const PointRec* pts;
  const double* sum_w;
  int n_data, dof;
  const double* params;
  int n;
  double* chi2;
  double* dV;
  // model curve: the plotting grid and the magnitudes on it
  const PointRec* grid;
  int n_plot;
  double* vmod;
  // RMSD: RmsdArgs<CF> (op 4: the input model, op 5: the shifted models)
  const void* rmsd;
};

template <class CF, int VC, int NF, bool MX>
hipError_t launch_nf(const LaunchArgs& a) {
  if (a.op == 0) {
    hipLaunchKernelGGL((fit_init_kernel<CF, VC>), dim3(a.n_groups), dim3(a.group_size), 0, a.stream,
                       (const GroupDesc<CF>*)a.descs, *(const SlotBufs<CF>*)a.bufs);
  } else if (a.op == 1) {
    const GroupDesc<CF>* d = (const GroupDesc<CF>*)a.descs;
    const SlotBufs<CF>& B = *(const SlotBufs<CF>*)a.bufs;
    const int w = a.group_size > 256 ? 4 : a.waves;  // W < 4 kernels take groups of up to 256
    if (w == 2)
      hipLaunchKernelGGL((fit_kernel<CF, VC, NF, MX, 2>), dim3(a.n_groups), dim3(a.group_size), 0, a.stream,
                         d, B);
    else if (w == 3)
      hipLaunchKernelGGL((fit_kernel<CF, VC, NF, MX, 3>), dim3(a.n_groups), dim3(a.group_size), 0, a.stream,
                         d, B);
    else
      hipLaunchKernelGGL((fit_kernel<CF, VC, NF, MX, 4>), dim3(a.n_groups), dim3(a.group_size), 0, a.stream,
                         d, B);
  } else if (a.op == 2) {
    int blocks = (a.n + 255) / 256;
    hipLaunchKernelGGL((chi2_batch_kernel<CF, VC, NF, MX>), dim3(blocks), dim3(256), 0, a.stream, a.pts,
                       a.sum_w, a.n_data, a.n_filters, a.dof, a.params, a.n, a.chi2, a.dV);
  } else if (a.op == 3) {
    int blocks = (a.n + 255) / 256;
    hipLaunchKernelGGL((model_curve_kernel<CF, VC, NF, MX>), dim3(blocks), dim3(256), 0, a.stream, a.pts,
                       a.sum_w, a.n_data, a.n_filters, a.dof, a.grid, a.n_plot, a.params, a.n, a.chi2, a.dV,
                       a.vmod);
  } else if (a.op == 4) {
    hipLaunchKernelGGL((rmsd_ref_kernel<CF, VC, NF, MX>), dim3(1), dim3(64), 0, a.stream,
                       *(const RmsdArgs<CF>*)a.rmsd);
  } else if (a.group_size > 256) {  // groups of up to 1024: 128 registers
    hipLaunchKernelGGL((rmsd_kernel<CF, VC, NF, MX, 4>), dim3(a.n_groups), dim3(a.group_size), 0, a.stream,
                       *(const RmsdArgs<CF>*)a.rmsd);
  } else {
    hipLaunchKernelGGL((rmsd_kernel<CF, VC, NF, MX, 2>), dim3(a.n_groups), dim3(a.group_size), 0, a.stream,
                       *(const RmsdArgs<CF>*)a.rmsd);
  }
  return hipGetLastError();
}

template <class CF, int VC, bool MX>
hipError_t launch_cf(const LaunchArgs& a) {
  if (a.n_filters <= 1) return launch_nf<CF, VC, 1, MX>(a);
  if (a.n_filters <= 4) return launch_nf<CF, VC, 4, MX>(a);
  return launch_nf<CF, VC, MAXF, MX>(a);
}

// Simplex float and evaluator form by precision: FP64 (double simplex, FP64
// evaluator), F32SIMPLEX (float simplex, FP64 evaluator: the reference's
// default CHI_FLOAT), MIXED (float simplex, mixed-precision evaluator).
template <int VC>
hipError_t launch_variant(const LaunchArgs& a) {
  if (a.precision == AST_PREC_FP64) return launch_cf<double, VC, false>(a);
  if (a.precision == AST_PREC_MIXED) return launch_cf<float, VC, true>(a);
  return launch_cf<float, VC, false>(a);
}

}  // namespace ast
