/* asteroid.h -- C ABI of libasteroid_hip.so, the MI355X (gfx950) fitting core.
 *
 * The reference has no plugin or FFI API: its boundary is the CUDA launch of
 * chi2_gpu plus the __device__ symbols the host fills with cudaMemcpyToSymbol.
 * Each entry point below replaces one of those pieces; the reference
 * file:line it stands in for is given with it.  Every function returns 0 on
 * success or a negative AST_E* code (ast_strerror names it) and never exits
 * the process (the reference's ERR macro printed and exited,
 * cuda_errors.h:4-14; the host driver keeps that behaviour on top of these
 * codes).  One context per host thread; calls on a context are synchronous
 * unless stated otherwise.
 */
#ifndef ASTEROID_ABI_H
#define ASTEROID_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ constants */
/* Columns of the Property table (asteroid.h:115-123). */
enum {
  AST_P_TYPE = 0,
  AST_P_INDEPENDENT = 1,
  AST_P_FROZEN = 2,
  AST_P_ISEG = 3,
  AST_P_MULTI_SEGMENT = 4,
  AST_P_PERIODIC = 5,
  AST_N_COLUMNS = 6
};
/* Values of the P_periodic column (asteroid.h:126-131). */
enum {
  AST_PERIODIC = 1,
  AST_PERIODIC_LAM = 0,
  AST_SOFT_BOTH = -1,
  AST_HARD_LEFT = -2,
  AST_HARD_RIGHT = -3,
  AST_HARD_BOTH = -4
};
enum { AST_MAX_PARAMS = 40 /* SEGMENT with torque and BC, four segments, no ONE_LE (asteroid.h:104-106) */, AST_MAX_TYPES = 16, AST_MAX_FILTERS = 8, AST_MAX_MINIMA = 10 /* NOBS_MAX, asteroid.h:219 */,
       AST_MAX_SEGMENTS = 4 /* N_SEG, asteroid.h:40-46 */,
       AST_NOALLOC_MAX_MODELS = 16384 /* models per ast_chi2_batch_noalloc call */ };

/* Error codes. */
enum {
  AST_OK = 0,
  AST_E_ARG = -1,        /* bad argument (null pointer, size out of range) */
  AST_E_STATE = -2,      /* call out of order (e.g. run before upload)    */
  AST_E_HIP = -3,        /* HIP runtime error (ast_last_hip_error)         */
  AST_E_NOMEM = -4,      /* device or host allocation failed              */
  AST_E_DATA = -5,       /* input data rejected (unsorted times, too many filters, outside ephemeris span) */
  AST_E_UNSUPPORTED = -6,/* variant combination not built                 */
  AST_E_NODEVICE = -7,   /* no HIP device                                 */
  AST_E_CANCELLED = -8,  /* job cancelled before it finished              */
  AST_E_BUSY = -9        /* context is serving: ast_run, ast_chi2_batch and ast_fetch_lanes are refused until ast_serve_stop */
};

/* Arithmetic forms of the evaluator and simplex (CHI_FLOAT, asteroid.h:24-28). */
enum {
  AST_PREC_FP64 = 0,     /* -DACC: simplex and chi^2 in double, FP64 model      */
  AST_PREC_F32SIMPLEX = 1,/* default build: simplex and chi^2 in float, FP64 model */
  AST_PREC_MIXED = 2     /* float simplex; FP64 Euler-angle integration and chi^2 sums; FP32 geometry and brightness */
};

/* Model variant: the compile-time macros of the reference (makefile:3-36),
 * chosen at run time here.  TORQUE2 implies TORQUE and RECT implies BC, as in
 * the reference (asteroid.h:32-38).  The kernels are built for LS, BC and
 * BW_BALL with and without torque, and for TREND with torque, TORQUE2,
 * ROTATE (with BC, with and without torque), RECT (with and without torque)
 * NUDGE (with torque, with and without BC) and SEGMENT (LS with and without
 * torque and BC, three or four segments, with and without ONE_LE: 16 to 40
 * parameters); other combinations return AST_E_UNSUPPORTED.  P_PHI and
 * P_BOTH are not built. */
typedef struct {
  int torque;     /* -DTORQUE : 6-ODE constant-torque rotation                 */
  int bc;         /* -DBC     : photometric b,c independent of kinematic ones  */
  int bw_ball;    /* -DBW_BALL: black-and-white ball brightness                */
  int p_psi;      /* -DP_PSI  : L generated from 1/P_psi in random search      */
  int interp;     /* -DINTERP : E,S from a 3-point ephemeris                   */
  int precision;  /* AST_PREC_*                                                */
  double time_step; /* TIME_STEP, days (asteroid.h:202)                        */
  int trend;      /* -DTREND  : magnitude de-trended by A times the phase angle (cuda.c:692-697) */
  int torque2;    /* -DTORQUE2: a second torque after the split time Tt (cuda.c:288-317)       */
  int rotate;     /* -DROTATE : brightness ellipsoid rotated by theta_R, phi_R, psi_R (cuda.c:475-516) */
  int rect;       /* -DRECT   : rectangular prism brightness, phase 0 (cuda.c:609-661)          */
  int nudge;      /* -DNUDGE  : chi^2 nudged by model minima near observed ones (cuda.c:716-758, 912-992);
                   *            needs ast_set_minima                                               */
  int v1s;        /* -DV1S    : with NUDGE, deeper model minima are not penalised (cuda.c:940-945) */
  int n_seg;      /* -DSEGMENT -DN_SEG=n: the curve split at n start times, each segment its own
                   *            rotation (cuda.c:162-813); 0 or 1 = one segment; needs ast_set_segments */
  int one_le;     /* -DONE_LE : with SEGMENT, L and Es shared by all segments (asteroid.c:68-72) */
} ast_variant;

/* Parameter type ids of a variant: the reference's __COUNTER__ order
 * (asteroid.h:137-178); -1 where the variant has no such type.  The -f and
 * -l options of the host driver take these ids. */
typedef struct {
  int theta_M, phi_M, phi_0, L, Ti, Ts, Tl, c_tumb, b_tumb, Es, psi_0;
  int c, b, theta_R, phi_R, kappa;
  int n_types;   /* N_TYPES  */
  int n_params;  /* N_PARAMS (asteroid.h:106-110) */
  int A;         /* TREND                */
  int T2i, T2s, T2l, Tt;  /* TORQUE2     */
  int psi_R;     /* ROTATE               */
} ast_type_ids;

static inline int ast_variant_type_ids(const ast_variant* v, ast_type_ids* t) {
  int id = 0;
  if (!v || !t) return AST_E_ARG;
  const int torque = v->torque || v->torque2, bc = v->bc || v->rect;
  t->theta_M = id++;
  t->phi_M = id++;
  t->phi_0 = id++;
  t->L = id++;
  t->A = v->trend ? id++ : -1;
  t->Ti = t->Ts = t->Tl = -1;
  if (torque) { t->Ti = id++; t->Ts = id++; t->Tl = id++; }
  t->T2i = t->T2s = t->T2l = t->Tt = -1;
  if (v->torque2) { t->T2i = id++; t->T2s = id++; t->T2l = id++; t->Tt = id++; }
  t->c_tumb = id++;
  t->b_tumb = id++;
  t->Es = id++;
  t->psi_0 = id++;
  t->c = t->b = -1;
  if (bc) { t->c = id++; t->b = id++; }
  t->theta_R = t->phi_R = t->kappa = t->psi_R = -1;
  if (v->rotate || v->bw_ball) { t->theta_R = id++; t->phi_R = id++; }
  if (v->rotate) t->psi_R = id++;
  if (v->bw_ball) t->kappa = id++;
  t->n_types = id;
  t->n_params = 8 + (bc ? 2 : 0) + (v->rotate ? 3 : 0) + (v->trend ? 1 : 0) + (v->torque2 ? 7 : (torque ? 3 : 0)) +
                (v->bw_ball ? 3 : 0);
  if (v->n_seg > 1)  /* each further segment repeats the non-shared parameters (asteroid.h:104-108) */
    t->n_params += (6 + (v->torque2 ? 7 : (torque ? 3 : 0)) - (v->one_le ? 2 : 0)) * (v->n_seg - 1);
  return AST_OK;
}

/* One observation after ingest (struct obs_data, asteroid.h:355-368):
 * V absolute magnitude, w = 1/sgm^2, mjd asteroid time relative to the first
 * observation, filter index.  E and S (unit asteroid->Earth and
 * asteroid->Sun vectors) are read only when the variant is not INTERP. */
typedef struct {
  double mjd;
  float V;
  float w;
  int filter;
  int pad_;
  double E[3];
  double S[3];
} ast_obs;

/* The INTERP 3-point ephemeris (E_x0..S_z0, MJD0; asteroid.h:444,
 * gpu_prepare.c:39-45): E[c][k] is component c (x,y,z) of the
 * asteroid->Earth vector at ephemeris time mjd[k], times relative like
 * ast_obs.mjd (read_data.c:298-302). */
typedef struct {
  double E[3][3];
  double S[3][3];
  double mjd[3];
} ast_ephem3;

/* struct x2_struct (asteroid.h:340-351).  ppsi1/ppsi2 are 1/P_psi bounds in
 * 1/day, stored as float as the reference does (asteroid.c:599-600). */
typedef struct {
  int reopt;
  float ppsi1;
  float ppsi2;
} ast_x2;

/* One launch of the fitting kernel (chi2_gpu<<<N_BLOCKS,BSIZE>>>,
 * asteroid.c:675): n_groups reseed groups (blocks) of group_size simplexes
 * (threads). */
typedef struct {
  int n_groups;        /* N_BLOCKS                                             */
  int group_size;      /* BSIZE: 64, 128, 256, 512 or 1024                     */
  int nstages;         /* -Nstages                                             */
  int n_steps;         /* N_STEPS simplex-step limit (asteroid.h:213)           */
  int64_t eval_budget; /* chi^2 evaluations per simplex and stage; 0 = none    */
  double size_min;     /* SIZE_MIN; 0 = the precision's default (asteroid.h:233-237) */
  int ticks_per_launch;/* evaluation rounds per kernel launch; 0 = automatic   */
  int use_graphs;      /* 1: replay each launch cycle from a captured HIP graph (same results) */
  int resident_groups; /* nstages = 1 only: run the n_groups*group_size fits on this many workgroups,
                          each lane starting the next unstarted fit when its fit ends (a device work
                          queue); 0 = one workgroup per group.  Results, random streams and per-fit
                          trajectories are the same either way. */
} ast_run_cfg;

typedef struct ast_ctx ast_ctx;

/* Create a context on HIP device `device` for one variant (replaces the
 * compile-time macro choice, makefile:38-53, and Is_GPU_present,
 * cuda_errors.h:18-39). */
int ast_ctx_create(ast_ctx** out, const ast_variant* variant, int device);
int ast_ctx_destroy(ast_ctx* ctx);

/* Parameter metadata: Property[N_PARAMS][6] (asteroid.c:86-125), Types
 * (asteroid.c:152-169), Limits[2][N_TYPES] (asteroid.c:473-589, limits[t] is
 * the lower and limits[n_types+t] the upper limit) and x2 (asteroid.c:591-601).
 * Replaces the four cudaMemcpyToSymbol calls at asteroid.c:603-606.  The
 * arrays are copied. */
int ast_set_model_space(ast_ctx* ctx, int n_params, const int* property,
                        int n_types, const int* types, const double* limits,
                        const ast_x2* x2);

/* Light curve: replaces gpu_prepare's dData upload (gpu_prepare.c:13-14) and
 * the INTERP symbols (gpu_prepare.c:38-46).  `eph` is required for INTERP
 * variants and ignored otherwise.  Times must increase (read_data.c:81-85);
 * at most AST_MAX_FILTERS filters.  The data are copied. */
int ast_upload_curve(ast_ctx* ctx, const ast_obs* obs, int n_data, int n_filters,
                     const ast_ephem3* eph);

/* NUDGE: the observed brightness minima, replacing prepare_chi2_params'
 * d_chi2_params upload (misc.c:279-323).  t_obs are times on the curve's
 * clock (the reference's MJD - hMJD0, kept as float) and V_obs magnitudes
 * in the converted system; 1 <= n_obs <= AST_MAX_MINIMA, else AST_E_ARG.
 * The caller keeps only minima inside the data range, as misc.c:294-296
 * does.  AST_E_STATE on a context whose variant is not NUDGE.  Every
 * evaluation of a NUDGE context needs it (AST_E_STATE until set); it may
 * come before or after ast_upload_curve. */
int ast_set_minima(ast_ctx* ctx, const float* t_obs, const float* V_obs, int n_obs);

/* SEGMENT: the segment start times, replacing the hard-coded T_START table
 * (asteroid.h:40-46) and the split in read_data.c:244-250.  t_start are
 * times on the curve's clock (the reference's T_START - hMJD0), n_seg of
 * them, n_seg equal to the variant's; AST_E_ARG otherwise, AST_E_STATE on a
 * context without SEGMENT.  Must come before ast_upload_curve.  As in the
 * reference, observation i opens the next segment s when its time is at or
 * after t_start[s]; observations before the first segment are left out of
 * the sums but still count in the degrees of freedom.  A curve on which a
 * segment gets no opening observation is refused (AST_E_DATA).
 * ast_model_curve splits its grid the same way (read_data.c:357-363) and
 * returns NaN for grid points before the first start, which the reference
 * never computes; ast_rmsd is refused (AST_E_UNSUPPORTED): the reference's
 * chi2_gpu_rms never loads the segment starts. */
int ast_set_segments(ast_ctx* ctx, const double* t_start, int n_seg);

/* SEGMENT with the observations' split given by the caller: first_obs[k] is
 * the index of the first observation of segment k (increasing).  The
 * reference splits the observations on the observer's clock, before the
 * light-time correction (read_data.c:244-250: MJD_obs[i] >= T_START[iseg]),
 * and the plotting grid on the light-time-corrected clock (read_data.c:357:
 * hPlot.MJD + hMJD0 >= T_START); so an observation within one light time
 * after a start opens that segment in the reference although its corrected
 * time is before the start.  t_start (on the curve's clock, as for
 * ast_set_segments) splits the plotting grid.  A segment whose first
 * observation is past the table is refused at ast_upload_curve
 * (AST_E_DATA).  Replaces an earlier ast_set_segments. */
int ast_set_segment_starts(ast_ctx* ctx, const int* first_obs, const double* t_start, int n_seg);

/* Seed the per-simplex random streams and reset each group's best chi^2 to
 * 1e30: setup_kernel(..., seed, d_f, 1) (cuda.c:2039-2054, asteroid.c:629-634).
 * The project's counter-based generator replaces cuRAND. */
int ast_rng_seed(ast_ctx* ctx, uint64_t seed);

/* Reset each group's best chi^2 to 1e30 without touching the random
 * streams: setup_kernel(..., 0, d_f, 0) (asteroid.c:816-818, -keep). */
int ast_reset_best(ast_ctx* ctx);

/* Starting model for reoptimisation, physical units: d_params0
 * (asteroid.c:636-637, 766). */
int ast_set_start(ast_ctx* ctx, const double* params0);

/* Run one launch of the fitting kernel to completion (all simplexes of all
 * groups finish their nstages stages): chi2_gpu (cuda.c:1325-1834) with the
 * launch at asteroid.c:675 and the sync at asteroid.c:688. */
int ast_run(ast_ctx* ctx, const ast_run_cfg* cfg);

/* ast_run in pieces, for callers that interleave other work or time a run
 * launch by launch (the reference's host loop, asteroid.c:655-688, launches
 * and waits on chi2_gpu once per pass).  ast_run_begin uploads the job and
 * starts its groups; ast_run_continue runs up to max_cycles launch cycles
 * (0 = until the job ends) and sets *finished once every fit has ended;
 * ast_run_status reports the fits started and ended so far.  While a run is
 * unfinished, ast_rng_seed, ast_reset_best, ast_rmsd, ast_run_jobs,
 * ast_autotune and ast_serve_start return AST_E_STATE; a new ast_run_begin
 * abandons it.  ast_last_counts and ast_last_timing cover the cycles since
 * ast_run_begin.  ast_run is ast_run_begin + ast_run_continue(ctx, 0, ...). */
int ast_run_begin(ast_ctx* ctx, const ast_run_cfg* cfg);
int ast_run_continue(ast_ctx* ctx, int max_cycles, int* finished);
int ast_run_status(ast_ctx* ctx, int64_t* fits_started, int64_t* fits_done, int* finished);
/* End an unfinished run early: its fits stop where they are; per-group
 * results are those written by groups that had finished. */
int ast_run_cancel(ast_ctx* ctx);

/* Per-group results of the runs since the last seed/reset: d_f, d_params,
 * d_dV (asteroid.c:728-730).  f has n_groups entries, params
 * n_groups*n_params, dV n_groups*n_filters (caller-owned). */
int ast_fetch(ast_ctx* ctx, double* f, double* params, double* dV);

/* Per-simplex diagnostics of the last ast_run (no counterpart in the
 * reference, which keeps them in registers): the stage-final best chi^2 of
 * each simplex (1e30 if it failed), its chi^2 evaluation count and simplex
 * step count summed over stages.  Arrays of n_groups*group_size; any may be
 * NULL. */
int ast_fetch_lanes(ast_ctx* ctx, double* f_last, int64_t* n_evals, int64_t* n_steps);

/* The evaluator alone (chi2one, cuda.c:81-1011): chi^2 and the per-filter
 * delta_V of n parameter vectors (physical units, n*n_params doubles),
 * computed one model per GPU lane.  chi2 has n entries, dV n*n_filters. */
int ast_chi2_batch(ast_ctx* ctx, const double* params, int n, double* chi2, double* dV);

/* ast_chi2_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The models, chi^2 and delta_V are staged in a fixed device
 * array of the library (one per GPU, 6.4 MB, held until the process ends)
 * instead of buffers allocated per call; results are those of
 * ast_chi2_batch.  At most AST_NOALLOC_MAX_MODELS models per call (more is
 * AST_E_ARG).  All contexts on one GPU share that array: calls that overlap
 * on one GPU (from several threads or contexts) race and must be
 * serialised by the caller.  The context must have its model space and
 * curve (ast_set_model_space, ast_upload_curve, which do allocate). */
int ast_chi2_batch_noalloc(ast_ctx* ctx, const double* params, int n, double* chi2, double* dV);

/* Device copy bandwidth: copies bytes (rounded down to 16-byte words)
 * between two device buffers that the call allocates and frees, reps times,
 * with a plain streaming copy kernel, and returns the bytes read plus
 * written per second of the timed copies (ast_last_timing gives their
 * milliseconds).  No counterpart in the reference: a measurement aid, the
 * achievable HBM rate beside the fitting kernel's roofline and a kernel of
 * known traffic that calibrates the HBM traffic counters.  AST_E_BUSY while
 * serving. */
int ast_copy_bandwidth(ast_ctx* ctx, int64_t bytes, int reps, double* bytes_per_s);

/* The -plot evaluation (chi2_plot, cuda.c:1865-1966, launched at
 * asteroid.c:880; its grid is built at read_data.c:304-360).  For each of n
 * models (physical parameters, n*n_params doubles): chi^2 and delta_V
 * against the uploaded curve (cuda.c:1939-1942), then the model magnitude
 * plus delta_V[0] (d_Vmod, cuda.c:699-703) at n_plot grid times, evenly
 * spaced from the first to the last observation (NPLOT = 20000 in the
 * reference, asteroid.h:266).  The grid's two ends are the first and last
 * observations; its interior E,S come from `eph`, the 3-point ephemeris
 * (read_data.c:345; required for every variant).  Outputs, caller-owned:
 * t (n_plot grid times, may be NULL), vmod (n*n_plot, model k's curve at
 * vmod[k*n_plot ...]), chi2 (n), dV (n*n_filters).  One model per GPU lane. */
int ast_model_curve(ast_ctx* ctx, const double* params, int n, const ast_ephem3* eph, int n_plot,
                    double* t, double* vmod, double* chi2, double* dV);

/* The periodogram of a model light curve's brightness minima (minima(),
 * misc.c:65-276, called at asteroid.c:972; the default branch: minima at
 * grid times, step-function histogram of 10000 frequency bins from
 * 0.3/day to 24/day).  t and vm are the n grid times and magnitudes
 * (ast_model_curve's t and one model's vmod).  The top n_clusters clusters
 * of the normalised histogram (cl_fr, cl_H of asteroid.c:974-979, NCL_MAX =
 * 5 in the reference) come back in cl_fr (frequency, 1/day; -1 where no
 * cluster reaches 3 sigma) and cl_H (height in sigmas; 0 where none);
 * n_minima (may be NULL) receives the number of minima.  More than 10000
 * minima is AST_E_DATA (misc.c:80-84). */
int ast_periodogram(ast_ctx* ctx, const double* t, const double* vm, int n, int n_clusters,
                    double* cl_fr, double* cl_H, int* n_minima);

/* One launch of the RMSD confidence-interval kernel (chi2_gpu_rms,
 * cuda.c:2148-2389, the -DRMSD build; launched at asteroid.c:673). */
typedef struct {
  int n_groups;      /* N_BLOCKS                                                  */
  int group_size;    /* BSIZE, at most 1024                                       */
  int n_samples;     /* -Nstages: random models per simplex lane (cuda.c:2255)    */
  float dx;          /* -dx: standard deviation of the shift of every scale-free
                        coordinate (dx_rand, asteroid.c:271-277)                 */
  int spherical_k;   /* -DSPHERICAL_K: torque ranges as K, theta_K, phi_K        */
} ast_rmsd_cfg;

/* Around the start model (ast_set_start, the d_params0 role) every lane
 * draws n_samples models: each scale-free coordinate shifted by dx times a
 * standard normal deviate (from the lane's random stream, which persists
 * across calls as the curandState array does), clamped at hard limits, and
 * mapped with the model space's reopt.  A model whose RMSD (the RMSD build's
 * merit, cuda.c:890-906) is below f1 = f0 + f0/sqrt(N_data - N_PARAMS -
 * N_filters), f0 the start model's RMSD, widens its lane's per-parameter
 * range; each group's ranges come back as float par_min, par_max
 * (n_groups*n_params: dpar_min, dpar_max, asteroid.c:692-693).  n_tot and
 * n_bad (may be NULL) receive this call's number of models evaluated and
 * rejected (d_Ntot, d_Nbad), f0 and f1 the two RMSD values (d_f0, d_f1). */
int ast_rmsd(ast_ctx* ctx, const ast_rmsd_cfg* cfg, float* par_min, float* par_max, int64_t* n_tot,
             int64_t* n_bad, double* f0, double* f1);

/* Timing of the last ast_run / ast_chi2_batch, from HIP events on the
 * context's stream: total kernel milliseconds and the number of launches. */
int ast_last_timing(ast_ctx* ctx, double* kernel_ms, int* launches);

/* Algorithmic counts of the last ast_run: chi^2 evaluations and
 * model-point evaluations (evaluations times observations). */
int ast_last_counts(ast_ctx* ctx, int64_t* evals, int64_t* model_points);

/* Default Property/Types/Limits of a variant, as the reference's main()
 * builds them (asteroid.c:86-169, 473-564), with the P_psi range
 * ppsi_hours[0..1] mapped as at asteroid.c:599-600.  property needs
 * AST_MAX_PARAMS*6 ints, types AST_MAX_TYPES ints, limits 2*AST_MAX_TYPES
 * doubles (lower limits first, then upper, each n_types long). */
int ast_default_model_space(const ast_variant* variant, const double* ppsi_hours,
                            int* n_params, int* property, int* n_types, int* types,
                            double* limits, ast_x2* x2);

/* ------------------------------------------------------------ serving
 * No counterpart in the reference, whose host loop runs one kernel at a time
 * (asteroid.c:655-825).  A serving context admits fit jobs as they arrive,
 * each with its own light curve, model space, variant, seed and number of
 * simplex starts, while earlier jobs are still on the GPU: a scheduler thread
 * places each job's groups into free simplex slots when its HBM footprint
 * fits, runs one launch per kernel class over everything resident (continuous
 * batching) and retires jobs as their groups finish.  A job's results depend
 * only on the job (its own random streams, indexed by its own simplex
 * numbers), not on what runs beside it.  Submission, polling and
 * cancellation are thread-safe. */
typedef struct {
  int64_t max_resident_fits;     /* simplex slots on the device                  */
  int group_size;                /* simplexes per group (workgroup); 0 = 256     */
  int64_t hbm_budget_bytes;      /* admission limit on resident job bytes; 0 = slots only */
  double work_points_per_launch; /* evaluation work per simplex per launch, in observation-
                                    equivalents (an RK4 step counts 1.4-2.8); 0 = 16000 */
  int max_params;                /* parameters a slot holds (the largest job's count);
                                    0 = AST_MAX_PARAMS; larger jobs are refused (AST_E_ARG) */
} ast_serve_cfg;

typedef struct {
  ast_variant variant;
  int n_params;
  const int* property;      /* as ast_set_model_space */
  int n_types;
  const int* types;
  const double* limits;
  ast_x2 x2;                /* x2.reopt = 1 starts from params0 */
  const ast_obs* obs;       /* as ast_upload_curve */
  int n_data;
  int n_filters;
  const ast_ephem3* eph;
  int n_fits;               /* simplex starts, a multiple of the group size */
  int nstages;
  int n_steps;
  int64_t eval_budget;
  double size_min;
  uint64_t seed;
  const double* params0;    /* start model (reoptimisation), or NULL */
  uint64_t tag;             /* returned with the result */
  int n_minima;             /* NUDGE: observed minima, as ast_set_minima (0 otherwise) */
  const float* t_minima;
  const float* V_minima;
  const double* t_segments; /* SEGMENT: the variant's n_seg start times, as ast_set_segments (else NULL) */
  const int* seg_first_obs; /* SEGMENT: first observation of each segment, as ast_set_segment_starts
                               (NULL: split the observations by t_segments) */
} ast_job;

typedef struct {
  uint64_t id;
  uint64_t tag;
  int status;               /* AST_OK, AST_E_CANCELLED or an error */
  int n_params, n_filters, n_groups;
  int best_group;
  double best_f;            /* best chi^2 over the job's groups */
  double best_params[AST_MAX_PARAMS];
  double best_dV[AST_MAX_FILTERS];
  int64_t n_evals;          /* chi^2 evaluations */
  int64_t model_points;     /* evaluations times observations */
  double submit_to_done_ms; /* latency */
  double queue_ms;          /* submit to admission */
  double run_ms;            /* wall time of the cycles the job took part in */
} ast_result;

typedef struct {
  int pending, resident;
  int64_t free_fits, resident_bytes, launches, evals, model_points;
  double kernel_ms;
} ast_serve_stats_t;

/* ------------------------------------------------------------ batches
 * No counterpart in the reference, which fits one light curve per process
 * (README.md:80-86 runs many processes).  Runs n_jobs fit jobs (each with
 * its own curve, variant, model space, seed and number of simplexes) to
 * completion in shared launch cycles: every cycle is one launch of the
 * fitting kernel per kernel class (variant, filter capacity, precision) over
 * every running group of every job of that class.  results[i] belongs to
 * jobs[i]; its latency fields are measured from the call's start.  Each
 * job's results are those it gets alone (its random streams are indexed by
 * its own simplex numbers).  Refused while the context is serving. */
typedef struct {
  int group_size;                /* simplexes per group; 0 = 256 */
  double work_points_per_launch; /* as in ast_serve_cfg; 0 = 16000 */
  int use_graphs;                /* 1: replay unchanged launch cycles from a captured HIP graph */
  int schedule;                  /* 0: heaviest jobs dispatched first, each job's launches sized so that
                                    all jobs end together (the GPU stays full to the end); 1: equal
                                    work per launch for every job; 2: as 1, heaviest first.  Results
                                    never depend on it. */
} ast_batch_cfg;

int ast_run_jobs(ast_ctx* ctx, const ast_job* jobs, int n_jobs, const ast_batch_cfg* cfg, ast_result* results);

/* ------------------------------------------------------------ autotuner
 * No counterpart in the reference (its block size and registers are fixed
 * at compile time, asteroid.h:183-199).  Every kernel class (variant,
 * filter capacity, precision) has three builds of the fitting kernel with
 * different register budgets: 4 waves per SIMD (128 VGPRs, groups of up to
 * 1024 simplexes), 3 (168 VGPRs) and 2 (256 VGPRs), the last two for groups
 * of up to 256.  They compute exactly the same results.  ast_autotune runs
 * `cycles` launch cycles of the given jobs (one class at a time, from the
 * same start for every budget, so the work is identical) at each budget,
 * twice, interleaved, keeping each budget's faster time; it keeps the
 * fastest budget for each class in the context and reports those kernel
 * times; ast_set_waves sets one class by hand.  Untuned classes use 4. */
typedef struct {
  int torque, bc, bw_ball, precision, filter_capacity;
  int waves;             /* chosen budget */
  double kernel_ms[3];   /* kernel time at 4, 3 and 2 waves per SIMD (0 = not run) */
} ast_tune_entry;

int ast_autotune(ast_ctx* ctx, const ast_job* jobs, int n_jobs, const ast_batch_cfg* cfg, int cycles,
                 ast_tune_entry* out, int max_out, int* n_out);
int ast_set_waves(ast_ctx* ctx, const ast_variant* variant, int n_filters, int waves);

int ast_serve_start(ast_ctx* ctx, const ast_serve_cfg* cfg);
/* Copies the job's inputs; returns its id. */
int ast_submit(ast_ctx* ctx, const ast_job* job, uint64_t* id);
/* Up to `max` finished jobs; waits up to timeout_ms for the first one. */
int ast_poll(ast_ctx* ctx, ast_result* out, int max, int* n_out, int timeout_ms);
int ast_cancel(ast_ctx* ctx, uint64_t id);
int ast_serve_stats(ast_ctx* ctx, ast_serve_stats_t* stats);
/* Cancels what is left and joins the scheduler. */
int ast_serve_stop(ast_ctx* ctx);

const char* ast_strerror(int code);
/* Text of the last HIP error seen by this thread (for AST_E_HIP). */
const char* ast_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif /* ASTEROID_ABI_H */
