// asteroid_main.cpp -- the `asteroid` command-line driver.
//
// Keeps the roles of the reference's host program: main() (asteroid.c:18-831,
// fitting branch), read_data() (read_data.c:6-377) and
// quadratic_interpolation() (misc.c:20-35), with the same flags and the same
// stdout and result-file lines, on top of the C ABI of libasteroid_hip.so.
// The reference chooses its model variant at compile time (makefile:3-36);
// here the same macros are run-time flags (-DTORQUE, -DTORQUE2, -DP_PSI,
// -DBC, -DRECT, -DROTATE, -DTREND, -DBW_BALL, -DNUDGE, -DV1S, -DSEGMENT, -DN_SEG=3|4,
// -DONE_LE, -DINTERP, -DACC),
// defaulting to -DP_PSI -DTORQUE -DINTERP; with -DNUDGE the observed minima
// come from observed.min in the working directory;
// -DMIXED (no counterpart in the reference) selects the mixed-precision form.
// A failing library call prints the error and exits with status 1 (the
// reference's ERR macro, cuda_errors.h:4-14).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "asteroid.h"

namespace {

const double PI = 3.141592653589793238;  // asteroid.h:16
const double LIGHT_SPEED = 173.144632674; // au/day, asteroid.h:287
const int MAX_LINE = 128;                 // asteroid.h:256

#define CHECK(call)                                                                                   \
  do {                                                                                                \
    int rc_ = (call);                                                                                 \
    if (rc_ != AST_OK) {                                                                              \
      fprintf(stderr, "ASTassert: %s %s %s %d\n", ast_strerror(rc_), ast_last_hip_error(), __FILE__, \
              __LINE__);                                                                              \
      exit(1);                                                                                        \
    }                                                                                                 \
  } while (0)

[[noreturn]] void die(const char* msg) {
  printf("%s\n", msg);
  exit(1);
}

struct Ingested {
  std::vector<ast_obs> obs;
  int n_filters = 0;
  std::vector<char> filters;
  ast_ephem3 eph{};
  double mjd0 = 0;  // hMJD0
  std::vector<double> mjd_raw;  // observer's MJD of each observation, before the light-time correction
};

// SEGMENT's start times (asteroid.h:40-46) and the first observation of each
// segment: read_data.c:244-250 compares the observer's MJD (MJD_obs[i],
// before the light-time correction) with T_START.  Returns the number of
// segments opened.
int segment_first_obs(const std::vector<double>& mjd_raw, int n_seg, const double* T, int* first) {
  int k = 0;
  for (size_t i = 0; i < mjd_raw.size(); ++i)
    if (k < n_seg && mjd_raw[i] >= T[k]) first[k++] = (int)i;
  return k;
}
const double T_START3[3] = {58051.044623, 58053.078872, 58055.234144};
const double T_START4[4] = {58051.044623, 58053.078872, 58054.093273, 58055.234144};

// 3-point Lagrange interpolation of E and S at time t (misc.c:20-35).
void quad_interp(const double M[3], const double E[3][3], const double S[3][3], double t, double* e, double* s) {
  double rr0 = (t - M[1]) * (t - M[2]) / (M[0] - M[1]) / (M[0] - M[2]);
  double rr1 = (t - M[0]) * (t - M[2]) / (M[1] - M[0]) / (M[1] - M[2]);
  double rr2 = (t - M[0]) * (t - M[1]) / (M[2] - M[0]) / (M[2] - M[1]);
  for (int c = 0; c < 3; ++c) {
    e[c] = E[c][0] * rr0 + E[c][1] * rr1 + E[c][2] * rr2;
    s[c] = S[c][0] * rr0 + S[c][1] * rr1 + S[c][2] * rr2;
  }
}

FILE* open_eph(const char* name) {
  FILE* f = fopen(name, "r");
  if (!f) {
    printf("Ephemeris file %s does not exist!\n", name);
    exit(1);
  }
  char line[MAX_LINE];
  while (fgets(line, sizeof line, f))
    if (strcmp(line, "$$SOE\n") == 0) return f;
  printf("No $$SOE marker in %s!\n", name);
  exit(1);
}

// read_data.c:6-302: light curve, the three ephemerides, corrections.
Ingested read_data(const char* data_file, bool interp) {
  Ingested D;
  FILE* fp = fopen(data_file, "r");
  if (!fp) {
    printf("Input file %s does not exist!\n", data_file);
    exit(1);
  }
  std::vector<double> mjd_obs;
  std::vector<float> V, w;
  std::vector<int> filt;
  char line[MAX_LINE];
  int W_filter = -1, D_filter = -1;
  printf("Filters:\n");
  while (fgets(line, sizeof line, fp)) {
    if (!strchr(line, '\n')) die("Error: every light-curve line must end with a newline!");
    char f;
    double MJD1, V1, sgm;
    if (sscanf(line, "%c %lf %lf %lf", &f, &MJD1, &V1, &sgm) != 4) die("Error: bad light-curve line!");
    if (!mjd_obs.empty() && MJD1 <= mjd_obs.back()) die("Error: the data have to be sorted chronologically!");
    mjd_obs.push_back(MJD1);
    V.push_back((float)V1);
    w.push_back((float)(1.0 / (sgm * sgm)));
    int k = 0;
    for (; k < (int)D.filters.size(); ++k)
      if (D.filters[k] == f) break;
    if (k == (int)D.filters.size()) {
      D.filters.push_back(f);
      if (f == 'W') W_filter = k;
      if (f == 'D') D_filter = k;
      printf("%d: %c\n", k + 1, f);
    }
    filt.push_back(k);
  }
  fclose(fp);
  const int N = (int)mjd_obs.size();
  if (N == 0) die("Error: empty light curve!");
  D.n_filters = (int)D.filters.size();
  if (D.n_filters > AST_MAX_FILTERS) {
    fprintf(stderr, "Too many filters - at most %d are supported! %d\n", AST_MAX_FILTERS, D.n_filters);
    exit(1);
  }
  // Ephemerides: 3-point sliding window; each observation is interpolated in
  // the first window [MJD0[0], MJD0[2]) that contains it (read_data.c:127-235).
  FILE* fA = open_eph("asteroid.eph");
  FILE* fE = open_eph("earth.eph");
  FILE* fS = open_eph("sun.eph");
  char lA[MAX_LINE], lE[MAX_LINE], lS[MAX_LINE];
  double E0[3][3] = {{0}}, S0[3][3] = {{0}}, M0[3] = {0};
  std::vector<double> Ex(N), Ey(N), Ez(N), Sx(N), Sy(N), Sz(N);
  int j = -1, l = 0, i = 0;
  double JD = 0;
  while (fgets(lA, sizeof lA, fA)) {
    if (strcmp(lA, "$$EOE\n") == 0) break;
    j++;
    if (!fgets(lE, sizeof lE, fE) || !fgets(lS, sizeof lS, fS)) die("Error: ephemeris files differ in length!");
    if (j % 2 == 0) {
      sscanf(lA, "%lf", &JD);
    } else {
      if (l > 2) {
        for (int m = 0; m < 2; ++m) {
          for (int c = 0; c < 3; ++c) { E0[c][m] = E0[c][m + 1]; S0[c][m] = S0[c][m + 1]; }
          M0[m] = M0[m + 1];
        }
        l = 2;
      }
      double Xa, Ya, Za, Xe, Ye, Ze, Xs, Ys, Zs;
      if (sscanf(lA, "%lE %lE %lE", &Xa, &Ya, &Za) != 3 || sscanf(lE, "%lE %lE %lE", &Xe, &Ye, &Ze) != 3 ||
          sscanf(lS, "%lE %lE %lE", &Xs, &Ys, &Zs) != 3)
        die("Error: bad ephemeris vector line!");
      E0[0][l] = Xe - Xa; E0[1][l] = Ye - Ya; E0[2][l] = Ze - Za;
      S0[0][l] = Xs - Xa; S0[1][l] = Ys - Ya; S0[2][l] = Zs - Za;
      M0[l] = JD - 2400000.5;
      l++;
      if (l > 2) {
        while (i < N && mjd_obs[i] >= M0[0] && mjd_obs[i] < M0[2]) {
          double e[3], s[3];
          quad_interp(M0, E0, S0, mjd_obs[i], e, s);
          Ex[i] = e[0]; Ey[i] = e[1]; Ez[i] = e[2];
          Sx[i] = s[0]; Sy[i] = s[1]; Sz[i] = s[2];
          i++;
        }
      }
    }
    if (i >= N) break;
  }
  fclose(fA);
  fclose(fE);
  fclose(fS);
  if (l < 3) die("Error: the ephemerides need at least three points!");
  if (i < N) die("Error: observations outside the ephemeris span!");
  // Corrections (read_data.c:243-285).
  D.obs.resize(N);
  for (int k = 0; k < N; ++k) {
    double E = std::sqrt(Ex[k] * Ex[k] + Ey[k] * Ey[k] + Ez[k] * Ez[k]);
    double S = std::sqrt(Sx[k] * Sx[k] + Sy[k] * Sy[k] + Sz[k] * Sz[k]);
    ast_obs& o = D.obs[k];
    std::memset(&o, 0, sizeof o);
    o.V = V[k];
    if (filt[k] != W_filter) o.V = (float)(o.V + 5.0 * std::log10(1.0 / E * 1.0 / S));
    o.w = w[k];
    o.filter = filt[k];
    double delay = E / LIGHT_SPEED;
    o.mjd = mjd_obs[k];
    D.mjd_raw.push_back(mjd_obs[k]);
    if (filt[k] != W_filter && filt[k] != D_filter) o.mjd = o.mjd - delay;
    if (k == 0) D.mjd0 = o.mjd;
    o.mjd = o.mjd - D.mjd0;
    o.E[0] = Ex[k] / E; o.E[1] = Ey[k] / E; o.E[2] = Ez[k] / E;
    o.S[0] = Sx[k] / S; o.S[1] = Sy[k] / S; o.S[2] = Sz[k] / S;
  }
  // The INTERP tables are the last window, in the same relative time
  // (read_data.c:298-302).
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 3; ++k) { D.eph.E[c][k] = E0[c][k]; D.eph.S[c][k] = S0[c][k]; }
  for (int k = 0; k < 3; ++k) D.eph.mjd[k] = M0[k] - D.mjd0;
  (void)interp;
  return D;
}

// The -plot branch of main() (asteroid.c:833-1059 without PROFILES,
// ANIMATE, MINIMA_TEST, PLOT_OMEGA, LAST, MY_L and NOPRINT): the model
// curve on the NPLOT-point grid, delta_V, model_params.dat, the minima
// periodogram line, the chi2_plot line, model.dat and data.dat.
// chi_float: the reference's CHI_FLOAT is float (every form but -DACC), so
// data.dat's converted magnitudes are computed in float as there.
int plot_model(ast_ctx* ctx, const ast_variant& v, const Ingested& D, const std::vector<double>& params,
               const int* property, int n_params, bool chi_float) {
  const int NPLOT = 20000;  // asteroid.h:266
  const int NCL_MAX = 5;    // asteroid.h:277
  const int N_data = (int)D.obs.size(), N_filters = D.n_filters;
  printf("\n*** Plotting ***\n\n");
  std::vector<double> t(NPLOT), vm(NPLOT), dV(N_filters);
  double chi2 = 0;
  CHECK(ast_model_curve(ctx, params.data(), 1, &D.eph, NPLOT, t.data(), vm.data(), &chi2, dV.data()));
  FILE* fV = fopen("delta_V", "w");
  if (!fV) die("Cannot open delta_V!");
  for (int m = 0; m < N_filters; ++m) fprintf(fV, "%8.4f\n", dV[m]);
  fclose(fV);
  // The initial model without the torque, photometric b,c, TREND's A and
  // BW_BALL parameters (asteroid.c:904-932); TORQUE2's second torque and
  // split time and ROTATE's angles are written.
  ast_type_ids T;
  ast_variant_type_ids(&v, &T);
  FILE* fmod = fopen("model_params.dat", "w");
  if (!fmod) die("Cannot open model_params.dat!");
  for (int k = 0; k < n_params; ++k) {
    const int ty = property[k * AST_N_COLUMNS + AST_P_TYPE];
    if (v.torque && (ty == T.Ti || ty == T.Ts || ty == T.Tl)) continue;
    if (v.bc && (ty == T.c || ty == T.b)) continue;
    if (v.trend && ty == T.A) continue;
    if (v.bw_ball && (ty == T.theta_R || ty == T.phi_R || ty == T.kappa)) continue;
    fprintf(fmod, "%15.11f ", params[k]);
  }
  fprintf(fmod, "\n");
  fclose(fmod);
  double cl_fr[NCL_MAX], cl_H[NCL_MAX];
  CHECK(ast_periodogram(ctx, t.data(), vm.data(), NPLOT, NCL_MAX, cl_fr, cl_H, nullptr));
  for (int j = 0; j < NCL_MAX; ++j) printf("%f ", cl_fr[j]);
  for (int j = 0; j < NCL_MAX; ++j) printf("%f ", cl_H[j]);
  printf("\n");
  // asteroid.c:984-1023 averages h_dlsq2 over six time bins, but no kernel
  // writes d_dlsq2 (chi2_plot takes it and leaves it alone) and it is never
  // copied back, so h_dlsq2 keeps the zeros of its fresh allocation: the
  // distance is 0 when every bin holds data and 0/0 otherwise.
  double d2[6], w[6];
  for (int i = 0; i < 6; ++i) { d2[i] = 0.0; w[i] = 0.0; }
  for (int i = 0; i < N_data; ++i) {
    const double tm = D.obs[i].mjd;
    const int ind = tm < 0.5 ? 0 : tm < 1.5 ? 1 : tm < 2.7 ? 2 : tm < 3.7 ? 3 : tm < 4.5 ? 4 : 5;
    d2[ind] = d2[ind] + 0.0;
    w[ind] = w[ind] + 1;
  }
  double sum = 0.0, SW = 0.0;
  for (int i = 0; i < 6; ++i) {
    const double W = (i == 2 || i == 3) ? 2 : 1;  // extra weight on bins 2 and 3
    sum = sum + W * d2[i] / w[i];
    SW = SW + W;
  }
  const double dist = std::sqrt(sum / SW);
  printf("chi2_plot = %13.6e, lsq = %13.6e\n", chi2, dist);
  FILE* fp = fopen("model.dat", "w");
  if (!fp) die("Cannot open model.dat!");
  for (int i = 0; i < NPLOT; ++i) fprintf(fp, "%13.7f %13.6e\n", D.mjd0 + t[i], vm[i]);
  fclose(fp);
  fp = fopen("data.dat", "w");
  if (!fp) die("Cannot open data.dat!");
  for (const ast_obs& o : D.obs) {
    // V converted to the first filter; the error bar 1/sqrt(w) in float
    double Vc;
    if (chi_float) Vc = (float)((float)(o.V - (float)dV[o.filter]) + (float)dV[0]);
    else Vc = (double)o.V - dV[o.filter] + dV[0];
    const float sg = 1 / std::sqrt(o.w);
    fprintf(fp, "%13.7f %13.6e %13.6e w\n", D.mjd0 + o.mjd, Vc, sg);
  }
  fclose(fp);
  return 0;
}

// The RMSD build's main loop (asteroid.c:643-725): one confidence-interval
// launch per cycle; the cumulative rejected fraction and the cumulative
// range of every parameter go to stdout, the ranges to the results file.
int rmsd_loop(ast_ctx* ctx, int n_blocks, int bsize, int Nstages, float dx_rand, int spherical_k, int n_params,
              int Ncases, const char* results) {
  std::vector<float> PAR_min(n_params, 1e30f), PAR_max(n_params, -1e30f);
  std::vector<float> hmin((size_t)n_blocks * n_params), hmax((size_t)n_blocks * n_params);
  printf("\n");
  FILE* fp = fopen(results, "w");
  if (!fp) die("Cannot open the output file!");
  ast_rmsd_cfg cfg{n_blocks, bsize, Nstages, dx_rand, spherical_k};
  int64_t Ntot = 0, Nbad = 0;  // d_Ntot, d_Nbad: never reset
  int loop_counter = 0;
  while (true) {
    loop_counter++;
    int64_t nt = 0, nb = 0;
    double f0 = 0, f1 = 0;
    CHECK(ast_rmsd(ctx, &cfg, hmin.data(), hmax.data(), &nt, &nb, &f0, &f1));
    Ntot += nt;
    Nbad += nb;
    if (loop_counter == 1) printf("f0=%f, f1=%f\n\n", (float)f0, (float)f1);
    printf("%e ", Nbad / (double)Ntot);
    for (int j = 0; j < n_params; ++j) {
      for (int i = 0; i < n_blocks; ++i) {
        const size_t k = (size_t)i * n_params + j;
        if (hmin[k] < PAR_min[j]) PAR_min[j] = hmin[k];
        if (hmax[k] > PAR_max[j]) PAR_max[j] = hmax[k];
      }
      printf("%13.6e %13.6e ; ", PAR_min[j], PAR_max[j]);
      fprintf(fp, "%13.6e %13.6e ; ", PAR_min[j], PAR_max[j]);
    }
    printf("\n");
    fflush(stdout);
    fprintf(fp, "\n");
    fflush(fp);
    if (loop_counter == Ncases) break;
  }
  fclose(fp);
  return 0;
}

void usage() {
  printf("\n Command line arguments:\n\n");
  printf("-best : only keep the best result\n");
  printf("-f type_constant value: forces the parameter with the type_constant to be frozen during optimization at \"value\" \n");
  printf("-i name : input (data) file name\n");
  printf("-keep : keep all intermediate results, not just the best ones\n");
  printf("-l type_constant limit1 limit2 : specify range for the parameter type_constant\n");
  printf("-m param1 param2 ... paramN : input model parameters, for plotting and re-optimization\n");
  printf("     If one of the parameters has a special value of \"v\", it is allowed to vary randomly within its full range.\n");
  printf("-N number : exit after \"number\" cycles\n");
  printf("-Nstages number : each initial optimization stage is followed by \"number-1\" reoptimization stages\n");
  printf("-o name : output (results) file name\n");
  printf("-Ppsi min max : minimum and maximum values for Ppsi period, in hours\n");
  printf("-reopt : reoptimize the model provided with -m switch\n");
  printf("-seed SEED : use the SEED number to initialize the random number generator\n");
  printf("-t : travelling reoptimization\n");
  printf("-dx value : maximum shift for parameters in scale-free units (0...1), with -DRMSD\n");
  printf("-D<MACRO> : model variant (TORQUE, TORQUE2, P_PSI, BC, RECT, ROTATE, TREND, BW_BALL, NUDGE, V1S, INTERP,\n"
         "            ACC, MIXED)\n"
         "            and mode (RMSD, SPHERICAL_K);\n"
         "            default -DP_PSI -DTORQUE -DINTERP\n");
  printf("-blocks n, -bsize n : reseed groups per launch and simplexes per group (N_BLOCKS, BSIZE)\n");
  printf("-steps n : simplex steps per stage (N_STEPS); -budget n : chi^2 evaluations per stage\n");
  printf("-time_step d : ODE time step in days (TIME_STEP); -device n : HIP device\n");
  printf("-dump_ingest name : write the ingested observations and exit\n");
  printf("\n");
}

}  // namespace

// NUDGE: the observed minima of observed.min (prepare_chi2_params,
// misc.c:279-323): "t V" per line, read as floats; only minima within 0.05
// days of the data range are kept, at most NOBS_MAX of them, with times on
// the curve's clock.
void load_minima(ast_ctx* ctx, const Ingested& D) {
  FILE* f1 = fopen("observed.min", "r");
  if (!f1) die("Cannot open observed.min!");
  char line[1024];
  std::vector<float> t_min, V_min;
  const double last = D.obs.back().mjd;
  while (fgets(line, sizeof line, f1)) {
    float t_obs = 0, V_obs = 0;
    if (sscanf(line, "%f %f", &t_obs, &V_obs) != 2) continue;
    const double small = 0.05;
    if (t_obs >= D.mjd0 - small && t_obs <= last + D.mjd0 + small) {
      if ((int)t_min.size() >= AST_MAX_MINIMA) {
        printf("Too many lines in observed.min file! (>NOBS_MAX)!\n");
        exit(1);
      }
      t_min.push_back((float)(t_obs - D.mjd0));
      V_min.push_back(V_obs);
    }
  }
  fclose(f1);
  if (t_min.empty()) {
    printf("No local minima in the data range!\n");
    exit(1);
  }
  printf("%d minima in observed.min\n", (int)t_min.size());
  CHECK(ast_set_minima(ctx, t_min.data(), V_min.data(), (int)t_min.size()));
}

int main(int argc, char** argv) {
  if (argc == 1) {
    usage();
    return 0;
  }
  ast_variant v{};
  bool any_macro = false;
  int acc = 0;
  v.time_step = 1e-2;
  int j_input = -1, j_results = -1;
  int travel = 0, keep = 0, best = 0, reopt = 0, model = 0, plot = 0, rmsd = 0, spherical_k = 0;
  float dx_rand = 0.0f;
  unsigned long seed = 0;
  int Ncases = -1, Nstages = 1;
  int n_blocks = 1024, bsize = 256, n_steps = 5000, device = 0;
  long budget = 0;
  double Ppsi1 = 2.0, Ppsi2 = 4800.0;
  std::vector<std::string> m_args;
  std::vector<int> Fi, Li;
  std::vector<double> Fv, L0, L1;
  const char* dump = nullptr;
  for (int j = 1; j < argc;) {
    const char* a = argv[j];
    auto need = [&](int k) {
      if (j + k >= argc) die("Bad arguments!");
    };
    if (a[0] != '-') die("Bad arguments!");
    if (!strncmp(a, "-D", 2) && strlen(a) > 2) {
      any_macro = true;
      std::string mac(a + 2);
      if (mac == "TORQUE") v.torque = 1;
      else if (mac == "P_PSI") v.p_psi = 1;
      else if (mac == "BC") v.bc = 1;
      else if (mac == "BW_BALL") v.bw_ball = 1;
      else if (mac == "TREND") v.trend = 1;
      else if (mac == "TORQUE2") v.torque2 = v.torque = 1;  // asteroid.h:32-34
      else if (mac == "ROTATE") v.rotate = 1;
      else if (mac == "RECT") v.rect = v.bc = 1;            // asteroid.h:36-38
      else if (mac == "NUDGE") v.nudge = 1;
      else if (mac == "V1S") v.v1s = 1;
      else if (mac == "SEGMENT") { if (v.n_seg < 2) v.n_seg = 4; }  // N_SEG defaults to 4 (asteroid.h:42-46)
      else if (mac == "N_SEG=3") v.n_seg = 3;
      else if (mac == "N_SEG=4") v.n_seg = 4;
      else if (mac == "ONE_LE") v.one_le = 1;
      else if (mac == "INTERP") v.interp = 1;
      else if (mac == "ACC") acc = 1;
      else if (mac == "MIXED") acc = 2;
      else if (mac == "RMSD") rmsd = 1;
      else if (mac == "SPHERICAL_K") spherical_k = 1;
      else {
        printf("Unsupported macro -D%s\n", mac.c_str());
        return 1;
      }
      j += 1;
    } else if (!strcmp(a, "-i")) { need(1); j_input = j + 1; j += 2; }
    else if (!strcmp(a, "-o")) { need(1); j_results = j + 1; j += 2; }
    else if (!strcmp(a, "-Ppsi")) { need(2); Ppsi1 = atof(argv[j + 1]); Ppsi2 = atof(argv[j + 2]); j += 3; }
    else if (!strcmp(a, "-m")) {
      // The parameter count depends on the variant, so -m takes every value up
      // to the next switch; checked against N_PARAMS below.
      j++;
      while (j < argc && !(argv[j][0] == '-' && (argv[j][1] < '0' || argv[j][1] > '9') && argv[j][1] != '.')) {
        m_args.push_back(argv[j]);
        j++;
      }
      model = 1;
    } else if (!strcmp(a, "-plot")) { plot = 1; j += 1; }
    else if (!strcmp(a, "-t")) { travel = 1; j += 1; }
    else if (!strcmp(a, "-reopt")) { reopt = 1; j += 1; }
    else if (!strcmp(a, "-keep")) { keep = 1; j += 1; }
    else if (!strcmp(a, "-best")) { best = 1; j += 1; }
    else if (!strcmp(a, "-seed")) { need(1); seed = strtoul(argv[j + 1], nullptr, 10); j += 2; }
    else if (!strcmp(a, "-N")) { need(1); Ncases = atoi(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-Nstages")) { need(1); Nstages = atoi(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-f")) { need(2); Fi.push_back(atoi(argv[j + 1])); Fv.push_back(atof(argv[j + 2])); j += 3; }
    else if (!strcmp(a, "-l")) {
      need(3);
      Li.push_back(atoi(argv[j + 1]));
      L0.push_back(atof(argv[j + 2]));
      L1.push_back(atof(argv[j + 3]));
      j += 4;
    } else if (!strcmp(a, "-blocks")) { need(1); n_blocks = atoi(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-bsize")) { need(1); bsize = atoi(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-steps")) { need(1); n_steps = atoi(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-budget")) { need(1); budget = atol(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-time_step")) { need(1); v.time_step = atof(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-device")) { need(1); device = atoi(argv[j + 1]); j += 2; }
    else if (!strcmp(a, "-dump_ingest")) { need(1); dump = argv[j + 1]; j += 2; }
    else if (!strcmp(a, "-dx")) { need(1); dx_rand = (float)atof(argv[j + 1]); j += 2; }
    else die("Bad arguments!");
  }
  // The default variant; -DRMSD and -DSPHERICAL_K alone keep it.
  if (!any_macro || !(v.torque || v.p_psi || v.bc || v.bw_ball || v.interp || v.trend || v.rotate || acc)) {
    v.torque = 1;
    v.p_psi = 1;
    v.interp = 1;
  }
  v.precision = acc == 2 ? AST_PREC_MIXED : (acc ? AST_PREC_FP64 : AST_PREC_F32SIMPLEX);
  if (j_input == -1) die("-i parameter is missing!");
  if ((reopt || plot) && !model) die("-reopt and -plot switches require -m switch!");
  if (!reopt && !plot && model) die("-m can only be used together with -reopt or -plot switches!");

  // Parameter tables of the variant (asteroid.c:86-169, 473-601).
  int n_params = 0, n_types = 0;
  int property[AST_MAX_PARAMS * AST_N_COLUMNS], types[AST_MAX_TYPES];
  double limits[2 * AST_MAX_TYPES];
  ast_x2 x2;
  double pp[2] = {Ppsi1, Ppsi2};
  CHECK(ast_default_model_space(&v, pp, &n_params, property, &n_types, types, limits, &x2));
  std::vector<double> params(n_params, 0.0);
  if (model) {
    if ((int)m_args.size() != n_params) die("Not enough of model parameters in -m switch!");
    for (int k = 0; k < n_params; ++k) {
      if (m_args[k][0] == 'v') {
        property[k * AST_N_COLUMNS + AST_P_FROZEN] = -1;
        params[k] = 0;
      } else {
        params[k] = atof(m_args[k].c_str());
      }
    }
  }
  for (size_t f = 0; f < Fi.size(); ++f) {  // asteroid.c:566-581
    int itype = Fi[f];
    if (itype < 0 || itype >= n_types) die("Bad type constant in -f!");
    for (int k = 0; k < n_params; ++k)
      if (property[k * AST_N_COLUMNS + AST_P_TYPE] == itype) {
        property[k * AST_N_COLUMNS + AST_P_FROZEN] = 1;
        limits[itype] = Fv[f];
        limits[n_types + itype] = Fv[f];
      }
  }
  for (size_t f = 0; f < Li.size(); ++f) {  // asteroid.c:583-589
    int itype = Li[f];
    if (itype < 0 || itype >= n_types) die("Bad type constant in -l!");
    limits[itype] = L0[f];
    limits[n_types + itype] = L1[f];
  }
  x2.reopt = reopt;

  Ingested D = read_data(argv[j_input], v.interp);
  const int N_data = (int)D.obs.size(), N_filters = D.n_filters;
  if (dump) {
    FILE* fd = fopen(dump, "w");
    if (!fd) die("Cannot open the dump file!");
    fprintf(fd, "# mjd0 %.17g\n", D.mjd0);
    for (const ast_obs& o : D.obs)
      fprintf(fd, "%.17g %.9g %.9g %d %.17g %.17g %.17g %.17g %.17g %.17g\n", o.mjd, o.V, o.w, o.filter, o.E[0],
              o.E[1], o.E[2], o.S[0], o.S[1], o.S[2]);
    for (int k = 0; k < 3; ++k)
      fprintf(fd, "# eph %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", D.eph.mjd[k], D.eph.E[0][k], D.eph.E[1][k],
              D.eph.E[2][k], D.eph.S[0][k], D.eph.S[1][k], D.eph.S[2][k]);
    if (v.n_seg > 1) {  // the segments' first observations
      int first[4] = {0, 0, 0, 0};
      const int n = segment_first_obs(D.mjd_raw, v.n_seg, v.n_seg == 3 ? T_START3 : T_START4, first);
      fprintf(fd, "# seg_first");
      for (int k = 0; k < n; ++k) fprintf(fd, " %d", first[k]);
      fprintf(fd, "\n");
    }
    fclose(fd);
    return 0;
  }

  ast_ctx* ctx = nullptr;
  int rc = ast_ctx_create(&ctx, &v, device);
  if (rc == AST_E_NODEVICE) die("No HIP devices!");
  CHECK(rc);
  CHECK(ast_set_model_space(ctx, n_params, property, n_types, types, limits, &x2));
  if (v.n_seg > 1) {
    // the observations split on the observer's clock (read_data.c:244-250),
    // the plotting grid at the start times on the curve's clock (read_data.c:357)
    const double* T = v.n_seg == 3 ? T_START3 : T_START4;
    double ts[4];
    int first[4] = {0, 0, 0, 0};
    for (int k = 0; k < v.n_seg; ++k) ts[k] = T[k] - D.mjd0;
    if (segment_first_obs(D.mjd_raw, v.n_seg, T, first) < v.n_seg) die("Error: a data segment has no observations!");
    CHECK(ast_set_segment_starts(ctx, first, ts, v.n_seg));
  }
  CHECK(ast_upload_curve(ctx, D.obs.data(), N_data, N_filters, &D.eph));
  if (v.nudge) load_minima(ctx, D);
  if (plot) {
    int rc_plot = plot_model(ctx, v, D, params, property, n_params, acc != 1);
    CHECK(ast_ctx_destroy(ctx));
    return rc_plot;
  }

  printf("\n*** Simplex optimization ***\n\n");
  printf(" N_threads = %d\n", n_blocks * bsize);
  if (j_results == -1) die("Bad output file name!");
  CHECK(ast_rng_seed(ctx, seed == 0 ? (uint64_t)time(nullptr) : (uint64_t)seed));
  if (reopt) CHECK(ast_set_start(ctx, params.data()));
  if (rmsd) {
    int rc_rmsd = rmsd_loop(ctx, n_blocks, bsize, Nstages, dx_rand, spherical_k, n_params, Ncases,
                            argv[j_results]);
    CHECK(ast_ctx_destroy(ctx));
    return rc_rmsd;
  }
  ast_run_cfg cfg{n_blocks, bsize, Nstages, n_steps, budget, 0.0, 0};
  std::vector<double> h_f(n_blocks), h_params((size_t)n_blocks * n_params), h_dV((size_t)n_blocks * N_filters);
  int loop_counter = 0;
  while (true) {  // asteroid.c:655-825
    loop_counter++;
    CHECK(ast_run(ctx, &cfg));
    CHECK(ast_fetch(ctx, h_f.data(), h_params.data(), h_dV.data()));
    int i_best = 0;
    double chi2_tot = 1e32;
    for (int i = 0; i < n_blocks; ++i)
      if (h_f[i] < chi2_tot) { chi2_tot = h_f[i]; i_best = i; }
    // asteroid.c:751-757 meant to wrap periodic angles into [0, 2pi), but its
    // index update `l = l++` leaves l at -1 under gcc, so angles are printed
    // raw; they are printed raw here too.
    if (reopt && travel) {
      for (int k = 0; k < n_params; ++k) params[k] = h_params[(size_t)i_best * n_params + k];
      CHECK(ast_set_start(ctx, params.data()));
    }
    printf("%13.6e ", h_f[i_best]);
    for (int k = 0; k < n_params; ++k) printf("%15.11f ", h_params[(size_t)i_best * n_params + k]);
    printf("\n");
    fflush(stdout);
    FILE* fp = fopen(argv[j_results], keep ? "a" : "w");
    if (!fp) die("Cannot open the output file!");
    int i1 = best ? i_best : 0, i2 = best ? i_best + 1 : n_blocks;
    for (int i = i1; i < i2; ++i) {
      fprintf(fp, "%13.6e ", h_f[i]);
      for (int m = 0; m < N_filters; ++m) fprintf(fp, "%13.6e ", h_dV[(size_t)i * N_filters + m]);
      for (int k = 0; k < n_params; ++k) fprintf(fp, "%15.11f ", h_params[(size_t)i * n_params + k]);
      fprintf(fp, "\n");
    }
    fclose(fp);
    if (keep) CHECK(ast_reset_best(ctx));
    if (loop_counter == Ncases) break;
  }
  CHECK(ast_ctx_destroy(ctx));
  return 0;
}
