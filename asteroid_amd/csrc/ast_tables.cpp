// ast_tables.cpp -- ast_default_model_space(): the parameter tables the
// reference's main() builds for a variant (asteroid.c:86-169 Property and
// Types, asteroid.c:473-564 hard-coded limits, asteroid.c:599-600 P_psi).
// The row budget is AST_MAX_PARAMS; variants beyond it are not built.
#include <cmath>
#include <cstring>

#include "../../include/asteroid.h"

extern "C" int ast_default_model_space(const ast_variant* v, const double* ppsi_hours, int* n_params, int* property,
                                       int* n_types, int* types, double* limits, ast_x2* x2) {
  if (!v || !n_params || !property || !n_types || !types || !limits || !x2) return AST_E_ARG;
  ast_type_ids T;
  ast_variant_type_ids(v, &T);
  if (T.n_params > AST_MAX_PARAMS || T.n_types > AST_MAX_TYPES) return AST_E_UNSUPPORTED;
  const double PI = 3.141592653589793238, RAD = 180.0 / PI, Tmax = 10.0;
  int rows[AST_MAX_PARAMS][AST_N_COLUMNS];
  int n = 0;
  auto add = [&](int type, int indep, int multi, int periodic) {
    int* r = rows[n++];
    r[AST_P_TYPE] = type;
    r[AST_P_INDEPENDENT] = indep;
    r[AST_P_FROZEN] = 0;
    r[AST_P_ISEG] = 0;
    r[AST_P_MULTI_SEGMENT] = multi;
    r[AST_P_PERIODIC] = periodic;
  };
  add(T.theta_M, 1, 0, AST_HARD_BOTH);
  add(T.phi_M, 1, 0, AST_PERIODIC);
  add(T.phi_0, 1, 0, AST_PERIODIC);
  if (v->trend) add(T.A, 1, 1, AST_HARD_BOTH);
  if (v->torque || v->torque2) {
    add(T.Ti, 1, 0, AST_SOFT_BOTH);
    add(T.Ts, 1, 0, AST_SOFT_BOTH);
    add(T.Tl, 1, 0, AST_SOFT_BOTH);
  }
  if (v->torque2) {
    add(T.T2i, 1, 0, AST_SOFT_BOTH);
    add(T.T2s, 1, 0, AST_SOFT_BOTH);
    add(T.T2l, 1, 0, AST_SOFT_BOTH);
    add(T.Tt, 1, 0, AST_HARD_BOTH);
  }
  add(T.c_tumb, 1, 1, AST_HARD_RIGHT);
  add(T.b_tumb, 0, 1, AST_HARD_BOTH);
  add(T.Es, 0, v->one_le ? 1 : 0, AST_HARD_BOTH);  // ONE_LE: shared by the segments (asteroid.c:68-72)
  add(T.L, 1, v->one_le ? 1 : 0, AST_HARD_LEFT);
  add(T.psi_0, 0, 0, AST_PERIODIC_LAM);
  if (v->bc || v->rect) {
    add(T.c, 1, 1, AST_HARD_RIGHT);
    add(T.b, 0, 1, AST_HARD_BOTH);
  }
  if (v->rotate || v->bw_ball) {
    add(T.theta_R, 1, 1, AST_HARD_BOTH);
    add(T.phi_R, 1, 1, AST_PERIODIC);
  }
  if (v->rotate) add(T.psi_R, 1, 1, AST_PERIODIC);
  if (v->bw_ball) add(T.kappa, 1, 1, AST_SOFT_BOTH);
  const int nt = T.n_types;
  double lo[AST_MAX_TYPES] = {0}, hi[AST_MAX_TYPES] = {0};
  lo[T.theta_M] = 0.001 / RAD;
  hi[T.theta_M] = 179.999 / RAD;
  lo[T.L] = 48.0 * PI / 10;
  hi[T.L] = 48.0 * PI / 0.1;
  if (v->trend) {  // asteroid.c:498-503
    lo[T.A] = -5;
    hi[T.A] = 5;
  }
  if (v->torque || v->torque2) {
    lo[T.Ti] = -Tmax; hi[T.Ti] = Tmax;
    lo[T.Ts] = -Tmax; hi[T.Ts] = Tmax;
    lo[T.Tl] = -Tmax; hi[T.Tl] = Tmax;
  }
  if (v->torque2) {  // asteroid.c:515-525
    lo[T.T2i] = -Tmax; hi[T.T2i] = Tmax;
    lo[T.T2s] = -Tmax; hi[T.T2s] = Tmax;
    lo[T.T2l] = -Tmax; hi[T.T2l] = Tmax;
    lo[T.Tt] = 0.001; hi[T.Tt] = 0.999;
  }
  lo[T.c_tumb] = std::log(0.01);
  hi[T.c_tumb] = std::log(1.0);
  lo[T.b_tumb] = 0;
  hi[T.b_tumb] = 1;
  if (v->bc || v->rect) {
    lo[T.c] = lo[T.c_tumb];
    hi[T.c] = hi[T.c_tumb];
    lo[T.b] = 0;
    hi[T.b] = 1;
  }
  if (v->rotate || v->bw_ball) {
    lo[T.theta_R] = 0.001 / RAD;
    hi[T.theta_R] = 179.999 / RAD;
  }
  if (v->bw_ball) {
    lo[T.kappa] = std::log(0.01);
    hi[T.kappa] = std::log(0.1);
  }
  const int n0 = n;  // SEGMENT: further segments repeat the unshared parameters (asteroid.c:128-150)
  for (int k = 1; k < v->n_seg; ++k)
    for (int j = 0; j < n0; ++j)
      if (rows[j][AST_P_MULTI_SEGMENT] == 0) {
        if (n >= AST_MAX_PARAMS) return AST_E_UNSUPPORTED;
        for (int c = 0; c < AST_N_COLUMNS; ++c) rows[n][c] = rows[j][c];
        rows[n++][AST_P_ISEG] = k;
      }
  if (n != T.n_params) return AST_E_UNSUPPORTED;
  *n_params = n;
  *n_types = nt;
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < AST_N_COLUMNS; ++k) property[i * AST_N_COLUMNS + k] = rows[i][k];
  for (int t = 0; t < nt; ++t) types[t] = -1;
  for (int i = 0; i < n0; ++i) types[rows[i][AST_P_TYPE]] = i;
  for (int t = 0; t < nt; ++t) {
    limits[t] = lo[t];
    limits[nt + t] = hi[t];
  }
  double p1 = ppsi_hours ? ppsi_hours[0] : 2.0, p2 = ppsi_hours ? ppsi_hours[1] : 4800.0;
  x2->reopt = 0;
  x2->ppsi1 = (float)(24.0 / p2);
  x2->ppsi2 = (float)(24.0 / p1);
  return AST_OK;
}
