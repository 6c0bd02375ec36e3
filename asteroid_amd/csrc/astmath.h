// astmath.h -- the project's own FP64 elementary functions.
//
// One header, compiled unchanged by hipcc for gfx950 device code, by hipcc's
// host pass and by g++.  Every function is a fixed sequence of IEEE-754
// binary64 operations (+ - * /, sqrt, fma, rint, floor, comparisons and
// exponent-field bit operations), all correctly rounded on both sides, so the
// host and the GPU produce bitwise-identical results as long as the including
// translation unit is compiled with -ffp-contract=off (fma is always spelled
// out).  No vendor math library is called.
//
// Methods (coefficients from gen_math_coeffs.py, Chebyshev fits at 60 digits):
//   sincos: Cody-Waite reduction by pi/2 (3-part constant, fma), odd sine and
//           even cosine kernels on |r| <= pi/4, quadrant swap/sign.
//   sincos_t: table-driven (Tang's method): x = k pi/512 + r (3-part pi/512,
//           |r| <= pi/1024), sin and cos of k pi/512 from a 1024-entry table,
//           short kernels for sin r and cos r - 1, angle addition.
//   atan2 : ratio t = min/max in [0,1]; for t > tan(pi/8) the identity
//           atan t = pi/4 + atan((t-1)/(t+1)); odd kernel on |u| <= tan(pi/8);
//           octant/quadrant fix-up with two-part pi constants.
//   log   : x = m 2^e, m in [sqrt(1/2), sqrt(2)); s = f/(2+f), f = m-1;
//           log(1+f) = f - (f^2/2 - s (f^2/2 + z R1(z))), z = s^2; e ln2 added
//           in two parts.
//   log10 : log(x) times 1/ln10 in two parts.
//   exp   : k = rint(x/ln2), r = x - k ln2 (two parts, fma), 1 + r + r^2 Q(r),
//           scaled by 2^k through the exponent field.
// Accuracy (tests/test_astmath.py against mpmath): about 1 ulp.
//
// The functions are templates over the scalar type R.  For double the
// primitives below are used; the oracle's operation-counting scalar supplies
// overloads of the same primitive names (found by argument-dependent lookup),
// which is how the algorithmic FLOP count of an evaluation is measured.
#pragma once
#include <stdint.h>
#include "astmath_coeffs.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AST_HD __host__ __device__ inline
#else
#define AST_HD inline
#endif

namespace astm {

// ---------------------------------------------------------------- primitives
AST_HD double p_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
AST_HD double p_sqrt(double a) { return __builtin_sqrt(a); }
AST_HD double p_rint(double a) { return __builtin_rint(a); }
AST_HD double p_floor(double a) { return __builtin_floor(a); }
AST_HD double p_fabs(double a) { return __builtin_fabs(a); }
AST_HD double p_copysign(double a, double b) { return __builtin_copysign(a, b); }
// Raw value of a scalar, for comparisons and exponent-field work (not counted).
AST_HD double p_val(double a) { return a; }

#if defined(__HIPCC__)
// n / d rounded as IEEE division rounds it, in 8 instructions where the
// compiler's division takes 11: the hardware reciprocal estimate refined by
// two Newton steps, the quotient and one residual correction, the same
// operations in the same order as the compiler's sequence without its range
// scaling (v_div_scale, v_div_fmas) and special-case fix-up (v_div_fixup).
// Those act only when d or the quotient is subnormal, |n| < 2^-969, the
// quotient overflows, or an operand is 0, infinite or NaN.  The evaluator's
// divisors (sin theta in the torque ODE, norms and products of norms in the
// brightness, 2 + f in log) never meet them with finite states, and n = +0
// gives +0 as the division does; a divisor of exactly 0 gives NaN where the
// division gives an infinity that the evaluation then turns into NaN.
__device__ inline double div_rn(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = n * r;
  return __builtin_fma(__builtin_fma(-d, q, n), r, q);
}
// sqrt(x) rounded as IEEE sqrt rounds it, in 13 instructions where the
// compiler's takes 18: the hardware reciprocal square root refined by the
// compiler's own Newton steps, in its order, without its scaling of
// x < 2^-767 (the evaluator's arguments, squared norms and 1 - cos^2, are 0
// or far above it); +-0 and +inf are returned as they are.
__device__ inline double sqrt_rn(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  double g = x * r;
  double h = r * 0.5;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  double d = __builtin_fma(-g, g, x);
  h = __builtin_fma(h, e, h);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return __builtin_amdgcn_class(x, 0x260) ? x : g;  // class bits of -0, +0, +inf
}
#endif
// a / b, the division of the canonical form: div_rn on the GPU, the IEEE
// division on the host (the same value).
AST_HD double p_div(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return div_rn(a, b);
#else
  return a / b;
#endif
}

AST_HD uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
AST_HD double of_bits(uint64_t b) { return __builtin_bit_cast(double, b); }
// -x if bit 1 of flag is set, else x (flag is 0 or 2): the bit is shifted
// into the sign position and xored in (exact, not counted).
AST_HD double p_flipsign(double x, int flag) {
  return of_bits(bits_of(x) ^ ((uint64_t)(uint32_t)(flag & 2) << 62));
}

// 1.5*2^52: for |v| < 2^51, v + RINT_SHIFT rounds v to the nearest integer
// k (ties to even) and holds k in two's complement in its low mantissa bits.
constexpr double RINT_SHIFT = 6755399441055744.0;
// Low word of an integer-valued double k, |k| < 2^51: k + RINT_SHIFT is exact.
// One add, where a range-checked float-to-int conversion takes three
// instructions.
AST_HD uint32_t low_word_of_int(double k) { return (uint32_t)bits_of(k + RINT_SHIFT); }
// x with its sign bit xored by bit 31 of m; m's other bits are ignored
// (exact, not counted).  On the GPU the mask and the xor into the high word
// are one v_bitop3_b32 (hi ^ (m & 0x80000000), table 0x78).
AST_HD double p_xorsign(double x, uint32_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t b = bits_of(x);
  const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(b >> 32), m, 0x80000000u, 0x78);
  return of_bits(((uint64_t)hi << 32) | (uint32_t)b);
#else
  return of_bits(bits_of(x) ^ ((uint64_t)(m & 0x80000000u) << 32));
#endif
}

// 2^k as a double, for -1022 <= k <= 1023.
AST_HD double pow2i(int k) { return of_bits((uint64_t)(k + 1023) << 52); }

// x * 2^k for |x| in [0.5, 2], any int k: two exact power-of-two factors, so
// results in the subnormal range round once, at the last multiply.
AST_HD double scale2(double x, int k) {
  if (k > 2046) k = 2046;
  if (k < -2100) k = -2100;
  int k1 = k / 2;
  int k2 = k - k1;
  if (k1 < -1022) { k2 += k1 + 1022; k1 = -1022; }
  if (k2 < -1022) k2 = -1022;
  if (k1 > 1023) { k2 += k1 - 1023; k1 = 1023; }
  if (k2 > 1023) k2 = 1023;
  return (x * pow2i(k1)) * pow2i(k2);
}

// ------------------------------------------------------------------- kernels
template <class R> AST_HD R k_sin(R z) {  // S(z), sin r = r + r^3 S(r^2)
  R p = p_fma(z, R(AST_SIN_5), R(AST_SIN_4));
  p = p_fma(z, p, R(AST_SIN_3));
  p = p_fma(z, p, R(AST_SIN_2));
  p = p_fma(z, p, R(AST_SIN_1));
  return p_fma(z, p, R(AST_SIN_0));
}
template <class R> AST_HD R k_cos(R z) {  // C(z), cos r = 1 - r^2/2 + r^4 C(r^2)
  R p = p_fma(z, R(AST_COS_5), R(AST_COS_4));
  p = p_fma(z, p, R(AST_COS_3));
  p = p_fma(z, p, R(AST_COS_2));
  p = p_fma(z, p, R(AST_COS_1));
  return p_fma(z, p, R(AST_COS_0));
}
template <class R> AST_HD R k_atan(R z) {  // A(z), atan u = u + u^3 A(u^2)
  R p = p_fma(z, R(AST_ATAN_10), R(AST_ATAN_9));
  p = p_fma(z, p, R(AST_ATAN_8));
  p = p_fma(z, p, R(AST_ATAN_7));
  p = p_fma(z, p, R(AST_ATAN_6));
  p = p_fma(z, p, R(AST_ATAN_5));
  p = p_fma(z, p, R(AST_ATAN_4));
  p = p_fma(z, p, R(AST_ATAN_3));
  p = p_fma(z, p, R(AST_ATAN_2));
  p = p_fma(z, p, R(AST_ATAN_1));
  return p_fma(z, p, R(AST_ATAN_0));
}
template <class R> AST_HD R k_log(R z) {  // R1(z)
  R p = p_fma(z, R(AST_LOG_6), R(AST_LOG_5));
  p = p_fma(z, p, R(AST_LOG_4));
  p = p_fma(z, p, R(AST_LOG_3));
  p = p_fma(z, p, R(AST_LOG_2));
  p = p_fma(z, p, R(AST_LOG_1));
  return p_fma(z, p, R(AST_LOG_0));
}
template <class R> AST_HD R k_exp(R r) {  // Q(r), exp r = 1 + r + r^2 Q(r)
  R p = p_fma(r, R(AST_EXP_10), R(AST_EXP_9));
  p = p_fma(r, p, R(AST_EXP_8));
  p = p_fma(r, p, R(AST_EXP_7));
  p = p_fma(r, p, R(AST_EXP_6));
  p = p_fma(r, p, R(AST_EXP_5));
  p = p_fma(r, p, R(AST_EXP_4));
  p = p_fma(r, p, R(AST_EXP_3));
  p = p_fma(r, p, R(AST_EXP_2));
  p = p_fma(r, p, R(AST_EXP_1));
  return p_fma(r, p, R(AST_EXP_0));
}

// ----------------------------------------------------------------- functions
// sin and cos of x together.  Non-finite x gives NaN for both.
template <class R> AST_HD void sincos(R x, R* s, R* c) {
  R k = p_rint(x * R(AST_TWO_OVER_PI));
  R r = p_fma(-k, R(AST_PIO2_1), x);
  r = p_fma(-k, R(AST_PIO2_2), r);
  r = p_fma(-k, R(AST_PIO2_3), r);
  R z = r * r;
  R sn = p_fma(r * z, k_sin(z), r);
  R hz = R(0.5) * z;
  R w = R(1.0) - hz;
  R cs = w + p_fma(z * z, k_cos(z), (R(1.0) - w) - hz);
  // quadrant q = k mod 4: odd q swaps sin and cos; sin takes the sign of bit
  // 1 of q, cos that of bit 1 of q + 1 (each shifted to bit 31)
  const uint32_t q = low_word_of_int(p_val(k));
  R s0 = (q & 1) ? cs : sn;
  R c0 = (q & 1) ? sn : cs;
  *s = p_xorsign(s0, q << 30);
  *c = p_xorsign(c0, (q + 1) << 30);
}

// Table of sin and cos of k pi/512, k = 0..1023, for sincos_t.  On the host a
// constant array; on the GPU a 16 KB copy in LDS, so that 64 lanes with 64
// different k read it in one ds_read_b128 each: a kernel that calls sincos_t
// must call sincos_table_load() first, with every thread of the workgroup
// (it ends with a barrier); until then the copy holds garbage.
#if defined(__HIP_DEVICE_COMPILE__)
static __shared__ double2 g_sctab[AST_SCTAB_N];
__device__ inline void sincos_table_load() {
  constexpr double ts[AST_SCTAB_N] = AST_SCTAB_SIN;
  constexpr double tc[AST_SCTAB_N] = AST_SCTAB_COS;
  for (int i = threadIdx.x; i < AST_SCTAB_N; i += blockDim.x) g_sctab[i] = make_double2(ts[i], tc[i]);
  __syncthreads();
}
__device__ inline void sctab_get(uint32_t i, double* s, double* c) {
  const double2 v = g_sctab[i];
  *s = v.x;
  *c = v.y;
}
#else
AST_HD void sincos_table_load() {}
inline void sctab_get(uint32_t i, double* s, double* c) {
  static const double ts[AST_SCTAB_N] = AST_SCTAB_SIN;
  static const double tc[AST_SCTAB_N] = AST_SCTAB_COS;
  *s = ts[i];
  *c = tc[i];
}
#endif

// sin and cos of x, table-driven (Tang's method): x = k pi/512 + r by a
// Cody-Waite reduction with a three-part pi/512 (|r| <= pi/1024), S = sin and
// C = cos of (k mod 1024) pi/512 from the table, sin r = r + r^3 ST(r^2) and
// cos r - 1 = r^2 (-1/2 + r^2 CT), then
//   sin x = (S + C sin r) + S (cos r - 1),  cos x = (C - S sin r) + C (cos r - 1),
// each as two fmas (within 1 ulp of the rounded value like the three-operation
// S + (S (cos r - 1) + C sin r), rounded correctly a little less often).
// k is x 512/pi rounded once, by an fma onto RINT_SHIFT (which also leaves k
// in the low word for the table index), and RINT_SHIFT taken off again.
// About 1 ulp for |x| < 2^51 pi/512, with no quadrant logic: 17 instructions
// and one table read on the GPU against 39 for sincos.  Non-finite x gives
// NaN for both.
template <class R> AST_HD void sincos_t(R x, R* s, R* c) {
  const R kb = p_fma(x, R(AST_SCTAB_INV_STEP), R(RINT_SHIFT));
  const R k = kb - R(RINT_SHIFT);
  R r = p_fma(-k, R(AST_SCTAB_STEP_1), x);
  r = p_fma(-k, R(AST_SCTAB_STEP_2), r);
  r = p_fma(-k, R(AST_SCTAB_STEP_3), r);
  double tsv, tcv;
  sctab_get((uint32_t)bits_of(p_val(kb)) & (AST_SCTAB_N - 1), &tsv, &tcv);
  const R S(tsv), C(tcv);
  R z = r * r;
  R sr = p_fma(r * z, p_fma(z, R(AST_SINT_1), R(AST_SINT_0)), r);
  R cm = z * p_fma(z, R(AST_COST_0), R(-0.5));
  *s = p_fma(S, cm, p_fma(C, sr, S));
  *c = p_fma(C, cm, p_fma(-S, sr, C));
}

template <class R> AST_HD R sin(R x) { R s, c; sincos(x, &s, &c); return s; }
template <class R> AST_HD R cos(R x) { R s, c; sincos(x, &s, &c); return c; }
template <class R> AST_HD R tan(R x) { R s, c; sincos(x, &s, &c); return s / c; }

// atan2(y, x) for finite or NaN arguments; atan2(0, 0) = 0 with the sign of y
// (x >= 0) or +-pi (x < 0).
template <class R> AST_HD R atan2(R y, R x) {
  R ax = p_fabs(x);
  R ay = p_fabs(y);
  bool sw = p_val(ay) > p_val(ax);
  R num = sw ? ax : ay;
  R den = sw ? ay : ax;
  R t = (p_val(den) == 0.0) ? R(0.0) : num / den;
  bool big = p_val(t) > AST_TAN_PI_8;
  R u = t;
  if (big) u = (t - R(1.0)) / (t + R(1.0));
  R z = u * u;
  R a = p_fma(u * z, k_atan(z), u);
  if (big) a = R(AST_PIO4_HI) + (a + R(AST_PIO4_LO));
  if (sw) a = (R(AST_PIO2_HI) - a) + R(AST_PIO2_LO);
  if (p_val(x) < 0.0) a = (R(AST_PI_HI) - a) + R(AST_PI_LO);
  return p_copysign(a, y);
}

template <class R> AST_HD R atan(R x) { return atan2(x, R(1.0)); }

// asin(x) = atan2(x, sqrt((1-x)(1+x))); NaN for |x| > 1.
template <class R> AST_HD R asin(R x) {
  return atan2(x, p_sqrt((R(1.0) - x) * (R(1.0) + x)));
}

// acos(x) = atan2(sqrt((1-x)(1+x)), x); NaN for |x| > 1.
template <class R> AST_HD R acos(R x) {
  return atan2(p_sqrt((R(1.0) - x) * (R(1.0) + x)), x);
}

// Natural logarithm: -inf at 0, NaN below 0 or for NaN, +inf at +inf.
// log(m 2^e) for m in (sqrt(1/2), sqrt(2)]: the kernel and the e ln2 terms.
template <class R> AST_HD R log_split(double mv, int e) {
  R m = R(mv);
  R f = m - R(1.0);
  R s = p_div(f, R(2.0) + f);
  R z = s * s;
  R rz = z * k_log(z);
  R hfsq = R(0.5) * f * f;
  R dk = R((double)e);
  return dk * R(AST_LN2_HI) - ((hfsq - p_fma(s, hfsq + rz, dk * R(AST_LN2_LO))) - f);
}

template <class R> AST_HD R log(R x) {
  double xv = p_val(x);
#if defined(__HIP_DEVICE_COMPILE__)
  // Without branches on the GPU: the hardware splits x = m 2^E with m in
  // [0.5, 1) (exactly, subnormals included), which gives the same m, e as
  // the exponent-field split below, and the special cases are selected at
  // the end.
  const double m = __builtin_amdgcn_frexp_mant(xv);
  const int E = __builtin_amdgcn_frexp_exp(xv);
  const double m2 = 2.0 * m;
  const bool big = m2 > 1.4142135623730951;
  const R r = log_split<R>(big ? m : m2, big ? E : E - 1);
  const double sp = xv == 0.0 ? -__builtin_inf() : (xv == __builtin_inf() ? xv : __builtin_nan(""));
  return __builtin_amdgcn_class(xv, 0x180) ? r : R(sp);  // class bits of +subnormal, +normal
#else
  if (!(xv > 0.0)) return R(xv == 0.0 ? -__builtin_inf() : __builtin_nan(""));
  if (xv == __builtin_inf()) return x;
  uint64_t b = bits_of(xv);
  int ex = (int)((b >> 52) & 0x7ff);
  int sub = 0;
  if (ex == 0) {  // subnormal: bring into the normal range first (exact)
    b = bits_of(xv * 18014398509481984.0);  // 2^54
    ex = (int)((b >> 52) & 0x7ff);
    sub = 54;
  }
  int e = ex - 1023 - sub;
  double mv = of_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1,2)
  if (mv > 1.4142135623730951) { mv *= 0.5; e += 1; }
  return log_split<R>(mv, e);
#endif
}

template <class R> AST_HD R log10(R x) {
  R y = log(x);
  return p_fma(y, R(AST_INV_LN10_HI), y * R(AST_INV_LN10_LO));
}

// Exponential: +inf above ~709.78, +0 below ~-745.13, NaN for NaN.
template <class R> AST_HD R exp(R x) {
  double xv = p_val(x);
  if (xv != xv) return x;
  if (xv > 709.782712893384) return R(__builtin_inf());
  if (xv < -745.1332191019412) return R(0.0);
  R k = p_rint(x * R(AST_INV_LN2));
  R r = p_fma(-k, R(AST_LN2_HI), x);
  r = p_fma(-k, R(AST_LN2_LO), r);
  R e = R(1.0) + p_fma(r * r, k_exp(r), r);
  return R(scale2(p_val(e), (int)p_val(k)));
}

// Single-precision exponential of the simplex set-up (float arithmetic there):
// the double result rounded to float.
AST_HD float expf(float x) { return (float)exp<double>((double)x); }

// Single-precision power for x >= 0 (the NUDGE merit's powf): exp(y log x)
// in double, rounded to float; 0 for x = 0 and y > 0, 1 for x = 1.
AST_HD float powf(float x, float y) { return (float)exp<double>((double)y * log<double>((double)x)); }

}  // namespace astm
