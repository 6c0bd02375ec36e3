// astrng.h -- the project's random stream for simplex start-up.
//
// The reference draws from one cuRAND XORWOW state per thread
// (curand_init(seed, id, 0), cuda.c:2046; curand_uniform at cuda.c:1462,
// 1519, 1523).  Here every simplex has a counter-based stream instead:
// draw k of simplex `fit` under seed `seed` is word (k mod 4) of
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) applied to the counter (k/4 low, k/4 high, fit
// low, fit high) with key (seed low, seed high).  A stream is a pure
// function of (seed, fit, k), so host and device produce identical draws and
// a simplex only has to remember how many draws it has taken (the role of
// the curandState written back at cuda.c:1829).
//
// uniform01 maps a draw to a float in (0, 1], as curand_uniform does:
// ((u >> 8) + 1) * 2^-24, exact in single precision.
#pragma once
#include <stdint.h>
#include "astmath.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AST_RNG_HD __host__ __device__ inline
#else
#define AST_RNG_HD inline
#endif

namespace astrng {

AST_RNG_HD void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * ctr[0];
    uint64_t p1 = (uint64_t)M1 * ctr[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ ctr[1] ^ k0;
    uint32_t n2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = n0;
    ctr[1] = lo1;
    ctr[2] = n2;
    ctr[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

AST_RNG_HD uint32_t draw_u32(uint64_t seed, uint64_t fit, uint32_t k) {
  uint32_t c[4] = {k >> 2, 0u, (uint32_t)fit, (uint32_t)(fit >> 32)};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  return c[k & 3];
}

AST_RNG_HD float uniform01(uint64_t seed, uint64_t fit, uint32_t k) {
  return (float)((draw_u32(seed, fit, k) >> 8) + 1u) * (1.0f / 16777216.0f);
}

// A standard normal deviate (the role of curand_normal, cuda.c:2276) from
// draws k and k+1 by the Box-Muller transform (Box & Muller, Ann. Math.
// Statist. 29, 1958): sqrt(-2 ln u1) cos(2 pi u2), evaluated in double with
// the project's own log and cos and rounded to float.
AST_RNG_HD float normal01(uint64_t seed, uint64_t fit, uint32_t k) {
  const double u1 = (double)uniform01(seed, fit, k);
  const double u2 = (double)uniform01(seed, fit, k + 1);
  const double two_pi = 6.283185307179586232;
  const double r = astm::p_sqrt(-2.0 * astm::log(u1));
  return (float)(r * astm::cos(two_pi * u2));
}

}  // namespace astrng
