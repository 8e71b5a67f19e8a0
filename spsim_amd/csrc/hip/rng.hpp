// Counter-based random numbers for the detector chain (device and host).
//
// Philox-4x32-10 as published by Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3" (SC'11): a 10-round bijection of a
// 128-bit counter under a 64-bit key. Every random draw of the detector chain
// is addressed by (pixel, pattern, purpose, attempt), so a result does not
// depend on the launch shape or on which kernel produced it.
//
// Samplers:
//  * Poisson, mean < 10: inversion by sequential search on one uniform.
//  * Poisson, 10 <= mean < 1000: PTRS, the transformed rejection method of
//    W. Hormann, "The transformed rejection method for generating Poisson
//    random variables", Insurance: Mathematics and Economics 12 (1993) 39-45.
//  * mean >= 1000: (int)(mean + sqrt(mean) N(0,1)), truncated toward zero, the
//    reference's Gaussian branch (src/noise.c:26-28,44-46).
//  * N(0,1): Box-Muller on two 53-bit uniforms.
#pragma once

#include <cstdint>
#include <cmath>

#ifdef __HIPCC__
#define SPSIM_HD __host__ __device__ __forceinline__
#else
#define SPSIM_HD inline
#endif

namespace spsim {

enum RngPurpose : uint32_t { kRngPoisson = 1u, kRngReadout = 2u };

struct U4 {
  uint32_t v[4];
};

SPSIM_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

SPSIM_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(M0, c.v[0], hi0, lo0);
    mulhilo32(M1, c.v[2], hi1, lo1);
    U4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Two uniforms in (0, 1) with 53-bit resolution from one Philox block.
SPSIM_HD void uniform2(const U4 &r, double &u0, double &u1) {
  const double s = 1.0 / 9007199254740992.0;  // 2^-53
  u0 = ((double)(((uint64_t)(r.v[0] >> 5) << 26) | (uint64_t)(r.v[1] >> 6)) + 0.5) * s;
  u1 = ((double)(((uint64_t)(r.v[2] >> 5) << 26) | (uint64_t)(r.v[3] >> 6)) + 0.5) * s;
}

SPSIM_HD U4 draw(uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose, uint32_t attempt) {
  U4 c;
  c.v[0] = idx;
  c.v[1] = pattern;
  c.v[2] = purpose;
  c.v[3] = attempt;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

SPSIM_HD double std_normal(uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose, uint32_t attempt) {
  double u0, u1;
  uniform2(draw(seed, idx, pattern, purpose, attempt), u0, u1);
  return sqrt(-2.0 * log(u0)) * cos(6.283185307179586476925286766559 * u1);
}

// log(k!) for integer k >= 0: exact table below 10, Stirling series above.
SPSIM_HD double log_factorial(double k) {
  if (k < 10.0) {
    const double t[10] = {0.0,
                          0.0,
                          0.69314718055994530942,
                          1.79175946922805500081,
                          3.17805383034794561964,
                          4.78749174278204599424,
                          6.57925121201010099506,
                          8.52516136106541430017,
                          10.60460290274525022842,
                          12.80182748008146961121};
    return t[(int)k];
  }
  const double n = k + 1.0;  // log Gamma(n)
  const double in = 1.0 / n, in2 = in * in;
  return (n - 0.5) * log(n) - n + 0.91893853320467274178 +
         in * (1.0 / 12.0 - in2 * (1.0 / 360.0 - in2 * (1.0 / 1260.0 - in2 * (1.0 / 1680.0))));
}

// Poisson(mean) as the reference's generate_poisson_noise would draw it, with
// the sampler addressed by (seed, idx, pattern). Bounded loops: inversion stops
// after 256 steps (probability of reaching them below 1e-100 for mean < 10)
// and PTRS after 64 rejections (probability < 0.2^64), returning the mode.
SPSIM_HD double poisson_sample(double mean, uint64_t seed, uint32_t idx, uint32_t pattern) {
  if (!(mean > 0.0)) return 0.0;
  if (mean >= 1000.0) {
    const double g = std_normal(seed, idx, pattern, kRngPoisson, 0u);
    // (int) conversion: truncation toward zero, saturated at INT_MAX where the
    // reference's conversion would overflow.
    const double v = mean + sqrt(mean) * g;
    return v >= 2147483647.0 ? 2147483647.0 : trunc(v);
  }
  if (mean < 10.0) {
    double u0, u1;
    uniform2(draw(seed, idx, pattern, kRngPoisson, 0u), u0, u1);
    (void)u1;
    double p = exp(-mean), F = p;
    int k = 0;
    while (u0 > F && k < 256) {
      ++k;
      p *= mean / k;
      F += p;
    }
    return (double)k;
  }
  const double slam = sqrt(mean), loglam = log(mean);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t att = 0; att < 64u; ++att) {
    double u, v;
    uniform2(draw(seed, idx, pattern, kRngPoisson, att), u, v);
    u -= 0.5;
    const double us = 0.5 - fabs(u);
    const double k = floor((2.0 * a / us + b) * u + mean + 0.43);
    if (us >= 0.07 && v <= vr) return k;
    if (k < 0.0 || (us < 0.013 && v > us)) continue;
    if (log(v) + log(invalpha) - log(a / (us * us) + b) <= -mean + k * loglam - log_factorial(k)) return k;
  }
  return floor(mean);
}

}  // namespace spsim
