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
//  * The count of a whole bin (binned outputs only): one Poisson draw with the
//    summed mean of the bin's pixels below 1000, inversion or PTRS at any mean.
//  * N(0,1): Box-Muller on two 53-bit uniforms (FP64; FP32 for the readout noise).
#pragma once

#include <cstdint>
#include <cmath>

#ifdef __HIPCC__
#define SPSIM_HD __host__ __device__ __forceinline__
// Rare branches (PTRS, the Gaussian branch) stay out of line, so their
// registers do not limit the occupancy of the kernels' common path.
#define SPSIM_HD_RARE __host__ __device__ __noinline__
#else
#define SPSIM_HD inline
#define SPSIM_HD_RARE inline
#endif

namespace spsim {

enum RngPurpose : uint32_t { kRngPoisson = 1u, kRngReadout = 2u, kRngPoissonBin = 3u };

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

// A uniform in (0, 1) from one 32-bit word, exact in FP64: (w + 1/2) 2^-32.
SPSIM_HD double uniform32(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

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

// PTRS for mean >= 10 on the attempts att0 .. att0 + 63 (64 rejections at
// most, probability < 0.2^64, then the mode).
SPSIM_HD_RARE double poisson_ptrs(double mean, uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose,
                                  uint32_t att0 = 0u) {
  const double slam = sqrt(mean), loglam = log(mean);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t att = 0; att < 64u; ++att) {
    double u, v;
    uniform2(draw(seed, idx, pattern, purpose, att0 + att), u, v);
    u -= 0.5;
    const double us = 0.5 - fabs(u);
    const double k = floor((2.0 * a / us + b) * u + mean + 0.43);
    if (us >= 0.07 && v <= vr) return k;
    if (k < 0.0 || (us < 0.013 && v > us)) continue;
    if (log(v) + log(invalpha) - log(a / (us * us) + b) <= -mean + k * loglam - log_factorial(k)) return k;
  }
  return floor(mean);
}

// Inversion by sequential search for mean < 10: the smallest k whose CDF
// reaches u (stops after 256 steps, whose probability is below 1e-100 there).
SPSIM_HD double poisson_inversion(double mean, double u) {
  double p = exp(-mean), F = p;
  int k = 0;
  while (u > F && k < 256) {
    ++k;
    p *= mean / k;
    F += p;
  }
  return (double)k;
}

// Poisson(mean) with no Gaussian branch, addressed by (seed, idx, pattern,
// purpose): inversion below 10 on the first uniform of attempt 0, PTRS above.
SPSIM_HD double poisson_exact(double mean, uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose) {
  if (!(mean > 0.0)) return 0.0;
  if (mean >= 10.0) return poisson_ptrs(mean, seed, idx, pattern, purpose);
  double u0, u1;
  uniform2(draw(seed, idx, pattern, purpose, 0u), u0, u1);
  (void)u1;
  return poisson_inversion(mean, u0);
}

// The reference's Gaussian branch, (int)(mean + sqrt(mean) N(0,1)) truncated
// toward zero (src/noise.c:26-28,44-46), saturated at INT_MAX where the
// reference's conversion would overflow.
SPSIM_HD_RARE double poisson_gaussian(double mean, uint64_t seed, uint32_t idx, uint32_t pattern) {
  const double g = std_normal(seed, idx, pattern, kRngPoisson, 0u);
  const double v = mean + sqrt(mean) * g;
  return v >= 2147483647.0 ? 2147483647.0 : trunc(v);
}

// Poisson(mean) as the reference's generate_poisson_noise would draw it, with
// the sampler addressed by (seed, idx, pattern): from mean 1000 up,
// (int)(mean + sqrt(mean) N(0,1)) truncated toward zero (src/noise.c:26-28,44-46),
// below that poisson_exact.
SPSIM_HD double poisson_sample(double mean, uint64_t seed, uint32_t idx, uint32_t pattern) {
  if (!(mean > 0.0)) return 0.0;
  if (mean >= 1000.0) return poisson_gaussian(mean, seed, idx, pattern);
  return poisson_exact(mean, seed, idx, pattern, kRngPoisson);
}

#ifdef __HIPCC__
// The same Box-Muller draw in FP32 on the hardware's log2, square root and
// cosine (argument in turns), from the same two 53-bit uniforms rounded to
// FP32: for the readout noise, where a relative error of about 1e-6 in a
// N(0, sigma) sample is far below any use of it.
__device__ __forceinline__ float std_normal_f32(uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose) {
  double u0, u1;
  uniform2(draw(seed, idx, pattern, purpose, 0u), u0, u1);
  const float l2 = __builtin_amdgcn_logf((float)u0);  // log2(u0)
  const float r = __builtin_sqrtf(-2.0f * 0.693147180559945309f * l2);
  return r * __builtin_amdgcn_cosf((float)u1);
}

// FP32 Box-Muller from two given uniforms.
__device__ __forceinline__ float box_muller_f32(double u0, double u1) {
  const float l2 = __builtin_amdgcn_logf((float)u0);
  return __builtin_sqrtf(-2.0f * 0.693147180559945309f * l2) * __builtin_amdgcn_cosf((float)u1);
}

// The photon count of a bin with summed mean < 10 from its block (seed, bin,
// pattern, kRngPoissonBin, 0): inversion on the 53-bit uniform u of words
// 0-1 (bins with larger means take PTRS on attempts 1..64, bin_count). The
// inversion runs in FP32 with
// certainty margins and returns the FP64 inversion's k exactly: the FP32
// exp2 of the rounded argument is within 2e-6 of e^-mean, each recurrence
// step p *= mean / k adds at most 2e-7, and the running sum and the rounded
// uniform 6e-8 each, so for k <= 40 the FP32 CDF is within 1.3e-5 of the
// FP64 one; k is accepted when u_f clears both neighbouring CDF values by
// 3e-5 relative, and any closer case (or k > 40) takes the FP64 inversion.
__device__ __forceinline__ double bin_count_small(float mean, const U4 &r) {
  if (!(mean > 0.0f)) return 0.0;
  double u0, u1;
  uniform2(r, u0, u1);
  (void)u1;
  const float uf = (float)u0;
  float p = __builtin_amdgcn_exp2f(-mean * 1.44269504088896341f);
  float F = p, lo = 0.0f;  // lo: CDF below the candidate k
  int k = 0;
  while (uf >= F * 0.99997f && k < 40) {
    if (uf <= F * 1.00003f) return poisson_inversion((double)mean, u0);  // too close to F_k
    ++k;
    lo = F;
    p *= mean / (float)k;
    F += p;
  }
  if (uf >= F * 0.99997f || (k > 0 && uf <= lo * 1.00003f)) return poisson_inversion((double)mean, u0);
  return (double)k;
}

// Any bin's count: PTRS on attempts 1..64 from mean 10 up.
__device__ __forceinline__ double bin_count(float mean, const U4 &r, uint64_t seed, uint32_t bin, uint32_t pattern) {
  if (mean >= 10.0f) return poisson_ptrs((double)mean, seed, bin, pattern, kRngPoissonBin, 1u);
  return bin_count_small(mean, r);
}
#endif

}  // namespace spsim
