// Detector chain on gfx950: Thomson x solid angle x beam -> photons, Poisson
// photon counting, electrons, binned readout with dark current, readout noise,
// clamp and ADC, and the noiseless binned output.
//
// Each stage restates the reference's arithmetic in the same precision and
// order (FP64 geometry rounded to FP32, FP32 products left to right, FP32
// bin sums in (y, x) order) with FMA contraction off, so the deterministic
// outputs agree with a host evaluation of the same expressions. The stage
// kernels back the reference-shaped host API; chain_binned_runs_kernel (bins
// of up to 32 raw pixels) and chain_binned_kernel fuse all stages into one
// pass over the intensities (one lane per binned pixel) and are what the
// batched pipeline uses. Random draws are addressed by
// (seed, pixel, pattern), so fused and staged runs produce identical noise.
#include "epilogue.hpp"

#include <algorithm>

#include "rng.hpp"

namespace spsim {
namespace {

constexpr int kT = 256;

__device__ __forceinline__ void pixel_xy(const DetConst &d, int x, int y, double &px, double &py) {
#pragma clang fp contract(off)
  px = ((x - (d.nx - 1.0) / 2.0) / d.nx) * d.width - d.center_x;
  py = ((y - (d.ny - 1.0) / 2.0) / d.ny) * d.height - d.center_y;
}

// src/diffraction.c:1109-1144
__device__ __forceinline__ float thomson_at(const DetConst &d, int x, int y) {
#pragma clang fp contract(off)
  double pf = 1.0;
  if (d.polarization != 0) {
    double px, py;
    pixel_xy(d, x, y, px, py);
    const double r = sqrt(px * px + py * py + d.dist2);
    if (d.polarization == 1) {
      const double chi = asin(py / r);
      pf = cos(chi) * cos(chi);
    } else if (d.polarization == 2) {
      const double chi = asin(px / r);
      pf = cos(chi) * cos(chi);
    } else if (d.polarization == 3) {
      const double chi = asin(sqrt(px * px + py * py) / r);
      pf = (1. + cos(chi) * cos(chi)) / 2.;
    }
  }
  const double r0 = 2.81794e-15;
  return (float)((r0 * r0) * pf);
}

// src/diffraction.c:1146-1187 (2-D detector)
__device__ __forceinline__ float solid_angle_at(const DetConst &d, int x, int y) {
#pragma clang fp contract(off)
  if (d.spherical) return (float)(d.pixel_area / (d.distance * d.distance));
  double px, py;
  pixel_xy(d, x, y, px, py);
  double inc = atan(sqrt(px * px + py * py) / d.distance);
  inc = cos(inc);
  inc = inc * inc * inc;
  return (float)(d.pixel_area / d.dist2 * inc);
}

struct RawPixel {
  float th, om, ph, cnt, el;
};

// Stages of one raw pixel from its Thomson factor and solid angle: photons
// (src/diffraction.c:1202), Poisson count (src/noise.c:42-46), electrons
// (src/amplification.c:66-68). idx is the raw pixel's flat index
// (z * ny + y) * nx + x, the Poisson draw's counter.
__device__ __forceinline__ RawPixel raw_from_geometry(const DetConst &d, float th, float om, uint32_t idx, float I,
                                                      uint64_t seed, uint32_t pattern) {
#pragma clang fp contract(off)
  RawPixel r;
  r.th = th;
  r.om = om;
  r.ph = r.th * r.om * I * d.beam_intensity;
  const float mean = r.ph * d.qe;
  r.cnt = (float)poisson_sample((double)mean, seed, idx, pattern);
  r.el = r.cnt * d.e_per_photon;
  return r;
}

__device__ __forceinline__ RawPixel raw_pixel(const DetConst &d, int x, int y, uint32_t idx, float I, uint64_t seed,
                                              uint32_t pattern) {
  return raw_from_geometry(d, thomson_at(d, x, y), solid_angle_at(d, x, y), idx, I, seed, pattern);
}

// Readout noise of one binned pixel: readout_noise x N(0, 1) (FP32 Box-Muller).
__device__ __forceinline__ float readout_noise(const DetConst &d, uint64_t seed, uint32_t bin, uint32_t pattern) {
  return d.readout_noise * std_normal_f32(seed, bin, pattern, kRngReadout);
}

// src/amplification.c:104-108: noise once per binned pixel, clamp at 0, ADC.
__device__ __forceinline__ float finish_real(const DetConst &d, float acc, uint64_t seed, uint32_t bin,
                                             uint32_t pattern) {
#pragma clang fp contract(off)
  acc += readout_noise(d, seed, bin, pattern);
  if (acc < 0) acc = 0;
  acc *= d.adc;
  return acc;
}

struct ChainOut {
  float *th, *om, *ph, *cnt, *el, *real, *noiseless;
};

__global__ __launch_bounds__(kT) void chain_binned_kernel(DetConst d, const float *__restrict__ I, int npatterns,
                                                          uint64_t seed, uint64_t pattern_base, ChainOut o) {
#pragma clang fp contract(off)
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y, nbz = d.nz / d.bin_z;
  const long long nb = (long long)nbx * nby * nbz;
  const long long gid = (long long)blockIdx.x * kT + threadIdx.x;
  if (gid >= nb * npatterns) return;
  const int p = (int)(gid / nb);
  const int bin = (int)(gid % nb);
  const int bx = bin % nbx, by = (bin / nbx) % nby, bz = bin / (nbx * nby);
  const uint32_t pat = (uint32_t)(pattern_base + p);
  const size_t npix = (size_t)d.nx * d.ny * d.nz;
  const float *Ip = I + (size_t)p * npix;
  const float dark = d.dark_current * d.exposure_time;
  float acc = 0.0f, accn = 0.0f;
  // bin sums in (z, y, x) order (src/amplification.c:94-102, 139-147)
  for (int zi = 0; zi < d.bin_z; ++zi) {
    for (int yi = 0; yi < d.bin_y; ++yi) {
      for (int xi = 0; xi < d.bin_x; ++xi) {
        const int x = bx * d.bin_x + xi, y = by * d.bin_y + yi, z = bz * d.bin_z + zi;
        const size_t i = ((size_t)z * d.ny + y) * d.nx + x;
        const RawPixel r = raw_pixel(d, x, y, (uint32_t)i, Ip[i], seed, pat);
        acc += (r.el + dark);
        accn += r.ph * d.qe * d.e_per_photon_noiseless * d.adc;
        const size_t oi = (size_t)p * npix + i;
        if (o.th) o.th[oi] = r.th;
        if (o.om) o.om[oi] = r.om;
        if (o.ph) o.ph[oi] = r.ph;
        if (o.cnt) o.cnt[oi] = r.cnt;
        if (o.el) o.el[oi] = r.el;
      }
    }
  }
  const size_t ob = (size_t)p * nb + bin;
  if (o.real) o.real[ob] = finish_real(d, acc, seed, (uint32_t)bin, pat);
  if (o.noiseless) o.noiseless[ob] = accn;
}

// The same chain for a run of patterns per lane: the Thomson factor and solid
// angle of a bin's raw pixels do not depend on the pattern (FP64 geometry,
// most of a pixel's arithmetic), so each lane forms them once into its own LDS
// slots, geo[t][lane] and geo[nraw + t][lane], and then takes its bin through
// patterns [blockIdx.y * run, + run). Same arithmetic and draws as
// chain_binned_kernel, so the outputs are bitwise those of one pattern at a time.
constexpr int kRunMaxRaw = 32;  // raw pixels per bin held in LDS (64 KiB per workgroup)

__global__ __launch_bounds__(kT) void chain_binned_runs_kernel(DetConst d, const float *__restrict__ I, int npatterns,
                                                               int run, uint64_t seed, uint64_t pattern_base,
                                                               ChainOut o) {
#pragma clang fp contract(off)
  extern __shared__ float geo[];
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y, nbz = d.nz / d.bin_z;
  const int nraw = d.bin_x * d.bin_y * d.bin_z;
  const long long nb = (long long)nbx * nby * nbz;
  const long long gid = (long long)blockIdx.x * kT + threadIdx.x;
  // lanes use only their own LDS slots: no barrier; a wave's first lane has a
  // bin whenever any lane does (kT is a multiple of 64), so it writes the word
  if (gid >= nb) return;
  const long long nw = (nb + 63) / 64;
  const int bin = (int)gid;
  const int bx = bin % nbx, by = (bin / nbx) % nby, bz = bin / (nbx * nby);
  {
    int t = 0;
    for (int zi = 0; zi < d.bin_z; ++zi)
      for (int yi = 0; yi < d.bin_y; ++yi)
        for (int xi = 0; xi < d.bin_x; ++xi, ++t) {
          const int x = bx * d.bin_x + xi, y = by * d.bin_y + yi;
          geo[t * kT + threadIdx.x] = thomson_at(d, x, y);
          geo[(nraw + t) * kT + threadIdx.x] = solid_angle_at(d, x, y);
        }
  }
  const size_t npix = (size_t)d.nx * d.ny * d.nz;
  const float dark = d.dark_current * d.exposure_time;
  const int p0 = (int)blockIdx.y * run, p1 = min(npatterns, p0 + run);
  for (int p = p0; p < p1; ++p) {
    const uint32_t pat = (uint32_t)(pattern_base + p);
    const float *Ip = I + (size_t)p * npix;
    float acc = 0.0f, accn = 0.0f;
    int t = 0;
    // bin sums in (z, y, x) order (src/amplification.c:94-102, 139-147)
    for (int zi = 0; zi < d.bin_z; ++zi) {
      for (int yi = 0; yi < d.bin_y; ++yi) {
        for (int xi = 0; xi < d.bin_x; ++xi, ++t) {
          const int x = bx * d.bin_x + xi, y = by * d.bin_y + yi, z = bz * d.bin_z + zi;
          const size_t i = ((size_t)z * d.ny + y) * d.nx + x;
          const RawPixel r = raw_from_geometry(d, geo[t * kT + threadIdx.x], geo[(nraw + t) * kT + threadIdx.x],
                                               (uint32_t)i, Ip[i], seed, pat);
          acc += (r.el + dark);
          accn += r.ph * d.qe * d.e_per_photon_noiseless * d.adc;
          const size_t oi = (size_t)p * npix + i;
          if (o.th) o.th[oi] = r.th;
          if (o.om) o.om[oi] = r.om;
          if (o.ph) o.ph[oi] = r.ph;
          if (o.cnt) o.cnt[oi] = r.cnt;
          if (o.el) o.el[oi] = r.el;
        }
      }
    }
    const size_t ob = (size_t)p * nb + bin;
    if (o.real) o.real[ob] = finish_real(d, acc, seed, (uint32_t)bin, pat);
    if (o.noiseless) o.noiseless[ob] = accn;
  }
}

// Binned outputs only, one count per bin. The pixels of a bin whose mean
// (photons x QE) is below 1000 have independent Poisson counts, and their sum
// is one Poisson variable with the summed mean (FP32 sum in (z, y, x) order):
// it is drawn once per bin from the block (seed, bin, pattern,
// kRngPoissonBin, 0), whose words 0-1 also feed the inversion and words 2-3
// the readout noise (bin_count, box_muller_f32). Pixels on the reference's
// Gaussian branch keep their own draw (src/noise.c:44-46). The real output's
// sum adds count x electrons-per-photon + dark for the Gaussian pixels and
// dark for the others in (z, y, x) order, then the joint count x
// electrons-per-photon and the readout noise; clamp and ADC as
// src/amplification.c:104-108. The noiseless output is the per-pixel sum, bit
// for bit that of the other chains.
//
// Two passes, so that the common pass carries no register-hungry branch:
// chain_bin_draws_kernel finishes every bin whose summed mean is below 10 and
// that has no Gaussian pixel, and marks the others: one bit per bin in a word
// per wave and pattern (marks[p][bin / 64], from a ballot), and a NaN in
// their real output; chain_bin_fixup_kernel reads the words and recomputes the
// marked bins with PTRS and the Gaussian branch. Each lane of the first pass keeps the Thomson factor x
// solid angle of its bin's raw pixels in LDS for a run of patterns; VEC4
// reads a bin row of 4 pixels as one 16-byte load (bin_x = 4, nx a multiple
// of 4, I 16-byte aligned), B44 fixes 4 x 4 x 1 bins at compile time.
constexpr int kBinDrawMaxRaw = 64;  // raw pixels per bin kept in LDS (64 KiB per workgroup)

template <bool VEC4, bool B44, bool GEO_LDS>
__global__ __launch_bounds__(kT) void chain_bin_draws_kernel(DetConst d, const float *__restrict__ I, int npatterns,
                                                             int run, uint64_t seed, uint64_t pattern_base,
                                                             float *__restrict__ real, float *__restrict__ noiseless,
                                                             unsigned long long *__restrict__ marks) {
#pragma clang fp contract(off)
  extern __shared__ float thom[];
  if constexpr (B44) {
    d.bin_x = 4;
    d.bin_y = 4;
    d.bin_z = 1;
  }
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y, nbz = d.nz / d.bin_z;
  const int nraw = d.bin_x * d.bin_y * d.bin_z;
  const long long nb = (long long)nbx * nby * nbz;
  const long long gid = (long long)blockIdx.x * kT + threadIdx.x;
  // lanes use only their own LDS slots: no barrier; a wave's first lane has a
  // bin whenever any lane does (kT is a multiple of 64), so it writes the word
  if (gid >= nb) return;
  const long long nw = (nb + 63) / 64;
  const int bin = (int)gid;
  const int bx = bin % nbx, by = (bin / nbx) % nby, bz = bin / (nbx * nby);
  constexpr bool in_lds = GEO_LDS;  // nraw <= kBinDrawMaxRaw
  const float dark = d.dark_current * d.exposure_time;
  float dark_sum = 0.0f;  // the real output's sum over a bin without Gaussian pixels
  {
    int t = 0;
    for (int zi = 0; zi < d.bin_z; ++zi)
      for (int yi = 0; yi < d.bin_y; ++yi)
        for (int xi = 0; xi < d.bin_x; ++xi, ++t) {
          dark_sum += dark;
          if constexpr (in_lds) {
            const int x = bx * d.bin_x + xi, y = by * d.bin_y + yi;
            thom[t * kT + threadIdx.x] = thomson_at(d, x, y) * solid_angle_at(d, x, y);
          }
        }
  }
  const size_t npix = (size_t)d.nx * d.ny * d.nz;
  const int p0 = (int)blockIdx.y * run, p1 = min(npatterns, p0 + run);
  // Two patterns of the bin per iteration: their loads, sums and Philox blocks
  // are independent, so the two chains interleave.
  for (int p = p0; p < p1; p += 2) {
    const bool two = p + 1 < p1;  // uniform
    const float *IpA = I + (size_t)p * npix;
    const float *IpB = two ? IpA + npix : IpA;
    float accn[2] = {0.0f, 0.0f}, lam[2] = {0.0f, 0.0f}, peak[2] = {0.0f, 0.0f};
    int t = 0;
    for (int zi = 0; zi < d.bin_z; ++zi) {
      for (int yi = 0; yi < d.bin_y; ++yi) {
        const int y = by * d.bin_y + yi, z = bz * d.bin_z + zi;
        const size_t row = ((size_t)z * d.ny + y) * d.nx + (size_t)bx * d.bin_x;
        float v4[2][4];
        if constexpr (VEC4) {
          const float4 qa = *reinterpret_cast<const float4 *>(IpA + row);
          const float4 qb = *reinterpret_cast<const float4 *>(IpB + row);
          v4[0][0] = qa.x;
          v4[0][1] = qa.y;
          v4[0][2] = qa.z;
          v4[0][3] = qa.w;
          v4[1][0] = qb.x;
          v4[1][1] = qb.y;
          v4[1][2] = qb.z;
          v4[1][3] = qb.w;
        }
        for (int xi = 0; xi < d.bin_x; ++xi, ++t) {
          const int x = bx * d.bin_x + xi;
          float g;
          if constexpr (in_lds)
            g = thom[t * kT + threadIdx.x];
          else
            g = thomson_at(d, x, y) * solid_angle_at(d, x, y);
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            const float Ii = VEC4 ? v4[k][xi & 3] : (k ? IpB : IpA)[row + xi];
            const float mean = g * Ii * d.beam_intensity * d.qe;  // (((th om) I) beam) QE
            accn[k] += mean * d.e_per_photon_noiseless * d.adc;
            lam[k] += mean;
            peak[k] = fmaxf(peak[k], mean);
          }
        }
      }
    }
    U4 r[2];
    if (real) {
      // a marked bin's block goes unused (its bin is redrawn by the second pass)
      r[0] = draw(seed, (uint32_t)bin, (uint32_t)(pattern_base + p), kRngPoissonBin, 0u);
      r[1] = draw(seed, (uint32_t)bin, (uint32_t)(pattern_base + p + 1), kRngPoissonBin, 0u);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      if (k == 1 && !two) break;
      const size_t ob = (size_t)(p + k) * nb + bin;
      if (noiseless) noiseless[ob] = accn[k];
      if (!real) continue;
      const bool marked = peak[k] >= 1000.0f || lam[k] >= 10.0f;
      const unsigned long long m = __ballot(marked);
      if (threadIdx.x % 64 == 0) marks[(size_t)(p + k) * nw + gid / 64] = m;
      if (marked) {
        real[ob] = __builtin_nanf("");  // recomputed by chain_bin_fixup_kernel
        continue;
      }
      float acc = dark_sum;
      acc += (float)bin_count_small(lam[k], r[k]) * d.e_per_photon;
      acc += d.readout_noise * box_muller_f32(uniform32(r[k].v[2]), uniform32(r[k].v[3]));
      if (acc < 0) acc = 0;
      acc *= d.adc;
      real[ob] = acc;
    }
  }
}

// Second pass: the bins the first pass marked (summed mean >= 10 or a
// Gaussian pixel), with every branch. One lane per mark word (nw words of 64
// bins per pattern, npatterns x nw in all) recomputes the bins whose bits are
// set.
__device__ __noinline__ void fixup_bin(DetConst d, const float *__restrict__ I, long long nb, long long gid,
                                       uint64_t seed, uint64_t pattern_base, float *__restrict__ real);

__global__ __launch_bounds__(kT) void chain_bin_fixup_kernel(DetConst d, const float *__restrict__ I, long long nb,
                                                             long long nw, long long nwords, uint64_t seed,
                                                             uint64_t pattern_base,
                                                             const unsigned long long *__restrict__ marks,
                                                             float *__restrict__ real) {
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  if (i >= nwords) return;
  unsigned long long m = marks[i];
  const long long p = i / nw, first = (i - p * nw) * 64;
  while (m) {
    const int k = __builtin_ctzll(m);
    m &= m - 1;
    fixup_bin(d, I, nb, p * nb + first + k, seed, pattern_base, real);
  }
}

__device__ __noinline__ void fixup_bin(DetConst d, const float *__restrict__ I, long long nb, long long gid,
                                       uint64_t seed, uint64_t pattern_base, float *__restrict__ real) {
#pragma clang fp contract(off)
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y;
  const int p = (int)(gid / nb), bin = (int)(gid % nb);
  const int bx = bin % nbx, by = (bin / nbx) % nby, bz = bin / (nbx * nby);
  const uint32_t pat = (uint32_t)(pattern_base + p);
  const float *Ip = I + (size_t)p * d.nx * d.ny * d.nz;
  const float dark = d.dark_current * d.exposure_time;
  float acc = 0.0f, lam = 0.0f;
  for (int zi = 0; zi < d.bin_z; ++zi)
    for (int yi = 0; yi < d.bin_y; ++yi)
      for (int xi = 0; xi < d.bin_x; ++xi) {
        const int x = bx * d.bin_x + xi, y = by * d.bin_y + yi, z = bz * d.bin_z + zi;
        const size_t i = ((size_t)z * d.ny + y) * d.nx + x;
        const float mean = thomson_at(d, x, y) * solid_angle_at(d, x, y) * Ip[i] * d.beam_intensity * d.qe;
        if (mean >= 1000.0f) {
          const float cnt = (float)poisson_gaussian((double)mean, seed, (uint32_t)i, pat);
          acc += (cnt * d.e_per_photon + dark);
        } else {
          lam += mean;
          acc += dark;
        }
      }
  const U4 r = draw(seed, (uint32_t)bin, pat, kRngPoissonBin, 0u);
  acc += (float)bin_count(lam, r, seed, (uint32_t)bin, pat) * d.e_per_photon;
  acc += d.readout_noise * box_muller_f32(uniform32(r.v[2]), uniform32(r.v[3]));
  if (acc < 0) acc = 0;
  acc *= d.adc;
  real[gid] = acc;
}

// Raw outputs for pixels outside complete bins (only when nx, ny or nz is not
// a multiple of the binning and raw outputs were requested).
__global__ __launch_bounds__(kT) void chain_leftover_kernel(DetConst d, const float *__restrict__ I, int npatterns,
                                                            uint64_t seed, uint64_t pattern_base, ChainOut o) {
  const size_t npix = (size_t)d.nx * d.ny * d.nz;
  const long long gid = (long long)blockIdx.x * kT + threadIdx.x;
  if (gid >= (long long)npix * npatterns) return;
  const int p = (int)(gid / (long long)npix);
  const size_t i = (size_t)(gid % (long long)npix);
  const int x = (int)(i % d.nx), y = (int)((i / d.nx) % d.ny), z = (int)(i / ((size_t)d.nx * d.ny));
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y, nbz = d.nz / d.bin_z;
  if (x < nbx * d.bin_x && y < nby * d.bin_y && z < nbz * d.bin_z) return;  // covered by the binned kernels
  const RawPixel r = raw_pixel(d, x, y, (uint32_t)i, I[(size_t)p * npix + i], seed, (uint32_t)(pattern_base + p));
  const size_t oi = (size_t)p * npix + i;
  if (o.th) o.th[oi] = r.th;
  if (o.om) o.om[oi] = r.om;
  if (o.ph) o.ph[oi] = r.ph;
  if (o.cnt) o.cnt[oi] = r.cnt;
  if (o.el) o.el[oi] = r.el;
}

// ---- stage kernels (one raw pixel or one binned pixel per lane) ----

__device__ __forceinline__ int raw_count(const DetConst &d) { return d.nx * d.ny * d.nz; }

__global__ __launch_bounds__(kT) void thomson_kernel(DetConst d, float *__restrict__ out) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= raw_count(d)) return;
  out[i] = thomson_at(d, i % d.nx, (i / d.nx) % d.ny);
}

__global__ __launch_bounds__(kT) void solid_angle_kernel(DetConst d, float *__restrict__ out) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= raw_count(d)) return;
  out[i] = solid_angle_at(d, i % d.nx, (i / d.nx) % d.ny);
}

__global__ __launch_bounds__(kT) void photons_kernel(DetConst d, const float *__restrict__ th,
                                                     const float *__restrict__ om, const float *__restrict__ I,
                                                     float *__restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= raw_count(d)) return;
  out[i] = th[i] * om[i] * I[i] * d.beam_intensity;
}

__global__ __launch_bounds__(kT) void poisson_kernel(DetConst d, const float *__restrict__ ph, uint64_t seed,
                                                     uint32_t pattern, float *__restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= raw_count(d)) return;
  const float mean = ph[i] * d.qe;
  out[i] = (float)poisson_sample((double)mean, seed, (uint32_t)i, pattern);
}

__global__ __launch_bounds__(kT) void electrons_kernel(DetConst d, const float *__restrict__ cnt,
                                                       float *__restrict__ out) {
  const int i = blockIdx.x * kT + threadIdx.x;
  if (i >= raw_count(d)) return;
  out[i] = cnt[i] * d.e_per_photon;
}

__global__ __launch_bounds__(kT) void real_output_kernel(DetConst d, const float *__restrict__ el, uint64_t seed,
                                                         uint32_t pattern, float *__restrict__ out) {
#pragma clang fp contract(off)
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y, nbz = d.nz / d.bin_z;
  const int bin = blockIdx.x * kT + threadIdx.x;
  if (bin >= nbx * nby * nbz) return;
  const int bx = bin % nbx, by = (bin / nbx) % nby, bz = bin / (nbx * nby);
  const float dark = d.dark_current * d.exposure_time;
  float acc = 0.0f;
  for (int zi = 0; zi < d.bin_z; ++zi)
    for (int yi = 0; yi < d.bin_y; ++yi)
      for (int xi = 0; xi < d.bin_x; ++xi)
        acc += (el[((bz * d.bin_z + zi) * d.ny + (by * d.bin_y + yi)) * d.nx + (bx * d.bin_x + xi)] + dark);
  out[bin] = finish_real(d, acc, seed, (uint32_t)bin, pattern);
}

__global__ __launch_bounds__(kT) void noiseless_output_kernel(DetConst d, const float *__restrict__ ph,
                                                              float *__restrict__ out) {
#pragma clang fp contract(off)
  const int nbx = d.nx / d.bin_x, nby = d.ny / d.bin_y, nbz = d.nz / d.bin_z;
  const int bin = blockIdx.x * kT + threadIdx.x;
  if (bin >= nbx * nby * nbz) return;
  const int bx = bin % nbx, by = (bin / nbx) % nby, bz = bin / (nbx * nby);
  float acc = 0.0f;
  for (int zi = 0; zi < d.bin_z; ++zi)
    for (int yi = 0; yi < d.bin_y; ++yi)
      for (int xi = 0; xi < d.bin_x; ++xi)
        acc += ph[((bz * d.bin_z + zi) * d.ny + (by * d.bin_y + yi)) * d.nx + (bx * d.bin_x + xi)] * d.qe *
               d.e_per_photon_noiseless * d.adc;
  out[bin] = acc;
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

int make_det_const(const spsim_hip_detector *p, DetConst *d) {
  if (!p || !d) return fail(SPSIM_HIP_ERR_INVALID, "detector is NULL");
  const int nz = p->nz > 0 ? p->nz : 1, bin_z = p->bin_z > 0 ? p->bin_z : 1;
  if (p->nx < 0 || p->ny < 0 || p->nz < 0 || (long long)p->nx * p->ny * nz > (1LL << 28))
    return fail(SPSIM_HIP_ERR_INVALID, "bad detector size %d x %d x %d", p->nx, p->ny, p->nz);
  if (p->bin_x < 1 || p->bin_y < 1 || p->bin_z < 0) return fail(SPSIM_HIP_ERR_INVALID, "binning must be >= 1");
  d->nx = p->nx;
  d->ny = p->ny;
  d->nz = nz;
  d->bin_x = p->bin_x;
  d->bin_y = p->bin_y;
  d->bin_z = bin_z;
  d->width = p->width;
  d->height = p->height;
  d->center_x = p->center_x;
  d->center_y = p->center_y;
  d->distance = p->distance;
  // `det->distance*det->distance` and `pixel_width*pixel_height` are float
  // products in the reference (src/diffraction.c:1119,1162,1172-1182).
  d->dist2 = (double)(p->distance * p->distance);
  d->pixel_area = (double)(p->pixel_width * p->pixel_height);
  d->spherical = p->spherical;
  d->polarization = p->polarization;
  d->beam_intensity = p->beam_intensity;
  d->qe = p->quantum_efficiency;
  // src/amplification.c:56-61 and :121-123
  const double h_c = 1.98644521e-25;
  const double photon_energy = h_c / (double)p->wavelength;
  d->e_per_photon = 1.0f;
  if (p->electron_hole_production_energy != 0.0f)
    d->e_per_photon = (float)(photon_energy / (double)p->electron_hole_production_energy);
  d->e_per_photon_noiseless = (float)(photon_energy / (double)p->electron_hole_production_energy);
  d->dark_current = p->dark_current;
  d->exposure_time = p->exposure_time;
  d->readout_noise = p->readout_noise;
  d->adc = 1.0f;
  if (p->linear_full_well != 0.0f && p->maximum_value != 0.0f) d->adc = p->maximum_value / p->linear_full_well;
  return SPSIM_HIP_OK;
}

int launch_stage(Stage st, const DetConst &d, const float *in0, const float *in1, const float *in2, uint64_t seed,
                 uint32_t pattern, float *out, hipStream_t s) {
  const long long npix = (long long)d.nx * d.ny * d.nz;
  const long long nbin = (long long)(d.nx / d.bin_x) * (d.ny / d.bin_y) * (d.nz / d.bin_z);
  switch (st) {
    case Stage::kThomson:
      if (npix) hipLaunchKernelGGL(thomson_kernel, dim3(blocks_for(npix)), dim3(kT), 0, s, d, out);
      break;
    case Stage::kSolidAngle:
      if (npix) hipLaunchKernelGGL(solid_angle_kernel, dim3(blocks_for(npix)), dim3(kT), 0, s, d, out);
      break;
    case Stage::kPhotons:
      if (npix) hipLaunchKernelGGL(photons_kernel, dim3(blocks_for(npix)), dim3(kT), 0, s, d, in0, in1, in2, out);
      break;
    case Stage::kPoisson:
      if (npix) hipLaunchKernelGGL(poisson_kernel, dim3(blocks_for(npix)), dim3(kT), 0, s, d, in0, seed, pattern, out);
      break;
    case Stage::kElectrons:
      if (npix) hipLaunchKernelGGL(electrons_kernel, dim3(blocks_for(npix)), dim3(kT), 0, s, d, in0, out);
      break;
    case Stage::kRealOutput:
      if (nbin) hipLaunchKernelGGL(real_output_kernel, dim3(blocks_for(nbin)), dim3(kT), 0, s, d, in0, seed, pattern, out);
      break;
    case Stage::kNoiselessOutput:
      if (nbin) hipLaunchKernelGGL(noiseless_output_kernel, dim3(blocks_for(nbin)), dim3(kT), 0, s, d, in0, out);
      break;
  }
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

int launch_chain(const DetConst &d, const float *I, int npatterns, uint64_t seed, uint64_t pattern_base, float *th,
                 float *om, float *ph, float *cnt, float *el, float *real, float *noiseless, hipStream_t s) {
  if (npatterns <= 0) return SPSIM_HIP_OK;
  ChainOut o{th, om, ph, cnt, el, real, noiseless};
  const long long nb = (long long)(d.nx / d.bin_x) * (d.ny / d.bin_y) * (d.nz / d.bin_z);
  const int nraw = d.bin_x * d.bin_y * d.bin_z;
  if (nb > 0 && nraw <= kRunMaxRaw) {
    // runs of patterns per lane: about 2048 workgroups in all
    const long long gx = (nb + kT - 1) / kT;
    long long groups = (2048 + gx - 1) / gx;
    if (groups > npatterns) groups = npatterns;
    const int run = (int)((npatterns + groups - 1) / groups);
    groups = (npatterns + run - 1) / run;
    if (gx > 0x7fffffffLL) return fail(SPSIM_HIP_ERR_INVALID, "too many bins");
    const size_t lds = 2 * (size_t)nraw * kT * sizeof(float);
    hipLaunchKernelGGL(chain_binned_runs_kernel, dim3((unsigned)gx, (unsigned)groups), dim3(kT), lds, s, d, I,
                       npatterns, run, seed, pattern_base, o);
    SPSIM_HIP_TRY(hipGetLastError());
  } else if (nb > 0) {
    hipLaunchKernelGGL(chain_binned_kernel, dim3(blocks_for(nb * npatterns)), dim3(kT), 0, s, d, I, npatterns, seed,
                       pattern_base, o);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  const bool raw = th || om || ph || cnt || el;
  const bool ragged = (d.nx % d.bin_x) != 0 || (d.ny % d.bin_y) != 0 || (d.nz % d.bin_z) != 0;
  const long long npix = (long long)d.nx * d.ny * d.nz;
  if (raw && ragged && npix > 0) {
    hipLaunchKernelGGL(chain_leftover_kernel, dim3(blocks_for(npix * npatterns)), dim3(kT), 0, s, d, I, npatterns, seed,
                       pattern_base, o);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  return SPSIM_HIP_OK;
}

int launch_chain_bin_draws(const DetConst &d, const float *I, int npatterns, uint64_t seed, uint64_t pattern_base,
                           float *real, float *noiseless, hipStream_t s) {
  const long long nb = (long long)(d.nx / d.bin_x) * (d.ny / d.bin_y) * (d.nz / d.bin_z);
  if (npatterns <= 0 || nb <= 0 || (!real && !noiseless)) return SPSIM_HIP_OK;
  const int nraw = d.bin_x * d.bin_y * d.bin_z;
  // runs of patterns per lane: about 2048 workgroups in all
  const long long gx = (nb + kT - 1) / kT;
  if (gx > 0x7fffffffLL) return fail(SPSIM_HIP_ERR_INVALID, "too many bins");
  long long groups = (2048 + gx - 1) / gx;
  if (groups > npatterns) groups = npatterns;
  const int run = (int)((npatterns + groups - 1) / groups);
  groups = (npatterns + run - 1) / run;
  const size_t lds = nraw <= kBinDrawMaxRaw ? (size_t)nraw * kT * sizeof(float) : 0;
  const bool vec4 = d.bin_x == 4 && d.nx % 4 == 0 && ((uintptr_t)I & 15) == 0;
  const bool b44 = vec4 && d.bin_y == 4 && d.bin_z == 1;
  const bool geo_lds = nraw <= kBinDrawMaxRaw;
  const dim3 grid((unsigned)gx, (unsigned)groups);
  // the mark words (real output only): from the library workspace
  const long long nw = (nb + 63) / 64, nwords = nw * npatterns;
  unsigned long long *marks = nullptr;
  if (real) {
    void *ws = nullptr;
    SPSIM_TRY(workspace(sizeof(unsigned long long) * (size_t)nwords, &ws));
    marks = (unsigned long long *)ws;
  }
#define SPSIM_BIN_DRAWS(V, B, G)                                                                                \
  hipLaunchKernelGGL((chain_bin_draws_kernel<V, B, G>), grid, dim3(kT), G ? lds : 0, s, d, I, npatterns, run, \
                     seed, pattern_base, real, noiseless, marks)
  if (b44)
    SPSIM_BIN_DRAWS(true, true, true);
  else if (vec4 && geo_lds)
    SPSIM_BIN_DRAWS(true, false, true);
  else if (vec4)
    SPSIM_BIN_DRAWS(true, false, false);
  else if (geo_lds)
    SPSIM_BIN_DRAWS(false, false, true);
  else
    SPSIM_BIN_DRAWS(false, false, false);
#undef SPSIM_BIN_DRAWS
  SPSIM_HIP_TRY(hipGetLastError());
  if (real) {
    hipLaunchKernelGGL(chain_bin_fixup_kernel, dim3(blocks_for(nwords)), dim3(kT), 0, s, d, I, nb, nw, nwords, seed,
                       pattern_base, marks, real);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  return SPSIM_HIP_OK;
}

}  // namespace spsim

extern "C" int spsim_hip_detector_chain_binned(const spsim_hip_detector *det, const float *d_I, int npatterns,
                                               uint64_t seed, uint64_t pattern_base, float *d_real_output,
                                               float *d_noiseless_output, void *stream) {
  using namespace spsim;
  DetConst d;
  SPSIM_TRY(make_det_const(det, &d));
  if (npatterns < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative npatterns");
  if ((long long)npatterns * d.nx * d.ny * d.nz > (1LL << 31))
    return fail(SPSIM_HIP_ERR_INVALID, "npatterns * npix exceeds 2^31");
  if (npatterns > 0 && (long long)d.nx * d.ny * d.nz > 0 && !d_I)
    return fail(SPSIM_HIP_ERR_INVALID, "d_I is NULL");
  SPSIM_TRY(ensure_device(nullptr));
  return launch_chain_bin_draws(d, d_I, npatterns, seed, pattern_base, d_real_output, d_noiseless_output,
                                (hipStream_t)stream);
}

extern "C" int spsim_hip_detector_chain(const spsim_hip_detector *det, const float *d_I, int npatterns, uint64_t seed,
                                        uint64_t pattern_base, float *d_thomson, float *d_solid_angle, float *d_photons,
                                        float *d_count, float *d_electrons, float *d_real_output,
                                        float *d_noiseless_output, void *stream) {
  using namespace spsim;
  DetConst d;
  SPSIM_TRY(make_det_const(det, &d));
  if (npatterns < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative npatterns");
  if ((long long)npatterns * d.nx * d.ny * d.nz > (1LL << 31))
    return fail(SPSIM_HIP_ERR_INVALID, "npatterns * npix exceeds 2^31");
  if (npatterns > 0 && (long long)d.nx * d.ny * d.nz > 0 && !d_I)
    return fail(SPSIM_HIP_ERR_INVALID, "d_I is NULL");
  SPSIM_TRY(ensure_device(nullptr));
  return launch_chain(d, d_I, npatterns, seed, pattern_base, d_thomson, d_solid_angle, d_photons, d_count, d_electrons,
                      d_real_output, d_noiseless_output, (hipStream_t)stream);
}
