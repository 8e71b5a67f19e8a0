// Structure-factor kernels for gfx950 (MI355X).
//
// F(q) = sum_j f_{Z_j}(|q|) w_j exp(-2 pi i q . r_j), I = |F|^2  (reference
// compute_pattern_on_list, src/diffraction.c:770-853).
//
// Layout of the work:
//  * A workgroup (4 waves, 256 lanes) owns a tile of pixels (per-pixel
//    kernels: K pixels per lane strided by 256) or of mirror quads (grid
//    kernels, below), so global loads and stores of q, F and I stay coalesced.
//  * Atoms arrive pre-grouped by species (one segment per element). For each
//    segment the form factor f_Z(|q|) is evaluated once per pixel, then the
//    segment's atoms are streamed through LDS in chunks of CHUNK records
//    (positions widened to FP64 when staged, and rotated by R^T when the
//    pattern carries an orientation). Every lane reads the same atom record
//    (an LDS broadcast) and applies it to all of its pixels.
//  * Phase: q.r is formed in FP64 (q and r are FP32 inputs, so each product is
//    exact) and reduced to a fraction of a turn with v_fract_f64 before the
//    conversion to FP32; v_sin_f32/v_cos_f32 take their argument in turns.
//  * Accumulation: FP32 partial sums over one LDS chunk, multiplied by f_Z and
//    added into FP64 per-pixel accumulators once per chunk (two-level summation).
//  * Per-atom Debye-Waller factor exp(-B_j s^2) (optional) is folded with the
//    amplitude into one exp2: w_j e^{-B_j s^2} = 2^(log2 w_j - B_j s^2 log2 e).
#include "sf_kernels.hpp"
#include "form_factor.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace spsim {

namespace {

constexpr int kThreads = 256;

struct __align__(16) AtomLds {
  double x, y;
  double z;
  float w;  // amplitude weight, or its log2 when per-atom B is present
  float b;  // per-atom B factor (A^2)
};
static_assert(sizeof(AtomLds) == 32, "LDS atom record must be 32 bytes");

// Loads q for one pixel (zero for pixels past the end) and s^2 = (|q| 1e-10 / 2)^2
// in A^-2. |q| is rotation invariant, so s^2 is the same with or without R.
__device__ __forceinline__ void load_q(const float *__restrict__ hkl, int pix, bool valid, double &qx, double &qy,
                                       double &qz, float &s2) {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (valid) {
    x = hkl[3 * pix + 0];
    y = hkl[3 * pix + 1];
    z = hkl[3 * pix + 2];
  }
  qx = x;
  qy = y;
  qz = z;
  s2 = (float)((qx * qx + qy * qy + qz * qz) * 2.5e-21);
}

// Stages atoms [c0, c0+n) into LDS: positions widened to FP64 and, with R,
// replaced by R^T r; with per-atom B the amplitude is stored as log2 w. Must be
// called by every thread of the workgroup; the first barrier protects the
// previous chunk's readers.
template <bool HAS_B>
__device__ __forceinline__ void stage_chunk(const spsim_hip_atom *__restrict__ atoms, const float *__restrict__ atom_b,
                                            int c0, int n, const double *R, AtomLds *__restrict__ sm) {
  __syncthreads();
  for (int t = threadIdx.x; t < n; t += kThreads) {
    const spsim_hip_atom a = atoms[c0 + t];
    AtomLds r;
    const double x = a.x, y = a.y, z = a.z;
    if (R) {
      r.x = __builtin_fma(R[6], z, __builtin_fma(R[3], y, R[0] * x));
      r.y = __builtin_fma(R[7], z, __builtin_fma(R[4], y, R[1] * x));
      r.z = __builtin_fma(R[8], z, __builtin_fma(R[5], y, R[2] * x));
    } else {
      r.x = x;
      r.y = y;
      r.z = z;
    }
    if constexpr (HAS_B) {
      r.w = log2f(a.w);
      r.b = atom_b[c0 + t];
    } else {
      r.w = a.w;
      r.b = 0.0f;
    }
    sm[t] = r;
  }
  __syncthreads();
}

// Sums the atoms [a_begin, a_end) of the given segments into Fr, Fi for the
// lane's K pixels (Fi accumulates minus the imaginary part). Must be called by
// every thread of the workgroup.
template <int K, int CHUNK, bool HAS_B>
__device__ __forceinline__ void tile_sum(const double (&qx)[K], const double (&qy)[K], const double (&qz)[K],
                                         const float (&s2)[K], const spsim_hip_atom *__restrict__ atoms,
                                         const float *__restrict__ atom_b, const spsim_hip_segment *__restrict__ segs,
                                         int nsegs, int a_begin, int a_end, const float *__restrict__ ff,
                                         float b_global, const double *R, AtomLds *__restrict__ sm, double (&Fr)[K],
                                         double (&Fi)[K]) {
  float nl2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    nl2[k] = -s2[k] * (float)kLog2e;
    Fr[k] = 0.0;
    Fi[k] = 0.0;
  }
  for (int sg = 0; sg < nsegs; ++sg) {
    const int s_start = segs[sg].start;
    const int s_end = s_start + segs[sg].count;
    const int species = segs[sg].species;
    const int lo = max(s_start, a_begin);
    const int hi = min(s_end, a_end);
    if (lo >= hi) continue;
    const float Bs = (species == SPSIM_HIP_UNIT_SPECIES) ? 0.0f : b_global;
    const float *c = ff + species * SPSIM_HIP_FF_STRIDE;
    float f[K];
#pragma unroll
    for (int k = 0; k < K; ++k) f[k] = form_factor(c, s2[k], Bs);

    for (int c0 = lo; c0 < hi; c0 += CHUNK) {
      const int n = min(CHUNK, hi - c0);
      stage_chunk<HAS_B>(atoms, atom_b, c0, n, R, sm);
      float sr[K], si[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        sr[k] = 0.0f;
        si[k] = 0.0f;
      }
      for (int j = 0; j < n; ++j) {
        const AtomLds a = sm[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double t = __builtin_fma(qz[k], a.z, __builtin_fma(qy[k], a.y, qx[k] * a.x));
          const float ph = (float)__builtin_amdgcn_fract(t);
          const float cs = __builtin_amdgcn_cosf(ph);
          const float sn = __builtin_amdgcn_sinf(ph);
          float wt;
          if constexpr (HAS_B) {
            wt = __builtin_amdgcn_exp2f(__builtin_fmaf(a.b, nl2[k], a.w));
          } else {
            wt = a.w;
          }
          sr[k] = __builtin_fmaf(wt, cs, sr[k]);
          si[k] = __builtin_fmaf(wt, sn, si[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        Fr[k] = __builtin_fma((double)f[k], (double)sr[k], Fr[k]);
        Fi[k] = __builtin_fma((double)f[k], (double)si[k], Fi[k]);
      }
    }
  }
}

// ---- Wavelength spectrum ---------------------------------------------------
//
// The reference's spectrum kernel (src/diffraction.cu:416-476) averages the
// amplitude over wavelength samples: each atom contributes
// f(|q|) w_j sum_k W_k e^{-2 pi i s_k q.r_j} with phase scales s_k = lambda/lambda_k
// and normalised weights W_k (the form factor stays at the unscaled |q|). The
// FP64 phase q.r is formed once per atom and pixel. With the centre scale sc,
// t = sc q.r and e_k = s_k / sc - 1, the sample sum factors as
//   sum_k W_k e^{-2 pi i s_k q.r} = e^{-2 pi i t} G(t),
//   G(t) = sum_k W_k e^{-2 pi i e_k t},
// where only the first factor needs the FP64 reduction of t to turns. G is a
// smooth function of t alone (|e_k| is about half the bandwidth), evaluated as
// two FP32 polynomials in (t / tlim)^2 fitted on the host (spectrum_fit), so a
// term costs one sine and cosine whatever the number of samples. Lanes with
// |t| > tlim, and spectra of fewer than 4 samples, take the per-sample sum:
// each s_k q.r formed and reduced to turns in FP64. The weighted sample sum is
// then folded into the atom's amplitude like a per-atom weight.
template <int K, int CHUNK, bool HAS_B>
__device__ __forceinline__ void spectrum_sum(const double (&qx)[K], const double (&qy)[K], const double (&qz)[K],
                                             const float (&s2)[K], const spsim_hip_atom *__restrict__ atoms,
                                             const float *__restrict__ atom_b,
                                             const spsim_hip_segment *__restrict__ segs, int nsegs, int a_begin,
                                             int a_end, const float *__restrict__ ff, float b_global, const double *R,
                                             const Spectrum &sp, AtomLds *__restrict__ sm, double (&Fr)[K],
                                             double (&Fi)[K]) {
  float nl2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    nl2[k] = -s2[k] * (float)kLog2e;
    Fr[k] = 0.0;
    Fi[k] = 0.0;
  }
  const int ns = sp.n;
  for (int sg = 0; sg < nsegs; ++sg) {
    const int s_start = segs[sg].start;
    const int s_end = s_start + segs[sg].count;
    const int species = segs[sg].species;
    const int lo = max(s_start, a_begin);
    const int hi = min(s_end, a_end);
    if (lo >= hi) continue;
    const float Bs = (species == SPSIM_HIP_UNIT_SPECIES) ? 0.0f : b_global;
    const float *c = ff + species * SPSIM_HIP_FF_STRIDE;
    float f[K];
#pragma unroll
    for (int k = 0; k < K; ++k) f[k] = form_factor(c, s2[k], Bs);
    for (int c0 = lo; c0 < hi; c0 += CHUNK) {
      const int n = min(CHUNK, hi - c0);
      stage_chunk<HAS_B>(atoms, atom_b, c0, n, R, sm);
      float sr[K], si[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        sr[k] = 0.0f;
        si[k] = 0.0f;
      }
      for (int j = 0; j < n; ++j) {
        const AtomLds a = sm[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double t = __builtin_fma(qz[k], a.z, __builtin_fma(qy[k], a.y, qx[k] * a.x));
          const double tc = sp.sc * t;
          const float tf = (float)tc;
          float ur = 0.0f, ui = 0.0f;
          if (__builtin_fabsf(tf) <= sp.tlim) {
            const float ph = (float)__builtin_amdgcn_fract(tc);
            const float cs = __builtin_amdgcn_cosf(ph);
            const float sn = __builtin_amdgcn_sinf(ph);
            const float v = tf * sp.inv_tlim;
            const float u = v * v;
            float gc = sp.p[kSpectrumDegree], gq = sp.q[kSpectrumDegree];
#pragma unroll
            for (int d = kSpectrumDegree - 1; d >= 0; --d) {
              gc = __builtin_fmaf(gc, u, sp.p[d]);
              gq = __builtin_fmaf(gq, u, sp.q[d]);
            }
            const float gs = v * gq;
            ur = __builtin_fmaf(cs, gc, -sn * gs);
            ui = __builtin_fmaf(sn, gc, cs * gs);
          } else {
            for (int m = 0; m < ns; ++m) {
              const float ph = (float)__builtin_amdgcn_fract(sp.s[m] * t);
              ur = __builtin_fmaf(sp.w[m], __builtin_amdgcn_cosf(ph), ur);
              ui = __builtin_fmaf(sp.w[m], __builtin_amdgcn_sinf(ph), ui);
            }
          }
          float wt;
          if constexpr (HAS_B) {
            wt = __builtin_amdgcn_exp2f(__builtin_fmaf(a.b, nl2[k], a.w));
          } else {
            wt = a.w;
          }
          sr[k] = __builtin_fmaf(wt, ur, sr[k]);
          si[k] = __builtin_fmaf(wt, ui, si[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        Fr[k] = __builtin_fma((double)f[k], (double)sr[k], Fr[k]);
        Fi[k] = __builtin_fma((double)f[k], (double)si[k], Fi[k]);
      }
    }
  }
}

// ---- Mirror-symmetric detectors ------------------------------------------
//
// A detector centred on the beam has an exactly mirror-symmetric q list: the
// pixel mirrored in x has q = (-qx, qy, qz), the one mirrored in y (qx, -qy, qz)
// (the geometry negates exactly and rounds identically). One lane then owns the
// four mirror pixels of a "quad" and, per atom, forms
//   a = qx rx, b = qy ry, c = qz rz   (FP64, exact products),
// the phases c, a + b and a - b reduced to turns, and their sines and cosines
// (6 transcendentals for 4 terms instead of 8). With u = w e^{-2 pi i c}:
//   pixel (x, y):   sum u e^{-2 pi i (a+b)},   pixel (x', y'): sum u e^{+2 pi i (a+b)}
//   pixel (x, y'):  sum u e^{-2 pi i (a-b)},   pixel (x', y):  sum u e^{+2 pi i (a-b)}
// which are assembled from the four real sums P = sum u cos, Q = sum u sin per
// angle. |q| is equal across the quad, so the form factor and the per-atom
// Debye-Waller factor are shared too. Each workgroup checks the symmetry of its
// quads bit for bit and falls back to the per-pixel sum if any quad is not
// symmetric, so results never depend on the assumption. An orientation R is
// applied to the atoms (R^T r), which keeps the q list symmetric.

// Quad index -> the four mirror pixels; the representative (x, y) lies in the
// lower half of each axis. Middle rows/columns of odd sizes repeat a pixel.
__device__ __forceinline__ void quad_pixels(int qidx, int nx, int ny, int &p0, int &p1, int &p2, int &p3) {
  const int hx = (nx + 1) / 2;
  const int x = qidx % hx, y = qidx / hx;
  const int xm = nx - 1 - x, ym = ny - 1 - y;
  p0 = y * nx + x;
  p1 = y * nx + xm;
  p2 = ym * nx + x;
  p3 = ym * nx + xm;
}

__device__ __forceinline__ bool mirrored(const float *__restrict__ hkl, int p0, int p, float sx, float sy) {
  return hkl[3 * p + 0] == sx * hkl[3 * p0 + 0] && hkl[3 * p + 1] == sy * hkl[3 * p0 + 1] &&
         hkl[3 * p + 2] == hkl[3 * p0 + 2];
}

template <int K, int CHUNK, bool HAS_B>
__device__ __forceinline__ void quad_sum(const double (&qx)[K], const double (&qy)[K], const double (&qz)[K],
                                         const float (&s2)[K], const spsim_hip_atom *__restrict__ atoms,
                                         const float *__restrict__ atom_b, const spsim_hip_segment *__restrict__ segs,
                                         int nsegs, int a_begin, int a_end, const float *__restrict__ ff,
                                         float b_global, const double *R, AtomLds *__restrict__ sm,
                                         double (&Fr)[4 * K], double (&Fi)[4 * K]) {
  float nl2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) nl2[k] = -s2[k] * (float)kLog2e;
#pragma unroll
  for (int i = 0; i < 4 * K; ++i) {
    Fr[i] = 0.0;
    Fi[i] = 0.0;
  }
  for (int sg = 0; sg < nsegs; ++sg) {
    const int s_start = segs[sg].start;
    const int s_end = s_start + segs[sg].count;
    const int species = segs[sg].species;
    const int lo = max(s_start, a_begin);
    const int hi = min(s_end, a_end);
    if (lo >= hi) continue;
    const float Bs = (species == SPSIM_HIP_UNIT_SPECIES) ? 0.0f : b_global;
    const float *c = ff + species * SPSIM_HIP_FF_STRIDE;
    float f[K];
#pragma unroll
    for (int k = 0; k < K; ++k) f[k] = form_factor(c, s2[k], Bs);
    for (int c0 = lo; c0 < hi; c0 += CHUNK) {
      const int n = min(CHUNK, hi - c0);
      stage_chunk<HAS_B>(atoms, atom_b, c0, n, R, sm);
      float pr[K], pi[K], qr[K], qi[K], p2r[K], p2i[K], q2r[K], q2i[K];
#pragma unroll
      for (int k = 0; k < K; ++k) pr[k] = pi[k] = qr[k] = qi[k] = p2r[k] = p2i[k] = q2r[k] = q2i[k] = 0.0f;
      for (int j = 0; j < n; ++j) {
        const AtomLds at = sm[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double ta = qx[k] * at.x;
          const double tb = qy[k] * at.y;
          const double tc = qz[k] * at.z;
          const float fc = (float)__builtin_amdgcn_fract(tc);
          const float fs = (float)__builtin_amdgcn_fract(ta + tb);
          const float fd = (float)__builtin_amdgcn_fract(ta - tb);
          float w;
          if constexpr (HAS_B) {
            w = __builtin_amdgcn_exp2f(__builtin_fmaf(at.b, nl2[k], at.w));
          } else {
            w = at.w;
          }
          const float ur = w * __builtin_amdgcn_cosf(fc);
          const float ui = w * __builtin_amdgcn_sinf(fc);
          const float cs = __builtin_amdgcn_cosf(fs), ss = __builtin_amdgcn_sinf(fs);
          const float cd = __builtin_amdgcn_cosf(fd), sd = __builtin_amdgcn_sinf(fd);
          pr[k] = __builtin_fmaf(ur, cs, pr[k]);
          pi[k] = __builtin_fmaf(ui, cs, pi[k]);
          qr[k] = __builtin_fmaf(ur, ss, qr[k]);
          qi[k] = __builtin_fmaf(ui, ss, qi[k]);
          p2r[k] = __builtin_fmaf(ur, cd, p2r[k]);
          p2i[k] = __builtin_fmaf(ui, cd, p2i[k]);
          q2r[k] = __builtin_fmaf(ur, sd, q2r[k]);
          q2i[k] = __builtin_fmaf(ui, sd, q2i[k]);
        }
      }
      // Fi accumulates minus the imaginary part, as in tile_sum.
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double fk = f[k];
        Fr[4 * k + 0] = __builtin_fma(fk, (double)pr[k] - (double)qi[k], Fr[4 * k + 0]);  // (x, y): c + (a+b)
        Fi[4 * k + 0] = __builtin_fma(fk, (double)pi[k] + (double)qr[k], Fi[4 * k + 0]);
        Fr[4 * k + 1] = __builtin_fma(fk, (double)p2r[k] + (double)q2i[k], Fr[4 * k + 1]);  // (x', y): c - (a-b)
        Fi[4 * k + 1] = __builtin_fma(fk, (double)p2i[k] - (double)q2r[k], Fi[4 * k + 1]);
        Fr[4 * k + 2] = __builtin_fma(fk, (double)p2r[k] - (double)q2i[k], Fr[4 * k + 2]);  // (x, y'): c + (a-b)
        Fi[4 * k + 2] = __builtin_fma(fk, (double)p2i[k] + (double)q2r[k], Fi[4 * k + 2]);
        Fr[4 * k + 3] = __builtin_fma(fk, (double)pr[k] + (double)qi[k], Fr[4 * k + 3]);  // (x', y'): c - (a+b)
        Fi[4 * k + 3] = __builtin_fma(fk, (double)pi[k] - (double)qr[k], Fi[4 * k + 3]);
      }
    }
  }
}

// Writes one pixel's result: a split partial (F, exact sign) or F and I.
__device__ __forceinline__ void store_pixel(int p, double Fr, double Fi_neg, int npix, int split, double2 *partial,
                                            float2 *F, float *I) {
  if (partial) {
    partial[(size_t)split * npix + p] = make_double2(Fr, -Fi_neg);
  } else {
    if (F) F[p] = make_float2((float)Fr, (float)(-Fi_neg));
    if (I) I[p] = (float)(Fr * Fr + Fi_neg * Fi_neg);
  }
}

// A tile of K quads per lane of an nx x ny detector list: per-quad symmetry
// check (workgroup vote), then the quad sum or the per-pixel fallback.
template <int K, int CHUNK, bool HAS_B>
__device__ __forceinline__ void grid_tile(const float *__restrict__ hkl, int nx, int ny, int quad_base,
                                          const spsim_hip_atom *__restrict__ atoms, const float *__restrict__ atom_b,
                                          const spsim_hip_segment *__restrict__ segs, int nsegs, int a_begin,
                                          int a_end, const float *__restrict__ ff, float b_global, const double *R,
                                          AtomLds *__restrict__ sm, int (&pix)[4 * K], double (&Fr)[4 * K],
                                          double (&Fi)[4 * K]) {
  const int nquads = grid_quads(nx, ny);
  bool sym = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int qd = quad_base + threadIdx.x + k * kThreads;
    if (qd < nquads) {
      quad_pixels(qd, nx, ny, pix[4 * k], pix[4 * k + 1], pix[4 * k + 2], pix[4 * k + 3]);
      const int p0 = pix[4 * k];
      sym = sym && mirrored(hkl, p0, pix[4 * k + 1], -1.0f, 1.0f) && mirrored(hkl, p0, pix[4 * k + 2], 1.0f, -1.0f) &&
            mirrored(hkl, p0, pix[4 * k + 3], -1.0f, -1.0f);
    } else {
      pix[4 * k] = pix[4 * k + 1] = pix[4 * k + 2] = pix[4 * k + 3] = -1;
    }
  }
  const bool fallback = __syncthreads_or(!sym);
  if (!fallback) {
    double qx[K], qy[K], qz[K];
    float s2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) load_q(hkl, pix[4 * k], pix[4 * k] >= 0, qx[k], qy[k], qz[k], s2[k]);
    quad_sum<K, CHUNK, HAS_B>(qx, qy, qz, s2, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm, Fr, Fi);
  } else {
    double qx[4 * K], qy[4 * K], qz[4 * K];
    float s2[4 * K];
#pragma unroll
    for (int i = 0; i < 4 * K; ++i) load_q(hkl, pix[i], pix[i] >= 0, qx[i], qy[i], qz[i], s2[i]);
    tile_sum<4 * K, CHUNK, HAS_B>(qx, qy, qz, s2, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm, Fr,
                                  Fi);
  }
}

// ---- Square mirror-symmetric detectors: octants -----------------------------
//
// On a square detector centred on the beam with square pixels, the transposed
// pixel (row and column swapped) holds (qy, qx, qz) bit for bit: the
// geometry forms px and py by the same expression, and |p| adds the same
// squares. One lane then owns an octant, the quad of (x, y) and the quad of
// its transpose (y, x), 8 pixels that share c = qz rz, the form factor and the
// Debye-Waller factor. With a' = qy rx, b' = qx ry for the transposed quad,
// the phases c, a + b, a - b, a' + b', a' - b' need 10 sines and cosines for 8
// terms (12 with two separate quads). Octants cover x <= y < (n + 1) / 2; on
// the diagonal both quads are the same pixels and store the same values.
// Workgroups whose octants are not symmetric fall back to two independent
// quads, and without mirror symmetry to the per-pixel sum.

// Octant index t -> (x, y), x <= y, t = y (y + 1) / 2 + x.
__device__ __forceinline__ void octant_xy(int t, int &x, int &y) {
  int yy = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((yy + 1) * (yy + 2) / 2 <= t) ++yy;
  while (yy * (yy + 1) / 2 > t) --yy;
  y = yy;
  x = t - yy * (yy + 1) / 2;
}

// pix[0..3]: quad of (x, y) in quad_pixels order; pix[4..7]: quad of (y, x).
__device__ __forceinline__ void octant_pixels(int t, int n, int (&pix)[8]) {
  int x, y;
  octant_xy(t, x, y);
  const int xm = n - 1 - x, ym = n - 1 - y;
  pix[0] = y * n + x;
  pix[1] = y * n + xm;
  pix[2] = ym * n + x;
  pix[3] = ym * n + xm;
  pix[4] = x * n + y;
  pix[5] = x * n + ym;
  pix[6] = xm * n + y;
  pix[7] = xm * n + ym;
}

// q at pixel p equals (sx * q0.y, sy * q0.x, q0.z) bit for bit.
__device__ __forceinline__ bool transposed(const float *__restrict__ hkl, int p0, int p, float sx, float sy) {
  return hkl[3 * p + 0] == sx * hkl[3 * p0 + 1] && hkl[3 * p + 1] == sy * hkl[3 * p0 + 0] &&
         hkl[3 * p + 2] == hkl[3 * p0 + 2];
}

template <int CHUNK, bool HAS_B>
__device__ __forceinline__ void octant_sum(double qx, double qy, double qz, float s2,
                                           const spsim_hip_atom *__restrict__ atoms, const float *__restrict__ atom_b,
                                           const spsim_hip_segment *__restrict__ segs, int nsegs, int a_begin,
                                           int a_end, const float *__restrict__ ff, float b_global, const double *R,
                                           AtomLds *__restrict__ sm, double (&Fr)[8], double (&Fi)[8]) {
  const float nl2 = -s2 * (float)kLog2e;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    Fr[i] = 0.0;
    Fi[i] = 0.0;
  }
  for (int sg = 0; sg < nsegs; ++sg) {
    const int s_start = segs[sg].start;
    const int s_end = s_start + segs[sg].count;
    const int species = segs[sg].species;
    const int lo = max(s_start, a_begin);
    const int hi = min(s_end, a_end);
    if (lo >= hi) continue;
    const float Bs = (species == SPSIM_HIP_UNIT_SPECIES) ? 0.0f : b_global;
    const float f = form_factor(ff + species * SPSIM_HIP_FF_STRIDE, s2, Bs);
    for (int c0 = lo; c0 < hi; c0 += CHUNK) {
      const int n = min(CHUNK, hi - c0);
      stage_chunk<HAS_B>(atoms, atom_b, c0, n, R, sm);
      // [0]: quad (x, y), [1]: transposed quad; P/Q over a+b, P2/Q2 over a-b
      float pr[2] = {0.f, 0.f}, pi[2] = {0.f, 0.f}, qr[2] = {0.f, 0.f}, qi[2] = {0.f, 0.f};
      float p2r[2] = {0.f, 0.f}, p2i[2] = {0.f, 0.f}, q2r[2] = {0.f, 0.f}, q2i[2] = {0.f, 0.f};
      for (int j = 0; j < n; ++j) {
        const AtomLds at = sm[j];
        const double tc = qz * at.z;
        const double ta[2] = {qx * at.x, qy * at.x};
        const double tb[2] = {qy * at.y, qx * at.y};
        const float fc = (float)__builtin_amdgcn_fract(tc);
        float w;
        if constexpr (HAS_B) {
          w = __builtin_amdgcn_exp2f(__builtin_fmaf(at.b, nl2, at.w));
        } else {
          w = at.w;
        }
        const float ur = w * __builtin_amdgcn_cosf(fc);
        const float ui = w * __builtin_amdgcn_sinf(fc);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float fs = (float)__builtin_amdgcn_fract(ta[h] + tb[h]);
          const float fd = (float)__builtin_amdgcn_fract(ta[h] - tb[h]);
          const float cs = __builtin_amdgcn_cosf(fs), ss = __builtin_amdgcn_sinf(fs);
          const float cd = __builtin_amdgcn_cosf(fd), sd = __builtin_amdgcn_sinf(fd);
          pr[h] = __builtin_fmaf(ur, cs, pr[h]);
          pi[h] = __builtin_fmaf(ui, cs, pi[h]);
          qr[h] = __builtin_fmaf(ur, ss, qr[h]);
          qi[h] = __builtin_fmaf(ui, ss, qi[h]);
          p2r[h] = __builtin_fmaf(ur, cd, p2r[h]);
          p2i[h] = __builtin_fmaf(ui, cd, p2i[h]);
          q2r[h] = __builtin_fmaf(ur, sd, q2r[h]);
          q2i[h] = __builtin_fmaf(ui, sd, q2i[h]);
        }
      }
      const double fk = f;
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // same assembly as quad_sum
        Fr[4 * h + 0] = __builtin_fma(fk, (double)pr[h] - (double)qi[h], Fr[4 * h + 0]);
        Fi[4 * h + 0] = __builtin_fma(fk, (double)pi[h] + (double)qr[h], Fi[4 * h + 0]);
        Fr[4 * h + 1] = __builtin_fma(fk, (double)p2r[h] + (double)q2i[h], Fr[4 * h + 1]);
        Fi[4 * h + 1] = __builtin_fma(fk, (double)p2i[h] - (double)q2r[h], Fi[4 * h + 1]);
        Fr[4 * h + 2] = __builtin_fma(fk, (double)p2r[h] - (double)q2i[h], Fr[4 * h + 2]);
        Fi[4 * h + 2] = __builtin_fma(fk, (double)p2i[h] + (double)q2r[h], Fi[4 * h + 2]);
        Fr[4 * h + 3] = __builtin_fma(fk, (double)pr[h] + (double)qi[h], Fr[4 * h + 3]);
        Fi[4 * h + 3] = __builtin_fma(fk, (double)pi[h] - (double)qr[h], Fi[4 * h + 3]);
      }
    }
  }
}

// One quad of an octant (pixels pix[first .. first + 3]) by the quad sum, or
// pixel by pixel when the list is not mirror-symmetric.
template <int CHUNK, bool HAS_B>
__device__ __forceinline__ void octant_half(const float *__restrict__ hkl, const int (&pix)[8], int first, bool mirror,
                                            const spsim_hip_atom *__restrict__ atoms, const float *__restrict__ atom_b,
                                            const spsim_hip_segment *__restrict__ segs, int nsegs, int a_begin,
                                            int a_end, const float *__restrict__ ff, float b_global, const double *R,
                                            AtomLds *__restrict__ sm, double (&Fr)[4], double (&Fi)[4]) {
  if (mirror) {
    double qx[1], qy[1], qz[1];
    float s2[1];
    load_q(hkl, pix[first], pix[first] >= 0, qx[0], qy[0], qz[0], s2[0]);
    quad_sum<1, CHUNK, HAS_B>(qx, qy, qz, s2, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm, Fr, Fi);
  } else {
    double qx[4], qy[4], qz[4];
    float s2[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) load_q(hkl, pix[first + i], pix[first + i] >= 0, qx[i], qy[i], qz[i], s2[i]);
    tile_sum<4, CHUNK, HAS_B>(qx, qy, qz, s2, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm, Fr, Fi);
  }
}

// One octant per lane of an n x n detector list: workgroup votes on mirror and
// transpose symmetry, then the octant sum, two quads, or eight pixels.
template <int CHUNK, bool HAS_B>
__device__ __forceinline__ void octant_tile(const float *__restrict__ hkl, int n, int oct_base,
                                            const spsim_hip_atom *__restrict__ atoms, const float *__restrict__ atom_b,
                                            const spsim_hip_segment *__restrict__ segs, int nsegs, int a_begin,
                                            int a_end, const float *__restrict__ ff, float b_global, const double *R,
                                            AtomLds *__restrict__ sm, int (&pix)[8], double (&Fr)[8],
                                            double (&Fi)[8]) {
  const int t = oct_base + threadIdx.x;
  bool mir = true, tra = true;
  if (t < grid_octants(n)) {
    octant_pixels(t, n, pix);
    const int p0 = pix[0], t0 = pix[4];
    mir = mirrored(hkl, p0, pix[1], -1.0f, 1.0f) && mirrored(hkl, p0, pix[2], 1.0f, -1.0f) &&
          mirrored(hkl, p0, pix[3], -1.0f, -1.0f) && mirrored(hkl, t0, pix[5], -1.0f, 1.0f) &&
          mirrored(hkl, t0, pix[6], 1.0f, -1.0f) && mirrored(hkl, t0, pix[7], -1.0f, -1.0f);
    tra = transposed(hkl, p0, t0, 1.0f, 1.0f);
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) pix[i] = -1;
  }
  const bool no_mirror = __syncthreads_or(!mir);
  const bool no_transpose = __syncthreads_or(!tra);
  if (!no_mirror && !no_transpose) {
    double qx, qy, qz;
    float s2;
    load_q(hkl, pix[0], pix[0] >= 0, qx, qy, qz, s2);
    octant_sum<CHUNK, HAS_B>(qx, qy, qz, s2, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm, Fr, Fi);
  } else {
    // Fallbacks run the two quads one after the other, so the kernel's
    // register budget is the octant path's.
    double hr[4], hi[4];
    octant_half<CHUNK, HAS_B>(hkl, pix, 0, !no_mirror, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm,
                              hr, hi);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Fr[i] = hr[i];
      Fi[i] = hi[i];
    }
    octant_half<CHUNK, HAS_B>(hkl, pix, 4, !no_mirror, atoms, atom_b, segs, nsegs, a_begin, a_end, ff, b_global, R, sm,
                              hr, hi);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      Fr[4 + i] = hr[i];
      Fi[4 + i] = hi[i];
    }
  }
}

// ---- single-pattern kernels (blockIdx.x = tile, blockIdx.y = atom split) ----

template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_pattern_kernel(PatternLaunch a, double2 *__restrict__ partial) {
  __shared__ AtomLds sm[CHUNK];
  const int split = blockIdx.y, nsplit = gridDim.y;
  const int a_begin = (int)(((long long)a.natoms * split) / nsplit);
  const int a_end = (int)(((long long)a.natoms * (split + 1)) / nsplit);
  const int base = blockIdx.x * (kThreads * K) + threadIdx.x;
  double qx[K], qy[K], qz[K];
  float s2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int pix = base + k * kThreads;
    load_q(a.hkl, pix, pix < a.npix, qx[k], qy[k], qz[k], s2[k]);
  }
  double Fr[K], Fi[K];
  tile_sum<K, CHUNK, HAS_B>(qx, qy, qz, s2, a.atoms, a.atom_b, a.segs, a.nsegs, a_begin, a_end, a.ff, a.b_global,
                            a.has_rot ? a.rot.r : nullptr, sm, Fr, Fi);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int pix = base + k * kThreads;
    if (pix < a.npix) store_pixel(pix, Fr[k], Fi[k], a.npix, split, partial, a.F, a.I);
  }
}

template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_grid_kernel(PatternLaunch a, int nx, int ny,
                                                           double2 *__restrict__ partial) {
  __shared__ AtomLds sm[CHUNK];
  const int split = blockIdx.y, nsplit = gridDim.y;
  const int a_begin = (int)(((long long)a.natoms * split) / nsplit);
  const int a_end = (int)(((long long)a.natoms * (split + 1)) / nsplit);
  int pix[4 * K];
  double Fr[4 * K], Fi[4 * K];
  grid_tile<K, CHUNK, HAS_B>(a.hkl, nx, ny, blockIdx.x * (kThreads * K), a.atoms, a.atom_b, a.segs, a.nsegs, a_begin,
                             a_end, a.ff, a.b_global, a.has_rot ? a.rot.r : nullptr, sm, pix, Fr, Fi);
#pragma unroll
  for (int i = 0; i < 4 * K; ++i)
    if (pix[i] >= 0) store_pixel(pix[i], Fr[i], Fi[i], a.npix, split, partial, a.F, a.I);
}

template <int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_octant_kernel(PatternLaunch a, int n, int ny_unused,
                                                             double2 *__restrict__ partial) {
  __shared__ AtomLds sm[CHUNK];
  const int split = blockIdx.y, nsplit = gridDim.y;
  const int a_begin = (int)(((long long)a.natoms * split) / nsplit);
  const int a_end = (int)(((long long)a.natoms * (split + 1)) / nsplit);
  int pix[8];
  double Fr[8], Fi[8];
  octant_tile<CHUNK, HAS_B>(a.hkl, n, blockIdx.x * kThreads, a.atoms, a.atom_b, a.segs, a.nsegs, a_begin, a_end, a.ff,
                            a.b_global, a.has_rot ? a.rot.r : nullptr, sm, pix, Fr, Fi);
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (pix[i] >= 0) store_pixel(pix[i], Fr[i], Fi[i], a.npix, split, partial, a.F, a.I);
}

template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_spectrum_kernel(PatternLaunch a, Spectrum sp,
                                                               double2 *__restrict__ partial) {
  __shared__ AtomLds sm[CHUNK];
  const int split = blockIdx.y, nsplit = gridDim.y;
  const int a_begin = (int)(((long long)a.natoms * split) / nsplit);
  const int a_end = (int)(((long long)a.natoms * (split + 1)) / nsplit);
  const int base = blockIdx.x * (kThreads * K) + threadIdx.x;
  double qx[K], qy[K], qz[K];
  float s2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int pix = base + k * kThreads;
    load_q(a.hkl, pix, pix < a.npix, qx[k], qy[k], qz[k], s2[k]);
  }
  double Fr[K], Fi[K];
  spectrum_sum<K, CHUNK, HAS_B>(qx, qy, qz, s2, a.atoms, a.atom_b, a.segs, a.nsegs, a_begin, a_end, a.ff, a.b_global,
                                a.has_rot ? a.rot.r : nullptr, sp, sm, Fr, Fi);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int pix = base + k * kThreads;
    if (pix < a.npix) store_pixel(pix, Fr[k], Fi[k], a.npix, split, partial, a.F, a.I);
  }
}

// Deterministic reduction of the atom splits of one pattern, in split order.
__global__ __launch_bounds__(kThreads) void sf_reduce_kernel(const double2 *__restrict__ partial, int nsplit, int npix,
                                                             float2 *__restrict__ F, float *__restrict__ I) {
  const int pix = blockIdx.x * kThreads + threadIdx.x;
  if (pix >= npix) return;
  double re = 0.0, im = 0.0;
  for (int s = 0; s < nsplit; ++s) {
    const double2 v = partial[(size_t)s * npix + pix];
    re += v.x;
    im += v.y;
  }
  if (F) F[pix] = make_float2((float)re, (float)im);
  if (I) I[pix] = (float)(re * re + im * im);
}

__global__ void reset_queue_kernel(unsigned *q) { *q = 0u; }

// ---- persistent batched kernels ----
//
// Workgroups take the next (pattern, tile, atom range) item from the queue
// until the list is exhausted. The item index is uniform across the
// workgroup, so every wave leaves the loop together. Each pattern's
// orientation R is applied to the atoms while they are staged.
//
// A tile whose pattern has its atoms split over several items (nchunks > 1)
// leaves one FP64 partial slot per item; the workgroup that finishes the
// tile's last item (counted with one atomic per tile) adds the slots in chunk
// order and writes F and I, so the result does not depend on which workgroup
// finished last. It then sets the tile's counter back to zero for the next
// run, so replays need no clearing of the counters.

__device__ __forceinline__ int next_item(unsigned *queue, int *s_item) {
  __syncthreads();
  if (threadIdx.x == 0) *s_item = (int)atomicAdd(queue, 1u);
  __syncthreads();
  return *s_item;
}

// Writes the NP pixels of each lane (pix[i] < 0: none) of one work item:
// directly, or through the item's partial slot and the tile's last-finisher
// reduction. Must be called by every thread of the workgroup.
template <int NP>
__device__ __forceinline__ void finish_item(const BatchLaunch &a, const BatchItem &it, const int (&pix)[NP],
                                            const double (&Fr)[NP], const double (&Fi)[NP], int *s_last) {
  const size_t off = (size_t)it.pattern * a.npix;
  float2 *F = a.F ? a.F + off : nullptr;
  float *I = a.I + off;
  if (it.nchunks <= 1) {
#pragma unroll
    for (int i = 0; i < NP; ++i)
      if (pix[i] >= 0) store_pixel(pix[i], Fr[i], Fi[i], a.npix, 0, nullptr, F, I);
    return;
  }
  constexpr int kSlot = NP * kThreads;  // double2 entries per slot, [i][lane]
  double2 *slots = a.partials + (size_t)it.slot * kSlot;
  double2 *mine = slots + (size_t)it.chunk * kSlot;
#pragma unroll
  for (int i = 0; i < NP; ++i) mine[i * kThreads + threadIdx.x] = make_double2(Fr[i], Fi[i]);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) *s_last = atomicAdd(a.counters + it.group, 1u) == (unsigned)(it.nchunks - 1);
  __syncthreads();
  if (!*s_last) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  double re[NP], im[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) re[i] = im[i] = 0.0;
  for (int c = 0; c < it.nchunks; ++c) {
    const double2 *sc = slots + (size_t)c * kSlot;
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const double2 v = sc[i * kThreads + threadIdx.x];
      re[i] += v.x;
      im[i] += v.y;
    }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i)
    if (pix[i] >= 0) store_pixel(pix[i], re[i], im[i], a.npix, 0, nullptr, F, I);
  if (threadIdx.x == 0) a.counters[it.group] = 0u;
}

template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_batch_kernel(BatchLaunch a) {
  __shared__ AtomLds sm[CHUNK];
  __shared__ int s_item, s_last;
  for (;;) {
    const int item = next_item(a.queue, &s_item);
    if (item >= a.nitems) break;
    const BatchItem it = a.items[item];
    const MolInfo m = a.mols[a.pattern_mol[it.pattern]];
    const int base = it.tile * (kThreads * K) + threadIdx.x;
    double qx[K], qy[K], qz[K];
    float s2[K];
    int pix[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int p = base + k * kThreads;
      pix[k] = p < a.npix ? p : -1;
      load_q(a.hkl, p, p < a.npix, qx[k], qy[k], qz[k], s2[k]);
    }
    double Fr[K], Fi[K];
    // segment starts are relative to the molecule's first packed atom
    tile_sum<K, CHUNK, HAS_B>(qx, qy, qz, s2, a.atoms + m.atom_begin, HAS_B ? a.atom_b + m.atom_begin : nullptr,
                              a.segs + m.seg_begin, m.nseg, it.a_begin, it.a_end, a.ff, a.b_global,
                              a.rot + 9 * (size_t)it.pattern, sm, Fr, Fi);
    finish_item<K>(a, it, pix, Fr, Fi, &s_last);
  }
}

template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_batch_grid_kernel(BatchLaunch a) {
  __shared__ AtomLds sm[CHUNK];
  __shared__ int s_item, s_last;
  for (;;) {
    const int item = next_item(a.queue, &s_item);
    if (item >= a.nitems) break;
    const BatchItem it = a.items[item];
    const MolInfo m = a.mols[a.pattern_mol[it.pattern]];
    int pix[4 * K];
    double Fr[4 * K], Fi[4 * K];
    grid_tile<K, CHUNK, HAS_B>(a.hkl, a.nx, a.ny, it.tile * (kThreads * K), a.atoms + m.atom_begin,
                               HAS_B ? a.atom_b + m.atom_begin : nullptr, a.segs + m.seg_begin, m.nseg, it.a_begin,
                               it.a_end, a.ff, a.b_global, a.rot + 9 * (size_t)it.pattern, sm, pix, Fr, Fi);
    finish_item<4 * K>(a, it, pix, Fr, Fi, &s_last);
  }
}

template <int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_batch_octant_kernel(BatchLaunch a) {
  __shared__ AtomLds sm[CHUNK];
  __shared__ int s_item, s_last;
  for (;;) {
    const int item = next_item(a.queue, &s_item);
    if (item >= a.nitems) break;
    const BatchItem it = a.items[item];
    const MolInfo m = a.mols[a.pattern_mol[it.pattern]];
    int pix[8];
    double Fr[8], Fi[8];
    octant_tile<CHUNK, HAS_B>(a.hkl, a.nx, it.tile * kThreads, a.atoms + m.atom_begin,
                              HAS_B ? a.atom_b + m.atom_begin : nullptr, a.segs + m.seg_begin, m.nseg, it.a_begin,
                              it.a_end, a.ff, a.b_global, a.rot + 9 * (size_t)it.pattern, sm, pix, Fr, Fi);
    finish_item<8>(a, it, pix, Fr, Fi, &s_last);
  }
}

// ---- instantiation tables ----

#define SPSIM_PIXEL_TILES(X) \
  X(1, 64) X(1, 128) X(1, 256) X(2, 64) X(2, 128) X(2, 256) X(4, 64) X(4, 128) X(4, 256) X(8, 64) X(8, 128) X(8, 256)
#define SPSIM_QUAD_TILES(X) X(1, 64) X(1, 128) X(1, 256) X(2, 64) X(2, 128) X(2, 256)

enum class Kind { kPattern, kGrid, kBatch, kBatchGrid, kSpectrum };

template <bool HAS_B>
const void *kernel_ptr(Kind kind, int K, int C) {
#define P(k_, c_) \
  if (K == k_ && C == c_) return kind == Kind::kPattern ? (const void *)&sf_pattern_kernel<k_, c_, HAS_B> \
                                                          : (const void *)&sf_batch_kernel<k_, c_, HAS_B>;
#define Q(k_, c_) \
  if (K == k_ && C == c_) return kind == Kind::kGrid ? (const void *)&sf_grid_kernel<k_, c_, HAS_B> \
                                                       : (const void *)&sf_batch_grid_kernel<k_, c_, HAS_B>;
#define S(k_, c_) \
  if (K == k_ && C == c_) return (const void *)&sf_spectrum_kernel<k_, c_, HAS_B>;
#define O(c_) \
  if (K == kOctantLane && C == c_) return kind == Kind::kGrid ? (const void *)&sf_octant_kernel<c_, HAS_B> \
                                                              : (const void *)&sf_batch_octant_kernel<c_, HAS_B>;
  if (kind == Kind::kGrid || kind == Kind::kBatchGrid) {
    O(64) O(128) O(256)
  }
  if (kind == Kind::kPattern || kind == Kind::kBatch) {
    SPSIM_PIXEL_TILES(P)
  } else if (kind == Kind::kSpectrum) {
    SPSIM_QUAD_TILES(S)
  } else {
    SPSIM_QUAD_TILES(Q)
  }
#undef P
#undef Q
#undef S
#undef O
  return nullptr;
}

const void *kernel_for(Kind kind, const spsim_hip_tile &t, bool has_b) {
  if (t.waves != 4) return nullptr;
  return has_b ? kernel_ptr<true>(kind, t.pixels_per_lane, t.chunk_atoms)
               : kernel_ptr<false>(kind, t.pixels_per_lane, t.chunk_atoms);
}

// partial_buf, when given, holds the atom-range partial sums (at least
// splits * npix entries) instead of the library workspace.
int launch_single(Kind kind, const PatternLaunch &a, int nx, int ny, const spsim_hip_tile &t, int splits,
                  hipStream_t s, const Spectrum *spec = nullptr, double2 *partial_buf = nullptr) {
  if (a.npix <= 0) return SPSIM_HIP_OK;
  const void *k = kernel_for(kind, t, a.atom_b != nullptr);
  if (!k) return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (per lane %d, chunk %d)", t.pixels_per_lane, t.chunk_atoms);
  const int units = kind == Kind::kGrid ? grid_units(t, nx, ny) : a.npix;
  const int per_tile = kind == Kind::kGrid ? grid_units_per_tile(t) : tile_units(t);
  const int tiles = (units + per_tile - 1) / per_tile;
  if (splits < 1) splits = 1;
  double2 *partial = nullptr;
  if (splits > 1 && partial_buf) {
    partial = partial_buf;
  } else if (splits > 1) {
    void *ws = nullptr;
    SPSIM_TRY(workspace(sizeof(double2) * (size_t)splits * a.npix, &ws));
    partial = (double2 *)ws;
  }
  PatternLaunch arg = a;
  if (kind == Kind::kGrid) {
    void *params[] = {(void *)&arg, (void *)&nx, (void *)&ny, (void *)&partial};
    SPSIM_HIP_TRY(hipLaunchKernel(k, dim3(tiles, splits), dim3(kThreads), params, 0, s));
  } else if (kind == Kind::kSpectrum) {
    Spectrum sp = *spec;
    void *params[] = {(void *)&arg, (void *)&sp, (void *)&partial};
    SPSIM_HIP_TRY(hipLaunchKernel(k, dim3(tiles, splits), dim3(kThreads), params, 0, s));
  } else {
    void *params[] = {(void *)&arg, (void *)&partial};
    SPSIM_HIP_TRY(hipLaunchKernel(k, dim3(tiles, splits), dim3(kThreads), params, 0, s));
  }
  if (splits > 1) {
    hipLaunchKernelGGL(sf_reduce_kernel, dim3((a.npix + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                       (const double2 *)partial, splits, a.npix, a.F, a.I);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  return SPSIM_HIP_OK;
}

int check_common(const float *d_hkl, long long npix, const spsim_hip_atom *d_atoms, const spsim_hip_segment *d_segs,
                 int nsegs, int natoms, const float *d_ff) {
  if (npix < 0 || nsegs < 0 || natoms < 0)
    return fail(SPSIM_HIP_ERR_INVALID, "negative size (npix=%lld nsegs=%d natoms=%d)", npix, nsegs, natoms);
  if (npix > (1LL << 28)) return fail(SPSIM_HIP_ERR_INVALID, "npix=%lld exceeds 2^28", npix);
  if (npix > 0 && !d_hkl) return fail(SPSIM_HIP_ERR_INVALID, "d_hkl is NULL");
  if (nsegs > 0 && (!d_atoms || !d_segs)) return fail(SPSIM_HIP_ERR_INVALID, "d_atoms/d_segs is NULL");
  if (nsegs == 0 && natoms != 0) return fail(SPSIM_HIP_ERR_INVALID, "natoms=%d but no segments", natoms);
  if (!d_ff) return fail(SPSIM_HIP_ERR_INVALID, "d_ff is NULL");
  return SPSIM_HIP_OK;
}

PatternLaunch make_launch(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                          const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
                          const double *rot9, float *d_F, float *d_I) {
  PatternLaunch a;
  a.hkl = d_hkl;
  a.npix = npix;
  a.atoms = d_atoms;
  a.atom_b = d_atom_b;
  a.segs = d_segs;
  a.nsegs = nsegs;
  a.natoms = natoms;
  a.ff = d_ff;
  a.b_global = b_global;
  a.F = (float2 *)d_F;
  a.I = d_I;
  a.has_rot = rot9 != nullptr;
  for (int i = 0; i < 9; ++i) a.rot.r[i] = rot9 ? rot9[i] : (i % 4 == 0 ? 1.0 : 0.0);
  return a;
}

}  // namespace

bool tile_supported(const spsim_hip_tile &t, bool grid) {
  return kernel_for(grid ? Kind::kGrid : Kind::kPattern, t, false) != nullptr && t.atom_splits >= 0;
}

spsim_hip_tile default_tile(int natoms, int npix, bool grid) {
  spsim_hip_tile t;
  t.waves = 4;
  if (grid)
    t.pixels_per_lane = 1;
  else
    t.pixels_per_lane = npix >= 4 * 256 * 256 ? 2 : 1;
  t.chunk_atoms = natoms >= 256 ? 256 : (natoms >= 128 ? 128 : 64);
  t.atom_splits = 0;
  return t;
}

int auto_splits(int tiles, int natoms, int chunk_atoms, int slots, int max_splits) {
  const int hi = std::max(1, std::min(std::min(64, max_splits), natoms / std::max(1, chunk_atoms)));
  int best = 1;
  double best_cost = 1e300;
  for (int s = 1; s <= hi; ++s) {
    // Rounds of resident workgroups x (atoms per range + a per-item cost of
    // about 64 atoms). Up to 4 rounds the items of a round start together and
    // the count is whole rounds; beyond that they finish at scattered times,
    // and the launch lasts the items' share per slot plus about half an item
    // (on 100,000 atoms x 1024^2 octants, 16 to 48 ranges measured
    // 18.25 + 14.4/s + 0.0073 s ms). The partial sums' write and read add
    // ~0.04 % per range.
    const double r = (double)tiles * s / std::max(slots, 1);
    const double rounds = r <= 4.0 ? std::ceil(r) : r + 0.5;
    const double cost = rounds * ((double)natoms / s + 64.0) * (1.0 + 4e-4 * s);
    if (cost < best_cost) {
      best_cost = cost;
      best = s;
    }
  }
  return best;
}

int resident_slots(const void *kernel, int cus) {
  int nb = 0;
  if (!kernel || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kThreads, 0) != hipSuccess || nb < 1)
    nb = 1;
  return nb * std::max(cus, 1);
}

bool host_transpose_symmetric(const float *hkl, int n) {
  for (int y = 0; y < n; ++y)
    for (int x = 0; x < y; ++x) {
      const float *q = hkl + 3 * ((size_t)y * n + x);
      const float *t = hkl + 3 * ((size_t)x * n + y);
      if (!(t[0] == q[1] && t[1] == q[0] && t[2] == q[2])) return false;
    }
  return true;
}

bool host_mirror_symmetric(const float *hkl, int nx, int ny) {
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) {
      const float *q = hkl + 3 * ((size_t)y * nx + x);
      const float *mx = hkl + 3 * ((size_t)y * nx + (nx - 1 - x));
      const float *my = hkl + 3 * ((size_t)(ny - 1 - y) * nx + x);
      if (!(mx[0] == -q[0] && mx[1] == q[1] && mx[2] == q[2])) return false;
      if (!(my[0] == q[0] && my[1] == -q[1] && my[2] == q[2])) return false;
    }
  }
  return true;
}

int launch_pattern(const PatternLaunch &a, const spsim_hip_tile &t, int splits, hipStream_t s) {
  return launch_single(Kind::kPattern, a, 0, 0, t, splits, s);
}

int launch_grid(const PatternLaunch &a, int nx, int ny, const spsim_hip_tile &t, int splits, hipStream_t s) {
  return launch_single(Kind::kGrid, a, nx, ny, t, splits, s);
}

int launch_spectrum(const PatternLaunch &a, const Spectrum &sp, const spsim_hip_tile &t, int splits, hipStream_t s) {
  return launch_single(Kind::kSpectrum, a, 0, 0, t, splits, s, &sp);
}

// Interpolation in u = v^2 on [0, 1] at the kSpectrumDegree + 1 Chebyshev-
// Lobatto points (u = 0 among them, so P(0) = sum_k W_k and the small phases
// of most terms carry no bias) of
//   P(u) = sum_k W_k cos(a_k v),  Q(u) = sum_k W_k sin(a_k v) / v  (a_k at v = 0),
// a_k = 2 pi e_k tlim, both entire in u; the monomial coefficients come from
// the Vandermonde system (Gaussian elimination with partial pivoting). With
// tlim = kSpectrumReach / (2 pi max|e_k|), the FP32 polynomials stay within
// 3e-7 sum_k |W_k| of the sample sums over |t| <= tlim, for any placement of
// the samples (1.2e-7 for Gaussian spectra, whose extreme samples weigh
// little). A reach of 2 radians left up to 1.1e-6 for samples weighted evenly
// out to the extremes.
void spectrum_fit(Spectrum &sp) {
  constexpr int N = kSpectrumDegree + 1;
  constexpr double kTwoPi = 6.283185307179586476925;
  for (int d = 0; d < N; ++d) sp.p[d] = sp.q[d] = 0.0f;
  sp.sc = 1.0;
  sp.tlim = -1.0f;
  sp.inv_tlim = 0.0f;
  if (sp.n < 4) return;
  double smin = sp.s[0], smax = sp.s[0];
  for (int k = 1; k < sp.n; ++k) {
    smin = std::min(smin, sp.s[k]);
    smax = std::max(smax, sp.s[k]);
  }
  const double sc = 0.5 * (smin + smax);
  if (!(sc > 0.0) || !std::isfinite(sc)) return;
  double emax = 0.0;
  for (int k = 0; k < sp.n; ++k) emax = std::max(emax, std::fabs(sp.s[k] / sc - 1.0));
  const double tlim = emax > 0.0 ? std::min(1e30, kSpectrumReach / (kTwoPi * emax)) : 1e30;
  double A[N][N + 2];
  for (int j = 0; j < N; ++j) {
    const double u = j == N - 1 ? 0.0 : 0.5 * (1.0 + std::cos(M_PI * j / (N - 1)));
    const double v = std::sqrt(u);
    double pc = 0.0, qc = 0.0;
    for (int k = 0; k < sp.n; ++k) {
      const double ak = kTwoPi * (sp.s[k] / sc - 1.0) * tlim;
      pc += (double)sp.w[k] * std::cos(ak * v);
      qc += (double)sp.w[k] * (v > 0.0 ? std::sin(ak * v) / v : ak);
    }
    double ud = 1.0;
    for (int d = 0; d < N; ++d, ud *= u) A[j][d] = ud;
    A[j][N] = pc;
    A[j][N + 1] = qc;
  }
  for (int c = 0; c < N; ++c) {
    int piv = c;
    for (int r = c + 1; r < N; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    for (int x = 0; x < N + 2; ++x) std::swap(A[c][x], A[piv][x]);
    for (int r = c + 1; r < N; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int x = c; x < N + 2; ++x) A[r][x] -= f * A[c][x];
    }
  }
  double cp[N], cq[N];
  for (int r = N - 1; r >= 0; --r) {
    double vp = A[r][N], vq = A[r][N + 1];
    for (int x = r + 1; x < N; ++x) {
      vp -= A[r][x] * cp[x];
      vq -= A[r][x] * cq[x];
    }
    cp[r] = vp / A[r][r];
    cq[r] = vq / A[r][r];
  }
  for (int d = 0; d < N; ++d) {
    sp.p[d] = (float)cp[d];
    sp.q[d] = (float)cq[d];
  }
  sp.sc = sc;
  sp.tlim = (float)tlim;
  sp.inv_tlim = (float)(1.0 / tlim);
}

int batch_blocks_per_cu(const spsim_hip_tile &t, bool has_b, bool grid, int *blocks) {
  const void *k = kernel_for(grid ? Kind::kBatchGrid : Kind::kBatch, t, has_b);
  if (!k) return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (per lane %d, chunk %d)", t.pixels_per_lane, t.chunk_atoms);
  int nb = 0;
  SPSIM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kThreads, 0));
  *blocks = std::max(1, nb);
  return SPSIM_HIP_OK;
}

int launch_batch(const BatchLaunch &a, const spsim_hip_tile &t, bool grid, int blocks, hipStream_t s) {
  const void *k = kernel_for(grid ? Kind::kBatchGrid : Kind::kBatch, t, a.atom_b != nullptr);
  if (!k) return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (per lane %d, chunk %d)", t.pixels_per_lane, t.chunk_atoms);
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, a.queue);
  SPSIM_HIP_TRY(hipGetLastError());
  if (a.nitems <= 0) return SPSIM_HIP_OK;
  BatchLaunch arg = a;
  void *params[] = {(void *)&arg};
  SPSIM_HIP_TRY(hipLaunchKernel(k, dim3(std::max(1, blocks)), dim3(kThreads), params, 0, s));
  return SPSIM_HIP_OK;
}

}  // namespace spsim

extern "C" int spsim_hip_sf_pattern(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                                    const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff,
                                    float b_global, const double *rot9, float *d_F, float *d_I,
                                    const spsim_hip_tile *tile, void *stream) {
  using namespace spsim;
  SPSIM_TRY(check_common(d_hkl, npix, d_atoms, d_segs, nsegs, natoms, d_ff));
  int cus = 0;
  SPSIM_TRY(ensure_device(&cus));
  if (npix == 0) return SPSIM_HIP_OK;
  const spsim_hip_tile t = tile ? *tile : default_tile(natoms, npix, false);
  if (!tile_supported(t, false))
    return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (pixels_per_lane=%d waves=%d chunk_atoms=%d)",
                t.pixels_per_lane, t.waves, t.chunk_atoms);
  const int tiles = (npix + tile_units(t) - 1) / tile_units(t);
  const int splits = t.atom_splits > 0
                         ? t.atom_splits
                         : auto_splits(tiles, natoms, t.chunk_atoms,
                                       resident_slots(kernel_for(Kind::kPattern, t, d_atom_b != nullptr), cus),
                                       (int)std::min<size_t>(64, kWorkspacePartials / (size_t)npix));
  const PatternLaunch a =
      make_launch(d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, rot9, d_F, d_I);
  return launch_pattern(a, t, splits, (hipStream_t)stream);
}

namespace spsim {
namespace {

// Checks the arguments of a grid call and picks its tile and atom-split count
// (explicit, or chosen to fill the GPU with at most max_partials partial sums,
// splits x npix). *npix_out = 0 means nothing to do.
int grid_setup(const float *d_hkl, int nx, int ny, const spsim_hip_atom *d_atoms, bool has_b,
               const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, const spsim_hip_tile *tile,
               size_t max_partials, spsim_hip_tile *t_out, int *splits_out, bool *explicit_splits, int *npix_out) {
  if (nx < 0 || ny < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative detector size %d x %d", nx, ny);
  const long long npix = (long long)nx * ny;
  SPSIM_TRY(check_common(d_hkl, npix, d_atoms, d_segs, nsegs, natoms, d_ff));
  int cus = 0;
  SPSIM_TRY(ensure_device(&cus));
  *npix_out = (int)npix;
  if (npix == 0) return SPSIM_HIP_OK;
  spsim_hip_tile t = tile ? *tile : default_tile(natoms, (int)npix, true);
  if (!tile && nx == ny) t.pixels_per_lane = kOctantLane;
  if (!tile_supported(t, true) || (t.pixels_per_lane == kOctantLane && nx != ny))
    return fail(SPSIM_HIP_ERR_INVALID,
                "unsupported grid tile (quads_per_lane=%d waves=%d chunk_atoms=%d; octants need nx == ny)",
                t.pixels_per_lane, t.waves, t.chunk_atoms);
  const int tiles = (grid_units(t, nx, ny) + grid_units_per_tile(t) - 1) / grid_units_per_tile(t);
  *explicit_splits = t.atom_splits > 0;
  *splits_out = t.atom_splits > 0
                    ? t.atom_splits
                    : auto_splits(tiles, natoms, t.chunk_atoms, resident_slots(kernel_for(Kind::kGrid, t, has_b), cus),
                                  (int)std::min<size_t>(64, max_partials / (size_t)npix));
  *t_out = t;
  return SPSIM_HIP_OK;
}

// ---- no-allocation variant of spsim_hip_sf_grid ----
//
// The atom-range partial sums live in this fixed device array instead of the
// library workspace, so the call allocates nothing. Host code never uses the
// array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device with that device current, and is kept
// in noalloc_ptr[device]. Every call on one device shares that device's array.
constexpr size_t kNoAllocPartials = (size_t)1 << 21;  // double2 entries: 32 MiB
static __device__ __align__(256) double2 scratch[kNoAllocPartials];

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
double2 *g_noalloc_ptr[kNoAllocDevices];

int noalloc_scratch(double2 **out) {
  int dev = 0;
  SPSIM_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) return fail(SPSIM_HIP_ERR_NODEVICE, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(SPSIM_HIP_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_noalloc_ptr[dev] = (double2 *)p;
  }
  *out = g_noalloc_ptr[dev];
  return SPSIM_HIP_OK;
}

}  // namespace
}  // namespace spsim

extern "C" int spsim_hip_sf_grid(const float *d_hkl, int nx, int ny, const spsim_hip_atom *d_atoms,
                                 const float *d_atom_b, const spsim_hip_segment *d_segs, int nsegs, int natoms,
                                 const float *d_ff, float b_global, const double *rot9, float *d_F, float *d_I,
                                 const spsim_hip_tile *tile, void *stream) {
  using namespace spsim;
  spsim_hip_tile t{};
  int splits = 1, npix = 0;
  bool explicit_splits = false;
  SPSIM_TRY(grid_setup(d_hkl, nx, ny, d_atoms, d_atom_b != nullptr, d_segs, nsegs, natoms, d_ff, tile,
                       kWorkspacePartials, &t, &splits, &explicit_splits, &npix));
  if (npix == 0) return SPSIM_HIP_OK;
  const PatternLaunch a =
      make_launch(d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, rot9, d_F, d_I);
  return launch_grid(a, nx, ny, t, splits, (hipStream_t)stream);
}

extern "C" int spsim_hip_sf_grid_noalloc(const float *d_hkl, int nx, int ny, const spsim_hip_atom *d_atoms,
                                         const float *d_atom_b, const spsim_hip_segment *d_segs, int nsegs,
                                         int natoms, const float *d_ff, float b_global, const double *rot9,
                                         float *d_F, float *d_I, const spsim_hip_tile *tile, void *stream) {
  using namespace spsim;
  spsim_hip_tile t{};
  int splits = 1, npix = 0;
  bool explicit_splits = false;
  SPSIM_TRY(grid_setup(d_hkl, nx, ny, d_atoms, d_atom_b != nullptr, d_segs, nsegs, natoms, d_ff, tile,
                       kNoAllocPartials, &t, &splits, &explicit_splits, &npix));
  if (npix == 0) return SPSIM_HIP_OK;
  const long long cap = (long long)(kNoAllocPartials / (size_t)npix);  // atom ranges that fit
  if (splits > 1 && (long long)splits > cap) {
    if (explicit_splits)
      return fail(SPSIM_HIP_ERR_INVALID,
                  "atom_splits=%d on %d pixels needs %lld bytes of partial sums; the no-allocation entry point "
                  "holds %zu (at most %lld atom ranges at this size)",
                  splits, npix, (long long)splits * npix * (long long)sizeof(double2),
                  kNoAllocPartials * sizeof(double2), cap < 1 ? 1LL : cap);
    splits = cap < 1 ? 1 : (int)cap;  // automatic choice: as many ranges as fit
  }
  double2 *partial = nullptr;
  if (splits > 1) SPSIM_TRY(noalloc_scratch(&partial));
  const PatternLaunch a =
      make_launch(d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, rot9, d_F, d_I);
  return launch_single(Kind::kGrid, a, nx, ny, t, splits, (hipStream_t)stream, nullptr, partial);
}

extern "C" int spsim_hip_sf_spectrum(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms,
                                     const float *d_atom_b, const spsim_hip_segment *d_segs, int nsegs, int natoms,
                                     const float *d_ff, float b_global, const double *rot9, int nsamples,
                                     const double *scales, const double *weights, float *d_F, float *d_I,
                                     const spsim_hip_tile *tile, void *stream) {
  using namespace spsim;
  SPSIM_TRY(check_common(d_hkl, npix, d_atoms, d_segs, nsegs, natoms, d_ff));
  if (nsamples < 1 || nsamples > SPSIM_HIP_MAX_SPECTRUM)
    return fail(SPSIM_HIP_ERR_INVALID, "nsamples=%d outside [1, %d]", nsamples, SPSIM_HIP_MAX_SPECTRUM);
  if (!scales || !weights) return fail(SPSIM_HIP_ERR_INVALID, "scales/weights is NULL");
  int cus = 0;
  SPSIM_TRY(ensure_device(&cus));
  if (npix == 0) return SPSIM_HIP_OK;
  spsim_hip_tile t = tile ? *tile : default_tile(natoms, npix, true);
  // two pixels per lane share the atom loads and loop overhead (3 % on a
  // 512 x 512 detector); smaller lists keep one for more workgroups
  if (!tile && npix >= 256 * 256) t.pixels_per_lane = 2;
  if (kernel_for(Kind::kSpectrum, t, false) == nullptr || t.atom_splits < 0)
    return fail(SPSIM_HIP_ERR_INVALID, "unsupported spectrum tile (pixels_per_lane=%d waves=%d chunk_atoms=%d)",
                t.pixels_per_lane, t.waves, t.chunk_atoms);
  Spectrum sp;
  sp.n = nsamples;
  for (int k = 0; k < SPSIM_HIP_MAX_SPECTRUM; ++k) {
    sp.s[k] = k < nsamples ? scales[k] : 0.0;
    sp.w[k] = k < nsamples ? (float)weights[k] : 0.0f;
  }
  spectrum_fit(sp);
  const int tiles = (npix + tile_units(t) - 1) / tile_units(t);
  const int splits = t.atom_splits > 0
                         ? t.atom_splits
                         : auto_splits(tiles, natoms, t.chunk_atoms,
                                       resident_slots(kernel_for(Kind::kSpectrum, t, d_atom_b != nullptr), cus),
                                       (int)std::min<size_t>(64, kWorkspacePartials / (size_t)npix));
  const PatternLaunch a =
      make_launch(d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, rot9, d_F, d_I);
  return launch_spectrum(a, sp, t, splits, (hipStream_t)stream);
}
