/*
 * ORACLE — test infrastructure only. CPU restatement of spsim's structure-factor
 * hot path, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg as the checker. Nothing in the product links or calls this file.
 *
 * Functions and what pins them (tests/test_oracle_*.py):
 *   oracle_form_factor        src/diffraction.c:183-193 (FP64). Pinned by f0(0)=sum(a)+c,
 *                             International Tables f0(sin(theta)/lambda) values and the
 *                             Debye-Waller identity.
 *   oracle_hkl_detector       src/diffraction.c:237-297 (same FP64 operations, FP32 store).
 *                             Pinned by the Ewald closed form of doc/UserManual.tex:159-179,
 *                             |q| = 2 sin(theta)/lambda and symmetry.
 *   oracle_rotate_hkl         q' = R q in FP64 (fixed fma order), rounded to FP32: the
 *                             library's convention for src/diffraction.c:300-323. Pinned by
 *                             exact quarter-turn rotations and norm preservation.
 *   oracle_sf_f64             src/diffraction.c:770-853 (plain) and :856-968 (vector quirks)
 *                             evaluated in FP64 on the FP32 inputs. Pinned by closed forms
 *                             (1 and 2 atoms, atoms at the origin, centrosymmetry, Friedel,
 *                             translation and rotation invariance, delta atoms) and an
 *                             arbitrary-precision (mpmath) sum.
 *   oracle_hkl_3d             src/diffraction.c:325-369. Pinned by its closed form (a
 *                             uniform grid of spacing extent/(n d lambda), zero at the
 *                             centre voxel, exactly antisymmetric about it) and 40-digit
 *                             mpmath values.
 *   oracle_spectrum_samples,
 *   oracle_sf_spectrum_f64    the wavelength-spectrum path of src/diffraction.cu:300-318,
 *                             416-476 in FP64. Pinned by closed forms (the weights
 *                             exp(-2 (2k/(S-1) - 1)^2), the scales 1/(1 - bw/2 + k bw/(S-1)),
 *                             one atom, atoms at the origin, the plain oracle for S = 1) and
 *                             by linearity over samples for point scatterers.
 *   oracle_sf_refsem          src/diffraction.c:770-853 in the reference's own FP32/FP64 mix.
 *                             Pinned against oracle_sf_f64 within the FP32 error bound.
 *                             The final |F|^2 assumes libspimage's sp_cabs is the FP32
 *                             modulus (libspimage is not available; parity unpinned for that
 *                             rounding step only).
 *   oracle_thomson, oracle_solid_angle, oracle_photons, oracle_noiseless_output,
 *   oracle_real_output        src/diffraction.c:1109-1204, src/amplification.c:54-153 in
 *                             the reference's precision. Pinned by closed forms (r0^2,
 *                             polarisation factors, Omega = A d / r^3, hand sums).
 *   oracle_philox4x32_10      Salmon et al., SC'11. Pinned by the Random123 known-answer
 *                             vectors.
 *   oracle_poisson            src/noise.c:32-51 restated with inversion (mean < 10), PTRS
 *                             (Hormann 1993; 10 <= mean < 1000) and the reference's Gaussian
 *                             branch. Pinned by mean/variance tests.
 *   oracle_poisson_exact,
 *   oracle_real_output_bin_draws  the binned chain with one count per bin: the sum of
 *                             independent Poisson counts is Poisson with the summed mean,
 *                             so a bin's Poisson pixels are drawn together; pixels on the
 *                             reference's Gaussian branch keep their own draw. Pinned by
 *                             mean/variance tests of the joint draw over bins of known mean.
 *   oracle_sf_grad_f64,
 *   oracle_sf_grad_given_f_f64 gradient of L = sum_p g_p I_p w.r.t. atom positions and B
 *                             factors in FP64 (no reference counterpart), for every atom or
 *                             for a listed subset given the forward F. Pinned by central
 *                             finite differences of oracle_sf_f64 (small problems, and the
 *                             Poisson NLL of the 20,000-atom differentiation workload on
 *                             64 atoms, tests/test_gpu_adjoint.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NSPEC 101
#define UNIT 100
#define FFS 9

static const double TWO_PI = 6.283185307179586476925286766559;

/* ---------------------------------------------------------------- form factor */

/* f_Z(d) with d = |q| in m^-1, global B in A^2; FP64 throughout. */
double oracle_form_factor(double d, int Z, double B, const float *ff) {
  const float *c = ff + (size_t)Z * FFS;
  const double dA = d * 1e-10;
  const double s2 = dA * dA * 0.25;
  double f = 0.0;
  for (int i = 0; i < 4; ++i) f += (double)c[i] * exp(-((double)c[4 + i] + B) * s2);
  f += (double)c[8] * exp(-B * s2);
  return f;
}

/* ---------------------------------------------------------------- geometry */

int oracle_hkl_detector(int nx, int ny, float width, float height, float distance, float center_x, float center_y,
                        int spherical, float wavelength, float *out) {
  const double ewald_radius = 1.0 / wavelength;
  const double pz = distance;
  const double px_size = width / nx;  /* float division, as in the reference */
  const double py_size = height / ny;
  const double fourier_px = 2 * sin(atan(px_size / pz) / 2) / wavelength;
  const double fourier_py = 2 * sin(atan(py_size / pz) / 2) / wavelength;
  size_t idx = 0;
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) {
      const double px = ((x - (nx - 1.0) / 2.0) / nx) * width - center_x;
      const double py = ((y - (ny - 1.0) / 2.0) / ny) * height - center_y;
      const double p = sqrt(px * px + py * py + pz * pz);
      const double rx = px / p, ry = py / p, rz = pz / p - 1.;
      if (!spherical) {
        out[idx++] = rx * ewald_radius;
        out[idx++] = ry * ewald_radius;
        out[idx++] = rz * ewald_radius;
      } else {
        out[idx++] = fourier_px * px / px_size;
        out[idx++] = fourier_py * py / py_size;
        out[idx++] = 0;
      }
    }
  }
  return 0;
}

/* src/diffraction.c:325-369: q = p / (distance * wavelength), the product in float. */
int oracle_hkl_3d(int nx, int ny, int nz, float width, float height, float depth, float center_x, float center_y,
                  float center_z, float distance, float wavelength, float *out) {
  const float dl = distance * wavelength;
  const double r2r = 1.0 / dl;
  size_t idx = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const double px = ((x - (nx - 1.0) / 2.0) / nx) * width - center_x;
        const double py = ((y - (ny - 1.0) / 2.0) / ny) * height - center_y;
        const double pz = ((z - (nz - 1.0) / 2.0) / nz) * depth - center_z;
        out[idx++] = (float)(px * r2r);
        out[idx++] = (float)(py * r2r);
        out[idx++] = (float)(pz * r2r);
      }
  return 0;
}

void oracle_rotate_hkl(const float *in, int n, const double *R, float *out) {
  for (int i = 0; i < n; ++i) {
    const double x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
    const float ox = (float)fma(R[2], z, fma(R[1], y, R[0] * x));
    const float oy = (float)fma(R[5], z, fma(R[4], y, R[3] * x));
    const float oz = (float)fma(R[8], z, fma(R[7], y, R[6] * x));
    out[3 * i] = ox;
    out[3 * i + 1] = oy;
    out[3 * i + 2] = oz;
  }
}

/* ---------------------------------------------------------------- structure factor */

/* Which form factor / amplitude an atom gets: -1 = skipped, UNIT = point scatterer. */
static int oracle_species(int z, int j, int n, int delta, int vector_mode) {
  if (vector_mode) return (delta && j < 4 * (n / 4)) ? UNIT : z;
  if (z == 0) return -1;
  return delta ? UNIT : z;
}

/* Phase 2 pi t with t reduced to [-1/2, 1/2] turns first (exact in FP64). */
static void cis_turns(double t, double *c, double *s) {
  const double r = t - nearbyint(t);
  *c = cos(TWO_PI * r);
  *s = sin(TWO_PI * r);
}

/*
 * FP64 evaluation on FP32 inputs. amp may be NULL (1), atom_b may be NULL.
 * vector_mode selects the vector path's semantics (src/diffraction.c:856-968).
 * R (row-major 3x3, may be NULL) rotates every q exactly in FP64 first.
 * F_out is [npix][2] (may be NULL), I_out [npix].
 */
int oracle_sf_f64(const float *hkl, int npix, const int *Z, const float *pos, const float *amp, const float *atom_b,
                  int natoms, int delta, int vector_mode, const float *ff, float B, double *F_out, double *I_out,
                  int nthreads, const double *R) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int i = 0; i < npix; ++i) {
    double qx = hkl[3 * i], qy = hkl[3 * i + 1], qz = hkl[3 * i + 2];
    if (R) { /* q' = R q, unrounded */
      const double x = qx, y = qy, z = qz;
      qx = R[0] * x + R[1] * y + R[2] * z;
      qy = R[3] * x + R[4] * y + R[5] * z;
      qz = R[6] * x + R[7] * y + R[8] * z;
    }
    const double d = sqrt(qx * qx + qy * qy + qz * qz);
    const double s2 = (d * 1e-10) * (d * 1e-10) * 0.25;
    double fcache[NSPEC];
    int have[NSPEC];
    memset(have, 0, sizeof have);
    double re = 0.0, im = 0.0;
    for (int j = 0; j < natoms; ++j) {
      const int sp = oracle_species(Z[j], j, natoms, delta, vector_mode);
      if (sp < 0) continue;
      if (!have[sp]) {
        fcache[sp] = (sp == UNIT) ? 1.0 : oracle_form_factor(d, sp, B, ff);
        have[sp] = 1;
      }
      double w = (sp == UNIT || !amp) ? 1.0 : (double)amp[j];
      if (atom_b) w *= exp(-(double)atom_b[j] * s2);
      const double t = qx * pos[3 * j] + qy * pos[3 * j + 1] + qz * pos[3 * j + 2];
      double c, s;
      cis_turns(t, &c, &s);
      re += fcache[sp] * w * c;
      im -= fcache[sp] * w * s;
    }
    if (F_out) {
      F_out[2 * i] = re;
      F_out[2 * i + 1] = im;
    }
    I_out[i] = re * re + im * im;
  }
  return 0;
}

/*
 * Wavelength samples of the reference's spectrum path
 * (src/diffraction.cu:300-303, 431-460), in FP64: an even sample count above 1
 * is raised by one; lambda_k = lambda (1 - bw/2) + k/(S-1) lambda bw spans
 * +-2 sigma with sigma = lambda bw / 4; Gaussian weights exp(-x^2/2) with
 * x = (lambda_k - lambda) / sigma, normalised to sum 1; the phase scale is
 * s_k = lambda / lambda_k. Returns the sample count used (1 when bw == 0).
 */
int oracle_spectrum_samples(double wavelength, double bandwidth, int nsamples, double *scale, double *weight) {
  if (nsamples > 1 && (nsamples & 1) == 0) nsamples++;
  if (bandwidth == 0.0 || nsamples <= 1) {
    scale[0] = 1.0;
    weight[0] = 1.0;
    return 1;
  }
  const double lo = wavelength * (1.0 - bandwidth / 2.0);
  const double sigma = (wavelength - lo) / 2.0;
  double total = 0.0;
  for (int k = 0; k < nsamples; ++k) {
    const double lk = lo + ((double)k / (nsamples - 1)) * wavelength * bandwidth;
    const double x = (lk - wavelength) / sigma;
    weight[k] = exp(-x * x / 2.0);
    scale[k] = wavelength / lk;
    total += weight[k];
  }
  for (int k = 0; k < nsamples; ++k) weight[k] /= total;
  return nsamples;
}

/*
 * Spectrum structure factor (src/diffraction.cu:416-476) in FP64 on FP32
 * inputs: F(q) = sum_j f_Zj(|q|) a_j sum_k w_k e^{-2 pi i s_k q.r_j}. The form
 * factor is taken at the unscaled |q| and follows the CPU formula (the GPU
 * path's c-term slip at :450 is not reproduced). Z = 0 atoms are skipped;
 * amp, atom_b and R as in oracle_sf_f64.
 */
int oracle_sf_spectrum_f64(const float *hkl, int npix, const int *Z, const float *pos, const float *amp,
                           const float *atom_b, int natoms, const float *ff, float B, int nsamples,
                           const double *scale, const double *weight, double *F_out, double *I_out, int nthreads,
                           const double *R) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int i = 0; i < npix; ++i) {
    double qx = hkl[3 * i], qy = hkl[3 * i + 1], qz = hkl[3 * i + 2];
    if (R) {
      const double x = qx, y = qy, z = qz;
      qx = R[0] * x + R[1] * y + R[2] * z;
      qy = R[3] * x + R[4] * y + R[5] * z;
      qz = R[6] * x + R[7] * y + R[8] * z;
    }
    const double d = sqrt(qx * qx + qy * qy + qz * qz);
    const double s2 = (d * 1e-10) * (d * 1e-10) * 0.25;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < natoms; ++j) {
      if (Z[j] == 0) continue;
      double a = oracle_form_factor(d, Z[j], B, ff) * (amp ? (double)amp[j] : 1.0);
      if (atom_b) a *= exp(-(double)atom_b[j] * s2);
      const double t = qx * pos[3 * j] + qy * pos[3 * j + 1] + qz * pos[3 * j + 2];
      double cr = 0.0, ci = 0.0;
      for (int k = 0; k < nsamples; ++k) {
        double c, sn;
        cis_turns(scale[k] * t, &c, &sn);
        cr += weight[k] * c;
        ci += weight[k] * sn;
      }
      re += a * cr;
      im -= a * ci;
    }
    if (F_out) {
      F_out[2 * i] = re;
      F_out[2 * i + 1] = im;
    }
    I_out[i] = re * re + im * im;
  }
  return 0;
}

/* scatt_factor in the reference's precision (src/diffraction.c:183-193). */
static float ref_scatt_factor(float d, int Z, float B, const float *ff) {
  const float *c = ff + (size_t)Z * FFS;
  float res = 0;
  d *= 1e-10;
  for (int i = 0; i < 4; ++i) res += c[i] * exp(-(c[i + 4] + B) * d * d * 0.25);
  res += c[8] * exp(-B * d * d * 0.25);
  return res;
}

/*
 * compute_pattern_on_list (src/diffraction.c:770-853) with the reference's own
 * rounding: FP32 |q|, FP32 dot product, FP64 2*pi scale stored to FP32, FP64
 * libm cos/sin, FP32 accumulation in atom order. illum is the per-atom
 * illumination (not its square root), may be NULL (1).
 */
int oracle_sf_refsem(const float *hkl, int npix, const int *Z, const float *pos, const float *illum, int natoms,
                     int delta, const float *ff, float B, float *F_out, float *I_out, int nthreads) {
  /* elements present, once per call (src/diffraction.c:794-800) */
  int present[NSPEC];
  memset(present, 0, sizeof present);
  for (int j = 0; j < natoms; ++j) present[Z[j]] = 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int i = 0; i < npix; ++i) {
    float cache[NSPEC];
    const float *h = hkl + 3 * (size_t)i;
    const float len = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    for (int z = 0; z < 100; ++z)
      if (present[z]) cache[z] = ref_scatt_factor(len, z, B, ff);
    float re = 0, im = 0;
    for (int j = 0; j < natoms; ++j) {
      if (!Z[j]) continue;
      const float il = illum ? illum[j] : 1.0f;
      const float sf = cache[Z[j]] * sqrt(il);
      const float tmp = -2 * M_PI * (h[0] * pos[3 * j] + h[1] * pos[3 * j + 1] + h[2] * pos[3 * j + 2]);
      if (!delta) {
        re += sf * cos(tmp);
        im += sf * sin(tmp);
      } else {
        re += cos(tmp);
        im += sin(tmp);
      }
    }
    if (F_out) {
      F_out[2 * i] = re;
      F_out[2 * i + 1] = im;
    }
    const float mod = sqrtf(re * re + im * im);
    I_out[i] = mod * mod;
  }
  return 0;
}

/*
 * Gradient of L = sum_p g_p I_p with I_p = |F_p|^2 from oracle_sf_f64's model
 * (plain mode, optional amplitudes and per-atom B), in FP64:
 *   dL/dr_j = -4 pi sum_p g_p a_pj q_p (Re F_p sin phi_pj + Im' ...)   (see below)
 * computed by direct differentiation of F_p = sum_j a_pj e^{-i phi_pj},
 * phi_pj = 2 pi q_p.r_j, a_pj = f_Zj(|q_p|) w_j exp(-B_j s_p^2):
 *   dI_p/dr_j = 2 Re(conj(F_p) dF_p/dr_j),  dF_p/dr_j = -2 pi i q_p a_pj e^{-i phi_pj}
 *   dI_p/dB_j = 2 Re(conj(F_p) (-s_p^2) a_pj e^{-i phi_pj})
 * grad_pos is [natoms][3] (original atom order), grad_b [natoms] (may be NULL).
 */
/*
 * The same gradient for the atoms listed in sub[0..nsub) only, given F_p
 * (F: [npix][2], FP64, the forward of every atom, e.g. from oracle_sf_f64).
 * grad_pos is [nsub][3] and grad_b [nsub] (may be NULL), in the order of sub.
 * Atoms with Z == 0 get zero (compute_pattern_on_list skips them).
 */
int oracle_sf_grad_given_f_f64(const float *hkl, int npix, const double *F, const int *Z, const float *pos,
                               const float *amp, const float *atom_b, const float *ff, float B, const double *g,
                               int nsub, const int *sub, double *grad_pos, double *grad_b, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int k = 0; k < nsub; ++k) {
    const int j = sub[k];
    double gx = 0, gy = 0, gz = 0, gb = 0;
    if (Z[j] != 0) {
      for (int p = 0; p < npix; ++p) {
        const double qx = hkl[3 * p], qy = hkl[3 * p + 1], qz = hkl[3 * p + 2];
        const double d = sqrt(qx * qx + qy * qy + qz * qz);
        const double s2 = (d * 1e-10) * (d * 1e-10) * 0.25;
        double a = oracle_form_factor(d, Z[j], B, ff) * (amp ? (double)amp[j] : 1.0);
        if (atom_b) a *= exp(-(double)atom_b[j] * s2);
        const double t = qx * pos[3 * j] + qy * pos[3 * j + 1] + qz * pos[3 * j + 2];
        double c, s;
        cis_turns(t, &c, &s);
        /* e^{-i phi} = c - i s ; conj(F) = Fr - i Fi */
        const double Fr = F[2 * p], Fi = F[2 * p + 1];
        /* Re(conj(F) * (-i) e^{-i phi}) = -(Fr s + Fi c) */
        const double u = -(Fr * s + Fi * c);
        /* Re(conj(F) e^{-i phi}) = Fr c - Fi s */
        const double v = Fr * c - Fi * s;
        const double w = 2.0 * g[p] * a;
        gx += w * TWO_PI * qx * u;
        gy += w * TWO_PI * qy * u;
        gz += w * TWO_PI * qz * u;
        gb += w * (-s2) * v;
      }
    }
    grad_pos[3 * k] = gx;
    grad_pos[3 * k + 1] = gy;
    grad_pos[3 * k + 2] = gz;
    if (grad_b) grad_b[k] = gb;
  }
  return 0;
}

int oracle_sf_grad_f64(const float *hkl, int npix, const int *Z, const float *pos, const float *amp,
                       const float *atom_b, int natoms, const float *ff, float B, const double *g, double *grad_pos,
                       double *grad_b, int nthreads) {
  double *F = (double *)malloc(sizeof(double) * 2 * (size_t)(npix > 0 ? npix : 1));
  double *I = (double *)malloc(sizeof(double) * (size_t)(npix > 0 ? npix : 1));
  int *all = (int *)malloc(sizeof(int) * (size_t)(natoms > 0 ? natoms : 1));
  if (!F || !I || !all) {
    free(F);
    free(I);
    free(all);
    return -1;
  }
  oracle_sf_f64(hkl, npix, Z, pos, amp, atom_b, natoms, 0, 0, ff, B, F, I, nthreads, NULL);
  for (int j = 0; j < natoms; ++j) all[j] = j;
  oracle_sf_grad_given_f_f64(hkl, npix, F, Z, pos, amp, atom_b, ff, B, g, natoms, all, grad_pos, grad_b, nthreads);
  free(F);
  free(I);
  free(all);
  return 0;
}

/* ---------------------------------------------------------------- detector chain */

static void pixel_xy(int x, int y, int nx, int ny, float width, float height, float cx, float cy, double *px,
                     double *py) {
  *px = ((x - (nx - 1.0) / 2.0) / nx) * width - cx;
  *py = ((y - (ny - 1.0) / 2.0) / ny) * height - cy;
}

/* src/diffraction.c:1109-1144 */
void oracle_thomson(int nx, int ny, float width, float height, float distance, float cx, float cy, int polarization,
                    float *out) {
  const double DD = distance * distance;
  const double r0 = 2.81794e-15;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      double pf = 1.;
      if (polarization != 0) {
        double px, py;
        pixel_xy(x, y, nx, ny, width, height, cx, cy, &px, &py);
        const double r = sqrt(px * px + py * py + DD);
        double chi;
        if (polarization == 1) {
          chi = asin(py / r);
          pf = cos(chi) * cos(chi);
        } else if (polarization == 2) {
          chi = asin(px / r);
          pf = cos(chi) * cos(chi);
        } else if (polarization == 3) {
          chi = asin(sqrt(px * px + py * py) / r);
          pf = (1. + cos(chi) * cos(chi)) / 2.;
        }
      }
      out[(size_t)y * nx + x] = (r0 * r0) * pf;
    }
}

/* src/diffraction.c:1146-1187 (2-D) */
void oracle_solid_angle(int nx, int ny, float width, float height, float distance, float pixel_width,
                        float pixel_height, float cx, float cy, int spherical, float *out) {
  const double d2 = distance * distance;
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      if (spherical) {
        const double r = distance;
        out[(size_t)y * nx + x] = pixel_width * pixel_height / (r * r);
      } else {
        double px, py;
        pixel_xy(x, y, nx, ny, width, height, cx, cy, &px, &py);
        double inc = atan(sqrt(px * px + py * py) / distance);
        inc = cos(inc);
        inc = inc * inc * inc;
        out[(size_t)y * nx + x] = pixel_width * pixel_height / d2 * inc;
      }
    }
}

/* src/diffraction.c:1201-1203 */
void oracle_photons(const float *th, const float *om, const float *I, float beam, size_t n, float *out) {
  for (size_t i = 0; i < n; ++i) out[i] = th[i] * om[i] * I[i] * beam;
}

/* electrons per photon, src/amplification.c:56-61 (noisy) and :121-123 (noiseless). */
float oracle_electrons_per_photon(float wavelength, float e_h, int noiseless) {
  const double h_c = 1.98644521e-25;
  const double photon_energy = h_c / wavelength;
  float epp = 1;
  if (noiseless || e_h) epp = photon_energy / e_h;
  return epp;
}

static float adc_constant(float full_well, float max_value) {
  float adc = 1;
  if (full_well && max_value) adc = max_value / full_well;
  return adc;
}

/* src/amplification.c:119-153; nz = bz = 1 for a 2-D detector */
void oracle_noiseless_output(const float *photons, int nx, int ny, int nz, int bx, int by, int bz, float qe,
                             float epp, float full_well, float max_value, float *out) {
  const float adc = adc_constant(full_well, max_value);
  int i = 0;
  for (int z = 0; z < nz / bz; ++z)
    for (int y = 0; y < ny / by; ++y)
      for (int x = 0; x < nx / bx; ++x) {
        float acc = 0;
        for (int zi = 0; zi < bz; ++zi)
          for (int yi = 0; yi < by; ++yi)
            for (int xi = 0; xi < bx; ++xi)
              acc += photons[((size_t)(z * bz + zi) * ny + (y * by + yi)) * nx + (x * bx + xi)] * qe * epp * adc;
        out[i++] = acc;
      }
}

/* src/amplification.c:77-114 (2-D) with the per-binned-pixel readout noise
   supplied by the caller (noise[i] = the Gaussian draw in electrons). */
void oracle_real_output(const float *electrons, int nx, int ny, int nz, int bx, int by, int bz, float dark_current,
                        float exposure, const float *noise, float full_well, float max_value, float *out) {
  const float adc = adc_constant(full_well, max_value);
  int i = 0;
  for (int z = 0; z < nz / bz; ++z)
    for (int y = 0; y < ny / by; ++y)
      for (int x = 0; x < nx / bx; ++x) {
        float acc = 0;
        for (int zi = 0; zi < bz; ++zi)
          for (int yi = 0; yi < by; ++yi)
            for (int xi = 0; xi < bx; ++xi)
              acc += (electrons[((size_t)(z * bz + zi) * ny + (y * by + yi)) * nx + (x * bx + xi)] +
                      dark_current * exposure);
        acc += noise[i];
        if (acc < 0) acc = 0;
        acc *= adc;
        out[i++] = acc;
      }
}

/* ---------------------------------------------------------------- random numbers */

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static void uniforms(uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose, uint32_t attempt, double *u0,
                     double *u1) {
  const uint32_t c[4] = {idx, pattern, purpose, attempt};
  const uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  oracle_philox4x32_10(c, k, r);
  const double s = 1.0 / 9007199254740992.0;
  *u0 = ((double)((((uint64_t)(r[0] >> 5)) << 26) | (r[1] >> 6)) + 0.5) * s;
  *u1 = ((double)((((uint64_t)(r[2] >> 5)) << 26) | (r[3] >> 6)) + 0.5) * s;
}

double oracle_std_normal(uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose) {
  double u0, u1;
  uniforms(seed, idx, pattern, purpose, 0, &u0, &u1);
  return sqrt(-2.0 * log(u0)) * cos(TWO_PI * u1);
}

/* Poisson draw for the counters (idx, pattern), purpose 1. */
static double inversion(double mean, double u) {
  double p = exp(-mean), F = p;
  int k = 0;
  while (u > F && k < 256) {
    ++k;
    p *= mean / k;
    F += p;
  }
  return k;
}

/* PTRS (Hormann 1993) on the attempts att0 .. att0 + 63. */
static double ptrs(double mean, uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose, uint32_t att0) {
  const double slam = sqrt(mean), loglam = log(mean);
  const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
  for (uint32_t att = 0; att < 64; ++att) {
    double u, v;
    uniforms(seed, idx, pattern, purpose, att0 + att, &u, &v);
    u -= 0.5;
    const double us = 0.5 - fabs(u);
    const double k = floor((2.0 * a / us + b) * u + mean + 0.43);
    if (us >= 0.07 && v <= vr) return k;
    if (k < 0.0 || (us < 0.013 && v > us)) continue;
    if (log(v) + log(invalpha) - log(a / (us * us) + b) <= -mean + k * loglam - lgamma(k + 1.0)) return k;
  }
  return floor(mean);
}

/* Poisson(mean) with no Gaussian branch: inversion below 10 on the first
   uniform of attempt 0, PTRS above, on the counters (idx, pattern, purpose,
   attempt). */
double oracle_poisson_exact(double mean, uint64_t seed, uint32_t idx, uint32_t pattern, uint32_t purpose) {
  if (!(mean > 0.0)) return 0.0;
  if (mean >= 10.0) return ptrs(mean, seed, idx, pattern, purpose, 0);
  double u, unused;
  uniforms(seed, idx, pattern, purpose, 0, &u, &unused);
  return inversion(mean, u);
}

/* One pixel's count as src/noise.c:32-51 draws it: (int)(mean + sqrt(mean) N(0,1))
   truncated toward zero from mean 1000 up, Poisson below, purpose 1. */
double oracle_poisson(double mean, uint64_t seed, uint32_t idx, uint32_t pattern) {
  if (!(mean > 0.0)) return 0.0;
  if (mean >= 1000.0) {
    const double v = mean + sqrt(mean) * oracle_std_normal(seed, idx, pattern, 1);
    return v >= 2147483647.0 ? 2147483647.0 : trunc(v);
  }
  return oracle_poisson_exact(mean, seed, idx, pattern, 1);
}

/* Binned real output with one count per bin (spsim_hip_detector_chain_binned):
   per bin, the pixels with mean photons x QE >= 1000 take their own Gaussian
   count (purpose 1, flat raw index) as oracle_poisson does, and the others'
   counts are drawn together as one Poisson variable with their summed mean
   (FP32 sum in (z, y, x) order) from the block (bin, pattern, purpose 3,
   attempt 0): inversion on the 53-bit uniform of its words 0-1 below 10,
   PTRS on attempts 1..64 above. acc sums, in (z, y, x) order, count x
   electrons-per-photon + dark for the Gaussian pixels and dark for the
   others, then adds the joint count x electrons-per-photon and the readout
   noise, readout_noise x Box-Muller of the block's words 2 and 3 as uniforms
   (w + 1/2) 2^-32; clamp and ADC as src/amplification.c:104-108. */
void oracle_real_output_bin_draws(const float *photons, int nx, int ny, int nz, int bx, int by, int bz, float qe,
                                  float e_per_photon, float dark_current, float exposure, float readout_noise,
                                  float full_well, float max_value, uint64_t seed, uint32_t pattern, float *out) {
  const float adc = adc_constant(full_well, max_value);
  const float dark = dark_current * exposure;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t bin = 0;
  for (int z = 0; z < nz / bz; ++z)
    for (int y = 0; y < ny / by; ++y)
      for (int x = 0; x < nx / bx; ++x, ++bin) {
        float acc = 0, lam = 0;
        for (int zi = 0; zi < bz; ++zi)
          for (int yi = 0; yi < by; ++yi)
            for (int xi = 0; xi < bx; ++xi) {
              const size_t i = ((size_t)(z * bz + zi) * ny + (y * by + yi)) * nx + (x * bx + xi);
              const float mean = photons[i] * qe;
              if (mean >= 1000.0f) {
                const float cnt = (float)oracle_poisson((double)mean, seed, (uint32_t)i, pattern);
                acc += (cnt * e_per_photon + dark);
              } else {
                lam += mean;
                acc += dark;
              }
            }
        const uint32_t ctr[4] = {bin, pattern, 3, 0};
        uint32_t r[4];
        oracle_philox4x32_10(ctr, key, r);
        double n = 0.0;
        if (lam >= 10.0f) {
          n = ptrs((double)lam, seed, bin, pattern, 3, 1);
        } else if (lam > 0.0f) {
          const double u = ((double)((((uint64_t)(r[0] >> 5)) << 26) | (r[1] >> 6)) + 0.5) / 9007199254740992.0;
          n = inversion((double)lam, u);
        }
        acc += (float)n * e_per_photon;
        const double ua = ((double)r[2] + 0.5) / 4294967296.0, ub = ((double)r[3] + 0.5) / 4294967296.0;
        acc += (float)((double)readout_noise * (sqrt(-2.0 * log(ua)) * cos(TWO_PI * ub)));
        if (acc < 0) acc = 0;
        acc *= adc;
        out[bin] = acc;
      }
}
