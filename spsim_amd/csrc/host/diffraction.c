/*
 * Drop-in structure-factor API (reference src/diffraction.c). The arithmetic
 * of every function here runs on the GPU (bridge.h -> spsim_hip.h); this file
 * keeps the reference's memory ownership (callee mallocs, caller frees), its
 * global form-factor table with the ./atomsf.lib override, and its stderr
 * reporting. A failed device call prints the reason and returns NULL (or
 * leaves the output array unset) instead of computing on the CPU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bridge.h"
#include "cromer_mann.h"
#include "diffraction.h"
#include "spsim_internal.h"

static float g_ff[SPSIM_HIP_NSPECIES][SPSIM_HIP_FF_STRIDE];
static int g_ff_initialized = 0;

void spsim_report(const char *what, int rc) {
  fprintf(stderr, "spsim: %s failed (%d): %s\n", what, rc, spsim_hip_last_error());
}

/*
 * ./atomsf.lib override (src/diffraction.c:118-177). Same record layout as the
 * reference reads it: comment lines starting with "AD"; then per element a name
 * line, a line whose first integer is Z with c at column 23, a line of a1..a4
 * and one of b1..b4 in 16-character columns, and one more line. Rows named in
 * the file are replaced; every other row keeps its coefficients except a1,
 * which is set to -1 as the reference does.
 */
static void read_atomsf_lib(void) {
  char line[1024];
  FILE *f = fopen("atomsf.lib", "r");
  if (!f) return;
  int seen[SPSIM_CM_ROWS];
  for (int z = 0; z < SPSIM_CM_ROWS; ++z) {
    g_ff[z][0] = -1.0f;
    seen[z] = 0;
  }
  line[0] = 0;
  while (fgets(line, sizeof line, f) && strncmp(line, "AD", 2) == 0) {
  }
  /* `line` now holds the name line of the first record (or is stale at EOF). */
  int more = !feof(f) || line[0];
  while (more) {
    char l2[1024] = {0}, la[1024] = {0}, lb[1024] = {0}, l5[1024] = {0};
    int z = -1;
    if (!fgets(l2, sizeof l2, f)) break;
    sscanf(l2, "%d", &z);
    if (!fgets(la, sizeof la, f) || !fgets(lb, sizeof lb, f)) break;
    if (!fgets(l5, sizeof l5, f)) l5[0] = 0;
    if (z >= 0 && z < SPSIM_CM_ROWS && !seen[z]) {
      float v = 0.0f;
      if (strlen(l2) > 23 && sscanf(l2 + 23, "%f", &v) == 1) g_ff[z][8] = v;
      for (int i = 0; i < 4; ++i) {
        if (strlen(la) > (size_t)(16 * i) && sscanf(la + 16 * i, "%f", &v) == 1) g_ff[z][i] = v;
        if (strlen(lb) > (size_t)(16 * i) && sscanf(lb + 16 * i, "%f", &v) == 1) g_ff[z][4 + i] = v;
      }
      seen[z] = 1;
    }
    line[0] = 0;
    more = fgets(line, sizeof line, f) != NULL && line[0];
  }
  fclose(f);
}

const float *spsim_form_factor_table(void) {
  if (!g_ff_initialized) {
    spsim_hip_default_form_factors(&g_ff[0][0]);
    read_atomsf_lib();
    g_ff_initialized = 1;
  }
  return &g_ff[0][0];
}

/* src/diffraction.c:223-235 */
static double illumination_function(const Experiment *exper, const float *pos) {
  if (!exper || !exper->beam_fwhm) return 1;
  const double dist2 = (pos[0] - exper->beam_center_x) * (pos[0] - exper->beam_center_x) +
                       (pos[1] - exper->beam_center_y) * (pos[1] - exper->beam_center_y);
  const double sigma = exper->beam_fwhm / 2.355;
  return exp(-dist2 / (2 * sigma * sigma));
}

float *get_HKL_list_for_detector(CCD *det, Experiment *exp, int *HKL_list_size) {
  if (!det || !exp || !HKL_list_size) return NULL;
  const int nx = det->nx, ny = det->ny;
  float *list = (float *)malloc(sizeof(float) * 3 * (size_t)(nx > 0 ? nx : 0) * (ny > 0 ? ny : 0) + 1);
  if (!list) return NULL;
  int rc = spsim_bridge_hkl_detector(nx, ny, det->width, det->height, det->distance, det->center_x, det->center_y,
                                     det->spherical, exp->wavelength, list);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("get_HKL_list_for_detector", rc);
    free(list);
    return NULL;
  }
  *HKL_list_size = nx * ny;
  return list;
}

float *get_HKL_list_for_3d_detector(CCD *det, Experiment *exp, int *HKL_list_size) {
  if (!det || !exp || !HKL_list_size) return NULL;
  const int nx = det->nx, ny = det->ny, nz = det->nz;
  const size_t n = (size_t)(nx > 0 ? nx : 0) * (ny > 0 ? ny : 0) * (nz > 0 ? nz : 0);
  float *list = (float *)malloc(sizeof(float) * 3 * n + 1);
  if (!list) return NULL;
  int rc = spsim_bridge_hkl_detector_3d(nx, ny, nz, det->width, det->height, det->depth, det->center_x,
                                        det->center_y, det->center_z, det->distance, exp->wavelength, list);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("get_HKL_list_for_3d_detector", rc);
    free(list);
    return NULL;
  }
  *HKL_list_size = (int)n;
  return list;
}

SpRotation *apply_orientation_to_HKL_list(float **HKL_list, int *HKL_list_size, Options *opts) {
  SpRotation *rot;
  double R[9];
  if (!HKL_list || !HKL_list_size || !opts) return NULL;
  if (opts->random_orientation) {
    rot = sp_rot_uniform();
  } else {
    /* intrinsic rotation: the coordinate system is rotated, not the sample */
    rot = sp_rot_euler(opts->euler_orientation[0], opts->euler_orientation[1], opts->euler_orientation[2]);
  }
  if (!rot) return NULL;
  for (int i = 0; i < 9; ++i) R[i] = (double)rot->data[i];
  int rc = spsim_bridge_rotate_hkl(*HKL_list, *HKL_list_size, R);
  if (rc != SPSIM_HIP_OK) spsim_report("apply_orientation_to_HKL_list", rc);
  return rot;
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Detector dimensions of a q list when they are known (nx*ny == size): the
   caller's CCD when it matches, otherwise a square list. The grid kernel
   checks the mirror symmetry it uses, so a wrong guess only costs speed. */
static void grid_dims(int size, const Options *opts, int *nx, int *ny) {
  *nx = *ny = 0;
  if (opts && opts->detector && opts->detector->nx > 0 && opts->detector->ny > 0 &&
      (long long)opts->detector->nx * opts->detector->ny == size) {
    *nx = opts->detector->nx;
    *ny = opts->detector->ny;
    return;
  }
  int n = (int)sqrt((double)size);
  while ((long long)n * n > size) --n;
  while ((long long)(n + 1) * (n + 1) <= size) ++n;
  if ((long long)n * n == size) *nx = *ny = n;
}

/* nsamples > 0 selects the wavelength-spectrum kernel with the given phase
   scales and weights (delta_atoms is then ignored, as the reference's spectrum
   kernel ignores it). */
static Diffraction_Pattern *compute_pattern(Molecule *mol, float *HKL_list, int HKL_list_size, float B,
                                            Experiment *exp, Options *opts, int mode, int nsamples,
                                            const double *scales, const double *weights) {
  const double t0 = now_seconds();
  if (!mol || HKL_list_size < 0 || (HKL_list_size > 0 && !HKL_list) || mol->natoms < 0) return NULL;
  const int N = mol->natoms;
  const int delta = (opts && nsamples == 0) ? opts->delta_atoms : 0;
  Diffraction_Pattern *res = (Diffraction_Pattern *)calloc(1, sizeof(Diffraction_Pattern));
  float *amp = (float *)malloc(sizeof(float) * (N > 0 ? N : 1));
  spsim_hip_atom *atoms = (spsim_hip_atom *)malloc(sizeof(spsim_hip_atom) * (N > 0 ? N : 1));
  spsim_hip_segment segs[SPSIM_HIP_NSPECIES];
  if (!res || !amp || !atoms) goto fail;
  res->F = (Complex *)malloc(sizeof(Complex) * (size_t)HKL_list_size + 1);
  res->ints = (float *)malloc(sizeof(float) * (size_t)HKL_list_size + 1);
  res->HKL_list = (float *)malloc(sizeof(float) * 3 * (size_t)HKL_list_size + 1);
  if (!res->F || !res->ints || !res->HKL_list) goto fail;
  if (HKL_list_size) memcpy(res->HKL_list, HKL_list, sizeof(float) * 3 * (size_t)HKL_list_size);
  res->HKL_list_size = HKL_list_size;
  res->rot = NULL;

  for (int j = 0; j < N; ++j) {
    const double il = illumination_function(exp, &mol->pos[3 * j]);
    /* plain path: sqrt(illumination) (src/diffraction.c:834); the vector path
       stores sqrt(illumination) and takes sqrt again (:886, :919-922). */
    amp[j] = (mode == SPSIM_HIP_PACK_VECTOR) ? (float)sqrt((float)sqrt(il)) : (float)sqrt((float)il);
  }
  int natoms = 0, nsegs = 0;
  int rc = spsim_hip_pack_atoms(mol->atomic_number, mol->pos, amp, NULL, N, delta, mode, atoms, NULL, NULL, segs,
                                &natoms, &nsegs);
  if (rc != SPSIM_HIP_OK) {
    fprintf(stderr, "spsim: compute_pattern_on_list: atomic numbers must lie in [0, 99]\n");
    goto fail;
  }
  int gx = 0, gy = 0;
  grid_dims(HKL_list_size, opts, &gx, &gy);
  if (nsamples > 0)
    rc = spsim_bridge_structure_factor_spectrum(HKL_list, HKL_list_size, atoms, NULL, segs, nsegs, natoms,
                                                spsim_form_factor_table(), B, nsamples, scales, weights, res->F,
                                                res->ints);
  else if (gx > 0)
    rc = spsim_bridge_structure_factor_grid(HKL_list, gx, gy, atoms, NULL, segs, nsegs, natoms,
                                            spsim_form_factor_table(), B, res->F, res->ints);
  else
    rc = spsim_bridge_structure_factor(HKL_list, HKL_list_size, atoms, NULL, segs, nsegs, natoms,
                                       spsim_form_factor_table(), B, res->F, res->ints);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("compute_pattern_on_list", rc);
    goto fail;
  }
  free(amp);
  free(atoms);
  if (opts && opts->verbosity_level > 0) {
    const double dt = now_seconds() - t0;
    printf("%g atoms.pixel/s\n", (double)HKL_list_size * N / (dt > 0 ? dt : 1e-9));
  }
  return res;
fail:
  if (res) {
    free(res->F);
    free(res->ints);
    free(res->HKL_list);
    free(res);
  }
  free(amp);
  free(atoms);
  return NULL;
}

Diffraction_Pattern *compute_pattern_on_list(Molecule *mol, float *HKL_list, int HKL_list_size, float B,
                                             Experiment *exp, Options *opts) {
  return compute_pattern(mol, HKL_list, HKL_list_size, B, exp, opts, SPSIM_HIP_PACK_PLAIN, 0, NULL, NULL);
}

Diffraction_Pattern *vector_compute_pattern_on_list(Molecule *mol, float *HKL_list, int HKL_list_size, float B,
                                                    Experiment *exp, Options *opts) {
  return compute_pattern(mol, HKL_list, HKL_list_size, B, exp, opts, SPSIM_HIP_PACK_VECTOR, 0, NULL, NULL);
}

Diffraction_Pattern *spsim_spectrum_pattern_on_list(Molecule *mol, float *HKL_list, int HKL_list_size, float B,
                                                      Experiment *exp, Options *opts) {
  double scales[SPSIM_HIP_MAX_SPECTRUM], weights[SPSIM_HIP_MAX_SPECTRUM];
  if (!exp || !opts) return NULL;
  /* the reference raises an even sample count in the caller's options (src/diffraction.cu:302-304) */
  if (opts->wavelength_samples > 1 && (opts->wavelength_samples & 1) == 0) opts->wavelength_samples++;
  const int n = spsim_hip_spectrum_samples(exp->wavelength, exp->bandwidth, opts->wavelength_samples, scales, weights);
  if (n < 1) {
    fprintf(stderr, "spsim: spsim_spectrum_pattern_on_list: wavelength_samples must lie in [1, %d] and the "
                    "wavelength must be positive\n", SPSIM_HIP_MAX_SPECTRUM);
    return NULL;
  }
  return compute_pattern(mol, HKL_list, HKL_list_size, B, exp, opts, SPSIM_HIP_PACK_PLAIN, n, scales, weights);
}

Diffraction_Pattern *compute_pattern_on_list_by_nfft(Molecule *mol, float *HKL_list, int HKL_list_size, CCD *det,
                                                     float B, Options *opts) {
  (void)det;
  if (!mol || HKL_list_size < 0 || (HKL_list_size > 0 && !HKL_list) || mol->natoms < 0) return NULL;
  const int N = mol->natoms;
  const int delta = opts ? opts->delta_atoms : 0;
  Diffraction_Pattern *res = (Diffraction_Pattern *)calloc(1, sizeof(Diffraction_Pattern));
  spsim_hip_atom *atoms = (spsim_hip_atom *)malloc(sizeof(spsim_hip_atom) * (N > 0 ? N : 1));
  spsim_hip_segment segs[SPSIM_HIP_NSPECIES];
  if (!res || !atoms) goto fail;
  res->F = (Complex *)malloc(sizeof(Complex) * (size_t)HKL_list_size + 1);
  res->ints = (float *)malloc(sizeof(float) * (size_t)HKL_list_size + 1);
  res->HKL_list = (float *)malloc(sizeof(float) * 3 * (size_t)HKL_list_size + 1);
  if (!res->F || !res->ints || !res->HKL_list) goto fail;
  if (HKL_list_size) memcpy(res->HKL_list, HKL_list, sizeof(float) * 3 * (size_t)HKL_list_size);
  res->HKL_list_size = HKL_list_size;
  res->rot = NULL;
  int natoms = 0, nsegs = 0;
  int rc = spsim_hip_pack_atoms(mol->atomic_number, mol->pos, NULL, NULL, N, delta, SPSIM_HIP_PACK_NFFT, atoms, NULL,
                                NULL, segs, &natoms, &nsegs);
  if (rc != SPSIM_HIP_OK) {
    fprintf(stderr, "spsim: compute_pattern_on_list_by_nfft: atomic numbers must lie in [0, 99]\n");
    goto fail;
  }
  rc = spsim_bridge_structure_factor_nfft(HKL_list, HKL_list_size, atoms, segs, nsegs, natoms,
                                          spsim_form_factor_table(), B, 1e-6, res->F, res->ints);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("compute_pattern_on_list_by_nfft", rc);
    goto fail;
  }
  free(atoms);
  return res;
fail:
  if (res) {
    free(res->F);
    free(res->ints);
    free(res->HKL_list);
    free(res);
  }
  free(atoms);
  return NULL;
}

void spsim_detector_params(const CCD *det, const Experiment *exp, spsim_hip_detector *p) {
  memset(p, 0, sizeof *p);
  p->nx = det->nx;
  p->ny = det->ny;
  p->bin_x = det->binning_x > 0 ? det->binning_x : 1;
  p->bin_y = det->binning_y > 0 ? det->binning_y : 1;
  p->width = det->width;
  p->height = det->height;
  p->distance = det->distance;
  p->pixel_width = det->pixel_width;
  p->pixel_height = det->pixel_height;
  p->center_x = det->center_x;
  p->center_y = det->center_y;
  /* a 3-D detector uses the spherical (constant) solid angle (src/diffraction.c:1171) */
  p->spherical = det->spherical || det->nz > 1;
  p->quantum_efficiency = det->quantum_efficiency;
  p->electron_hole_production_energy = det->electron_hole_production_energy;
  p->dark_current = det->dark_current;
  p->readout_noise = det->readout_noise;
  p->linear_full_well = det->linear_full_well;
  p->maximum_value = det->maximum_value;
  p->nz = det->nz > 0 ? det->nz : 1;
  p->bin_z = det->binning_z > 0 ? det->binning_z : 1;
  if (exp) {
    p->polarization = exp->polarization;
    p->beam_intensity = exp->beam_intensity;
    p->wavelength = exp->wavelength;
    p->exposure_time = exp->exposure_time;
  }
}

static size_t raw_pixels(const CCD *det) {
  const int nz = det->nz > 0 ? det->nz : 1;
  return (size_t)(det->nx > 0 ? det->nx : 0) * (det->ny > 0 ? det->ny : 0) * nz;
}

void calculate_thomson_correction(CCD *det, Experiment *exp) {
  spsim_hip_detector p;
  if (!det) return;
  spsim_detector_params(det, exp, &p);
  float *out = (float *)malloc(sizeof(float) * raw_pixels(det) + 1);
  int rc = spsim_bridge_thomson(&p, out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("calculate_thomson_correction", rc);
    free(out);
    return;
  }
  det->thomson_correction = out;
}

void calculate_pixel_solid_angle(CCD *det) {
  spsim_hip_detector p;
  if (!det) return;
  spsim_detector_params(det, NULL, &p);
  float *out = (float *)malloc(sizeof(float) * raw_pixels(det) + 1);
  int rc = spsim_bridge_solid_angle(&p, out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("calculate_pixel_solid_angle", rc);
    free(out);
    return;
  }
  det->solid_angle = out;
}

void calculate_photons_per_pixel(Diffraction_Pattern *pattern, Options *opts) {
  spsim_hip_detector p;
  if (!pattern || !opts || !opts->detector) return;
  CCD *det = opts->detector;
  if (!det->thomson_correction) calculate_thomson_correction(det, opts->experiment);
  if (!det->solid_angle) calculate_pixel_solid_angle(det);
  if (!det->thomson_correction || !det->solid_angle) return;
  spsim_detector_params(det, opts->experiment, &p);
  float *out = (float *)malloc(sizeof(float) * raw_pixels(det) + 1);
  int rc = spsim_bridge_photons(&p, det->thomson_correction, det->solid_angle, pattern->ints, out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("calculate_photons_per_pixel", rc);
    free(out);
    return;
  }
  det->photons_per_pixel = out;
}
