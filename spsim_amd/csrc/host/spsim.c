/*
 * Drop-in pipeline driver (reference src/spsim.c:84-371).
 *
 * simulate_shot keeps the reference's order (q list, orientation, structure
 * factor, detector chain) and fills the same CCD arrays, but runs the whole
 * detector chain as one fused GPU pass (spsim_hip_detector_chain) that is
 * bit-identical to calling the stages one by one with the same noise counters.
 * As in the reference, the previous shot's CCD arrays are replaced without
 * being freed; callers release them with free_output_in_options.
 */
#include <stdio.h>
#include <stdlib.h>

#include "bridge.h"
#include "spsim.h"
#include "spsim_internal.h"

/* src/spsim.c:84-97: a random formula molecule (also written to molout.pdb
 * in the working directory) or a PDB file; exits when input_type is neither. */
Molecule *get_molecule(Options *opts) {
  Molecule *mol = NULL;
  if (!opts) return NULL;
  if (opts->input_type == CHEM_FORMULA) {
    mol = get_Molecule_from_formula(opts->chem_formula, opts);
    if (mol) write_pdb_from_mol("molout.pdb", mol);
  } else if (opts->input_type == PDB) {
    mol = get_Molecule_from_pdb(opts->pdb_filename);
  } else {
    fprintf(stderr, "input_type not specified!\n");
    exit(0);
  }
  return mol;
}

float *get_HKL_list(Options *opts, int *HKL_list_size) {
  if (!opts || !opts->detector) return NULL;
  if (opts->detector->nz > 1) return get_HKL_list_for_3d_detector(opts->detector, opts->experiment, HKL_list_size);
  return get_HKL_list_for_detector(opts->detector, opts->experiment, HKL_list_size);
}

Diffraction_Pattern *compute_sf(Molecule *mol, float *HKL_list, int HKL_list_size, Options *opts) {
  Diffraction_Pattern *pattern = NULL;
  if (!opts) return NULL;
  if (opts->use_fft_for_sf) {
    fprintf(stderr, "spsim: compute_sf: the real-space FFT structure factor is not supported\n");
    return NULL;
  }
  /* src/spsim.c:119-122 (a build with NFFT_SUPPORT) */
  if (opts->use_nfft_for_sf) {
    pattern = compute_pattern_on_list_by_nfft(mol, HKL_list, HKL_list_size, opts->detector, opts->b_factor, opts);
    if (pattern) pattern->rot = NULL;
    return pattern;
  }
  /* use_cuda / vectorize select the same GPU kernel; vectorize keeps the
     vector path's semantics (src/spsim.c:127-140). With use_cuda, a non-zero
     bandwidth and more than one wavelength sample, the reference's GPU path
     averages over the spectrum (src/diffraction.cu:313-318). */
  if (opts->use_cuda && opts->experiment && opts->experiment->bandwidth != 0 && opts->wavelength_samples > 1)
    pattern = spsim_spectrum_pattern_on_list(mol, HKL_list, HKL_list_size, opts->b_factor, opts->experiment, opts);
  else if (opts->vectorize && !opts->use_cuda)
    pattern = vector_compute_pattern_on_list(mol, HKL_list, HKL_list_size, opts->b_factor, opts->experiment, opts);
  else
    pattern = compute_pattern_on_list(mol, HKL_list, HKL_list_size, opts->b_factor, opts->experiment, opts);
  if (pattern) pattern->rot = NULL;
  return pattern;
}

static Image *fill_image_meta(Image *img, SpRotation *rot, Options *opts) {
  img->detector->wavelength = opts->experiment->wavelength;
  img->detector->pixel_size[0] = opts->detector->pixel_width * opts->detector->binning_x;
  img->detector->pixel_size[1] = opts->detector->pixel_height * opts->detector->binning_y;
  img->detector->pixel_size[2] = opts->detector->pixel_depth * opts->detector->binning_z;
  img->detector->detector_distance = opts->detector->distance;
  if (rot) img->detector->orientation = rot;
  return img;
}

static int image_dims(const Options *opts, int *x, int *y, int *z) {
  if (!opts || !opts->detector || !opts->experiment) return 0;
  *x = opts->detector->nx;
  *y = opts->detector->ny;
  *z = opts->detector->nz > 0 ? opts->detector->nz : 1;
  return 1;
}

Image *make_image(float *array, SpRotation *rot, Options *opts) {
  int x, y, z;
  if (!array || !image_dims(opts, &x, &y, &z)) return NULL;
  Image *img = sp_image_alloc(x, y, z);
  if (!img) return NULL;
  array_to_image(array, img);
  return fill_image_meta(img, rot, opts);
}

Image *make_cimage(Complex *array, SpRotation *rot, Options *opts) {
  int x, y, z;
  if (!array || !image_dims(opts, &x, &y, &z)) return NULL;
  Image *img = sp_image_alloc(x, y, z);
  if (!img) return NULL;
  const size_t n = (size_t)x * y * z;
  for (size_t i = 0; i < n; ++i) {
    img->image->data[i] = array[i];
    img->mask->data[i] = 1;
  }
  return fill_image_meta(img, rot, opts);
}

void array_to_image(float *array, Image *img) {
  if (!array || !img || !img->image) return;
  const size_t n = (size_t)img->image->x * img->image->y * img->image->z;
  for (size_t i = 0; i < n; ++i) {
    img->image->data[i].re = array[i];
    img->image->data[i].im = 0.0f;
    img->mask->data[i] = 1;
  }
}

void iarray_to_image(int *array, Image *img) {
  if (!array || !img || !img->image) return;
  const size_t n = (size_t)img->image->x * img->image->y * img->image->z;
  for (size_t i = 0; i < n; ++i) {
    img->image->data[i].re = (float)array[i];
    img->image->data[i].im = 0.0f;
    img->mask->data[i] = 1;
  }
}

void gaussian_blur_pattern(Options *opts, Diffraction_Pattern *pattern) {
  (void)pattern;
  if (opts && opts->detector && opts->detector->gaussian_blurring)
    fprintf(stderr, "spsim: gaussian_blurring is not supported; the pattern is left unblurred\n");
}

Diffraction_Pattern *simulate_shot(Molecule *mol, Options *opts) {
  int HKL_list_size = 0;
  if (!mol || !opts || !opts->detector || !opts->experiment) return NULL;
  if (opts->sf_filename[0]) {
    fprintf(stderr, "spsim: simulate_shot: loading patterns from sf_filename is not supported\n");
    return NULL;
  }
  float *HKL_list = get_HKL_list(opts, &HKL_list_size);
  if (!HKL_list) return NULL;
  SpRotation *rot = apply_orientation_to_HKL_list(&HKL_list, &HKL_list_size, opts);
  Diffraction_Pattern *pattern = compute_sf(mol, HKL_list, HKL_list_size, opts);
  free(HKL_list);
  if (!pattern) {
    free(rot);
    return NULL;
  }
  pattern->rot = rot;
  gaussian_blur_pattern(opts, pattern);

  CCD *det = opts->detector;
  spsim_hip_detector p;
  spsim_detector_params(det, opts->experiment, &p);
  const int nz = p.nz > 0 ? p.nz : 1, bz = p.bin_z > 0 ? p.bin_z : 1;
  const size_t n = (size_t)det->nx * det->ny * nz;
  const size_t nb = (size_t)(det->nx / p.bin_x) * (det->ny / p.bin_y) * (nz / bz);
  float *th = (float *)malloc(sizeof(float) * n + 1), *om = (float *)malloc(sizeof(float) * n + 1);
  float *ph = (float *)malloc(sizeof(float) * n + 1), *cnt = (float *)malloc(sizeof(float) * n + 1);
  float *el = (float *)malloc(sizeof(float) * n + 1);
  float *real = (float *)malloc(sizeof(float) * nb + 1), *nl = (float *)malloc(sizeof(float) * nb + 1);
  int rc = SPSIM_HIP_ERR_NOMEM;
  if (th && om && ph && cnt && el && real && nl)
    rc = spsim_bridge_detector_chain(&p, pattern->ints, spsim_noise_seed(), spsim_next_shot_pattern(), th, om, ph, cnt,
                                     el, real, nl);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("simulate_shot (detector chain)", rc);
    free(th);
    free(om);
    free(ph);
    free(cnt);
    free(el);
    free(real);
    free(nl);
    return pattern;
  }
  det->thomson_correction = th;
  det->solid_angle = om;
  det->photons_per_pixel = ph;
  det->photon_count = cnt;
  det->electrons_per_pixel = el;
  det->real_output = real;
  det->noiseless_output = nl;
  return pattern;
}

void free_diffraction_pattern(Diffraction_Pattern *pattern) {
  if (!pattern) return;
  free(pattern->F);
  free(pattern->ints);
  free(pattern->HKL_list);
  free(pattern->rot);
  free(pattern);
}

void free_output_in_options(Options *opts) {
  if (!opts || !opts->detector) return;
  CCD *d = opts->detector;
  free(d->photons_per_pixel);
  free(d->thomson_correction);
  free(d->solid_angle);
  free(d->photon_count);
  free(d->electrons_per_pixel);
  free(d->real_output);
  free(d->noiseless_output);
  d->photons_per_pixel = d->thomson_correction = d->solid_angle = d->photon_count = NULL;
  d->electrons_per_pixel = d->real_output = d->noiseless_output = NULL;
}

void free_stuff(Diffraction_Pattern *pattern, Options *opts) {
  free_diffraction_pattern(pattern);
  free_output_in_options(opts);
}
