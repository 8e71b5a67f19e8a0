/*
 * Drop-in detector stages (reference src/noise.c:32-51 and
 * src/amplification.c:54-153). Each stage reads the previous stage's array
 * from the CCD struct, runs on the GPU and stores a newly malloc'd result, as
 * the reference does. Noise is drawn from the library's Philox stream keyed by
 * sp_srand(); each call uses the next pattern counter of its stage.
 */
#include <stdio.h>
#include <stdlib.h>

#include "amplification.h"
#include "bridge.h"
#include "noise.h"
#include "spsim_internal.h"

static size_t raw_pixels(const CCD *det) {
  return (size_t)det->nx * det->ny * (det->nz > 0 ? det->nz : 1);
}

/* (nx/bx)*(ny/by)*(nz/bz), src/amplification.c:88,133 */
static size_t binned_pixels(const CCD *det) {
  const int bx = det->binning_x > 0 ? det->binning_x : 1, by = det->binning_y > 0 ? det->binning_y : 1;
  const int nz = det->nz > 0 ? det->nz : 1, bz = det->binning_z > 0 ? det->binning_z : 1;
  return (size_t)(det->nx / bx) * (det->ny / by) * (nz / bz);
}

void generate_poisson_noise(CCD *det) {
  spsim_hip_detector p;
  if (!det) return;
  if (!det->photons_per_pixel) {
    fprintf(stderr, "Please calculate photons per pixel first\n");
    return;
  }
  spsim_detector_params(det, NULL, &p);
  float *out = (float *)malloc(sizeof(float) * raw_pixels(det) + 1);
  int rc = spsim_bridge_poisson(&p, det->photons_per_pixel, spsim_noise_seed(), spsim_next_poisson_pattern(), out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("generate_poisson_noise", rc);
    free(out);
    return;
  }
  det->photon_count = out;
}

void calculate_electrons_per_pixel(CCD *det, Experiment *experiment) {
  spsim_hip_detector p;
  if (!det) return;
  if (!det->photon_count) {
    fprintf(stderr, "Calculate photon_count first\n");
    return;
  }
  spsim_detector_params(det, experiment, &p);
  float *out = (float *)malloc(sizeof(float) * raw_pixels(det) + 1);
  int rc = spsim_bridge_electrons(&p, det->photon_count, out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("calculate_electrons_per_pixel", rc);
    free(out);
    return;
  }
  det->electrons_per_pixel = out;
}

void calculate_real_detector_output(CCD *det, Experiment *experiment) {
  spsim_hip_detector p;
  if (!det) return;
  if (!det->electrons_per_pixel) {
    fprintf(stderr, "Calculate electrons_per_pixel first\n");
    return;
  }
  spsim_detector_params(det, experiment, &p);
  float *out = (float *)malloc(sizeof(float) * binned_pixels(det) + 1);
  int rc = spsim_bridge_real_output(&p, det->electrons_per_pixel, spsim_noise_seed(), spsim_next_readout_pattern(), out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("calculate_real_detector_output", rc);
    free(out);
    return;
  }
  det->real_output = out;
}

void calculate_noiseless_detector_output(CCD *det, Experiment *experiment) {
  spsim_hip_detector p;
  if (!det) return;
  if (!det->photons_per_pixel) {
    fprintf(stderr, "Calculate photons_per_pixel first\n");
    return;
  }
  spsim_detector_params(det, experiment, &p);
  float *out = (float *)malloc(sizeof(float) * binned_pixels(det) + 1);
  int rc = spsim_bridge_noiseless_output(&p, det->photons_per_pixel, out);
  if (rc != SPSIM_HIP_OK) {
    spsim_report("calculate_noiseless_detector_output", rc);
    free(out);
    return;
  }
  det->noiseless_output = out;
}
