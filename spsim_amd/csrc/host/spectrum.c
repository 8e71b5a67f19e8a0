/*
 * Wavelength samples of a Gaussian spectrum (spsim_hip_spectrum_samples in
 * spsim_hip.h), following the reference's spectrum path
 * (src/diffraction.cu:302-303 raises an even sample count by one; :431-460
 * place the samples and weight them).
 */
#include <math.h>

#include "spsim_hip.h"

int spsim_hip_spectrum_samples(double wavelength, double bandwidth, int nsamples, double *scales, double *weights) {
  if (!scales || !weights) return SPSIM_HIP_ERR_INVALID;
  if (nsamples > 1 && (nsamples % 2) == 0) nsamples++;
  if (bandwidth == 0.0 || nsamples <= 1) {
    scales[0] = 1.0;
    weights[0] = 1.0;
    return 1;
  }
  if (nsamples > SPSIM_HIP_MAX_SPECTRUM || !(wavelength > 0.0)) return SPSIM_HIP_ERR_INVALID;
  /* the sampled range lambda (1 +- bandwidth/2) spans +-2 sigma */
  const double first = wavelength * (1.0 - 0.5 * bandwidth);
  const double sigma = 0.25 * wavelength * bandwidth;
  double sum = 0.0;
  for (int k = 0; k < nsamples; ++k) {
    const double lk = first + wavelength * bandwidth * ((double)k / (double)(nsamples - 1));
    const double x = (lk - wavelength) / sigma;
    weights[k] = exp(-0.5 * x * x);
    scales[k] = wavelength / lk;
    sum += weights[k];
  }
  for (int k = 0; k < nsamples; ++k) weights[k] /= sum;
  return nsamples;
}
