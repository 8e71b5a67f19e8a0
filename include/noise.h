/*
 * spsim-mi355x drop-in: detector photon noise (reference include/noise.h:32-38).
 */
#ifndef _NOISE_H_
#define _NOISE_H_ 1

#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/noise.c:32-51: photon_count = Poisson(photons_per_pixel * QE) for means
   below 1000, int(mean + N(0, sqrt(mean))) truncated toward zero above. */
void generate_poisson_noise(CCD *det);

#ifdef __cplusplus
}
#endif

#endif
