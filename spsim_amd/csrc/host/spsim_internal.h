/* Internal helpers shared by the host C files of the drop-in library. */
#ifndef SPSIM_INTERNAL_H
#define SPSIM_INTERNAL_H 1

#include <stdint.h>

#include "config.h"
#include "spsim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* random.c */
uint64_t spsim_noise_seed(void);
uint64_t spsim_next_poisson_pattern(void);
uint64_t spsim_next_readout_pattern(void);
uint64_t spsim_next_shot_pattern(void);

/* diffraction.c: the form-factor table in use ([101][9], ./atomsf.lib applied). */
const float *spsim_form_factor_table(void);
/* Fills the detector-chain parameters from the reference structs; exp may be NULL. */
void spsim_detector_params(const CCD *det, const Experiment *exp, spsim_hip_detector *p);
/* Reports a failed device call on stderr (the reference API has no status codes). */
void spsim_report(const char *what, int rc);

/* rotation.c: row-major 3x3 rotation of a unit quaternion (w, x, y, z). */
void spsim_quaternion_to_matrix(const double q[4], double R[9]);
/* Shoemake's uniform random unit quaternion from three uniforms in [0,1). */
void spsim_uniform_quaternion(double u1, double u2, double u3, double q[4]);

#ifdef __cplusplus
}
#endif

#endif
