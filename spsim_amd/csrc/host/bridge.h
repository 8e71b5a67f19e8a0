/*
 * Host-array bridge between the reference-shaped C API (diffraction.c,
 * detector.c, spsim.c) and the device entry points of spsim_hip.h. Each call
 * copies its inputs to HBM, runs the kernels on an internal stream, copies the
 * results back and returns 0 or a negative SPSIM_HIP_ERR_* code; the message is
 * available from spsim_hip_last_error().
 */
#ifndef SPSIM_BRIDGE_H
#define SPSIM_BRIDGE_H 1

#include "spsim_hip.h"
#include "spsim_types.h"

#ifdef __cplusplus
extern "C" {
#endif

int spsim_bridge_hkl_detector(int nx, int ny, float width, float height, float distance, float center_x,
                              float center_y, int spherical, float wavelength, float *out);

int spsim_bridge_hkl_detector_3d(int nx, int ny, int nz, float width, float height, float depth, float center_x,
                                 float center_y, float center_z, float distance, float wavelength, float *out);

int spsim_bridge_rotate_hkl(float *list, int npix, const double *rot9);

/* F (may be NULL) and I receive the pattern of the packed atoms on the q list. */
int spsim_bridge_structure_factor(const float *hkl, int npix, const spsim_hip_atom *atoms, const float *atom_b,
                                  const spsim_hip_segment *segs, int nsegs, int natoms, const float *ff, float b_global,
                                  Complex *F, float *I);

/* Same under a wavelength spectrum (spsim_hip_sf_spectrum); scales and weights
   are nsamples host values. */
int spsim_bridge_structure_factor_spectrum(const float *hkl, int npix, const spsim_hip_atom *atoms,
                                           const float *atom_b, const spsim_hip_segment *segs, int nsegs,
                                           int natoms, const float *ff, float b_global, int nsamples,
                                           const double *scales, const double *weights, Complex *F, float *I);

/* Same on a full nx x ny detector list ([ny][nx][3]); uses the mirror-symmetric
   kernel where the list allows it. */
int spsim_bridge_structure_factor_grid(const float *hkl, int nx, int ny, const spsim_hip_atom *atoms,
                                       const float *atom_b, const spsim_hip_segment *segs, int nsegs, int natoms,
                                       const float *ff, float b_global, Complex *F, float *I);

/* Same by the non-uniform FFT (spsim_hip_sf_nfft) with relative accuracy eps. */
int spsim_bridge_structure_factor_nfft(const float *hkl, int npix, const spsim_hip_atom *atoms,
                                       const spsim_hip_segment *segs, int nsegs, int natoms, const float *ff,
                                       float b_global, double eps, Complex *F, float *I);

/* Detector stages. Inputs are host arrays of nx*ny raw pixels; binned outputs
   have (nx/bin_x)*(ny/bin_y) entries. */
int spsim_bridge_thomson(const spsim_hip_detector *det, float *out);
int spsim_bridge_solid_angle(const spsim_hip_detector *det, float *out);
int spsim_bridge_photons(const spsim_hip_detector *det, const float *thomson, const float *solid_angle,
                         const float *I, float *out);
int spsim_bridge_poisson(const spsim_hip_detector *det, const float *photons, uint64_t seed, uint64_t pattern,
                         float *out);
int spsim_bridge_electrons(const spsim_hip_detector *det, const float *count, float *out);
int spsim_bridge_real_output(const spsim_hip_detector *det, const float *electrons, uint64_t seed, uint64_t pattern,
                             float *out);
int spsim_bridge_noiseless_output(const spsim_hip_detector *det, const float *photons, float *out);
/* All stages in one pass (spsim_hip_detector_chain); any output may be NULL. */
int spsim_bridge_detector_chain(const spsim_hip_detector *det, const float *I, uint64_t seed, uint64_t pattern,
                                float *thomson, float *solid_angle, float *photons, float *count, float *electrons,
                                float *real_output, float *noiseless_output);

#ifdef __cplusplus
}
#endif

#endif
