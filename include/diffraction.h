/*
 * spsim-mi355x drop-in: the structure-factor API (reference
 * include/diffraction.h:46-107). Every compute entry point here runs its
 * arithmetic on the GPU through the C-ABI declared in spsim_hip.h; the
 * host side only marshals the reference's malloc-owned arrays.
 */
#ifndef _DIFFRACTION_H_
#define _DIFFRACTION_H_ 1

#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* include/diffraction.h:46-62. Field order is part of the ABI. */
typedef struct {
  float *ints;        /* |F|^2 per HKL point */
  Complex *F;         /* structure factor per HKL point */
  int N_pix;
  float max_s[3];
  float d_s[3];
  float *HKL_list;    /* copy of the [HKL_list_size][3] list the pattern was computed on */
  int HKL_list_size;
  SpRotation *rot;
} Diffraction_Pattern;

/* Reference src/diffraction.c:237-297. Returns a malloc'd [ny][nx][3] float list of
   scattering vectors (m^-1) on the Ewald sphere; *HKL_list_size = nx*ny. */
float *get_HKL_list_for_detector(CCD *det, Experiment *exp, int *HKL_list_size);

/* Reference src/diffraction.c:325-369 (include/diffraction.h:73). Returns a
   malloc'd [nz][ny][nx][3] uniform grid q = p / (distance * wavelength) of a
   3-D detector volume; *HKL_list_size = nx*ny*nz. */
float *get_HKL_list_for_3d_detector(CCD *det, Experiment *exp, int *HKL_list_size);

/* Reference src/diffraction.c:770-853. F_i = sum_j f_Zj(|q_i|) sqrt(illum_j) exp(-2 pi i q_i.r_j),
   ints_i = |F_i|^2. Atoms with Z == 0 are skipped; delta_atoms forces f = 1. */
Diffraction_Pattern *compute_pattern_on_list(Molecule *mol, float *HKL_list, int HKL_list_size,
                                             float B, Experiment *exp, Options *opts);

/* Reference src/diffraction.c:856-968: same prototype, with the vector path's
   semantics (illumination amplitude illum^(1/4), Z == 0 atoms not skipped,
   delta_atoms applied to the first 4*floor(N/4) atoms only). */
Diffraction_Pattern *vector_compute_pattern_on_list(Molecule *mol, float *HKL_list, int HKL_list_size,
                                                    float B, Experiment *exp, Options *opts);

/* No counterpart in the reference's CPU API: the wavelength-spectrum branch of
   its GPU path (cuda_compute_pattern_on_list2 with
   CUDA_spectrum_scattering_from_all_atoms, src/diffraction.cu:300-318, 416-476),
   which compute_sf takes when use_cuda is set, exp->bandwidth != 0 and
   opts->wavelength_samples > 1. F_i = sum_j f_Zj(|q_i|) sqrt(illum_j)
   sum_k W_k exp(-2 pi i (lambda/lambda_k) q_i.r_j) over the samples of
   spsim_hip_spectrum_samples. As in the reference, an even
   opts->wavelength_samples is raised by one in place and delta_atoms is not
   applied. */
Diffraction_Pattern *spsim_spectrum_pattern_on_list(Molecule *mol, float *HKL_list, int HKL_list_size, float B,
                                                    Experiment *exp, Options *opts);

/* Reference src/diffraction.c:528-675 (include/diffraction.h:101, declared there
   only with NFFT_SUPPORT), the structure factor of the use_nfft_for_sf option:
   F_i = sum over the elements present (Z == 0 included, with the form factor of
   table row 0) of f_Z(|q_i|) sum_{j in Z} exp(-2 pi i q_i.r_j), unit strengths
   (no illumination), delta_atoms -> f = 1 for every atom, ints_i = |F_i|^2.
   The reference evaluates the sums with NFFT3's NNFFT (window parameter m = 2);
   this one with its own GPU type-3 NUFFT at relative accuracy 1e-6
   (spsim_hip_sf_nfft). det is unused, as in the reference. HKL_list is copied
   into the result (the reference leaves that field unset). */
Diffraction_Pattern *compute_pattern_on_list_by_nfft(Molecule *mol, float *HKL_list, int HKL_list_size, CCD *det,
                                                     float B, Options *opts);

/* Reference src/diffraction.c:1146-1187 */
void calculate_pixel_solid_angle(CCD *det);
/* Reference src/diffraction.c:1109-1144 */
void calculate_thomson_correction(CCD *det, Experiment *exp);
/* Reference src/diffraction.c:1190-1204 */
void calculate_photons_per_pixel(Diffraction_Pattern *pattern, Options *opts);

/* Reference src/diffraction.c:300-323. Rotates the list in place and returns the
   rotation (euler_orientation, or uniform random when random_orientation). */
SpRotation *apply_orientation_to_HKL_list(float **HKL_list, int *HKL_list_size, Options *opts);

#ifdef __cplusplus
}
#endif

#endif
