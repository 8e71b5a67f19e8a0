/*
 * spsim-mi355x drop-in: pipeline driver (reference include/spsim.h:40-59).
 */
#ifndef _SPSIM_H_
#define _SPSIM_H_ 1

#include "config.h"
#include "diffraction.h"
#include "molecule.h"
#include "noise.h"
#include "amplification.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/spsim.c:84-96 (chemical-formula input) */
Molecule *get_molecule(Options *opts);
/* src/spsim.c:98-108 (2-D detectors only) */
float *get_HKL_list(Options *opts, int *HKL_list_size);
/* src/spsim.c:110-144: the backend dispatcher. */
Diffraction_Pattern *compute_sf(Molecule *mol, float *HKL_list, int HKL_list_size, Options *opts);

/* src/spsim.c:154-228 */
Image *make_image(float *array, SpRotation *rot, Options *opts);
Image *make_cimage(Complex *array, SpRotation *rot, Options *opts);
void array_to_image(float *array, Image *img);
void iarray_to_image(int *array, Image *img);

/* src/spsim.c:146-152, 176-182 (include/spsim.h:51-52): one image per file,
   "<basename>-%05d.cxi" (HDF5, CXI 1.5 layout; HDF5 is loaded at the first
   write). */
void output_array(char *basename, int index, float *array, SpRotation *rot, Options *opts);
void output_carray(char *basename, int index, Complex *array, SpRotation *rot, Options *opts);
/* src/spsim.c:230-279 (include/spsim.h:53): the files the output_* flags select. */
void output_files(Diffraction_Pattern *pattern, Options *opts, int index);

/* src/spsim.c:57-81 */
void gaussian_blur_pattern(Options *opts, Diffraction_Pattern *pattern);

/* src/spsim.c:281-346: q list, orientation, structure factor, then the detector chain. */
Diffraction_Pattern *simulate_shot(Molecule *mol, Options *opts);

/* src/spsim.c:348-371 */
void free_stuff(Diffraction_Pattern *pattern, Options *opts);
void free_diffraction_pattern(Diffraction_Pattern *pattern);
void free_output_in_options(Options *opts);

#ifdef __cplusplus
}
#endif

#endif
