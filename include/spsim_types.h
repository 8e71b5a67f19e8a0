/*
 * spsim-mi355x: the small set of libspimage value types that spsim's public
 * structs and prototypes name.
 *
 * spsim's headers take these from libspimage (spimage.h). libspimage is not
 * vendored in the reference and is not present here, so the layouts below are
 * the ones spsim's own sources rely on:
 *   - Complex {re, im} of `real` (float): src/spsim_pybackend.i:26-33 reads
 *     `$1.re/$1.im`, src/mpi.c:103-107 branches on sizeof(Complex)==8.
 *   - SpRotation is a 3x3 matrix with rows/cols/data (the commented typemap in
 *     src/spsim_pybackend.i:127-135 reads rot->cols, rot->rows, rot->data) and
 *     is released with a single free() (src/spsim.c:357-359). This library
 *     allocates the struct and its 9 floats in one block so that free(rot)
 *     releases both.
 *   - Image and the sp_*3matrix views are used by make_image/make_cimage and
 *     by spsim.py's get_atoms_from_molecule (src/spsim.py:29-41).
 */
#ifndef SPSIM_TYPES_H
#define SPSIM_TYPES_H 1

#ifdef __cplusplus
extern "C" {
#endif

typedef float real;

typedef struct {
  real re;
  real im;
} Complex;

/* Dense row-major matrix; a rotation is rows = cols = 3. */
typedef struct {
  unsigned int rows;
  unsigned int cols;
  real *data;
} sp_matrix;

typedef sp_matrix SpRotation;

typedef struct {
  unsigned int x;
  unsigned int y;
  unsigned int z;
  int *data;
} sp_i3matrix;

typedef struct {
  unsigned int x;
  unsigned int y;
  unsigned int z;
  real *data;
} sp_3matrix;

typedef struct {
  unsigned int x;
  unsigned int y;
  unsigned int z;
  Complex *data;
} sp_c3matrix;

typedef struct {
  real image_center[3];
  real pixel_size[3];
  real detector_distance;
  real wavelength;
  SpRotation *orientation;
} Detector;

typedef struct {
  int phased;
  int shifted;
  int scaled;
  sp_c3matrix *image;
  sp_i3matrix *mask;
  Detector *detector;
  int num_dimensions;
} Image;

/* Image helpers used by spsim.py (src/spsim.py:30-40). */
Image *sp_image_alloc(int x, int y, int z);
void sp_image_free(Image *img);

/* Rotation helpers (libspimage names used by src/diffraction.c:307-310). */
SpRotation *sp_rot_alloc(void);
SpRotation *sp_rot_euler(real a1, real a2, real a3);
SpRotation *sp_rot_uniform(void);

/* Seeds the generator behind sp_rot_uniform and the molecule builders
   (src/spsim.c:385 seeds libspimage's generator with random_seed). */
void sp_srand(int seed);
double p_drand48(void);

#ifdef __cplusplus
}
#endif

#endif /* SPSIM_TYPES_H */
