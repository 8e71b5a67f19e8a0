/*
 * spsim-mi355x drop-in: simulation parameters.
 *
 * Field order, types and constants follow the reference header
 * include/config.h:31-181 exactly, because callers (the spsim CLI driver, the
 * Python binding, Condor) poke these structs by offset. All quantities are SI
 * (seconds, metres, joules) unless a comment says otherwise.
 */
#ifndef _CONFIG_H_
#define _CONFIG_H_ 1

#include "spsim_types.h"

#define CHEM_FORMULA 1
#define PDB 2

#define BOX_SPHERICAL 1
#define BOX_PARALLEL 2

#define POLARIZATION_IGNORE 0
#define POLARIZATION_VERTICAL 1
#define POLARIZATION_HORIZONTAL 2
#define POLARIZATION_UNPOLARIZED 3

#define VERSION "2.0"

#ifdef __cplusplus
extern "C" {
#endif

/* include/config.h:44-48 */
typedef struct {
  int n_elements;
  int *atomic_number;
  int *quantity;
} Chem_Formula;

/* include/config.h:50-58. pos is [natoms][3] in metres. */
typedef struct {
  int natoms;
  int *atomic_number;
  float *pos;
  int delta_atoms;
} Molecule;

/* include/config.h:60-76 */
typedef struct {
  float wavelength;      /* m */
  float bandwidth;       /* relative width holding 95% (2 sigma) of the spectrum */
  float exposure_time;   /* s */
  float beam_intensity;  /* photons per m^2, integrated over the exposure */
  float beam_center_x;   /* m */
  float beam_center_y;   /* m */
  float beam_fwhm;       /* m; Gaussian profile, 0 = uniform illumination */
  float beam_energy;     /* J, total pulse energy */
  float photon_energy;   /* eV, alternative to wavelength */
  float focal_diameter;  /* m */
  int polarization;      /* POLARIZATION_* */
} Experiment;

/* include/config.h:78-128 */
typedef struct {
  float *photons_per_pixel;
  float *thomson_correction;
  float *solid_angle;
  float *photon_count;
  float *electrons_per_pixel;
  float *real_output;
  float *noiseless_output;
  float distance;
  float width;
  float height;
  float depth;
  float pixel_width;
  float pixel_height;
  float pixel_depth;
  float quantum_efficiency;
  float electron_hole_production_energy; /* J */
  float readout_noise;                   /* electrons rms per binned pixel */
  float dark_current;                    /* electrons / (pixel s) */
  float linear_full_well;                /* electrons */
  int binning_x;
  int binning_y;
  int binning_z;
  int nx;
  int ny;
  int nz;
  float maximum_value;
  int spherical;
  float gaussian_blurring;
  float real_space_blurring;
  float center_x; /* detector centre relative to the beam axis (z) */
  float center_y;
  float center_z;
  int n_patterns;
} CCD;

/* include/config.h:131-181 */
typedef struct {
  int input_type;
  Chem_Formula *chem_formula;
  Experiment *experiment;
  CCD *detector;
  char pdb_filename[1024];
  char sf_filename[1024];
  char hkl_grid_filename[1024];
  int box_type;
  float box_dimension;
  int use_fft_for_sf;
  int use_nfft_for_sf;
  float b_factor;
  int origin_to_com;
  float euler_orientation[3];
  int random_orientation;
  int n_patterns;
  int vectorize;
  int delta_atoms;
  int fast_exit;
  int crystal_size[3];
  float crystal_cell[6];
  int use_cuda; /* here: 1 = run the structure factor on the MI355X */
  int wavelength_samples;
  int random_seed;
  int output_sf_vtk;
  int output_scatt_int;
  int output_ewald_vtk;
  int output_intensities;
  int output_noiseless_photons;
  int output_photons;
  int output_electron_pixel_vtk;
  int output_noiseless_count;
  int output_count;
  int output_realspace_histogram;
  int output_scattering_factors;
  int output_solid_angles;
  int output_real_space;
  int verbosity_level;
} Options;

/* src/config.c:30-87 */
Options *set_defaults(void);

/* include/config.h:184, src/config.c:89-426. Reads a libconfig-format file
 * (e.g. spsim.conf) into opt and fills the derived fields (nx, ny, nz,
 * wavelength or photon_energy, beam_intensity). Exits with status 1 on a
 * syntax error and with status 0 for input_type "pdb" without pdb_filename,
 * as the reference does; spsim_read_options_file reports the same conditions
 * as a status instead. */
void read_options_file(char *filename, Options *opt);

/* include/config.h:187, src/config.c:428-642. Writes the options as a
 * libconfig-format file (the spsim CLI writes spsim.confout). */
void write_options_file(char *filename, Options *res);

/* No counterpart in the reference: read_options_file for callers that cannot
 * exit. Returns 0, -1 (unreadable file, syntax error or a path longer than
 * 1023 characters) or -2 (input_type "pdb" without pdb_filename); err, when
 * not NULL, receives "<file>:<line>: <message>". */
int spsim_read_options_file(const char *filename, Options *opt, char *err, int errlen);

/* No counterpart in the reference: write_options_file with a status
 * (0, or -1 when the file cannot be written). */
int spsim_write_options_file(const char *filename, const Options *opt);

#ifdef __cplusplus
}
#endif

#endif
