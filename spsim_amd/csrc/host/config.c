/*
 * Options defaults (reference src/config.c:30-87). The values are the
 * reference's; use_cuda defaults to 1 because this library always computes the
 * structure factor on the MI355X.
 */
#include <stdlib.h>

#include "config.h"

Options *set_defaults(void) {
  Options *opt = (Options *)calloc(1, sizeof(Options));
  if (!opt) return NULL;
  opt->chem_formula = (Chem_Formula *)calloc(1, sizeof(Chem_Formula));
  opt->experiment = (Experiment *)calloc(1, sizeof(Experiment));
  opt->detector = (CCD *)calloc(1, sizeof(CCD));
  if (!opt->chem_formula || !opt->experiment || !opt->detector) {
    free(opt->chem_formula);
    free(opt->experiment);
    free(opt->detector);
    free(opt);
    return NULL;
  }
  opt->experiment->polarization = POLARIZATION_IGNORE;
  opt->detector->binning_x = 1;
  opt->detector->binning_y = 1;
  opt->detector->binning_z = 1;
  opt->detector->center_x = 0;
  opt->detector->center_y = 0;
  opt->detector->center_z = 0;
  opt->box_type = BOX_SPHERICAL;
  opt->box_dimension = 1e-9f;
  opt->use_fft_for_sf = 0;
  opt->use_nfft_for_sf = 0;
  opt->b_factor = 0;
  opt->n_patterns = 1;
  opt->origin_to_com = 0;
  opt->euler_orientation[0] = 0;
  opt->euler_orientation[1] = 0;
  opt->euler_orientation[2] = 0;
  opt->random_orientation = 0;
  opt->vectorize = 1;
  opt->use_cuda = 1;
  opt->delta_atoms = 0;
  opt->fast_exit = 0;
  opt->crystal_size[0] = 1;
  opt->crystal_size[1] = 1;
  opt->crystal_size[2] = 1;
  opt->crystal_cell[0] = 0;
  opt->crystal_cell[1] = 0;
  opt->crystal_cell[2] = 0;
  opt->crystal_cell[3] = 90;
  opt->crystal_cell[4] = 90;
  opt->crystal_cell[5] = 90;
  opt->wavelength_samples = 5;
  opt->random_seed = -1;
  opt->output_sf_vtk = 0;
  opt->output_scatt_int = 0;
  opt->output_ewald_vtk = 0;
  opt->output_intensities = 0;
  opt->output_noiseless_photons = 1;
  opt->output_photons = 1;
  opt->output_electron_pixel_vtk = 0;
  opt->output_noiseless_count = 0;
  opt->output_count = 0;
  opt->output_realspace_histogram = 0;
  opt->output_scattering_factors = 0;
  opt->output_solid_angles = 0;
  opt->output_real_space = 0;
  opt->verbosity_level = 1;
  return opt;
}
