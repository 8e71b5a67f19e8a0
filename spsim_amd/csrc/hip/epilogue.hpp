// Internal launch interface of the detector-chain kernels (epilogue.hip).
#pragma once

#include "common.hpp"

namespace spsim {

// Detector constants in the precision the reference computes them.
struct DetConst {
  int nx, ny, bin_x, bin_y;
  int nz, bin_z;  // 1 for a 2-D detector
  double width, height, center_x, center_y;  // float inputs widened
  double distance;
  double dist2;       // (double)(float distance * float distance)
  double pixel_area;  // (double)(float pixel_width * float pixel_height)
  int spherical, polarization;
  float beam_intensity, qe;
  float e_per_photon;            // noisy branch (1 when the e-h energy is 0)
  float e_per_photon_noiseless;  // noiseless branch (no zero check)
  float dark_current, exposure_time, readout_noise;
  float adc;
};

int make_det_const(const spsim_hip_detector *p, DetConst *d);

enum class Stage { kThomson, kSolidAngle, kPhotons, kPoisson, kElectrons, kRealOutput, kNoiselessOutput };

// in0..in2 per stage: photons(thomson, solid_angle, I); Poisson(photons);
// electrons(count); real output(electrons); noiseless(photons).
int launch_stage(Stage st, const DetConst &d, const float *in0, const float *in1, const float *in2, uint64_t seed,
                 uint32_t pattern, float *out, hipStream_t s);

int launch_chain(const DetConst &d, const float *I, int npatterns, uint64_t seed, uint64_t pattern_base, float *th,
                 float *om, float *ph, float *cnt, float *el, float *real, float *noiseless, hipStream_t s);

}  // namespace spsim
