// Host-array bridge used by the reference-shaped C API (see host/bridge.h).
// Every call stages host arrays into HBM on a private stream, runs the device
// entry points and copies the results back before returning.
#include <cstring>
#include <vector>

#include "../host/bridge.h"
#include "common.hpp"
#include "epilogue.hpp"
#include "sf_kernels.hpp"

namespace spsim {
namespace {

// Device buffers released on scope exit.
struct DevBufs {
  std::vector<void *> ptrs;
  ~DevBufs() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) return SPSIM_HIP_OK;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return fail(SPSIM_HIP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", n * sizeof(T), hipGetErrorString(e));
    ptrs.push_back(*p);
    return SPSIM_HIP_OK;
  }
  template <class T>
  int upload(T **p, const T *host, size_t n, hipStream_t s) {
    SPSIM_TRY(alloc(p, n));
    if (n && host) SPSIM_HIP_TRY(hipMemcpyAsync(*p, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    return SPSIM_HIP_OK;
  }
};

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
  int create() {
    SPSIM_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return SPSIM_HIP_OK;
  }
};

template <class T>
int download(T *host, const T *dev, size_t n, hipStream_t s) {
  if (n && host) SPSIM_HIP_TRY(hipMemcpyAsync(host, dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
  return SPSIM_HIP_OK;
}

int run_stage(const spsim_hip_detector *det, Stage st, const float *h0, const float *h1, const float *h2,
              size_t n_in, uint64_t seed, uint64_t pattern, float *out, size_t n_out) {
  DetConst d;
  SPSIM_TRY(make_det_const(det, &d));
  SPSIM_TRY(ensure_device(nullptr));
  if (!out) return fail(SPSIM_HIP_ERR_INVALID, "output is NULL");
  Stream st_;
  SPSIM_TRY(st_.create());
  DevBufs b;
  float *d0 = nullptr, *d1 = nullptr, *d2 = nullptr, *dout = nullptr;
  if (h0) SPSIM_TRY(b.upload(&d0, h0, n_in, st_.s));
  if (h1) SPSIM_TRY(b.upload(&d1, h1, n_in, st_.s));
  if (h2) SPSIM_TRY(b.upload(&d2, h2, n_in, st_.s));
  SPSIM_TRY(b.alloc(&dout, n_out));
  SPSIM_TRY(launch_stage(st, d, d0, d1, d2, seed, (uint32_t)pattern, dout, st_.s));
  SPSIM_TRY(download(out, (const float *)dout, n_out, st_.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st_.s));
  return SPSIM_HIP_OK;
}

size_t raw_count(const spsim_hip_detector *det) {
  return det ? (size_t)det->nx * det->ny * (det->nz > 0 ? det->nz : 1) : 0;
}
size_t bin_count(const spsim_hip_detector *det) {
  if (!det || det->bin_x < 1 || det->bin_y < 1) return 0;
  const int nz = det->nz > 0 ? det->nz : 1, bz = det->bin_z > 0 ? det->bin_z : 1;
  return (size_t)(det->nx / det->bin_x) * (det->ny / det->bin_y) * (nz / bz);
}

}  // namespace
}  // namespace spsim

extern "C" {

int spsim_bridge_hkl_detector(int nx, int ny, float width, float height, float distance, float center_x,
                              float center_y, int spherical, float wavelength, float *out) {
  using namespace spsim;
  SPSIM_TRY(ensure_device(nullptr));
  const size_t n = (size_t)(nx > 0 ? nx : 0) * (ny > 0 ? ny : 0) * 3;
  Stream st;
  SPSIM_TRY(st.create());
  DevBufs b;
  float *d = nullptr;
  SPSIM_TRY(b.alloc(&d, n));
  SPSIM_TRY(spsim_hip_hkl_detector(nx, ny, width, height, distance, center_x, center_y, spherical, wavelength, d, st.s));
  SPSIM_TRY(download(out, (const float *)d, n, st.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st.s));
  return SPSIM_HIP_OK;
}

int spsim_bridge_hkl_detector_3d(int nx, int ny, int nz, float width, float height, float depth, float center_x,
                                 float center_y, float center_z, float distance, float wavelength, float *out) {
  using namespace spsim;
  SPSIM_TRY(ensure_device(nullptr));
  const size_t n = (size_t)(nx > 0 ? nx : 0) * (ny > 0 ? ny : 0) * (nz > 0 ? nz : 0) * 3;
  Stream st;
  SPSIM_TRY(st.create());
  DevBufs b;
  float *d = nullptr;
  SPSIM_TRY(b.alloc(&d, n));
  SPSIM_TRY(spsim_hip_hkl_detector_3d(nx, ny, nz, width, height, depth, center_x, center_y, center_z, distance,
                                      wavelength, d, st.s));
  SPSIM_TRY(download(out, (const float *)d, n, st.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st.s));
  return SPSIM_HIP_OK;
}

int spsim_bridge_rotate_hkl(float *list, int npix, const double *rot9) {
  using namespace spsim;
  SPSIM_TRY(ensure_device(nullptr));
  if (npix <= 0) return SPSIM_HIP_OK;
  Stream st;
  SPSIM_TRY(st.create());
  DevBufs b;
  float *d = nullptr;
  SPSIM_TRY(b.upload(&d, (const float *)list, (size_t)npix * 3, st.s));
  SPSIM_TRY(spsim_hip_rotate_hkl(d, npix, rot9, d, st.s));
  SPSIM_TRY(download(list, (const float *)d, (size_t)npix * 3, st.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st.s));
  return SPSIM_HIP_OK;
}

// One structure-factor launch on host arrays: the spectrum kernel when
// nsamples > 0, else the grid kernel when nx > 0, else the per-pixel kernel.
static int structure_factor_impl(const float *hkl, int npix, int nx, int ny, const spsim_hip_atom *atoms,
                                 const float *atom_b, const spsim_hip_segment *segs, int nsegs, int natoms,
                                 const float *ff, float b_global, Complex *F, float *I, int nsamples = 0,
                                 const double *scales = nullptr, const double *weights = nullptr) {
  using namespace spsim;
  SPSIM_TRY(ensure_device(nullptr));
  if (npix <= 0) return SPSIM_HIP_OK;
  Stream st;
  SPSIM_TRY(st.create());
  DevBufs b;
  float *dq = nullptr, *dff = nullptr, *db = nullptr, *dF = nullptr, *dI = nullptr;
  spsim_hip_atom *da = nullptr;
  spsim_hip_segment *ds = nullptr;
  SPSIM_TRY(b.upload(&dq, hkl, (size_t)npix * 3, st.s));
  SPSIM_TRY(b.upload(&da, atoms, (size_t)natoms, st.s));
  if (atom_b) SPSIM_TRY(b.upload(&db, atom_b, (size_t)natoms, st.s));
  SPSIM_TRY(b.upload(&ds, segs, (size_t)nsegs, st.s));
  SPSIM_TRY(b.upload(&dff, ff, (size_t)SPSIM_HIP_NSPECIES * SPSIM_HIP_FF_STRIDE, st.s));
  if (F) SPSIM_TRY(b.alloc(&dF, (size_t)npix * 2));
  SPSIM_TRY(b.alloc(&dI, (size_t)npix));
  if (nsamples > 0)
    SPSIM_TRY(spsim_hip_sf_spectrum(dq, npix, da, db, ds, nsegs, natoms, dff, b_global, nullptr, nsamples, scales,
                                    weights, dF, dI, nullptr, st.s));
  else if (nx > 0)
    SPSIM_TRY(spsim_hip_sf_grid(dq, nx, ny, da, db, ds, nsegs, natoms, dff, b_global, nullptr, dF, dI, nullptr, st.s));
  else
    SPSIM_TRY(spsim_hip_sf_pattern(dq, npix, da, db, ds, nsegs, natoms, dff, b_global, nullptr, dF, dI, nullptr, st.s));
  if (F) SPSIM_TRY(download((float *)F, (const float *)dF, (size_t)npix * 2, st.s));
  SPSIM_TRY(download(I, (const float *)dI, (size_t)npix, st.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st.s));
  return SPSIM_HIP_OK;
}

int spsim_bridge_structure_factor(const float *hkl, int npix, const spsim_hip_atom *atoms, const float *atom_b,
                                  const spsim_hip_segment *segs, int nsegs, int natoms, const float *ff, float b_global,
                                  Complex *F, float *I) {
  return structure_factor_impl(hkl, npix, 0, 0, atoms, atom_b, segs, nsegs, natoms, ff, b_global, F, I);
}

int spsim_bridge_structure_factor_spectrum(const float *hkl, int npix, const spsim_hip_atom *atoms,
                                           const float *atom_b, const spsim_hip_segment *segs, int nsegs,
                                           int natoms, const float *ff, float b_global, int nsamples,
                                           const double *scales, const double *weights, Complex *F, float *I) {
  if (nsamples < 1) return spsim::fail(SPSIM_HIP_ERR_INVALID, "nsamples=%d", nsamples);
  return structure_factor_impl(hkl, npix, 0, 0, atoms, atom_b, segs, nsegs, natoms, ff, b_global, F, I, nsamples,
                               scales, weights);
}

int spsim_bridge_structure_factor_grid(const float *hkl, int nx, int ny, const spsim_hip_atom *atoms,
                                       const float *atom_b, const spsim_hip_segment *segs, int nsegs, int natoms,
                                       const float *ff, float b_global, Complex *F, float *I) {
  if (nx < 0 || ny < 0) return spsim::fail(SPSIM_HIP_ERR_INVALID, "negative detector size");
  return structure_factor_impl(hkl, nx * ny, nx, ny, atoms, atom_b, segs, nsegs, natoms, ff, b_global, F, I);
}

int spsim_bridge_structure_factor_nfft(const float *hkl, int npix, const spsim_hip_atom *atoms,
                                       const spsim_hip_segment *segs, int nsegs, int natoms, const float *ff,
                                       float b_global, double eps, Complex *F, float *I) {
  using namespace spsim;
  SPSIM_TRY(ensure_device(nullptr));
  if (npix <= 0) return SPSIM_HIP_OK;
  Stream st;
  SPSIM_TRY(st.create());
  DevBufs b;
  float *dq = nullptr, *dff = nullptr, *dF = nullptr, *dI = nullptr;
  spsim_hip_atom *da = nullptr;
  spsim_hip_segment *ds = nullptr;
  SPSIM_TRY(b.upload(&dq, hkl, (size_t)npix * 3, st.s));
  SPSIM_TRY(b.upload(&da, atoms, (size_t)natoms, st.s));
  SPSIM_TRY(b.upload(&ds, segs, (size_t)nsegs, st.s));
  SPSIM_TRY(b.upload(&dff, ff, (size_t)SPSIM_HIP_NSPECIES * SPSIM_HIP_FF_STRIDE, st.s));
  if (F) SPSIM_TRY(b.alloc(&dF, (size_t)npix * 2));
  SPSIM_TRY(b.alloc(&dI, (size_t)npix));
  SPSIM_TRY(spsim_hip_sf_nfft(dq, npix, da, nullptr, ds, nsegs, natoms, dff, b_global, nullptr, eps, dF, dI, nullptr, st.s));
  if (F) SPSIM_TRY(download((float *)F, (const float *)dF, (size_t)npix * 2, st.s));
  SPSIM_TRY(download(I, (const float *)dI, (size_t)npix, st.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st.s));
  return SPSIM_HIP_OK;
}

int spsim_bridge_thomson(const spsim_hip_detector *det, float *out) {
  using namespace spsim;
  return run_stage(det, Stage::kThomson, nullptr, nullptr, nullptr, 0, 0, 0, out, raw_count(det));
}

int spsim_bridge_solid_angle(const spsim_hip_detector *det, float *out) {
  using namespace spsim;
  return run_stage(det, Stage::kSolidAngle, nullptr, nullptr, nullptr, 0, 0, 0, out, raw_count(det));
}

int spsim_bridge_photons(const spsim_hip_detector *det, const float *thomson, const float *solid_angle,
                         const float *I, float *out) {
  using namespace spsim;
  if (!thomson || !solid_angle || !I) return fail(SPSIM_HIP_ERR_INVALID, "photons: input is NULL");
  return run_stage(det, Stage::kPhotons, thomson, solid_angle, I, raw_count(det), 0, 0, out, raw_count(det));
}

int spsim_bridge_poisson(const spsim_hip_detector *det, const float *photons, uint64_t seed, uint64_t pattern,
                         float *out) {
  using namespace spsim;
  if (!photons) return fail(SPSIM_HIP_ERR_INVALID, "poisson: photons is NULL");
  return run_stage(det, Stage::kPoisson, photons, nullptr, nullptr, raw_count(det), seed, pattern, out, raw_count(det));
}

int spsim_bridge_electrons(const spsim_hip_detector *det, const float *count, float *out) {
  using namespace spsim;
  if (!count) return fail(SPSIM_HIP_ERR_INVALID, "electrons: count is NULL");
  return run_stage(det, Stage::kElectrons, count, nullptr, nullptr, raw_count(det), 0, 0, out, raw_count(det));
}

int spsim_bridge_real_output(const spsim_hip_detector *det, const float *electrons, uint64_t seed, uint64_t pattern,
                             float *out) {
  using namespace spsim;
  if (!electrons) return fail(SPSIM_HIP_ERR_INVALID, "real output: electrons is NULL");
  return run_stage(det, Stage::kRealOutput, electrons, nullptr, nullptr, raw_count(det), seed, pattern, out,
                   bin_count(det));
}

int spsim_bridge_noiseless_output(const spsim_hip_detector *det, const float *photons, float *out) {
  using namespace spsim;
  if (!photons) return fail(SPSIM_HIP_ERR_INVALID, "noiseless output: photons is NULL");
  return run_stage(det, Stage::kNoiselessOutput, photons, nullptr, nullptr, raw_count(det), 0, 0, out, bin_count(det));
}

int spsim_bridge_detector_chain(const spsim_hip_detector *det, const float *I, uint64_t seed, uint64_t pattern,
                                float *thomson, float *solid_angle, float *photons, float *count, float *electrons,
                                float *real_output, float *noiseless_output) {
  using namespace spsim;
  DetConst d;
  SPSIM_TRY(make_det_const(det, &d));
  SPSIM_TRY(ensure_device(nullptr));
  const size_t n = raw_count(det), nb = bin_count(det);
  Stream st;
  SPSIM_TRY(st.create());
  DevBufs b;
  float *dI = nullptr, *o[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  float *host[7] = {thomson, solid_angle, photons, count, electrons, real_output, noiseless_output};
  const size_t sizes[7] = {n, n, n, n, n, nb, nb};
  SPSIM_TRY(b.upload(&dI, I, n, st.s));
  for (int i = 0; i < 7; ++i)
    if (host[i]) SPSIM_TRY(b.alloc(&o[i], sizes[i]));
  SPSIM_TRY(launch_chain(d, dI, 1, seed, pattern, o[0], o[1], o[2], o[3], o[4], o[5], o[6], st.s));
  for (int i = 0; i < 7; ++i)
    if (host[i]) SPSIM_TRY(download(host[i], (const float *)o[i], sizes[i], st.s));
  SPSIM_HIP_TRY(hipStreamSynchronize(st.s));
  return SPSIM_HIP_OK;
}

}  // extern "C"
