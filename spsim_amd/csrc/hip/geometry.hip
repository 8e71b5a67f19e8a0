// Detector geometry kernels: scattering vectors per pixel and their rotation.
//
// spsim_hip_hkl_detector reproduces get_HKL_list_for_detector
// (src/diffraction.c:237-297) and spsim_hip_hkl_detector_3d
// get_HKL_list_for_3d_detector (:325-369) operation for operation in FP64 with
// FMA contraction disabled, so the FP32 lists they store are bit-identical to
// the host computation of the same expressions.
#include "common.hpp"

namespace spsim {
namespace {

struct HklGeom {
  int nx, ny;
  double width, height;  // detector size (float inputs widened)
  double center_x, center_y;
  double pz;             // distance
  double ewald_radius;   // 1 / wavelength
  double px_size, py_size;
  double fourier_px, fourier_py;
  int spherical;
};

__global__ __launch_bounds__(256) void hkl_detector_kernel(HklGeom g, float *__restrict__ out) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)g.nx * g.ny) return;
  const int x = (int)(idx % g.nx);
  const int y = (int)(idx / g.nx);
  const double px = ((x - (g.nx - 1.0) / 2.0) / g.nx) * g.width - g.center_x;
  const double py = ((y - (g.ny - 1.0) / 2.0) / g.ny) * g.height - g.center_y;
  float *o = out + 3 * idx;
  if (!g.spherical) {
    const double p = sqrt(px * px + py * py + g.pz * g.pz);
    const double rx = px / p;
    const double ry = py / p;
    const double rz = g.pz / p - 1.0;
    o[0] = (float)(rx * g.ewald_radius);
    o[1] = (float)(ry * g.ewald_radius);
    o[2] = (float)(rz * g.ewald_radius);
  } else {
    o[0] = (float)(g.fourier_px * px / g.px_size);
    o[1] = (float)(g.fourier_py * py / g.py_size);
    o[2] = 0.0f;
  }
}

// Uniform reciprocal-space grid of a 3-D detector (get_HKL_list_for_3d_detector,
// src/diffraction.c:325-369): q = p / (distance * wavelength), with the product
// distance * wavelength rounded to float as the reference forms it.
struct Hkl3Geom {
  int nx, ny, nz;
  double width, height, depth;
  double center_x, center_y, center_z;
  double real_to_reciprocal;
};

__global__ __launch_bounds__(256) void hkl_3d_kernel(Hkl3Geom g, float *__restrict__ out) {
#pragma clang fp contract(off)
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long plane = (long long)g.nx * g.ny;
  if (idx >= plane * g.nz) return;
  const int x = (int)(idx % g.nx);
  const int y = (int)((idx / g.nx) % g.ny);
  const int z = (int)(idx / plane);
  const double px = ((x - (g.nx - 1.0) / 2.0) / g.nx) * g.width - g.center_x;
  const double py = ((y - (g.ny - 1.0) / 2.0) / g.ny) * g.height - g.center_y;
  const double pz = ((z - (g.nz - 1.0) / 2.0) / g.nz) * g.depth - g.center_z;
  float *o = out + 3 * idx;
  o[0] = (float)(px * g.real_to_reciprocal);
  o[1] = (float)(py * g.real_to_reciprocal);
  o[2] = (float)(pz * g.real_to_reciprocal);
}

struct Rot9 {
  double r[9];
};

__global__ __launch_bounds__(256) void rotate_hkl_kernel(const float *__restrict__ in, int npix, Rot9 R,
                                                         float *__restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npix) return;
  const double x = in[3 * i + 0], y = in[3 * i + 1], z = in[3 * i + 2];
  const float ox = (float)__builtin_fma(R.r[2], z, __builtin_fma(R.r[1], y, R.r[0] * x));
  const float oy = (float)__builtin_fma(R.r[5], z, __builtin_fma(R.r[4], y, R.r[3] * x));
  const float oz = (float)__builtin_fma(R.r[8], z, __builtin_fma(R.r[7], y, R.r[6] * x));
  out[3 * i + 0] = ox;
  out[3 * i + 1] = oy;
  out[3 * i + 2] = oz;
}

}  // namespace
}  // namespace spsim

extern "C" int spsim_hip_hkl_detector(int nx, int ny, float width, float height, float distance, float center_x,
                                      float center_y, int spherical, float wavelength, float *d_hkl, void *stream) {
  using namespace spsim;
  if (nx < 0 || ny < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative detector size %d x %d", nx, ny);
  if ((long long)nx * ny > (1LL << 28)) return fail(SPSIM_HIP_ERR_INVALID, "detector %d x %d too large", nx, ny);
  if (nx * ny > 0 && !d_hkl) return fail(SPSIM_HIP_ERR_INVALID, "d_hkl is NULL");
  SPSIM_TRY(ensure_device(nullptr));
  if (nx == 0 || ny == 0) return SPSIM_HIP_OK;
  HklGeom g;
  g.nx = nx;
  g.ny = ny;
  g.width = width;
  g.height = height;
  g.center_x = center_x;
  g.center_y = center_y;
  g.pz = distance;
  g.ewald_radius = 1.0 / (double)wavelength;
  // `det->width/nx` is a float division in the reference (float / int).
  g.px_size = (double)(width / (float)nx);
  g.py_size = (double)(height / (float)ny);
  g.fourier_px = 2 * sin(atan(g.px_size / g.pz) / 2) / (double)wavelength;
  g.fourier_py = 2 * sin(atan(g.py_size / g.pz) / 2) / (double)wavelength;
  g.spherical = spherical;
  const long long n = (long long)nx * ny;
  hipLaunchKernelGGL(hkl_detector_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, g,
                     d_hkl);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

extern "C" int spsim_hip_rotate_hkl(const float *d_in, int npix, const double *rot9, float *d_out, void *stream) {
  using namespace spsim;
  if (npix < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative npix");
  if (!rot9) return fail(SPSIM_HIP_ERR_INVALID, "rot9 is NULL");
  if (npix > 0 && (!d_in || !d_out)) return fail(SPSIM_HIP_ERR_INVALID, "d_in/d_out is NULL");
  SPSIM_TRY(ensure_device(nullptr));
  if (npix == 0) return SPSIM_HIP_OK;
  Rot9 R;
  for (int i = 0; i < 9; ++i) R.r[i] = rot9[i];
  hipLaunchKernelGGL(rotate_hkl_kernel, dim3((npix + 255) / 256), dim3(256), 0, (hipStream_t)stream, d_in, npix, R,
                     d_out);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

extern "C" int spsim_hip_hkl_detector_3d(int nx, int ny, int nz, float width, float height, float depth,
                                         float center_x, float center_y, float center_z, float distance,
                                         float wavelength, float *d_hkl, void *stream) {
  using namespace spsim;
  if (nx < 0 || ny < 0 || nz < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative detector size %d x %d x %d", nx, ny, nz);
  const long long n = (long long)nx * ny * nz;
  if (n > (1LL << 28)) return fail(SPSIM_HIP_ERR_INVALID, "detector %d x %d x %d too large", nx, ny, nz);
  if (n > 0 && !d_hkl) return fail(SPSIM_HIP_ERR_INVALID, "d_hkl is NULL");
  SPSIM_TRY(ensure_device(nullptr));
  if (n == 0) return SPSIM_HIP_OK;
  Hkl3Geom g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.width = width;
  g.height = height;
  g.depth = depth;
  g.center_x = center_x;
  g.center_y = center_y;
  g.center_z = center_z;
  const float dl = distance * wavelength;  // a float product in the reference
  g.real_to_reciprocal = 1.0 / (double)dl;
  hipLaunchKernelGGL(hkl_3d_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, g, d_hkl);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}
