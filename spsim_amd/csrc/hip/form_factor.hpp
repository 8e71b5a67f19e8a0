// Cromer-Mann form factor shared by the structure-factor, adjoint and NUFFT kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace spsim {

// f(s) = sum_i a_i e^{-(b_i + B) s^2} + c e^{-B s^2} for one table row c = {a1..a4, b1..b4, c}
// (b in A^2), with s2 = s^2 in A^-2 and B the global Debye-Waller factor. Terms
// are added in the reference's order (src/diffraction.c:183-193).
__device__ __forceinline__ float form_factor(const float *__restrict__ c, float s2, float B) {
  float f = 0.0f;
  f += c[0] * expf(-(c[4] + B) * s2);
  f += c[1] * expf(-(c[5] + B) * s2);
  f += c[2] * expf(-(c[6] + B) * s2);
  f += c[3] * expf(-(c[7] + B) * s2);
  f += c[8] * expf(-B * s2);
  return f;
}

}  // namespace spsim
