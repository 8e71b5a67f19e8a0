/*
 * Rotations for apply_orientation_to_HKL_list (src/diffraction.c:300-323).
 *
 * libspimage's sp_rot_euler / sp_rot_uniform are not vendored with the
 * reference, so their conventions are restated here:
 *  - sp_rot_euler(a1, a2, a3) = Rz(a1) Ry(a2) Rz(a3) (intrinsic z-y-z angles, radians);
 *  - sp_rot_uniform() draws a uniform unit quaternion with Shoemake's method
 *    (K. Shoemake, "Uniform random rotations", Graphics Gems III, 1992) from
 *    p_drand48() and converts it to a matrix.
 * The matrix is stored in one malloc block with its 9 floats, so free(rot)
 * (src/spsim.c:357-359) releases everything.
 */
#include <math.h>
#include <stdlib.h>

#include "spsim_internal.h"

SpRotation *sp_rot_alloc(void) {
  SpRotation *r = (SpRotation *)malloc(sizeof(SpRotation) + 9 * sizeof(real));
  if (!r) return NULL;
  r->rows = 3;
  r->cols = 3;
  r->data = (real *)(r + 1);
  for (int i = 0; i < 9; ++i) r->data[i] = (i % 4 == 0) ? 1.0f : 0.0f;
  return r;
}

static void matmul3(const double A[9], const double B[9], double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i + 0] * B[0 + j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

static void rot_z(double a, double R[9]) {
  const double c = cos(a), s = sin(a);
  R[0] = c; R[1] = -s; R[2] = 0;
  R[3] = s; R[4] = c;  R[5] = 0;
  R[6] = 0; R[7] = 0;  R[8] = 1;
}

static void rot_y(double a, double R[9]) {
  const double c = cos(a), s = sin(a);
  R[0] = c;  R[1] = 0; R[2] = s;
  R[3] = 0;  R[4] = 1; R[5] = 0;
  R[6] = -s; R[7] = 0; R[8] = c;
}

SpRotation *sp_rot_euler(real a1, real a2, real a3) {
  double A[9], B[9], C[9], T[9], R[9];
  SpRotation *r = sp_rot_alloc();
  if (!r) return NULL;
  rot_z(a1, A);
  rot_y(a2, B);
  rot_z(a3, C);
  matmul3(A, B, T);
  matmul3(T, C, R);
  for (int i = 0; i < 9; ++i) r->data[i] = (real)R[i];
  return r;
}

void spsim_quaternion_to_matrix(const double q[4], double R[9]) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = 1 - 2 * (y * y + z * z);
  R[1] = 2 * (x * y - w * z);
  R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z);
  R[4] = 1 - 2 * (x * x + z * z);
  R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y);
  R[7] = 2 * (y * z + w * x);
  R[8] = 1 - 2 * (x * x + y * y);
}

void spsim_uniform_quaternion(double u1, double u2, double u3, double q[4]) {
  const double two_pi = 6.283185307179586476925286766559;
  const double a = sqrt(1.0 - u1), b = sqrt(u1);
  q[0] = b * cos(two_pi * u3);
  q[1] = a * sin(two_pi * u2);
  q[2] = a * cos(two_pi * u2);
  q[3] = b * sin(two_pi * u3);
}

SpRotation *sp_rot_uniform(void) {
  double q[4], R[9];
  SpRotation *r = sp_rot_alloc();
  if (!r) return NULL;
  const double u1 = p_drand48(), u2 = p_drand48(), u3 = p_drand48();
  spsim_uniform_quaternion(u1, u2, u3, q);
  spsim_quaternion_to_matrix(q, R);
  for (int i = 0; i < 9; ++i) r->data[i] = (real)R[i];
  return r;
}
