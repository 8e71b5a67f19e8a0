// Microbenchmark for gfx950: v_sin/v_cos semantics (turns) and accuracy, transcendental
// vs FP32/FP64 FMA throughput. Not part of the library. Build and run on one MI355X:
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       -o valu_rates spsim_amd/csrc/microbench/valu_rates.hip && ./valu_rates
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cmath>
#include <vector>

#define CHK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

constexpr int ITERS = 4096;

__global__ void k_fma32(float* out, float a, float b) {
  float x[8];
  for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 1e-3f + i;
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = __builtin_fmaf(x[i], a, b);
  }
  float s = 0; for (int i = 0; i < 8; ++i) s += x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_fma64(double* out, double a, double b) {
  double x[8];
  for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 1e-3 + i;
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = __builtin_fma(x[i], a, b);
  }
  double s = 0; for (int i = 0; i < 8; ++i) s += x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_sin(float* out) {
  float x[8];
  for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 1e-3f + i * 0.1f;
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = __builtin_amdgcn_sinf(x[i]);
  }
  float s = 0; for (int i = 0; i < 8; ++i) s += x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_exp(float* out) {
  float x[8];
  for (int i = 0; i < 8; ++i) x[i] = -(threadIdx.x * 1e-3f + i * 0.1f);
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = -__builtin_amdgcn_exp2f(x[i]);
  }
  float s = 0; for (int i = 0; i < 8; ++i) s += x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void k_fract64(double* out) {
  double x[8];
  for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 1e-3 + i * 0.1;
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = __builtin_amdgcn_fract(x[i] + 0.3);
  }
  double s = 0; for (int i = 0; i < 8; ++i) s += x[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// sin and fma interleaved: does the transcendental overlap the FMA pipe?
__global__ void k_mix(float* out, float a, float b) {
  float x[8], y[8];
  for (int i = 0; i < 8; ++i) { x[i] = threadIdx.x * 1e-3f + i; y[i] = i * 0.01f; }
  for (int it = 0; it < ITERS; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      y[i] = __builtin_amdgcn_sinf(y[i]);
      x[i] = __builtin_fmaf(x[i], a, b);
      x[i] = __builtin_fmaf(x[i], a, b);
      x[i] = __builtin_fmaf(x[i], a, b);
      x[i] = __builtin_fmaf(x[i], a, b);
    }
  }
  float s = 0; for (int i = 0; i < 8; ++i) s += x[i] + y[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// The candidate inner loop: fp64 phase, fract, cvt, sin, cos, 2 adds.
__global__ void k_term(float* out, const double* q, int natoms) {
  double qx = q[0] + threadIdx.x * 1e-6, qy = q[1], qz = q[2];
  float sr = 0, si = 0;
  for (int j = 0; j < natoms; ++j) {
    double rx = 1e-9 * j, ry = 2e-9 * j, rz = 3e-9 * j;
    double t = qx * rx + qy * ry + qz * rz;
    float tf = (float)__builtin_amdgcn_fract(t);
    sr += __builtin_amdgcn_cosf(tf);
    si += __builtin_amdgcn_sinf(tf);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = sr + si;
}

// Accuracy: evaluate v_sin/v_cos on x in turns, compare with fp64 sin(2 pi x).
__global__ void k_acc(const float* x, float* s_hw, float* c_hw, double* s_ref, double* c_ref, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float v = x[i];
    s_hw[i] = __builtin_amdgcn_sinf(v);
    c_hw[i] = __builtin_amdgcn_cosf(v);
    double ang = 6.283185307179586476925286766559 * (double)v;
    s_ref[i] = sin(ang);
    c_ref[i] = cos(ang);
  }
}

template <class K, class... A>
float time_kernel(K k, int grid, int block, A... a) {
  hipEvent_t e0, e1; hipEventCreate(&e0); hipEventCreate(&e1);
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, 0, a...);
  hipDeviceSynchronize();
  hipEventRecord(e0);
  for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, 0, a...);
  hipEventRecord(e1); hipEventSynchronize(e1);
  float ms; hipEventElapsedTime(&ms, e0, e1);
  return ms / 3;
}

int main() {
  hipDeviceProp_t p; CHK(hipGetDeviceProperties(&p, 0));
  printf("device %s CUs %d clock %d kHz gcn %s\n", p.name, p.multiProcessorCount, p.clockRate, p.gcnArchName);
  const int grid = p.multiProcessorCount * 16, block = 256;
  const size_t nthr = (size_t)grid * block;
  float* outf; double* outd;
  CHK(hipMalloc(&outf, nthr * sizeof(float)));
  CHK(hipMalloc(&outd, nthr * sizeof(double)));
  double ops = (double)nthr * ITERS * 8;
  float ms;
  ms = time_kernel(k_fma32, grid, block, outf, 0.999f, 0.001f);
  printf("fma32   : %.3f ms  %.2f Tops/s (%.2f TFLOP/s)\n", ms, ops / ms / 1e9, 2 * ops / ms / 1e9);
  ms = time_kernel(k_fma64, grid, block, outd, 0.999, 0.001);
  printf("fma64   : %.3f ms  %.2f Tops/s\n", ms, ops / ms / 1e9);
  ms = time_kernel(k_sin, grid, block, outf);
  printf("sin32   : %.3f ms  %.2f Tops/s\n", ms, ops / ms / 1e9);
  ms = time_kernel(k_exp, grid, block, outf);
  printf("exp2_32 : %.3f ms  %.2f Tops/s (includes a negate)\n", ms, ops / ms / 1e9);
  ms = time_kernel(k_fract64, grid, block, outd);
  printf("fract64 : %.3f ms  %.2f Tops/s (includes an add64)\n", ms, ops / ms / 1e9);
  ms = time_kernel(k_mix, grid, block, outf, 0.999f, 0.001f);
  printf("mix 1sin+4fma: %.3f ms  %.2f Tsin/s\n", ms, ops / ms / 1e9);
  double hq[3] = {3.1e9, 2.2e9, -0.4e9};
  double* dq; CHK(hipMalloc(&dq, sizeof(hq))); CHK(hipMemcpy(dq, hq, sizeof(hq), hipMemcpyHostToDevice));
  const int nat = 4096;
  ms = time_kernel(k_term, grid, block, outf, (const double*)dq, nat);
  printf("term loop: %.3f ms  %.3f Tterms/s\n", ms, (double)nthr * nat / ms / 1e9);

  // accuracy
  const int n = 1 << 22;
  std::vector<float> hx(n);
  for (int i = 0; i < n; ++i) hx[i] = -1.0f + 2.0f * ((float)i + 0.5f) / n;  // [-1,1) turns
  float *dx, *ds, *dc; double *drs, *drc;
  CHK(hipMalloc(&dx, n * 4)); CHK(hipMalloc(&ds, n * 4)); CHK(hipMalloc(&dc, n * 4));
  CHK(hipMalloc(&drs, n * 8)); CHK(hipMalloc(&drc, n * 8));
  CHK(hipMemcpy(dx, hx.data(), n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_acc, dim3((n + 255) / 256), dim3(256), 0, 0, dx, ds, dc, drs, drc, n);
  CHK(hipDeviceSynchronize());
  std::vector<float> hs(n), hc(n); std::vector<double> hrs(n), hrc(n);
  CHK(hipMemcpy(hs.data(), ds, n * 4, hipMemcpyDeviceToHost));
  CHK(hipMemcpy(hc.data(), dc, n * 4, hipMemcpyDeviceToHost));
  CHK(hipMemcpy(hrs.data(), drs, n * 8, hipMemcpyDeviceToHost));
  CHK(hipMemcpy(hrc.data(), drc, n * 8, hipMemcpyDeviceToHost));
  // bands of |x|
  const double bands[5] = {0.0, 0.125, 0.25, 0.5, 1.0};
  for (int b = 0; b < 4; ++b) {
    double maxs = 0, maxc = 0, ss = 0, sc = 0, ms_ = 0, mc_ = 0; long cnt = 0;
    for (int i = 0; i < n; ++i) {
      double ax = fabs((double)hx[i]);
      if (ax < bands[b] || ax >= bands[b + 1]) continue;
      double es = hs[i] - hrs[i], ec = hc[i] - hrc[i];
      maxs = fmax(maxs, fabs(es)); maxc = fmax(maxc, fabs(ec));
      ss += es * es; sc += ec * ec; ms_ += es; mc_ += ec; ++cnt;
    }
    printf("|x| in [%.3f,%.3f): sin max %.3e rms %.3e mean %.3e | cos max %.3e rms %.3e mean %.3e (n=%ld)\n",
           bands[b], bands[b + 1], maxs, sqrt(ss / cnt), ms_ / cnt, maxc, sqrt(sc / cnt), mc_ / cnt, cnt);
  }
  printf("sin(0.25 turn)=%.9f cos(0.25)=%.9f sin(0.125)=%.9f\n", hs[(int)((0.25 + 1.0) / 2.0 * n)], hc[(int)((0.25 + 1.0) / 2.0 * n)], hs[(int)((0.125 + 1.0) / 2.0 * n)]);
  printf("DONE\n");
  return 0;
}
