// Adjoint of the structure factor on gfx950: gradients of a loss L(I) with
// dL/dI_p = g_p given, with respect to atom positions r_j and per-atom
// Debye-Waller factors B_j, plus the Poisson negative log-likelihood that
// produces g.
//
// With F_p = sum_j a_pj e^{-i phi_pj}, phi_pj = 2 pi q_p.r_j and
// a_pj = f_s(|q_p|) w_j e_pj, e_pj = exp(-B_j s_p^2), I_p = |F_p|^2:
//   dL/dr_j = -2 pi w_j sum_p e_pj q_p (Ar_p sin phi + Ai_p cos phi)
//   dL/dB_j =     -w_j sum_p e_pj s_p^2 (Ar_p cos phi - Ai_p sin phi)
// where (Ar_p, Ai_p) = 2 g_p f_s(|q_p|) (Re F_p, Im F_p).
//
// Layout: a workgroup owns up to 256*KA atoms of one species (each lane KA
// atoms, positions in FP64 registers) and a range of pixels, which it streams
// through LDS in chunks of 256 records {q (FP64 and FP32), Ar, Ai, s^2}; the
// per-pixel form factor and F-weights are formed once per chunk while staging.
// Phases use the forward kernel's arithmetic (FP64 q.r, fract, FP32 sincos).
// Sums are FP32 within a chunk and FP64 across chunks; the pixel ranges of one
// atom tile are reduced in a fixed order by a second kernel (deterministic).
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "form_factor.hpp"

namespace spsim {
namespace {

constexpr int kT = 256;
constexpr int kCP = 256;  // pixels per LDS chunk

struct __align__(16) PixRec {
  double qx, qy;
  double qz;
  float qxf, qyf;
  float qzf, ar;
  float ai, s2;
  float nl2, pad;
};
static_assert(sizeof(PixRec) == 56 || sizeof(PixRec) == 64, "pixel record size");

struct AtomTile {
  int start;    // first packed atom
  int count;    // atoms in the tile (<= 256 * KA)
  int species;
  int pad;
};

struct AdjArgs {
  const float *hkl;
  int npix;
  const spsim_hip_atom *atoms;
  const float *atom_b;  // null when no per-atom B
  const AtomTile *tiles;
  const float *ff;
  float b_global;
  const float2 *F;
  const float *gI;
  double4 *partial;  // [nsplit][natoms] {Ux, Uy, Uz, V}
  int natoms;
};

template <int KA, bool HAS_B, bool WANT_B>
__global__ __launch_bounds__(kT) void adjoint_kernel(AdjArgs a) {
  __shared__ PixRec sm[kCP];
  const AtomTile tile = a.tiles[blockIdx.x];
  const int split = blockIdx.y, nsplit = gridDim.y;
  const int p_begin = (int)(((long long)a.npix * split) / nsplit);
  const int p_end = (int)(((long long)a.npix * (split + 1)) / nsplit);
  const float Bs = (tile.species == SPSIM_HIP_UNIT_SPECIES) ? 0.0f : a.b_global;
  const float *c = a.ff + tile.species * SPSIM_HIP_FF_STRIDE;

  double rx[KA], ry[KA], rz[KA];
  float bj[KA];
  double Ux[KA], Uy[KA], Uz[KA], V[KA];
#pragma unroll
  for (int k = 0; k < KA; ++k) {
    const int i = threadIdx.x + k * kT;
    const bool ok = i < tile.count;
    const spsim_hip_atom at = ok ? a.atoms[tile.start + i] : spsim_hip_atom{0.f, 0.f, 0.f, 0.f};
    rx[k] = at.x;
    ry[k] = at.y;
    rz[k] = at.z;
    bj[k] = (HAS_B && ok) ? a.atom_b[tile.start + i] : 0.0f;
    Ux[k] = Uy[k] = Uz[k] = V[k] = 0.0;
  }

  for (int p0 = p_begin; p0 < p_end; p0 += kCP) {
    const int n = min(kCP, p_end - p0);
    __syncthreads();
    if (threadIdx.x < n) {
      const int p = p0 + threadIdx.x;
      const float x = a.hkl[3 * p], y = a.hkl[3 * p + 1], z = a.hkl[3 * p + 2];
      PixRec r;
      r.qx = x;
      r.qy = y;
      r.qz = z;
      r.qxf = x;
      r.qyf = y;
      r.qzf = z;
      const float s2 = (float)(((double)x * x + (double)y * y + (double)z * z) * 2.5e-21);
      const float G = 2.0f * a.gI[p] * form_factor(c, s2, Bs);
      const float2 Fp = a.F[p];
      r.ar = G * Fp.x;
      r.ai = G * Fp.y;
      r.s2 = s2;
      r.nl2 = -s2 * (float)kLog2e;
      r.pad = 0.0f;
      sm[threadIdx.x] = r;
    }
    __syncthreads();
    float ux[KA], uy[KA], uz[KA], v[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) ux[k] = uy[k] = uz[k] = v[k] = 0.0f;
    for (int i = 0; i < n; ++i) {
      const PixRec r = sm[i];
#pragma unroll
      for (int k = 0; k < KA; ++k) {
        const double t = __builtin_fma(r.qz, rz[k], __builtin_fma(r.qy, ry[k], r.qx * rx[k]));
        const float ph = (float)__builtin_amdgcn_fract(t);
        const float cs = __builtin_amdgcn_cosf(ph);
        const float sn = __builtin_amdgcn_sinf(ph);
        float u = __builtin_fmaf(r.ar, sn, r.ai * cs);
        float e = 1.0f;
        if constexpr (HAS_B) {
          e = __builtin_amdgcn_exp2f(bj[k] * r.nl2);
          u *= e;
        }
        ux[k] = __builtin_fmaf(u, r.qxf, ux[k]);
        uy[k] = __builtin_fmaf(u, r.qyf, uy[k]);
        uz[k] = __builtin_fmaf(u, r.qzf, uz[k]);
        if constexpr (WANT_B) {
          const float w = __builtin_fmaf(r.ar, cs, -(r.ai * sn));
          v[k] = __builtin_fmaf(e * r.s2, w, v[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      Ux[k] += ux[k];
      Uy[k] += uy[k];
      Uz[k] += uz[k];
      V[k] += v[k];
    }
  }
#pragma unroll
  for (int k = 0; k < KA; ++k) {
    const int i = threadIdx.x + k * kT;
    if (i < tile.count) a.partial[(size_t)split * a.natoms + tile.start + i] = make_double4(Ux[k], Uy[k], Uz[k], V[k]);
  }
}

__global__ __launch_bounds__(kT) void adjoint_reduce_kernel(const double4 *__restrict__ partial, int nsplit, int natoms,
                                                            const spsim_hip_atom *__restrict__ atoms,
                                                            float *__restrict__ grad_pos, float *__restrict__ grad_b) {
  const int j = blockIdx.x * kT + threadIdx.x;
  if (j >= natoms) return;
  double ux = 0, uy = 0, uz = 0, v = 0;
  for (int s = 0; s < nsplit; ++s) {
    const double4 p = partial[(size_t)s * natoms + j];
    ux += p.x;
    uy += p.y;
    uz += p.z;
    v += p.w;
  }
  const double w = atoms[j].w;
  const double k = -6.283185307179586476925286766559 * w;
  grad_pos[3 * j + 0] = (float)(k * ux);
  grad_pos[3 * j + 1] = (float)(k * uy);
  grad_pos[3 * j + 2] = (float)(k * uz);
  if (grad_b) grad_b[j] = (float)(-w * v);
}

template <int KA, bool HAS_B, bool WANT_B>
const void *adjoint_ptr() {
  return (const void *)&adjoint_kernel<KA, HAS_B, WANT_B>;
}

// Poisson NLL: per-block partial sums of the loss (FP64) and g = dL/dI.
__global__ __launch_bounds__(kT) void nll_kernel(const float *__restrict__ I, const float *__restrict__ y,
                                                 const float *__restrict__ K, int npix, double *__restrict__ block_sums,
                                                 float *__restrict__ gI) {
  __shared__ double red[kT];
  const int p = blockIdx.x * kT + threadIdx.x;
  double l = 0.0;
  if (p < npix) {
    const double k = K ? (double)K[p] : 1.0;
    const double lam = fmax(k * (double)I[p], 1e-30);
    const double yy = y[p];
    l = lam - (yy > 0.0 ? yy * log(lam) : 0.0);
    if (gI) gI[p] = (float)(k * (1.0 - yy / lam));
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kT) void sum_kernel(const double *__restrict__ v, int n, double *__restrict__ out) {
  __shared__ double red[kT];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) acc += v[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace
}  // namespace spsim

extern "C" int spsim_hip_sf_adjoint(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                                    const spsim_hip_segment *h_segs, int nsegs, int natoms, const float *d_ff,
                                    float b_global, const float *d_F, const float *d_gI, float *d_grad_pos,
                                    float *d_grad_b, void *stream) {
  using namespace spsim;
  if (npix < 0 || nsegs < 0 || natoms < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative size");
  if (natoms > 0 && (!d_atoms || !h_segs || !d_grad_pos)) return fail(SPSIM_HIP_ERR_INVALID, "NULL atom arrays");
  if (npix > 0 && (!d_hkl || !d_F || !d_gI)) return fail(SPSIM_HIP_ERR_INVALID, "NULL pixel arrays");
  if (!d_ff) return fail(SPSIM_HIP_ERR_INVALID, "d_ff is NULL");
  int cus = 0;
  SPSIM_TRY(ensure_device(&cus));
  hipStream_t s = (hipStream_t)stream;
  if (natoms == 0) return SPSIM_HIP_OK;
  constexpr int KA = 2;
  const int per_tile = kT * KA;
  std::vector<AtomTile> tiles;
  int covered = 0;
  for (int i = 0; i < nsegs; ++i) {
    const spsim_hip_segment &g = h_segs[i];
    if (g.start != covered || g.count < 0 || g.species < 0 || g.species >= SPSIM_HIP_NSPECIES)
      return fail(SPSIM_HIP_ERR_INVALID, "segments must tile [0, natoms) in order");
    for (int a0 = 0; a0 < g.count; a0 += per_tile)
      tiles.push_back(AtomTile{g.start + a0, std::min(per_tile, g.count - a0), g.species, 0});
    covered += g.count;
  }
  if (covered != natoms) return fail(SPSIM_HIP_ERR_INVALID, "segments cover %d atoms, natoms=%d", covered, natoms);
  const bool hb = d_atom_b != nullptr, wb = d_grad_b != nullptr;
  const void *k = hb ? (wb ? adjoint_ptr<KA, true, true>() : adjoint_ptr<KA, true, false>())
                     : (wb ? adjoint_ptr<KA, false, true>() : adjoint_ptr<KA, false, false>());
  // Enough pixel ranges to fill every CU to the kernel's occupancy (5 or 6
  // workgroups at 2 atoms per lane; 4 left a third of the wave slots empty),
  // each range at least a few chunks.
  int per_cu = 4;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kT, 0) != hipSuccess || per_cu < 1) per_cu = 4;
  const long long want = (long long)per_cu * cus;
  int nsplit = (int)std::max<long long>(1, (want + (long long)tiles.size() - 1) / (long long)tiles.size());
  nsplit = std::max(1, std::min(nsplit, std::max(1, npix / (4 * kCP))));
  const size_t tiles_bytes = ((tiles.size() * sizeof(AtomTile) + 255) / 256) * 256;
  void *ws = nullptr;
  SPSIM_TRY(workspace(tiles_bytes + sizeof(double4) * (size_t)nsplit * natoms, &ws));
  AtomTile *d_tiles = (AtomTile *)ws;
  double4 *partial = (double4 *)((char *)ws + tiles_bytes);
  SPSIM_HIP_TRY(hipMemcpyAsync(d_tiles, tiles.data(), tiles.size() * sizeof(AtomTile), hipMemcpyHostToDevice, s));
  if (npix == 0) {
    SPSIM_HIP_TRY(hipMemsetAsync(d_grad_pos, 0, sizeof(float) * 3 * (size_t)natoms, s));
    if (d_grad_b) SPSIM_HIP_TRY(hipMemsetAsync(d_grad_b, 0, sizeof(float) * (size_t)natoms, s));
    return SPSIM_HIP_OK;
  }
  AdjArgs a{d_hkl, npix, d_atoms, d_atom_b, d_tiles, d_ff, b_global, (const float2 *)d_F, d_gI, partial, natoms};
  void *params[] = {(void *)&a};
  SPSIM_HIP_TRY(hipLaunchKernel(k, dim3((unsigned)tiles.size(), nsplit), dim3(kT), params, 0, s));
  hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((natoms + kT - 1) / kT), dim3(kT), 0, s, (const double4 *)partial,
                     nsplit, natoms, d_atoms, d_grad_pos, d_grad_b);
  SPSIM_HIP_TRY(hipGetLastError());
  // The tile list was copied from host memory that goes out of scope.
  SPSIM_HIP_TRY(hipStreamSynchronize(s));
  return SPSIM_HIP_OK;
}

extern "C" int spsim_hip_poisson_nll(const float *d_I, const float *d_y, const float *d_K, int npix, double *d_loss,
                                     float *d_gI, void *stream) {
  using namespace spsim;
  if (npix < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative npix");
  if (npix > 0 && (!d_I || !d_y)) return fail(SPSIM_HIP_ERR_INVALID, "NULL input");
  SPSIM_TRY(ensure_device(nullptr));
  hipStream_t s = (hipStream_t)stream;
  if (npix == 0) {
    if (d_loss) SPSIM_HIP_TRY(hipMemsetAsync(d_loss, 0, sizeof(double), s));
    return SPSIM_HIP_OK;
  }
  const int nb = (npix + kT - 1) / kT;
  void *ws = nullptr;
  SPSIM_TRY(workspace(sizeof(double) * (size_t)nb, &ws));
  hipLaunchKernelGGL(nll_kernel, dim3(nb), dim3(kT), 0, s, d_I, d_y, d_K, npix, (double *)ws, d_gI);
  if (d_loss) hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kT), 0, s, (const double *)ws, nb, d_loss);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}
