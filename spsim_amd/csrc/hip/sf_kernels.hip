// Structure-factor kernels for gfx950 (MI355X).
//
// F(q) = sum_j f_{Z_j}(|q|) w_j exp(-2 pi i q . r_j), I = |F|^2  (reference
// compute_pattern_on_list, src/diffraction.c:770-853).
//
// Layout of the work:
//  * A workgroup (4 waves, 256 lanes) owns a pixel tile; each lane owns K
//    pixels strided by 256, so the tile is 256*K pixels and global loads and
//    stores of q, F and I stay coalesced.
//  * Atoms arrive pre-grouped by species (one segment per element). For each
//    segment the form factor f_Z(|q|) is evaluated once per pixel, then the
//    segment's atoms are streamed through LDS in chunks of CHUNK records
//    (positions widened to FP64 when staged). Every lane reads the same atom
//    record (an LDS broadcast) and applies it to its K pixels.
//  * Phase: q.r is formed in FP64 (q and r are FP32 inputs, so each product is
//    exact) and reduced to a fraction of a turn with v_fract_f64 before the
//    conversion to FP32; v_sin_f32/v_cos_f32 take their argument in turns.
//  * Accumulation: FP32 partial sums over one LDS chunk, multiplied by f_Z and
//    added into FP64 per-pixel accumulators once per chunk (two-level summation).
//  * Per-atom Debye-Waller factor exp(-B_j s^2) (optional) is folded with the
//    amplitude into one exp2: w_j e^{-B_j s^2} = 2^(log2 w_j - B_j s^2 log2 e).
#include "sf_kernels.hpp"

#include <algorithm>

namespace spsim {

namespace {

constexpr int kThreads = 256;

struct __align__(16) AtomLds {
  double x, y;
  double z;
  float w;  // amplitude weight, or its log2 when per-atom B is present
  float b;  // per-atom B factor (A^2)
};
static_assert(sizeof(AtomLds) == 32, "LDS atom record must be 32 bytes");

// Cromer-Mann form factor times the global Debye-Waller factor, with s2 = s^2 in
// A^-2. Terms are added in the reference's order (src/diffraction.c:183-193).
__device__ __forceinline__ float form_factor(const float *__restrict__ c, float s2, float B) {
  float f = 0.0f;
  f += c[0] * expf(-(c[4] + B) * s2);
  f += c[1] * expf(-(c[5] + B) * s2);
  f += c[2] * expf(-(c[6] + B) * s2);
  f += c[3] * expf(-(c[7] + B) * s2);
  f += c[8] * expf(-B * s2);
  return f;
}

// Loads q for one pixel (zero for pixels past the end). With ROT the vector is
// rotated by R in FP64 with a fixed fma order and rounded to FP32, which is
// bit-identical to spsim_hip_rotate_hkl and the host rotation.
template <bool ROT>
__device__ __forceinline__ void load_q(const float *__restrict__ hkl, int pix, bool valid, const double *R, double &qx,
                                       double &qy, double &qz, float &s2) {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (valid) {
    x = hkl[3 * pix + 0];
    y = hkl[3 * pix + 1];
    z = hkl[3 * pix + 2];
  }
  if constexpr (ROT) {
    const double dx = x, dy = y, dz = z;
    const float rx = (float)__builtin_fma(R[2], dz, __builtin_fma(R[1], dy, R[0] * dx));
    const float ry = (float)__builtin_fma(R[5], dz, __builtin_fma(R[4], dy, R[3] * dx));
    const float rz = (float)__builtin_fma(R[8], dz, __builtin_fma(R[7], dy, R[6] * dx));
    x = rx;
    y = ry;
    z = rz;
  }
  qx = x;
  qy = y;
  qz = z;
  // s = |q| * 1e-10 / 2 in 1/A  ->  s^2 = |q|^2 * 2.5e-21
  s2 = (float)((qx * qx + qy * qy + qz * qz) * 2.5e-21);
}

// Sums the atoms [a_begin, a_end) of the given segments into Fr, Fi for the
// lane's K pixels. Must be called by every thread of the workgroup.
template <int K, int CHUNK, bool HAS_B>
__device__ __forceinline__ void tile_sum(const double (&qx)[K], const double (&qy)[K], const double (&qz)[K],
                                         const float (&s2)[K], const spsim_hip_atom *__restrict__ atoms,
                                         const float *__restrict__ atom_b, const spsim_hip_segment *__restrict__ segs,
                                         int nsegs, int a_begin, int a_end, const float *__restrict__ ff,
                                         float b_global, AtomLds *__restrict__ sm, double (&Fr)[K], double (&Fi)[K]) {
  float nl2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    nl2[k] = -s2[k] * (float)kLog2e;
    Fr[k] = 0.0;
    Fi[k] = 0.0;
  }
  for (int sg = 0; sg < nsegs; ++sg) {
    const int s_start = segs[sg].start;
    const int s_end = s_start + segs[sg].count;
    const int species = segs[sg].species;
    const int lo = max(s_start, a_begin);
    const int hi = min(s_end, a_end);
    if (lo >= hi) continue;
    const float Bs = (species == SPSIM_HIP_UNIT_SPECIES) ? 0.0f : b_global;
    const float *c = ff + species * SPSIM_HIP_FF_STRIDE;
    float f[K];
#pragma unroll
    for (int k = 0; k < K; ++k) f[k] = form_factor(c, s2[k], Bs);

    for (int c0 = lo; c0 < hi; c0 += CHUNK) {
      const int n = min(CHUNK, hi - c0);
      __syncthreads();  // previous chunk fully consumed
      for (int t = threadIdx.x; t < n; t += kThreads) {
        const spsim_hip_atom a = atoms[c0 + t];
        AtomLds r;
        r.x = a.x;
        r.y = a.y;
        r.z = a.z;
        if constexpr (HAS_B) {
          r.w = log2f(a.w);
          r.b = atom_b[c0 + t];
        } else {
          r.w = a.w;
          r.b = 0.0f;
        }
        sm[t] = r;
      }
      __syncthreads();

      float sr[K], si[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        sr[k] = 0.0f;
        si[k] = 0.0f;
      }
#pragma unroll 2
      for (int j = 0; j < n; ++j) {
        const AtomLds a = sm[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double t = __builtin_fma(qz[k], a.z, __builtin_fma(qy[k], a.y, qx[k] * a.x));
          const float ph = (float)__builtin_amdgcn_fract(t);
          const float cs = __builtin_amdgcn_cosf(ph);
          const float sn = __builtin_amdgcn_sinf(ph);
          float wt;
          if constexpr (HAS_B) {
            wt = __builtin_amdgcn_exp2f(__builtin_fmaf(a.b, nl2[k], a.w));
          } else {
            wt = a.w;
          }
          sr[k] = __builtin_fmaf(wt, cs, sr[k]);
          si[k] = __builtin_fmaf(wt, sn, si[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        Fr[k] = __builtin_fma((double)f[k], (double)sr[k], Fr[k]);
        Fi[k] = __builtin_fma((double)f[k], (double)si[k], Fi[k]);
      }
    }
  }
}

// One pattern: blockIdx.x = pixel tile, blockIdx.y = atom split.
template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_pattern_kernel(PatternLaunch a, double2 *__restrict__ partial) {
  __shared__ AtomLds sm[CHUNK];
  const int tile = blockIdx.x;
  const int split = blockIdx.y;
  const int nsplit = gridDim.y;
  const int a_begin = (int)(((long long)a.natoms * split) / nsplit);
  const int a_end = (int)(((long long)a.natoms * (split + 1)) / nsplit);
  const int base = tile * (kThreads * K) + threadIdx.x;

  double qx[K], qy[K], qz[K];
  float s2[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int pix = base + k * kThreads;
    load_q<false>(a.hkl, pix, pix < a.npix, nullptr, qx[k], qy[k], qz[k], s2[k]);
  }
  double Fr[K], Fi[K];
  tile_sum<K, CHUNK, HAS_B>(qx, qy, qz, s2, a.atoms, a.atom_b, a.segs, a.nsegs, a_begin, a_end, a.ff, a.b_global,
                            sm, Fr, Fi);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int pix = base + k * kThreads;
    if (pix >= a.npix) continue;
    // exp(-i phi): imaginary part is minus the sine sum.
    if (partial) {
      partial[(size_t)split * a.npix + pix] = make_double2(Fr[k], -Fi[k]);
    } else {
      if (a.F) a.F[pix] = make_float2((float)Fr[k], (float)(-Fi[k]));
      if (a.I) a.I[pix] = (float)(Fr[k] * Fr[k] + Fi[k] * Fi[k]);
    }
  }
}

// Deterministic reduction of the atom splits of one pattern, in split order.
__global__ __launch_bounds__(kThreads) void sf_reduce_kernel(const double2 *__restrict__ partial, int nsplit, int npix,
                                                             float2 *__restrict__ F, float *__restrict__ I) {
  const int pix = blockIdx.x * kThreads + threadIdx.x;
  if (pix >= npix) return;
  double re = 0.0, im = 0.0;
  for (int s = 0; s < nsplit; ++s) {
    const double2 v = partial[(size_t)s * npix + pix];
    re += v.x;
    im += v.y;
  }
  if (F) F[pix] = make_float2((float)re, (float)im);
  if (I) I[pix] = (float)(re * re + im * im);
}

__global__ void reset_queue_kernel(unsigned *q) { *q = 0u; }

// Persistent kernel over the (pattern, pixel-tile) work list. Each workgroup
// takes the next item from the queue until the list is exhausted; the item
// index is uniform across the workgroup, so every wave leaves the loop together.
template <int K, int CHUNK, bool HAS_B>
__global__ __launch_bounds__(kThreads) void sf_batch_kernel(BatchLaunch a) {
  __shared__ AtomLds sm[CHUNK];
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int item = s_item;
    if (item >= a.nitems) break;
    const int2 it = a.items[item];
    const int p = it.x;
    const MolInfo m = a.mols[a.pattern_mol[p]];
    const double *R = a.rot + 9 * (size_t)p;
    const int base = it.y * (kThreads * K) + threadIdx.x;

    double qx[K], qy[K], qz[K];
    float s2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int pix = base + k * kThreads;
      load_q<true>(a.hkl, pix, pix < a.npix, R, qx[k], qy[k], qz[k], s2[k]);
    }
    double Fr[K], Fi[K];
    // segment starts are relative to the molecule's first packed atom
    tile_sum<K, CHUNK, HAS_B>(qx, qy, qz, s2, a.atoms + m.atom_begin, HAS_B ? a.atom_b + m.atom_begin : nullptr,
                              a.segs + m.seg_begin, m.nseg, 0, m.atom_end - m.atom_begin, a.ff, a.b_global, sm, Fr,
                              Fi);
    const size_t off = (size_t)p * a.npix;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int pix = base + k * kThreads;
      if (pix >= a.npix) continue;
      if (a.F) a.F[off + pix] = make_float2((float)Fr[k], (float)(-Fi[k]));
      a.I[off + pix] = (float)(Fr[k] * Fr[k] + Fi[k] * Fi[k]);
    }
  }
}

// Instantiation table: (K, CHUNK) pairs x HAS_B.
#define SPSIM_FOR_EACH_TILE(X) \
  X(1, 64) X(1, 128) X(1, 256) \
  X(2, 64) X(2, 128) X(2, 256) \
  X(4, 64) X(4, 128) X(4, 256) \
  X(8, 64) X(8, 128) X(8, 256)

template <bool HAS_B>
const void *pattern_kernel_ptr(int K, int C) {
#define X(k_, c_) \
  if (K == k_ && C == c_) return (const void *)&sf_pattern_kernel<k_, c_, HAS_B>;
  SPSIM_FOR_EACH_TILE(X)
#undef X
  return nullptr;
}

template <bool HAS_B>
const void *batch_kernel_ptr(int K, int C) {
#define X(k_, c_) \
  if (K == k_ && C == c_) return (const void *)&sf_batch_kernel<k_, c_, HAS_B>;
  SPSIM_FOR_EACH_TILE(X)
#undef X
  return nullptr;
}

const void *pattern_kernel(const spsim_hip_tile &t, bool has_b) {
  return has_b ? pattern_kernel_ptr<true>(t.pixels_per_lane, t.chunk_atoms)
               : pattern_kernel_ptr<false>(t.pixels_per_lane, t.chunk_atoms);
}
const void *batch_kernel(const spsim_hip_tile &t, bool has_b) {
  return has_b ? batch_kernel_ptr<true>(t.pixels_per_lane, t.chunk_atoms)
               : batch_kernel_ptr<false>(t.pixels_per_lane, t.chunk_atoms);
}

}  // namespace

bool tile_supported(const spsim_hip_tile &t) {
  return t.waves == 4 && pattern_kernel(t, false) != nullptr && t.atom_splits >= 0;
}

spsim_hip_tile default_tile(int natoms, int npix, bool has_b) {
  (void)has_b;
  spsim_hip_tile t;
  t.waves = 4;
  t.pixels_per_lane = npix >= 4 * 256 * 256 ? 4 : (npix >= 2 * 256 * 64 ? 2 : 1);
  t.chunk_atoms = natoms >= 256 ? 256 : (natoms >= 128 ? 128 : 64);
  t.atom_splits = 0;
  return t;
}

int auto_splits(const spsim_hip_tile &t, int natoms, int npix, int cus) {
  const int tiles = (npix + tile_pixels(t) - 1) / tile_pixels(t);
  const long long want_items = 4LL * cus * 2;  // ~4 rounds of 2 workgroups per CU
  long long s = (want_items + tiles - 1) / std::max(tiles, 1);
  const long long max_s = std::max(1, natoms / std::max(1, 2 * t.chunk_atoms));
  s = std::min(s, max_s);
  s = std::max(1LL, std::min(s, 64LL));
  return (int)s;
}

int launch_pattern(const PatternLaunch &a, const spsim_hip_tile &t, int splits, hipStream_t s) {
  if (a.npix <= 0) return SPSIM_HIP_OK;
  const bool has_b = a.atom_b != nullptr;
  const void *k = pattern_kernel(t, has_b);
  if (!k) return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (K=%d chunk=%d)", t.pixels_per_lane, t.chunk_atoms);
  const int tiles = (a.npix + tile_pixels(t) - 1) / tile_pixels(t);
  if (splits < 1) splits = 1;
  double2 *partial = nullptr;
  if (splits > 1) {
    void *ws = nullptr;
    SPSIM_TRY(workspace(sizeof(double2) * (size_t)splits * a.npix, &ws));
    partial = (double2 *)ws;
  }
  PatternLaunch arg = a;
  void *params[] = {(void *)&arg, (void *)&partial};
  SPSIM_HIP_TRY(hipLaunchKernel(k, dim3(tiles, splits), dim3(kThreads), params, 0, s));
  if (splits > 1) {
    hipLaunchKernelGGL(sf_reduce_kernel, dim3((a.npix + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                       (const double2 *)partial, splits, a.npix, a.F, a.I);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  return SPSIM_HIP_OK;
}

int batch_blocks_per_cu(const spsim_hip_tile &t, bool has_b, int *blocks) {
  const void *k = batch_kernel(t, has_b);
  if (!k) return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (K=%d chunk=%d)", t.pixels_per_lane, t.chunk_atoms);
  int nb = 0;
  SPSIM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, kThreads, 0));
  *blocks = std::max(1, nb);
  return SPSIM_HIP_OK;
}

int launch_batch(const BatchLaunch &a, const spsim_hip_tile &t, int blocks, hipStream_t s) {
  const bool has_b = a.atom_b != nullptr;
  const void *k = batch_kernel(t, has_b);
  if (!k) return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (K=%d chunk=%d)", t.pixels_per_lane, t.chunk_atoms);
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, a.queue);
  SPSIM_HIP_TRY(hipGetLastError());
  if (a.nitems <= 0) return SPSIM_HIP_OK;
  BatchLaunch arg = a;
  void *params[] = {(void *)&arg};
  SPSIM_HIP_TRY(hipLaunchKernel(k, dim3(std::max(1, blocks)), dim3(kThreads), params, 0, s));
  return SPSIM_HIP_OK;
}

}  // namespace spsim

extern "C" int spsim_hip_sf_pattern(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                                    const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff,
                                    float b_global, float *d_F, float *d_I, const spsim_hip_tile *tile, void *stream) {
  using namespace spsim;
  if (npix < 0 || nsegs < 0 || natoms < 0)
    return fail(SPSIM_HIP_ERR_INVALID, "negative size (npix=%d nsegs=%d natoms=%d)", npix, nsegs, natoms);
  if (npix > 0 && !d_hkl) return fail(SPSIM_HIP_ERR_INVALID, "d_hkl is NULL");
  if (nsegs > 0 && (!d_atoms || !d_segs)) return fail(SPSIM_HIP_ERR_INVALID, "d_atoms/d_segs is NULL");
  if (nsegs == 0 && natoms != 0) return fail(SPSIM_HIP_ERR_INVALID, "natoms=%d but no segments", natoms);
  if (!d_ff) return fail(SPSIM_HIP_ERR_INVALID, "d_ff is NULL");
  if (npix > (1 << 28)) return fail(SPSIM_HIP_ERR_INVALID, "npix=%d exceeds 2^28", npix);
  int cus = 0;
  SPSIM_TRY(ensure_device(&cus));
  if (npix == 0) return SPSIM_HIP_OK;
  spsim_hip_tile t = tile ? *tile : default_tile(natoms, npix, d_atom_b != nullptr);
  if (!tile_supported(t))
    return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (pixels_per_lane=%d waves=%d chunk_atoms=%d)",
                t.pixels_per_lane, t.waves, t.chunk_atoms);
  const int splits = t.atom_splits > 0 ? t.atom_splits : auto_splits(t, natoms, npix, cus);
  PatternLaunch a;
  a.hkl = d_hkl;
  a.npix = npix;
  a.atoms = d_atoms;
  a.atom_b = d_atom_b;
  a.segs = d_segs;
  a.nsegs = nsegs;
  a.natoms = natoms;
  a.ff = d_ff;
  a.b_global = b_global;
  a.F = (float2 *)d_F;
  a.I = d_I;
  return launch_pattern(a, t, splits, (hipStream_t)stream);
}
