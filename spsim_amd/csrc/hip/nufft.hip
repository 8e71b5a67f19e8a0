// Type-3 non-uniform FFT structure factor for gfx950 (MI355X): the
// asymptotic alternative to the direct sum, replacing the NNFFT path of
// compute_pattern_on_list_by_nfft (src/diffraction.c:528-675).
//
// The reference runs one NNFFT (non-equispaced nodes in space and frequency,
// NFFT3 library) per element present, with unit strengths at the atoms,
// multiplies each element's result by f_Z(|q|) (global B only) and sums:
//   F(q) = sum_Z f_Z(|q|) sum_{j in Z} exp(-2 pi i q.r_j),  I = |F|^2.
// This file computes the same sums with its own type-3 NUFFT (the method of
// Lee & Greengard, SIAM J. Sci. Comput. 2005, with the "exponential of
// semicircle" window of Barnett, Magland & af Klinteberg, SIAM J. Sci.
// Comput. 41 (2019) C479-C504), per axis d and with window width W:
//
//  1. Centre: x' = x - cx, s' = q - cs (cx, cs = centres of the atom and q
//     boxes; half-widths X, S). exp(-2 pi i q.x) = e^{-2 pi i cs.x'}
//     e^{-2 pi i s'.x'} e^{-2 pi i q.cx}: a phase per atom, the centred sum,
//     a phase per q.
//  2. Spread: grid spacing h = 1/(2 sigma1 S) (sigma1 = 2); every atom adds
//     c_j phi1((m - x'_j/h) 2/W) D_m to W^3 grid points m of its species'
//     grid, where phi(z) = exp(beta (sqrt(1 - z^2) - 1)) (beta1 = 2.30 W) and
//     D_m deconvolves the second window in advance.
//  3. Pruned FFT: along x, y, then z, each grid line (M points) is zero-padded
//     to nf >= sigma2 M points, transformed in LDS (Stockham) and only the L
//     outputs the targets can reach are kept. Lengths up to 1920 with a plan
//     of at most 3 stages run a kernel compiled for that length (fft_fixed,
//     composite radices up to 16); others the general kernel (fft_lines,
//     radices 8/4/2/3/5 chosen at run time).
//  4. Interpolate: each q gathers W2^3 FFT values per species with the
//     second window phi2 (width W2, beta2), divides by the first window's
//     transform at s', multiplies by f_Z(|q|), sums the species and applies
//     the q phase.
// The second window is either the first one (sigma2 = 2, W2 = W) or, when
// the interpolation dominates the cost (many q or many orientations per
// grid), a narrower one on a finer frequency grid: sigma2 = 4, W2 = W - 1,
// beta2 = 2.65 W2, which keeps the error of W = 7..10 (the first window's
// aliasing dominates) at 0.63 of the gathers.
// With per-atom B factors the spread window is instead a Gaussian per atom
// whose transform carries the atom's Debye-Waller factor (see GaussSpread).
// The error of the centred sum is about 10^{1-W} relative to the number of
// atoms (W = 7 by default); the grids and FFT run in FP32, the phases in FP64.
// The spread sums each grid cell's atoms in registers without FP atomics, in
// an order fixed by the input (a stable sort bins the atoms by index; the
// candidates of a block are taken in lane order), so F repeats bit for bit.
#include <array>
#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "form_factor.hpp"

namespace spsim {
namespace {

constexpr int kThreads = 256;
constexpr int kMinW = 4, kMaxW = 10;
constexpr int kMaxQuad = 2 * kMaxW + 6;
// LDS per FFT workgroup: one buffer of T lines (T nf <= kFftLds complex values,
// 15 KiB; kFixedLds, 16 KiB, for the fixed-length kernels, so that 4 lines of
// 500 points fit) and the twiddle table (up to 15 KiB), or one line of up to
// kMaxNf points (30 KiB, twiddles read from global memory); plus two pointers
// per line.
constexpr int kMaxNf = 3840;      // longest FFT line
constexpr int kFftLds = 1920;     // complex values per LDS buffer for several lines
constexpr int kFixedLds = 2048;   // the same for the fixed-length kernels (16 KiB)
constexpr int kMaxStages = 16;
constexpr int kMaxT = 256;          // lines per FFT workgroup
constexpr double kSigma = 2.0;     // oversampling of the spread window (and of the default second window)
constexpr double kSigmaFine = 4.0;  // oversampling of the narrow second window
constexpr size_t kMaxGridBytes = (size_t)64 << 30;

// Row-major R with q' = R q; the atoms are rotated instead (R^T r, FP64), which
// gives the same phase (R q).r = q.(R^T r) and keeps the q list untouched.
struct AtomRot {
  int on;
  double r[9];
};

__device__ __forceinline__ void rotate_atom(const AtomRot &R, double x, double y, double z, double p[3]) {
  if (R.on) {
    p[0] = __builtin_fma(R.r[6], z, __builtin_fma(R.r[3], y, R.r[0] * x));
    p[1] = __builtin_fma(R.r[7], z, __builtin_fma(R.r[4], y, R.r[1] * x));
    p[2] = __builtin_fma(R.r[8], z, __builtin_fma(R.r[5], y, R.r[2] * x));
  } else {
    p[0] = x;
    p[1] = y;
    p[2] = z;
  }
}

struct NufftDim {
  double cx, cs, h;  // atom centre, q centre (1/m), grid spacing (m)
  int m_lo, M;       // spread grid: m in [m_lo, m_lo + M)
  int nf;            // FFT length
  int l_lo, L;       // kept FFT outputs: l in [l_lo, l_lo + L)
};

constexpr int kCorrDeg = 8;  // degree of the correction polynomial in u

struct NufftGeom {
  NufftDim d[3];
  float beta1, beta2;  // spread window (width W), interpolation window (width W2)
  int nq;
  float qz[kMaxQuad];  // Gauss-Legendre nodes on [0, 1]
  float qc[kMaxQuad];  // 2 * weight * phi2(node): phi2_hat(xi) = sum qc cos(2 pi xi qz)
  // 2 / (W phi1_hat(xi)) for |xi| <= corr_xmax as a polynomial in
  // u = 2 (xi / corr_xmax)^2 - 1 (monomial coefficients, constant first);
  // xi = t W / 2 = t * corr_scale * corr_xmax for a target at t = s' h
  float corr_xmax, corr_scale;
  // > 0: Gaussian spreading; the correction is exp(2 pi^2 gauss_v0 t^2) per axis
  float gauss_v0;
  int lds_interp;  // host: interpolate with nufft_interp_lds (else nufft_interp)
  int box_cap;      // grid points per wave of the staged interpolation's LDS (<= kBoxCap)
  // > 0: the interpolation reads the q list as a tile_n x tile_n image, one
  // 16 x 16 pixel tile per workgroup (16 x 4 per wave)
  int tile_n;
  float corr[kCorrDeg + 1];
};

// Exact n / d for n, d < 2^16 as q = trunc(fma(n, fi, fh)) in FP32, with
// fi = 1/d rounded and fh = fi / 2: (n + 1/2) / d lies at least 1/(2d) from an
// integer, and the FP32 result's error, below (2^16 / d + 1) 2^-23 in absolute
// terms, is smaller than that. Full-rate instructions only, for the FFT's
// indices.
struct FastDiv {
  unsigned d;
  float fi, fh;
};

struct FftStage {
  int R;              // radix
  int stride;         // nf / (span R): twiddle table step
  FastDiv nb, span;   // butterflies per line (nf / R), product of the earlier radices
};

struct FftPass {
  long long nlines;  // lines = outer * inner
  long long inner;   // stride between consecutive points of a line (1: lines are contiguous)
  int M, nf, L, m_lo, l_lo;
  int m0, l0;        // slots of point 0 and output 0: m_lo, l_lo mod nf
  // the fast modes' offsets in points: inner, first-stage nb, m0, M, last-stage
  // span, l0, L and nf, each times inner
  int inner32, nb_in, m0_in, M_in, span_in, l0_in, L_in, nf_in;
  int T;             // lines per workgroup (<= kMaxT)
  int nsum;          // input = the sum of nsum grids sum_stride apart (fixed order; 1: one grid)
  long long sum_stride;
  FastDiv dT;
  int nstages;
  FftStage st[kMaxStages];
};

__device__ __forceinline__ float es_window(float z, float beta) {
  const float r = 1.0f - z * z;
  return r > 0.0f ? expf(beta * (sqrtf(r) - 1.0f)) : 0.0f;
}

// Fourier transform of the interpolation window at xi (cycles per half-width), by quadrature.
__device__ __forceinline__ float es_hat(const NufftGeom &g, float xi) {
  float s = 0.0f;
  for (int i = 0; i < g.nq; ++i) s += g.qc[i] * cospif(2.0f * xi * g.qz[i]);
  return s;
}

__device__ __forceinline__ double nan_min(double a, double b) { return (a != a || b != b) ? a + b : fmin(a, b); }
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }

// Per-block minima and maxima of the atom positions (0..2) and q (3..5), the
// largest |q|^2 (12), and, when atom_b is given, the largest per-atom B (13)
// and the largest -B (14): part[blk * kNb + v], v < 6 minima, 6..14 maxima.
// NaN propagates.
constexpr int kNb = 15;
constexpr int kBoundBlocks = 256;  // workgroups of bounds_partial (records folded on the host)

__device__ __forceinline__ double bound_combine(int v, double a, double b) {
  return v < 6 ? nan_min(a, b) : nan_max(a, b);
}

// One workgroup's record over atoms (and q) blk, blk + nblk, ... in steps of
// kThreads-wide strips, written to rec[0 .. kNb).
__device__ __forceinline__ void bounds_block(const spsim_hip_atom *__restrict__ atoms,
                                             const float *__restrict__ atom_b, int natoms,
                                             const float *__restrict__ hkl, int npix, const AtomRot &R, int blk,
                                             int nblk, double *__restrict__ rec) {
  double acc[kNb];
  for (int v = 0; v < kNb; ++v) acc[v] = v < 6 ? INFINITY : -INFINITY;
  const int stride = nblk * kThreads;
  for (int i = blk * kThreads + threadIdx.x; i < natoms; i += stride) {
    const spsim_hip_atom a = atoms[i];
    double p[3];
    rotate_atom(R, a.x, a.y, a.z, p);
    for (int v = 0; v < 3; ++v) {
      acc[v] = nan_min(acc[v], p[v]);
      acc[6 + v] = nan_max(acc[6 + v], p[v]);
    }
    if (atom_b) {
      acc[13] = nan_max(acc[13], (double)atom_b[i]);
      acc[14] = nan_max(acc[14], -(double)atom_b[i]);
    }
  }
  for (int i = blk * kThreads + threadIdx.x; i < npix; i += stride) {
    double n2 = 0.0;
    for (int v = 0; v < 3; ++v) {
      const double q = hkl[3 * (size_t)i + v];
      acc[3 + v] = nan_min(acc[3 + v], q);
      acc[9 + v] = nan_max(acc[9 + v], q);
      n2 += q * q;
    }
    acc[12] = nan_max(acc[12], n2);
  }
  __shared__ double red[kNb][kThreads];
  for (int v = 0; v < kNb; ++v) red[v][threadIdx.x] = acc[v];
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int v = 0; v < kNb; ++v) red[v][threadIdx.x] = bound_combine(v, red[v][threadIdx.x], red[v][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < kNb) rec[threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(kThreads) void bounds_partial(const spsim_hip_atom *__restrict__ atoms,
                                                          const float *__restrict__ atom_b, int natoms,
                                                          const float *__restrict__ hkl, int npix, AtomRot R,
                                                          double *__restrict__ part) {
  bounds_block(atoms, atom_b, natoms, hkl, npix, R, blockIdx.x, gridDim.x, part + (size_t)blockIdx.x * kNb);
}

// The atoms' records of many molecules in one launch: workgroup (x, m) covers
// molecule m's atoms [off[m], off[m + 1]) rotated by rot[9 m ..] (rot NULL: no
// rotation), into part[(m gridDim.x + x) kNb ..]. The q entries stay at their
// empty values (+inf minima, -inf maxima).
__global__ __launch_bounds__(kThreads) void bounds_molecules(const spsim_hip_atom *__restrict__ atoms,
                                                            const float *__restrict__ atom_b,
                                                            const int *__restrict__ off,
                                                            const double *__restrict__ rot,
                                                            double *__restrict__ part) {
  const int m = blockIdx.y;
  AtomRot R{};
  if (rot) {
    R.on = 1;
    for (int i = 0; i < 9; ++i) R.r[i] = rot[9 * (size_t)m + i];
  }
  const int a0 = off[m];
  bounds_block(atoms + a0, atom_b ? atom_b + a0 : nullptr, off[m + 1] - a0, nullptr, 0, R, blockIdx.x, gridDim.x,
               part + ((size_t)m * gridDim.x + blockIdx.x) * kNb);
}

// The per-call tables in one launch, thread i filling entry i of each:
//  - the deconvolution table D_m = 2 / (W phi2_hat(m W / (2 nf))) (W: the
//    interpolation window's width) for the spread grid of every axis, stored
//    back to back: dec[0 .. Mx) | dec[Mx .. Mx+My) | dec[.. + Mz);
//  - the twiddles tw[j] = exp(-2 pi i j / nf) of the three FFT lengths, back to
//    back (nfx | nfy | nfz);
//  - n4 zeroed 16-byte words at zero (the bins' start and end), grid-strided.
__device__ __forceinline__ void setup_tables(const NufftGeom &g, int W, float *__restrict__ dec,
                                             float2 *__restrict__ tw, float4 *__restrict__ zero, size_t n4, int i,
                                             int stride) {
  int off = 0, toff = 0;
  for (int d = 0; d < 3; ++d) {
    const NufftDim &D = g.d[d];
    if (i < D.M) {
      const float xi = (float)((double)(D.m_lo + i) * W / (2.0 * D.nf));
      dec[off + i] = 2.0f / ((float)W * es_hat(g, xi));
    }
    if (i < D.nf) {
      double sn, cs;
      sincospi(-2.0 * (double)i / (double)D.nf, &sn, &cs);
      tw[toff + i] = make_float2((float)cs, (float)sn);
    }
    off += D.M;
    toff += D.nf;
  }
  for (size_t k = i; k < n4; k += (size_t)stride) zero[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Per-atom spreading record: species grid, first grid point per axis, the
// window offset (first point - u, in grid units) per axis, the window's width
// in points, the phased strength and, for Gaussian windows, 1 / (2 v) per axis.
struct __align__(16) SpreadAtom {
  int ch, bx, by, bz;
  float fx, fy, fz;
  int w;
  float cr, ci, ax, ay;
  float az, kz, pad1, pad2;
};

// Gaussian spreading (atoms with their own B factors): the atom is spread as
// the normalised Gaussian of variance v_d = v0 + sigma^2 / h_d^2 grid units^2
// per axis, whose transform exp(-2 pi^2 v_d t^2) at t = q_d h_d is
// exp(-2 pi^2 v0 t^2) (divided out after the interpolation) times the atom's
// Debye-Waller factor exp(-2 pi^2 sigma^2 q_d^2), with sigma^2 =
// B 2.5e-21 / (2 pi^2) m^2 for B in A^2 (exp(-B s^2), s = |q| / 2 in 1/A).
// On an axis centred in frequency (t = (q - cs) h) the factor
// exp(-2 pi^2 sigma^2 (cs + t / h)^2) comes from the same Gaussian moved by the
// imaginary offset -2 pi i sigma^2 cs / h: its weights gain the phase
// exp(-2 pi i k (m - u)), k = sigma^2 cs / (h v), and the atom the factor
// exp(-2 pi^2 sigma^2 cs^2 (1 - sigma^2 / (h^2 v))). Only z is centred in
// this mode (the grids of x and y are symmetric about q = 0).
// The window keeps the points within sqrt(2 v tcut) of the atom
// (truncation exp(-tcut)); v0 = tcut / pi^2 balances that against the aliasing
// of the sigma = 2 grid, exp(-pi^2 v0).
struct GaussSpread {
  float v0, tcut;
};
constexpr double kBToSigma2 = 2.5e-21 / (2.0 * M_PI * M_PI);  // m^2 per A^2 of B

// Atoms are binned by the kSpreadBlock^3 grid block that holds the first point
// of their window (their home block). One 64-lane workgroup per output block
// then gathers the atoms of the home blocks whose windows can reach it; each
// lane owns one (x, y) column of the block and keeps its kSpreadBlock cells
// in registers, so the species grids need no zeroing and no atomics.
constexpr int kSpreadBlock = 8;
static_assert(kSpreadBlock * kSpreadBlock == 64, "one lane per (x, y) column of a spread block");

struct SpreadBins {
  int nbx, nby, nbz;  // spread blocks per axis
  long long nbins;    // nch * nbz * nby * nbx
};

__device__ __forceinline__ long long home_bin(const SpreadAtom &r, const SpreadBins &b) {
  return (((long long)r.ch * b.nbz + r.bz / kSpreadBlock) * b.nby + r.by / kSpreadBlock) * b.nbx + r.bx / kSpreadBlock;
}

// One thread per packed atom: the record above (ch = -1 marks an atom outside
// the grid, which the bounds exclude), and the binning's sort key (home block;
// nbins for an atom outside the grid) and value (the atom's index).
// atom_b (packed order, may be NULL): Gaussian windows of per-atom width
// (at most W points); else the exponential-of-semicircle window of width W.
__device__ __forceinline__ void spread_prep(const spsim_hip_atom *__restrict__ atoms, int natoms,
                                            const spsim_hip_segment *__restrict__ segs, int nsegs,
                                            const AtomRot &R, const NufftGeom &g, int W,
                                            const float *__restrict__ atom_b, GaussSpread gs,
                                            SpreadAtom *__restrict__ rec, SpreadBins bins,
                                            unsigned *__restrict__ keys, int *__restrict__ index, int j) {
  if (j >= natoms) return;
  int ch = 0;
  for (int s = 1; s < nsegs; ++s)
    if (segs[s].start <= j) ch = s;
  const spsim_hip_atom a = atoms[j];
  double pr[3];
  rotate_atom(R, a.x, a.y, a.z, pr);
  const double xc[3] = {pr[0] - g.d[0].cx, pr[1] - g.d[1].cx, pr[2] - g.d[2].cx};
  double ph = -(xc[0] * g.d[0].cs + xc[1] * g.d[1].cs + xc[2] * g.d[2].cs);
  ph -= rint(ph);
  float sn, cs;
  sincospif((float)(2.0 * ph), &sn, &cs);
  SpreadAtom r;
  int b[3];
  float f[3], ia[3] = {0.0f, 0.0f, 0.0f};
  // window: W points centred on u (exponential of semicircle), or for a
  // Gaussian the points within sqrt(2 v tcut) of u, the widest axis's count
  // on every axis (so the record holds one width)
  int w = W;
  double half = 0.5 * W, norm = 1.0, kz = 0.0;
  if (atom_b) {
    const double sig2 = (double)atom_b[j] * kBToSigma2;
    int hmax = 0;
    for (int d = 0; d < 3; ++d) {
      const double h = g.d[d].h;
      const double v = (double)gs.v0 + sig2 / (h * h);
      const int hw = (int)ceil(sqrt(2.0 * v * (double)gs.tcut));
      hmax = hw > hmax ? hw : hmax;
      ia[d] = (float)(0.5 / v);
      norm /= sqrt(2.0 * M_PI * v);
      if (d == 2) {
        const double cs = g.d[2].cs;
        kz = sig2 * cs / (h * v);
        norm *= exp(-2.0 * M_PI * M_PI * sig2 * cs * cs * (1.0 - sig2 / (h * h * v)));
      }
    }
    w = 2 * hmax + 1;
    half = (double)hmax;
    if (w > W) ch = -1;  // excluded by the host's choice of W from the largest B
  }
  for (int d = 0; d < 3; ++d) {
    const NufftDim &D = g.d[d];
    const double u = xc[d] / D.h;
    const double m0 = ceil(u - half);
    const double i0 = m0 - D.m_lo;
    if (!(i0 >= 0.0 && i0 + w <= (double)D.M)) ch = -1;
    b[d] = ch < 0 ? 0 : (int)i0;
    f[d] = (float)(m0 - u);
  }
  r.ch = ch;
  r.bx = b[0];
  r.by = b[1];
  r.bz = b[2];
  r.fx = f[0];
  r.fy = f[1];
  r.fz = f[2];
  r.w = w;
  r.cr = (float)(a.w * norm) * cs;
  r.ci = (float)(a.w * norm) * sn;
  r.ax = ia[0];
  r.ay = ia[1];
  r.az = ia[2];
  r.kz = (float)kz;
  r.pad1 = r.pad2 = 0.0f;
  rec[j] = r;
  keys[j] = r.ch < 0 ? (unsigned)bins.nbins : (unsigned)home_bin(r, bins);
  index[j] = j;
}

// The tables and the atom records in one launch (they are independent): the
// first tb workgroups fill the tables (grid-strided over tb workgroups), the
// rest take one atom per thread.
__global__ __launch_bounds__(kThreads) void tables_and_prep(NufftGeom g, int W2, float *__restrict__ dec,
                                                           float2 *__restrict__ tw, float4 *__restrict__ zero,
                                                           size_t n4, int tb,
                                                           const spsim_hip_atom *__restrict__ atoms, int natoms,
                                                           const spsim_hip_segment *__restrict__ segs, int nsegs,
                                                           AtomRot R, int Wsp, const float *__restrict__ atom_b,
                                                           GaussSpread gs, SpreadAtom *__restrict__ rec,
                                                           SpreadBins bins, unsigned *__restrict__ keys,
                                                           int *__restrict__ index) {
  if ((int)blockIdx.x < tb)
    setup_tables(g, W2, dec, tw, zero, n4, blockIdx.x * kThreads + threadIdx.x, tb * kThreads);
  else
    spread_prep(atoms, natoms, segs, nsegs, R, g, Wsp, atom_b, gs, rec, bins, keys, index,
                ((int)blockIdx.x - tb) * kThreads + threadIdx.x);
}

// The binning is a stable radix sort of (home bin, atom index) pairs (the keys
// come from spread_prep): every bin lists its atoms in index order, so the
// spread sums them in an order that depends on the input alone and the results
// repeat bit for bit. Atoms outside the grid take the key nbins and sort after
// every bin.
using BinSortConfig = rocprim::default_config;
//
// Grids of at most kLocalBins spread blocks (every small-grid call) take a
// counting sort of the same order instead, in as many launches as an atomic
// one: each run of kBinRun atoms counts its atoms per bin in LDS
// (bin_count_runs, a bin-major (bin, run) matrix), an exclusive scan of that
// matrix gives every (bin, run) its first slot, and bin_place_runs places each
// run's atoms chunk by chunk in index order.
constexpr int kLocalBins = 4096;
constexpr int kBinRun = 256;

__global__ __launch_bounds__(kThreads) void bin_count_runs(const unsigned *__restrict__ keys, int natoms, int nbins,
                                                          int runs, int *__restrict__ cnt) {
  __shared__ int hist[kLocalBins];
  for (int i = threadIdx.x; i < nbins; i += kThreads) hist[i] = 0;
  __syncthreads();
  const int r = blockIdx.x, j0 = r * kBinRun, j1 = min(natoms, j0 + kBinRun);
  for (int j = j0 + threadIdx.x; j < j1; j += kThreads) {
    const unsigned k = keys[j];
    if (k < (unsigned)nbins) atomicAdd(hist + k, 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += kThreads) cnt[(size_t)i * runs + r] = hist[i];
  if (r == 0 && threadIdx.x == 0) cnt[(size_t)nbins * runs] = 0;  // the scan's total
}

// One wave per run: chunks of 64 atoms in order; an atom goes to its bin's
// cursor plus its rank among the chunk's atoms of that bin (lanes in order,
// by ballots over the chunk's distinct keys).
__global__ __launch_bounds__(64) void bin_place_runs(const SpreadAtom *__restrict__ rec,
                                                    const unsigned *__restrict__ keys, int natoms, int nbins,
                                                    int runs, const int *__restrict__ pos,
                                                    SpreadAtom *__restrict__ sorted) {
  __shared__ int cur[kLocalBins];
  const int lane = threadIdx.x, r = blockIdx.x;
  for (int i = lane; i < nbins; i += 64) cur[i] = pos[(size_t)i * runs + r];
  __syncthreads();
  const int j0 = r * kBinRun, j1 = min(natoms, j0 + kBinRun);
  for (int base = j0; base < j1; base += 64) {
    const int j = base + lane;
    const unsigned k = j < j1 ? keys[j] : (unsigned)nbins;
    const bool valid = k < (unsigned)nbins;
    int rank = 0, count = 0;
    unsigned long long todo = __ballot(valid);
    while (todo) {  // one round per distinct key of the chunk
      const int leader = __ffsll((long long)todo) - 1;
      const unsigned kl = (unsigned)__shfl((int)k, leader);
      const unsigned long long m = __ballot(valid && k == kl);
      if (valid && k == kl) {
        rank = __popcll(m & ((1ull << lane) - 1ull));
        count = __popcll(m);
      }
      todo &= ~m;
    }
    const int p = valid ? cur[k] + rank : 0;
    __syncthreads();
    if (valid && rank == 0) cur[k] += count;
    if (valid) sorted[p] = rec[j];
    __syncthreads();
  }
}


// After the sort: the records in bin order, and each bin's atoms at
// [start[bin], end[bin]) from the boundaries of the sorted keys' runs (bins
// without atoms keep their cleared 0, 0).
__global__ __launch_bounds__(kThreads) void bin_gather(const SpreadAtom *__restrict__ rec, int natoms, unsigned nbins,
                                                      const unsigned *__restrict__ keys,
                                                      const int *__restrict__ index, SpreadAtom *__restrict__ sorted,
                                                      int *__restrict__ start, int *__restrict__ end) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= natoms) return;
  const unsigned k = keys[i];
  if (k >= nbins) return;
  sorted[i] = rec[index[i]];
  if (i == 0 || keys[i - 1] != k) start[k] = i;
  if (i == natoms - 1 || keys[i + 1] != k) end[k] = i + 1;
}

// One workgroup: out[i] = sum_{k < i} in[k] over n entries, in chunks of
// kScanPer * kThreads (each thread a contiguous run) with a running carry.
constexpr int kScanPer = 16;

__global__ __launch_bounds__(kThreads) void scan_exclusive(const int *__restrict__ in, long long n,
                                                          int *__restrict__ out) {
  __shared__ int part[kThreads];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (long long base = 0; base < n; base += (long long)kScanPer * kThreads) {
    int v[kScanPer], sum = 0;
    const long long first = base + (long long)kScanPer * threadIdx.x;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      v[k] = first + k < n ? in[first + k] : 0;
      sum += v[k];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
      const int t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    int run = carry + (threadIdx.x ? part[threadIdx.x - 1] : 0);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      if (first + k < n) out[first + k] = run;
      run += v[k];
    }
    __syncthreads();
    if (threadIdx.x == kThreads - 1) carry += part[kThreads - 1];
    __syncthreads();
  }
}

// Many bins: the same scan over chunks of kScanChunk entries in parallel.
// chunk_sums adds up each chunk (one workgroup per chunk); scan_exclusive then
// scans those sums into chunk offsets, and scan_chunks scans every chunk from
// its offset (one workgroup per chunk).
constexpr int kScanChunk = kScanPer * kThreads;

__global__ __launch_bounds__(kThreads) void chunk_sums(const int *__restrict__ in, long long n,
                                                      int *__restrict__ sums) {
  __shared__ int part[kThreads];
  const long long first = (long long)blockIdx.x * kScanChunk + (long long)kScanPer * threadIdx.x;
  int sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) sum += first + k < n ? in[first + k] : 0;
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < (unsigned)w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(kThreads) void scan_chunks(const int *__restrict__ in, long long n,
                                                       const int *__restrict__ offsets, int *__restrict__ out) {
  __shared__ int part[kThreads];
  const long long first = (long long)blockIdx.x * kScanChunk + (long long)kScanPer * threadIdx.x;
  int v[kScanPer], sum = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = first + k < n ? in[first + k] : 0;
    sum += v[k];
  }
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    const int t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += t;
    __syncthreads();
  }
  int run = offsets[blockIdx.x] + (threadIdx.x ? part[threadIdx.x - 1] : 0);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (first + k < n) out[first + k] = run;
    run += v[k];
  }
}

// A staged candidate: its window weights (times the deconvolution factor) at
// the block's kSpreadBlock points per axis (zero outside the window) and its
// phased strength.
struct __align__(16) SpreadCand {
  float wx[kSpreadBlock], wy[kSpreadBlock], wz[kSpreadBlock], wzi[kSpreadBlock];
  float cr, ci, pad0, pad1;
};

// One 64-lane workgroup per output block (channel, kz, ky, kx). The candidate
// atoms (home blocks up to NB = ceil((W - 1) / kSpreadBlock) behind per axis)
// are staged 64 at a time: each lane loads one, keeps it if its window reaches
// into the block and writes its window weights there. Then every lane adds
// every kept atom into its column (x0 + lane % 8, y0 + lane / 8) for the
// block's 8 z-planes, in registers.
// With few output blocks (small grids) the candidates of a block are split
// round-robin by 64-atom batch over gridDim.y workgroups, each writing its own
// partial grid (grid + blockIdx.y * grid_stride), which the first FFT pass adds
// up as it loads its lines.
// GAUSS: Gaussian windows of per-atom width r.w <= W (else every window is the
// exponential of semicircle of width W). CZ: complex z weights (the phase
// exp(-2 pi i r.kz (m - u)) of a Gaussian on a z axis centred in frequency).
template <int W, bool GAUSS, bool CZ>
__global__ __launch_bounds__(64) void spread_gather(const SpreadAtom *__restrict__ sorted,
                                                   const int *__restrict__ start, const int *__restrict__ end,
                                                   int bin_stride, SpreadBins b, NufftGeom g,
                                                   const float *__restrict__ dec, float2 *__restrict__ grid,
                                                   size_t grid_stride) {
  constexpr int kB = kSpreadBlock;
  constexpr int NB = (W + kB - 2) / kB;
  constexpr int NH = (NB + 1) * (NB + 1) * (NB + 1);
  __shared__ SpreadCand cand[64];
  __shared__ int hs[NH], hp[NH + 1];
  long long blk = blockIdx.x;
  const int kx = (int)(blk % b.nbx);
  blk /= b.nbx;
  const int ky = (int)(blk % b.nby);
  blk /= b.nby;
  const int kz = (int)(blk % b.nbz);
  const int ch = (int)(blk / b.nbz);
  const int Mx = g.d[0].M, My = g.d[1].M, Mz = g.d[2].M;
  const int x0 = kx * kB, y0 = ky * kB, z0 = kz * kB;
  const int lane = threadIdx.x, tx = lane % kB, ty = lane / kB;
  const float k2 = 2.0f / W;
  for (int i = lane; i < NH; i += 64) {
    const int dx = i % (NB + 1), dy = (i / (NB + 1)) % (NB + 1), dz = i / ((NB + 1) * (NB + 1));
    int st = 0, n = 0;
    if (kx >= dx && ky >= dy && kz >= dz) {
      const long long h = (((long long)ch * b.nbz + kz - dz) * b.nby + ky - dy) * b.nbx + kx - dx;
      st = start[h * bin_stride];
      n = end[h * bin_stride] - st;
    }
    hs[i] = st;
    hp[i + 1] = n;
  }
  __syncthreads();
  if (lane == 0) {
    hp[0] = 0;
    for (int i = 0; i < NH; ++i) hp[i + 1] += hp[i];
  }
  __syncthreads();
  const int total = hp[NH];
  float ar[kB], ai[kB];
#pragma unroll
  for (int z = 0; z < kB; ++z) ar[z] = ai[z] = 0.0f;
  for (int base = 64 * (int)blockIdx.y; base < total; base += 64 * (int)gridDim.y) {
    const int v = base + lane;
    SpreadAtom r;
    bool keep = false;
    if (v < total) {
      int bi = 0;
      while (v >= hp[bi + 1]) ++bi;
      r = sorted[hs[bi] + v - hp[bi]];
      const int rw = GAUSS ? r.w : W;
      keep = r.bx + rw > x0 && r.bx < x0 + kB && r.by + rw > y0 && r.by < y0 + kB && r.bz + rw > z0 &&
             r.bz < z0 + kB;
    }
    // the kept candidates in lane order (the workgroup is one wave)
    const unsigned long long kept = __ballot(keep);
    {
      if (keep) {
        [[maybe_unused]] const int rw = GAUSS ? r.w : W;
        SpreadCand &c = cand[__popcll(kept & ((1ull << lane) - 1ull))];
#pragma unroll
        for (int i = 0; i < kB; ++i) {
          const int ix = x0 + i - r.bx, iy = y0 + i - r.by, iz = z0 + i - r.bz;
          // table indices clamped into the grid: points past its end have zero weight
          const float dx = dec[min(x0 + i, Mx - 1)], dy = dec[Mx + min(y0 + i, My - 1)];
          const float dz = dec[Mx + My + min(z0 + i, Mz - 1)];
          if constexpr (GAUSS) {
            const float ex = r.fx + (float)ix, ey = r.fy + (float)iy, ez = r.fz + (float)iz;
            c.wx[i] = (ix >= 0 && ix < rw) ? expf(-ex * ex * r.ax) * dx : 0.0f;
            c.wy[i] = (iy >= 0 && iy < rw) ? expf(-ey * ey * r.ay) * dy : 0.0f;
            const float gz = (iz >= 0 && iz < rw) ? expf(-ez * ez * r.az) * dz : 0.0f;
            if constexpr (CZ) {
              float sn, cs;
              sincospif(-2.0f * r.kz * ez, &sn, &cs);
              c.wz[i] = gz * cs;
              c.wzi[i] = gz * sn;
            } else {
              c.wz[i] = gz;
            }
          } else {
            c.wx[i] = (ix >= 0 && ix < W) ? es_window((r.fx + (float)ix) * k2, g.beta1) * dx : 0.0f;
            c.wy[i] = (iy >= 0 && iy < W) ? es_window((r.fy + (float)iy) * k2, g.beta1) * dy : 0.0f;
            c.wz[i] = (iz >= 0 && iz < W) ? es_window((r.fz + (float)iz) * k2, g.beta1) * dz : 0.0f;
          }
        }
        c.cr = r.cr;
        c.ci = r.ci;
      }
    }
    __syncthreads();
    const int nc = __popcll(kept);
    for (int a = 0; a < nc; ++a) {
      const float w = cand[a].wx[tx] * cand[a].wy[ty];
      const float vr = cand[a].cr * w, vi = cand[a].ci * w;
#pragma unroll
      for (int z = 0; z < kB; ++z) {
        const float wz = cand[a].wz[z];
        if constexpr (CZ) {
          const float wzi = cand[a].wzi[z];
          ar[z] += vr * wz - vi * wzi;
          ai[z] += vr * wzi + vi * wz;
        } else {
          ar[z] += vr * wz;
          ai[z] += vi * wz;
        }
      }
    }
    __syncthreads();
  }
  const int gx = x0 + tx, gy = y0 + ty;
  float2 *out = grid + blockIdx.y * grid_stride;
  if (gx < Mx && gy < My) {
#pragma unroll
    for (int z = 0; z < kB; ++z)
      if (z0 + z < Mz) out[(((size_t)ch * Mz + z0 + z) * My + gy) * Mx + gx] = make_float2(ar[z], ai[z]);
  }
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// -i * a
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// n / d for n < 2^16 (FastDiv).
__device__ __forceinline__ int sdiv(int n, const FastDiv &f) { return (int)__builtin_fmaf((float)n, f.fi, f.fh); }
// n - (n / d) d for n < 2^16.
__device__ __forceinline__ int smod(int n, const FastDiv &f) { return n - (int)__umul24((unsigned)sdiv(n, f), f.d); }
// Hides v from the optimiser, so that a running index stays a chain of adds
// instead of being rewritten as multiplies by the unrolled step count.
__device__ __forceinline__ void opaque(int &v) { asm("" : "+v"(v)); }

__device__ __forceinline__ void dft4(float2 &u0, float2 &u1, float2 &u2, float2 &u3) {
  const float2 a = cadd(u0, u2), b = csub(u0, u2);
  const float2 c = cadd(u1, u3), d = mul_mi(csub(u1, u3));
  u0 = cadd(a, c);
  u2 = csub(a, c);
  u1 = cadd(b, d);
  u3 = csub(b, d);
}

// The R-point DFT of u in place, exp(-2 pi i n k / R), R in {2, 3, 4, 5, 8}.
template <int R>
__device__ __forceinline__ void dft_small(float2 (&u)[R]) {
  if constexpr (R == 2) {
    const float2 a = u[0];
    u[0] = cadd(a, u[1]);
    u[1] = csub(a, u[1]);
  } else if constexpr (R == 4) {
    dft4(u[0], u[1], u[2], u[3]);
  } else if constexpr (R == 8) {
    constexpr float h = 0.70710678118654752440f;  // cos(pi / 4)
    float2 e0 = u[0], e1 = u[2], e2 = u[4], e3 = u[6];
    float2 o0 = u[1], o1 = u[3], o2 = u[5], o3 = u[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    // odd half times exp(-2 pi i s / 8), s = 0..3
    o1 = make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
    o2 = mul_mi(o2);
    o3 = make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    u[0] = cadd(e0, o0);
    u[4] = csub(e0, o0);
    u[1] = cadd(e1, o1);
    u[5] = csub(e1, o1);
    u[2] = cadd(e2, o2);
    u[6] = csub(e2, o2);
    u[3] = cadd(e3, o3);
    u[7] = csub(e3, o3);
  } else if constexpr (R == 3) {
    constexpr float s3 = 0.86602540378443864676f;  // sin(2 pi / 3)
    const float2 t1 = cadd(u[1], u[2]), t2 = csub(u[1], u[2]);
    const float2 m = make_float2(u[0].x - 0.5f * t1.x, u[0].y - 0.5f * t1.y);
    const float2 r = mul_mi(make_float2(s3 * t2.x, s3 * t2.y));
    u[0] = cadd(u[0], t1);
    u[1] = cadd(m, r);
    u[2] = csub(m, r);
  } else {
    static_assert(R == 5, "radix");
    constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
    constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    const float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
    const float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
    const float2 m1 = make_float2(u[0].x + c1 * a1.x + c2 * a2.x, u[0].y + c1 * a1.y + c2 * a2.y);
    const float2 m2 = make_float2(u[0].x + c2 * a1.x + c1 * a2.x, u[0].y + c2 * a1.y + c1 * a2.y);
    const float2 r1 = mul_mi(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    const float2 r2 = mul_mi(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    u[0] = make_float2(u[0].x + a1.x + a2.x, u[0].y + a1.y + a2.y);
    u[1] = cadd(m1, r1);
    u[4] = csub(m1, r1);
    u[2] = cadd(m2, r2);
    u[3] = csub(m2, r2);
  }
}

// exp(-2 pi i j / N) at compile time: the angle reduced to [-pi, pi] and a
// Taylor series to double precision.
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double ct_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 16; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double ct_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 16; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
constexpr int ct_mod(int j, int n) { return ((j % n) + n) % n; }
constexpr double ct_angle(int j, int n) {
  const int r = ct_mod(j, n);
  return kTwoPi * (2 * r <= n ? r : r - n) / n;  // in [-pi, pi]
}

// a exp(-2 pi i J / N) with J, N known at compile time: exact for multiples
// of pi/2, two multiplies for odd multiples of pi/4, four otherwise.
template <int J, int N>
__device__ __forceinline__ float2 twiddle_ct(float2 a) {
  constexpr int j = ct_mod(J, N);
  if constexpr (j == 0) {
    return a;
  } else if constexpr (4 * j == N) {
    return mul_mi(a);
  } else if constexpr (2 * j == N) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * j == 3 * N) {
    return make_float2(-a.y, a.x);
  } else if constexpr (8 * j == N || 8 * j == 3 * N || 8 * j == 5 * N || 8 * j == 7 * N) {
    constexpr float h = 0.70710678118654752440f;
    // (1 - i) h, (-1 - i) h, (-1 + i) h, (1 + i) h
    constexpr float cr = (8 * j == N || 8 * j == 7 * N) ? h : -h;
    constexpr float ci = (8 * j == N || 8 * j == 3 * N) ? -h : h;
    return make_float2(cr * a.x - ci * a.y, cr * a.y + ci * a.x);
  } else {
    constexpr float c = (float)ct_cos(ct_angle(j, N)), s = (float)ct_sin(ct_angle(j, N));
    return make_float2(a.x * c + a.y * s, a.y * c - a.x * s);  // a (c - i s)
  }
}

template <int... I, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F &&f) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(std::integral_constant<int, i>) for i = 0 .. N-1, each index a constant.
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// First factor of a composite radix: 4 for multiples of 4 above 8 (R = 4 R2),
// else the smallest prime factor.
constexpr int radix_split(int R) {
  if (R % 4 == 0 && R > 8) return 4;
  for (int p = 2; p <= R; ++p)
    if (R % p == 0) return p;
  return R;
}

template <int R>
__device__ __forceinline__ void dft_r(float2 (&u)[R]);

// The R1 R2-point DFT by Cooley-Tukey in registers: R1 DFTs of size R2 over
// u[n1 + R1 n2], twiddles exp(-2 pi i n1 k2 / R), R2 DFTs of size R1; output
// k2 + R2 k1.
template <int R1, int R2>
__device__ __forceinline__ void dft_composite(float2 (&u)[R1 * R2]) {
  constexpr int R = R1 * R2;
  float2 a[R1][R2];
  static_for<R1>([&](auto n1c) {
    constexpr int n1 = decltype(n1c)::value;
    float2 v[R2];
    static_for<R2>([&](auto n2c) { v[decltype(n2c)::value] = u[n1 + R1 * decltype(n2c)::value]; });
    dft_r<R2>(v);
    static_for<R2>([&](auto k2c) {
      constexpr int k2 = decltype(k2c)::value;
      a[n1][k2] = twiddle_ct<n1 * k2, R>(v[k2]);
    });
  });
  static_for<R2>([&](auto k2c) {
    constexpr int k2 = decltype(k2c)::value;
    float2 v[R1];
    static_for<R1>([&](auto n1c) { v[decltype(n1c)::value] = a[decltype(n1c)::value][k2]; });
    dft_r<R1>(v);
    static_for<R1>([&](auto k1c) { u[k2 + R2 * decltype(k1c)::value] = v[decltype(k1c)::value]; });
  });
}

// The R-point DFT of u in place: hand-written for 2, 3, 4, 5 and 8, composite
// (radix_split) for other 2,3,5-smooth R.
template <int R>
__device__ __forceinline__ void dft_r(float2 (&u)[R]) {
  if constexpr (R == 2 || R == 3 || R == 4 || R == 5 || R == 8)
    dft_small<R>(u);
  else
    dft_composite<radix_split(R), R / radix_split(R)>(u);
}

// How the global stages of a pass address memory. kFftGeneral: 64-bit point
// offsets, an input summed over nsum grids, outputs that wrap (L > nf).
// kFftContig / kFftStrided (one grid, L <= nf, every offset of a line below
// 2^31): 32-bit offsets formed incrementally from per-pass products, for
// contiguous lines (inner == 1) and strided ones.
enum FftMode { kFftGeneral = 0, kFftContig = 1, kFftStrided = 2 };

// One Stockham stage of radix R on a line: butterfly i (of nb = nf/R) reads
// x[i + t nb], twiddles element t by exp(-2 pi i t k / (span R)) =
// tw[t k stride] with k = i mod span, applies the R-point DFT and writes
// y[(i - k) R + k + s span]. GLOAD: the first stage (span 1, no twiddles)
// reads the line's points straight from global memory instead, slot
// (m0 + m) mod nf holding point m (zero outside the M points); GSTORE: the
// last stage (k = i, so output slot i + s span) writes only the kept outputs,
// slot (l0 + l) mod nf to output l (every such l < L), to global memory
// instead of LDS. lin / lout: the line's point 0 and output 0; has: false for
// the padding lines of the last workgroup, which load zeros and store nothing.
template <int R, bool GLOAD, bool GSTORE, int MODE>
__device__ __forceinline__ void stockham_bfly(const FftPass &p, const FftStage &st, int i,
                                              const float2 *__restrict__ x, float2 *__restrict__ y,
                                              const float2 *__restrict__ tw, const float2 *__restrict__ lin,
                                              float2 *__restrict__ lout, bool has) {
  const int nf = p.nf, nb = (int)st.nb.d, span = (int)st.span.d;
  float2 u[R];
  int k = 0;
  if constexpr (GLOAD) {
    if constexpr (MODE == kFftGeneral) {
#pragma unroll
      for (int t = 0; t < R; ++t) {
        int m = i + t * nb - p.m0;
        m += m < 0 ? nf : 0;
        float2 v = make_float2(0.0f, 0.0f);
        if (m < p.M && has) {
          const long long at = (long long)m * p.inner;
          v = lin[at];
          for (int g = 1; g < p.nsum; ++g) {
            const float2 w = lin[at + g * p.sum_stride];
            v.x += w.x;
            v.y += w.y;
          }
        }
        u[t] = v;
      }
    } else {
      // slot offset o in (-nf_in, nf_in) wraps by one add
      const int base = (MODE == kFftContig ? i : (int)__umul24((unsigned)i, (unsigned)p.inner32)) - p.m0_in;
#pragma unroll
      for (int t = 0; t < R; ++t) {
        int o = base + t * p.nb_in;
        o += o < 0 ? p.nf_in : 0;
        u[t] = o < p.M_in && has ? lin[(unsigned)o] : make_float2(0.0f, 0.0f);
      }
    }
  } else {
    // tw[0] = (1, -0) leaves k == 0 exact, so every butterfly twiddles
    k = smod(i, st.span);
    int a = i;
#pragma unroll
    for (int t = 0; t < R; ++t, a += nb) u[t] = x[a];
    const int step = (int)__umul24((unsigned)k, (unsigned)st.stride);
    int w = step;
#pragma unroll
    for (int t = 1; t < R; ++t, w += step) {
      opaque(w);
      u[t] = cmul(u[t], tw[w]);
    }
  }
  dft_r<R>(u);
  if constexpr (GSTORE) {
    if (!has) return;
    if constexpr (MODE == kFftGeneral) {
#pragma unroll
      for (int s2 = 0; s2 < R; ++s2) {
        int l = i + s2 * span - p.l0;
        l += l < 0 ? nf : 0;
        for (; l < p.L; l += nf) lout[(long long)l * p.inner] = u[s2];
      }
    } else {
      const int base = (MODE == kFftContig ? i : (int)__umul24((unsigned)i, (unsigned)p.inner32)) - p.l0_in;
#pragma unroll
      for (int s2 = 0; s2 < R; ++s2) {
        int o = base + s2 * p.span_in;
        o += o < 0 ? p.nf_in : 0;
        if (o < p.L_in) lout[(unsigned)o] = u[s2];
      }
    }
  } else {
    const int j = (i - k) * R + k;
#pragma unroll
    for (int s2 = 0; s2 < R; ++s2) y[j + s2 * span] = u[s2];
  }
}

template <bool GLOAD, bool GSTORE, int MODE>
__device__ __forceinline__ void stockham_stage(const FftPass &p, const FftStage &st, int i, const float2 *x,
                                               float2 *y, const float2 *tw, const float2 *lin, float2 *lout,
                                               bool has) {
  switch (st.R) {
    case 8: stockham_bfly<8, GLOAD, GSTORE, MODE>(p, st, i, x, y, tw, lin, lout, has); break;
    case 4: stockham_bfly<4, GLOAD, GSTORE, MODE>(p, st, i, x, y, tw, lin, lout, has); break;
    case 2: stockham_bfly<2, GLOAD, GSTORE, MODE>(p, st, i, x, y, tw, lin, lout, has); break;
    case 3: stockham_bfly<3, GLOAD, GSTORE, MODE>(p, st, i, x, y, tw, lin, lout, has); break;
    default: stockham_bfly<5, GLOAD, GSTORE, MODE>(p, st, i, x, y, tw, lin, lout, has); break;
  }
}

// tw[j] = exp(-2 pi i j / nf), j < nf.
__global__ __launch_bounds__(kThreads) void twiddle_table(int nf, float2 *__restrict__ tw) {
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= nf) return;
  double sn, cs;
  sincospi(-2.0 * (double)j / (double)nf, &sn, &cs);
  tw[j] = make_float2((float)cs, (float)sn);
}

// Workgroup b of G in an order that keeps neighbouring workgroups on one XCD:
// the hardware deals workgroup IDs round-robin over the 8 XCDs, so ID b runs
// on XCD b mod 8 as that XCD's (b / 8)-th; giving each XCD a contiguous range
// of workgroups lets the few that share a cache line of a strided view meet
// in one L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G) {
  const unsigned x = b & 7u, s = b >> 3, q = G >> 3, r = G & 7u;
  return x < r ? x * (q + 1u) + s : r * (q + 1u) + (x - r) * q + s;
}

// A middle Stockham stage in place on the one LDS buffer of T lines: every
// thread first loads and twiddles the inputs of its butterflies (at most KB,
// kept in registers), then, after a barrier, applies the DFTs and writes the
// outputs over the same buffer. Same indexing as stockham_bfly.
template <int R, int KB>
__device__ __forceinline__ void stockham_inplace(const FftPass &p, const FftStage &st, float2 *__restrict__ buf,
                                                 const float2 *__restrict__ tw) {
  const int nf = p.nf, nb = (int)st.nb.d, span = (int)st.span.d, nbt = p.T * nb;
  float2 u[KB][R];
  int at[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    const int b = threadIdx.x + kb * kThreads;
    at[kb] = -1;
    if (b < nbt) {
      const int t = sdiv(b, st.nb);
      const int i = b - (int)__umul24((unsigned)t, (unsigned)nb);
      const int k = smod(i, st.span);
      const int line = (int)__umul24((unsigned)t, (unsigned)nf);
      int a = line + i;
#pragma unroll
      for (int r = 0; r < R; ++r, a += nb) u[kb][r] = buf[a];
      const int step = (int)__umul24((unsigned)k, (unsigned)st.stride);
      int w = step;
#pragma unroll
      for (int r = 1; r < R; ++r, w += step) {
        opaque(w);
        u[kb][r] = cmul(u[kb][r], tw[w]);
      }
      at[kb] = line + (i - k) * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    if (at[kb] >= 0) {
      dft_r<R>(u[kb]);
      int a = at[kb];
#pragma unroll
      for (int s2 = 0; s2 < R; ++s2, a += span) buf[a] = u[kb][s2];
    }
  }
  __syncthreads();
}

// Butterflies per thread of a radix-R stage on lines of at most P points in
// all (T nf <= P): ceil(P / (R kThreads)).
template <int R, int P>
constexpr int fft_kb() {
  return (P + R * kThreads - 1) / (R * kThreads);
}

// Pruned forward FFT of T lines per workgroup. A line of the input view
// [outer][M][inner] holds points m = m_lo + i at FFT slot m mod nf (the rest
// zero); the output view [outer][L][inner] receives FFT values at l = l_lo + i
// (slot l mod nf).
// The first Stockham stage reads the points from global memory and the last
// one writes the kept outputs there; the stages between run in place on one
// LDS buffer of T lines. Lines of a strided view (inner > 1) are adjacent in
// memory, so the global stages give consecutive lanes consecutive lines.
// TWL: lines of up to kFftLds points in all, the twiddle table copied into LDS
// after the buffer; otherwise one line of up to kMaxNf points, twiddles read
// from global memory. The T line offsets (input, output; -1: no line) follow
// in LDS.
template <bool TWL, int MODE>
__global__ __launch_bounds__(kThreads) void fft_lines(FftPass p, const float2 *__restrict__ in,
                                                     float2 *__restrict__ out, const float2 *__restrict__ tw_g) {
  constexpr int P = TWL ? kFftLds : kMaxNf;
  extern __shared__ float2 lds[];
  const int nf = p.nf, T = p.T;
  float2 *x = lds;
  float2 *tw_l = lds + T * nf;
  long long *lin_s = (long long *)(lds + T * nf + (TWL ? nf : 0));
  long long *lout_s = lin_s + T;
  if constexpr (TWL)
    for (int e = threadIdx.x; e < nf; e += kThreads) tw_l[e] = tw_g[e];
  const float2 *tw = TWL ? (const float2 *)tw_l : tw_g;
  const unsigned blk = MODE == kFftStrided ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const long long line0 = (long long)blk * T;
  for (int t = threadIdx.x; t < T; t += kThreads) {
    const long long line = line0 + t;
    if (line < p.nlines) {
      const long long o = line / p.inner;
      const long long c = line - o * p.inner;
      lin_s[t] = o * p.M * p.inner + c;
      lout_s[t] = o * p.L * p.inner + c;
    } else {
      lin_s[t] = lout_s[t] = -1;
    }
  }
  __syncthreads();
  const int ns = p.nstages;
  // Global stages map butterfly b to (line t, butterfly i): line-major for
  // contiguous lines, line-minor for strided views (coalesced across lines).
  const bool contiguous = MODE == kFftGeneral ? p.inner == 1 : MODE == kFftContig;
  {
    const FftStage &st = p.st[0];
    const int nb = (int)st.nb.d;
    for (int b = threadIdx.x; b < T * nb; b += kThreads) {
      const int q = sdiv(b, contiguous ? st.nb : p.dT);
      const int r = b - (int)__umul24((unsigned)q, contiguous ? (unsigned)nb : (unsigned)T);
      const int t = contiguous ? q : r, i = contiguous ? r : q;
      const long long li = lin_s[t], lo = lout_s[t];
      const bool has = li >= 0;
      if (ns == 1)
        stockham_stage<true, true, MODE>(p, st, i, nullptr, nullptr, tw, in + (has ? li : 0), out + (has ? lo : 0),
                                         has);
      else
        stockham_stage<true, false, MODE>(p, st, i, nullptr, x + __umul24((unsigned)t, (unsigned)nf), tw,
                                          in + (has ? li : 0), nullptr, has);
    }
  }
  if (ns == 1) return;
  __syncthreads();
  for (int s = 1; s < ns - 1; ++s) {
    const FftStage &st = p.st[s];
    switch (st.R) {
      case 8: stockham_inplace<8, fft_kb<8, P>()>(p, st, x, tw); break;
      case 4: stockham_inplace<4, fft_kb<4, P>()>(p, st, x, tw); break;
      case 2: stockham_inplace<2, fft_kb<2, P>()>(p, st, x, tw); break;
      case 3: stockham_inplace<3, fft_kb<3, P>()>(p, st, x, tw); break;
      default: stockham_inplace<5, fft_kb<5, P>()>(p, st, x, tw); break;
    }
  }
  {
    const FftStage &st = p.st[ns - 1];
    const int nb = (int)st.nb.d;
    for (int b = threadIdx.x; b < T * nb; b += kThreads) {
      const int q = sdiv(b, contiguous ? st.nb : p.dT);
      const int r = b - (int)__umul24((unsigned)q, contiguous ? (unsigned)nb : (unsigned)T);
      const int t = contiguous ? q : r, i = contiguous ? r : q;
      const long long lo = lout_s[t];
      const bool has = lo >= 0;
      stockham_stage<false, true, MODE>(p, st, i, x + __umul24((unsigned)t, (unsigned)nf), nullptr, tw, nullptr,
                                        out + (has ? lo : 0), has);
    }
  }
}

// Line lengths with a kernel of their own (fft_fixed): every size and index
// step is a compile-time constant, and the radices (up to kFixedMaxRadix,
// composite ones included) are chosen for the fewest stages. Radices above 16
// cost more in registers (fewer waves per SIMD) than their saved stage gains.
constexpr int kFixedMaxRadix = 16;

struct FixedPlan {
  int ns;
  int R[4];
};

// 2,3,5-smooth radices up to kFixedMaxRadix, largest first.
constexpr int kRadixCount = 11;
constexpr int kRadices[kRadixCount] = {16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};

constexpr bool smooth_ct(int n) {
  for (int p : {2, 3, 5})
    while (n > 1 && n % p == 0) n /= p;
  return n == 1;
}

// The plan of nf with the fewest stages (at most 4), then the smallest largest
// radix, then radices in decreasing order; ns == 0 when there is none.
constexpr FixedPlan fixed_plan(int nf) {
  FixedPlan best{0, {0, 0, 0, 0}};
  if (!smooth_ct(nf)) return best;
  for (int ns = 1; ns <= 4; ++ns) {
    // radices in decreasing order a >= b >= c >= d; the first hit in a
    // decreasing scan of a from the smallest feasible one is the best plan
    for (int ia = kRadixCount - 1; ia >= 0; --ia) {
      const int a = kRadices[ia];
      if (a > kFixedMaxRadix || nf % a) continue;
      const int n1 = nf / a;
      if (ns == 1) {
        if (n1 == 1) return {1, {a, 0, 0, 0}};
        continue;
      }
      for (int ib = ia; ib < kRadixCount; ++ib) {
        const int b = kRadices[ib];
        if (n1 % b) continue;
        const int n2 = n1 / b;
        if (ns == 2) {
          if (n2 == 1) return {2, {a, b, 0, 0}};
          continue;
        }
        for (int ic = ib; ic < kRadixCount; ++ic) {
          const int c = kRadices[ic];
          if (n2 % c) continue;
          const int n3 = n2 / c;
          if (ns == 3) {
            if (n3 == 1) return {3, {a, b, c, 0}};
            continue;
          }
          for (int id = ic; id < kRadixCount; ++id)
            if (kRadices[id] == n3) return {4, {a, b, c, n3}};
        }
      }
    }
  }
  return best;
}

template <int NF>
struct FixedPlanOf {
  static constexpr FixedPlan plan = fixed_plan(NF);
  static constexpr int span(int s) { return s == 0 ? 1 : span(s - 1) * plan.R[s - 1]; }
};

// The lengths with a fixed kernel: even, 2,3,5-smooth, 16 .. kFftLds (several
// lines per workgroup), at most 3 stages.
constexpr bool fixed_length(int n) {
  if (n < 16 || n > kFftLds || n % 2 || !smooth_ct(n)) return false;
  const FixedPlan p = fixed_plan(n);
  return p.ns >= 1 && p.ns <= 3;
}
constexpr int count_fixed_lengths() {
  int c = 0;
  for (int n = 16; n <= kFftLds; ++n) c += fixed_length(n) ? 1 : 0;
  return c;
}
constexpr int kNumFixed = count_fixed_lengths();
struct FixedLengths {
  int n[kNumFixed];
};
constexpr FixedLengths list_fixed_lengths() {
  FixedLengths f{};
  int c = 0;
  for (int n = 16; n <= kFftLds; ++n)
    if (fixed_length(n)) f.n[c++] = n;
  return f;
}
constexpr FixedLengths kFixed = list_fixed_lengths();

// q = n / D for 0 <= n < 2^16 (the FP32 division of FastDiv with constant
// operands).
template <int D>
__device__ __forceinline__ int cdiv(int n) {
  return (int)__builtin_fmaf((float)n, 1.0f / D, 0.5f / D);
}

// One stage of radix R (span SPAN) of fft_fixed on the workgroup's T lines of
// NF points, PITCH = NF + 1 apart in LDS. FIRST: reads the points from global memory (like stockham_bfly's
// GLOAD) and writes the LDS buffer; LAST: reads the buffer and writes the
// kept outputs to global memory; otherwise in place on the buffer.
template <int NF, int PITCH, int T, int MODE, int R, int SPAN, bool FIRST, bool LAST>
__device__ __forceinline__ void fixed_stage(const FftPass &p, float2 *__restrict__ x, const float2 *__restrict__ tw,
                                            const long long *lin_s, const long long *lout_s,
                                            const float2 *__restrict__ in, float2 *__restrict__ out) {
  constexpr int nb = NF / R, nbt = T * nb, stride = NF / (SPAN * R);
  constexpr bool contiguous = MODE == kFftContig;
  // butterfly b -> (line t, butterfly i): line-major for contiguous lines,
  // line-minor for strided views
  auto split = [](int b, int &t, int &i) {
    if constexpr (contiguous) {
      t = cdiv<nb>(b);
      i = b - t * nb;
    } else {
      i = cdiv<T>(b);
      t = b - i * T;
    }
  };
  auto twiddled = [&](float2 (&u)[R], const float2 *xl, int i) {
    int a = i;
#pragma unroll
    for (int r = 0; r < R; ++r, a += nb) u[r] = xl[a];
    if constexpr (SPAN > 1) {
      const int k = i - cdiv<SPAN>(i) * SPAN;
      const int step = k * stride;
      int w = step;
#pragma unroll
      for (int r = 1; r < R; ++r, w += step) {
        opaque(w);
        u[r] = cmul(u[r], tw[w]);
      }
    }
  };
  auto store = [&](const float2 (&u)[R], int i, bool has, long long lo) {
    if (!has) return;
    float2 *lout = out + lo;
    const int base = (contiguous ? i : i * p.inner32) - p.l0_in;
    const int span_in = SPAN * p.inner32;
#pragma unroll
    for (int s2 = 0; s2 < R; ++s2) {
      int o = base + s2 * span_in;
      o += o < 0 ? p.nf_in : 0;
      if (o < p.L_in) lout[(unsigned)o] = u[s2];
    }
  };
  if constexpr (FIRST) {
    const int nb_in = nb * p.inner32;
    for (int b = threadIdx.x; b < nbt; b += kThreads) {
      int t, i;
      split(b, t, i);
      const long long li = lin_s[t];
      const bool has = li >= 0;
      const float2 *lin = in + (has ? li : 0);
      float2 u[R];
      const int base = (contiguous ? i : i * p.inner32) - p.m0_in;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        int o = base + r * nb_in;
        o += o < 0 ? p.nf_in : 0;
        u[r] = o < p.M_in && has ? lin[(unsigned)o] : make_float2(0.0f, 0.0f);
      }
      dft_r<R>(u);
      if constexpr (LAST) {
        store(u, i, has, lout_s[t]);
      } else {
        float2 *yl = x + t * PITCH + i * R;
#pragma unroll
        for (int s2 = 0; s2 < R; ++s2) yl[s2] = u[s2];
      }
    }
    if constexpr (!LAST) __syncthreads();
  } else if constexpr (LAST) {
    for (int b = threadIdx.x; b < nbt; b += kThreads) {
      int t, i;
      split(b, t, i);
      float2 u[R];
      twiddled(u, x + t * PITCH, i);
      dft_r<R>(u);
      const long long lo = lout_s[t];
      store(u, i, lo >= 0, lo);
    }
  } else {
    constexpr int KB = (nbt + kThreads - 1) / kThreads;
    float2 u[KB][R];
    int at[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int b = threadIdx.x + kb * kThreads;
      at[kb] = -1;
      if (b < nbt) {
        const int t = cdiv<nb>(b), i = b - t * nb;
        twiddled(u[kb], x + t * PITCH, i);
        const int k = i - cdiv<SPAN>(i) * SPAN;
        at[kb] = t * PITCH + (i - k) * R + k;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      if (at[kb] >= 0) {
        dft_r<R>(u[kb]);
        float2 *yl = x + at[kb];
#pragma unroll
        for (int s2 = 0; s2 < R; ++s2) yl[s2 * SPAN] = u[kb][s2];
      }
    }
    __syncthreads();
  }
}

// fft_lines for lines of exactly NF points (a single grid, L <= NF, 32-bit
// offsets, MODE kFftContig or kFftStrided): the same layout and results, with
// the stages of FixedPlanOf<NF> unrolled at compile time. Lines sit NF + 1
// points apart in LDS: an odd pitch puts the same point of consecutive lines
// (consecutive lanes of a strided view's global stages) in different banks.
template <int NF, int MODE>
__global__ __launch_bounds__(kThreads) void fft_fixed(FftPass p, const float2 *__restrict__ in,
                                                     float2 *__restrict__ out, const float2 *__restrict__ tw_g) {
  using Plan = FixedPlanOf<NF>;
  constexpr int T = kFixedLds / NF < kMaxT ? kFixedLds / NF : kMaxT, PITCH = NF + 1;
  static_assert(Plan::plan.ns > 0 && T >= 1 && T <= kMaxT, "fixed FFT length");
  extern __shared__ float2 lds[];
  float2 *x = lds;
  float2 *tw = lds + T * PITCH;
  long long *lin_s = (long long *)(lds + T * PITCH + NF);
  long long *lout_s = lin_s + T;
  for (int e = threadIdx.x; e < NF; e += kThreads) tw[e] = tw_g[e];
  const unsigned blk = MODE == kFftStrided ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const long long line0 = (long long)blk * T;
  for (int t = threadIdx.x; t < T; t += kThreads) {
    const long long line = line0 + t;
    if (line < p.nlines) {
      const long long o = line / p.inner;
      const long long c = line - o * p.inner;
      lin_s[t] = o * p.M * p.inner + c;
      lout_s[t] = o * p.L * p.inner + c;
    } else {
      lin_s[t] = lout_s[t] = -1;
    }
  }
  __syncthreads();
  static_for<Plan::plan.ns>([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    fixed_stage<NF, PITCH, T, MODE, Plan::plan.R[s], Plan::span(s), s == 0, s == Plan::plan.ns - 1>(
        p, x, tw, lin_s, lout_s, in, out);
  });
}

// Per-q state of the interpolation: the (rotated) q in FP64, the window
// weights and first grid point per axis, the first window's correction.
template <int W>
struct InterpQ {
  double q0[3], q[3];
  float wt[3][W];
  int base[3];
  float corr;
  bool ok;
};

// Pattern pat interpolates at q' = R_pat q (rots: row-major [p][9], FP64);
// rots == nullptr means q' = q. ok is false for a q outside the grids
// (unreachable for q inside the bounds the grids were sized for).
template <int W>
__device__ __forceinline__ void interp_prepare(const float *__restrict__ hkl, int k, const NufftGeom &g,
                                               const double *__restrict__ rots, int pat, InterpQ<W> &s) {
  for (int i = 0; i < 3; ++i) s.q0[i] = s.q[i] = hkl[3 * (size_t)k + i];
  if (rots) {
    const double *R = rots + 9 * (size_t)pat;
    for (int i = 0; i < 3; ++i)
      s.q[i] = __builtin_fma(R[3 * i + 2], s.q0[2], __builtin_fma(R[3 * i + 1], s.q0[1], R[3 * i] * s.q0[0]));
  }
  s.corr = 1.0f;
  s.ok = true;
  for (int d = 0; d < 3; ++d) {
    const NufftDim &D = g.d[d];
    const double t = (s.q[d] - D.cs) * D.h;
    const double v = t * D.nf;
    const double l0 = ceil(v - 0.5 * W);
    const double i0 = l0 - D.l_lo;
    s.ok = s.ok && (i0 >= 0.0 && i0 + W <= (double)D.L);
    s.base[d] = s.ok ? (int)i0 : 0;
    const float fb = (float)(l0 - v);
#pragma unroll
    for (int i = 0; i < W; ++i) s.wt[d][i] = es_window((fb + (float)i) * (2.0f / W), g.beta2);
    float c;
    if (g.gauss_v0 > 0.0f) {
      const float tf = (float)t;
      c = expf(19.739208802178716f * g.gauss_v0 * tf * tf);  // 2 pi^2 v0 t^2
    } else {
      const float xr = (float)t * g.corr_scale;
      const float u = 2.0f * xr * xr - 1.0f;
      c = g.corr[kCorrDeg];
#pragma unroll
      for (int j = kCorrDeg - 1; j >= 0; --j) c = __builtin_fmaf(c, u, g.corr[j]);
    }
    s.corr *= c;
  }
}

// The separable window sum over one species' W^3 values; row(iz, iy) points at
// the W values of an x-run. kRollZ keeps the z loop rolled (fewer loads in
// flight, for LDS rows whose latency is short).
template <int W, bool kRollZ = false, typename Row>
__device__ __forceinline__ void window_sum(const InterpQ<W> &s, Row row, float &ar, float &ai) {
  ar = ai = 0.0f;
  // Every weight index is a constant after unrolling, so the weights stay in
  // registers: the z weights shift down one place per z iteration (the
  // current one is wz[0]) when that loop stays rolled.
  float wz[W];
#pragma unroll
  for (int j = 0; j < W; ++j) wz[j] = s.wt[2][j];
  constexpr int kUnrollZ = kRollZ ? 1 : W;
#pragma unroll kUnrollZ
  for (int iz = 0; iz < W; ++iz) {
    const float w2 = wz[0];
#pragma unroll
    for (int j = 0; j + 1 < W; ++j) wz[j] = wz[j + 1];
#pragma unroll
    for (int iy = 0; iy < W; ++iy) {
      const float2 *r = row(iz, iy);
      float rr = 0.0f, ri = 0.0f;
#pragma unroll
      for (int ix = 0; ix < W; ++ix) {
        const float2 h = r[ix];
        rr += h.x * s.wt[0][ix];
        ri += h.y * s.wt[0][ix];
      }
      const float wzy = w2 * s.wt[1][iy];
      ar += wzy * rr;
      ai += wzy * ri;
    }
  }
}

// Correction, q phase and the stores of F and I for one q.
template <int W>
__device__ __forceinline__ void interp_store(const NufftGeom &g, const InterpQ<W> &s, float fr, float fi, size_t out,
                                             float2 *__restrict__ F, float *__restrict__ I) {
  float Fr, Fi;
  if (!s.ok) {
    Fr = Fi = __builtin_nanf("");
  } else {
    fr *= s.corr;
    fi *= s.corr;
    double ph = -(s.q[0] * g.d[0].cx + s.q[1] * g.d[1].cx + s.q[2] * g.d[2].cx);
    ph -= rint(ph);
    float sn, cs;
    sincospif((float)(2.0 * ph), &sn, &cs);
    Fr = fr * cs - fi * sn;
    Fi = fr * sn + fi * cs;
  }
  if (F) F[out] = make_float2(Fr, Fi);
  if (I) I[out] = Fr * Fr + Fi * Fi;
}

__device__ __forceinline__ float q_s2(const double q0[3]) {
  return (float)((q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2]) * 2.5e-21);
}

// The q of a lane: with g.tile_n > 0 the list is a tile_n x tile_n image read
// in 16 x 16 pixel tiles per workgroup (16 x 4 per wave, so neighbouring lanes
// have neighbouring q in both directions); otherwise consecutive q. Returns
// npix for a lane past the image.
__device__ __forceinline__ int interp_q_index(const NufftGeom &g, int npix) {
  if (g.tile_n > 0) {
    const int n = g.tile_n, ntx = (n + 15) / 16, lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int x = (int)(blockIdx.x % ntx) * 16 + lane % 16, y = (int)(blockIdx.x / ntx) * 16 + wave * 4 + lane / 16;
    return (x < n && y < n) ? y * n + x : npix;
  }
  const int k = blockIdx.x * kThreads + threadIdx.x;
  return k < npix ? k : npix;
}

// One thread per q: W^3 gathers (W: the interpolation window's width) per
// species grid [ch][Lz][Ly][Lx] from global memory, window correction, form
// factor, species sum and the q phase. Patterns of a batch (one species grid
// set, many orientations) are blockIdx.y: pattern p interpolates at R_p q and
// writes F + p npix, I + p npix.
template <int W>
__global__ __launch_bounds__(kThreads) void nufft_interp(const float *__restrict__ hkl, int npix, NufftGeom g,
                                                        const float2 *__restrict__ H, int nch,
                                                        const spsim_hip_segment *__restrict__ segs,
                                                        const float *__restrict__ ff, float b_global,
                                                        const double *__restrict__ rots, int p0,
                                                        float2 *__restrict__ F, float *__restrict__ I) {
  const int k = interp_q_index(g, npix);
  if (k >= npix) return;
  const int pat = p0 + (int)blockIdx.y;
  InterpQ<W> s;
  interp_prepare<W>(hkl, k, g, rots, pat, s);
  float fr = 0.0f, fi = 0.0f;
  if (s.ok) {
    const float s2 = q_s2(s.q0);
    const int Lx = g.d[0].L, Ly = g.d[1].L, Lz = g.d[2].L;
    const size_t chs = (size_t)Lz * Ly * Lx;
    for (int ch = 0; ch < nch; ++ch) {
      const int sp = segs[ch].species;
      const float f = form_factor(ff + sp * SPSIM_HIP_FF_STRIDE, s2, sp == SPSIM_HIP_UNIT_SPECIES ? 0.0f : b_global);
      const float2 *Hc = H + ch * chs;
      float ar, ai;
      window_sum<W>(s, [&](int iz, int iy) { return Hc + ((size_t)(s.base[2] + iz) * Ly + s.base[1] + iy) * Lx + s.base[0]; },
                    ar, ai);
      fr += f * ar;
      fi += f * ai;
    }
  }
  interp_store<W>(g, s, fr, fi, (size_t)pat * npix + k, F, I);
}

// The same interpolation with each wave's footprint staged in LDS, one
// species at a time: neighbouring q (consecutive pixels) share most of their
// windows, so a wave's W^3 x 64 gathers touch a small box of grid points.
// The wave's lanes copy that box (the span of their first grid points plus W
// per axis, at most kBoxCap points) from global memory into their own LDS
// region and then gather from LDS, which serves twice the bytes per clock of
// the vector L1. A wave whose box is larger gathers from global memory.
// At least 4 waves per SIMD (at most 128 VGPRs): measured faster than the 3
// that the unconstrained allocation (136 VGPRs at W = 7) leaves. At W = 6, 5
// waves (76 VGPRs, no scratch) measured 7 % faster on the orientation batch;
// at W = 7 (84 VGPRs) 5 % slower, and W = 5 would spill.
constexpr int kBoxCap = 1024;  // grid points per wave (8 KiB of LDS per species)

template <int W>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(W == 6 ? 5 : 4))) void nufft_interp_lds(
    const float *__restrict__ hkl, int npix, NufftGeom g, const float2 *__restrict__ H, int nch,
    const spsim_hip_segment *__restrict__ segs, const float *__restrict__ ff, float b_global,
    const double *__restrict__ rots, int p0, float2 *__restrict__ F, float *__restrict__ I) {
  extern __shared__ float2 box[];  // g.box_cap points per wave
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int k = interp_q_index(g, npix);  // npix: no q
  const int pat = p0 + (int)blockIdx.y;
  InterpQ<W> s;
  if (k < npix) {
    interp_prepare<W>(hkl, k, g, rots, pat, s);
  } else {
    s.ok = false;
  }
  // the wave's box: [lo, hi + W) per axis over its lanes with ok
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = s.ok ? s.base[d] : 0x7fffffff;
    hi[d] = s.ok ? s.base[d] : -1;
  }
  for (int m = 32; m >= 1; m >>= 1)
    for (int d = 0; d < 3; ++d) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], m, 64));
      hi[d] = max(hi[d], __shfl_xor(hi[d], m, 64));
    }
  const int ex = hi[0] - lo[0] + W, ey = hi[1] - lo[1] + W, ez = hi[2] - lo[2] + W;
  // wave-uniform: some lane has ok (hi >= 0) and the box fits
  const bool staged = hi[0] >= 0 && (long long)ex * ey * ez <= g.box_cap;
  const int vol = staged ? ex * ey * ez : 0;
  const int lb = staged && s.ok ? ((s.base[2] - lo[2]) * ey + (s.base[1] - lo[1])) * ex + (s.base[0] - lo[0]) : 0;
  const float rex = 1.0f / (float)ex, rexy = 1.0f / (float)(ex * ey);
  const float s2 = k < npix ? q_s2(s.q0) : 0.0f;
  const int Lx = g.d[0].L, Ly = g.d[1].L, Lz = g.d[2].L;
  const size_t chs = (size_t)Lz * Ly * Lx;
  float2 *B = box + (size_t)wave * g.box_cap;
  float fr = 0.0f, fi = 0.0f;
  for (int ch = 0; ch < nch; ++ch) {
    const float2 *Hc = H + ch * chs;
    // i = (iz ey + iy) ex + ix < vol <= 1024: the float quotients are exact
    for (int i = lane; i < vol; i += 64) {
      const int iz = (int)(((float)i + 0.5f) * rexy);
      const int r = i - iz * ex * ey;
      const int iy = (int)(((float)r + 0.5f) * rex);
      const int ix = r - iy * ex;
      B[i] = Hc[((size_t)(lo[2] + iz) * Ly + lo[1] + iy) * Lx + lo[0] + ix];
    }
    __syncthreads();
    if (s.ok) {
      const int sp = segs[ch].species;
      const float f = form_factor(ff + sp * SPSIM_HIP_FF_STRIDE, s2, sp == SPSIM_HIP_UNIT_SPECIES ? 0.0f : b_global);
      float ar, ai;
      if (staged)
        window_sum<W, true>(s, [&](int iz, int iy) { return B + lb + (iz * ey + iy) * ex; }, ar, ai);
      else
        window_sum<W, true>(
            s, [&](int iz, int iy) { return Hc + ((size_t)(s.base[2] + iz) * Ly + s.base[1] + iy) * Lx + s.base[0]; },
            ar, ai);
      fr += f * ar;
      fi += f * ai;
    }
    __syncthreads();
  }
  if (k < npix) interp_store<W>(g, s, fr, fi, (size_t)pat * npix + k, F, I);
}

// The spectrum's amplitude average: F[i] = sum_k w[k] Fk[k][i] (k ascending),
// I[i] = |F[i]|^2; F or I may be NULL.
struct SpectrumWeights {
  float w[SPSIM_HIP_MAX_SPECTRUM];
};
__global__ __launch_bounds__(kThreads) void spectrum_combine(size_t npix, int ns, SpectrumWeights wt,
                                                            const float2 *__restrict__ Fk, float2 *__restrict__ F,
                                                            float *__restrict__ I) {
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < npix; i += (size_t)gridDim.x * kThreads) {
    float re = 0.0f, im = 0.0f;
    for (int k = 0; k < ns; ++k) {
      const float2 v = Fk[(size_t)k * npix + i];
      re = __builtin_fmaf(wt.w[k], v.x, re);
      im = __builtin_fmaf(wt.w[k], v.y, im);
    }
    if (F) F[i] = make_float2(re, im);
    if (I) I[i] = __builtin_fmaf(re, re, im * im);
  }
}

__global__ __launch_bounds__(kThreads) void zero_pattern(size_t n, float2 *__restrict__ F, float *__restrict__ I) {
  for (size_t k = (size_t)blockIdx.x * kThreads + threadIdx.x; k < n; k += (size_t)gridDim.x * kThreads) {
    if (F) F[k] = make_float2(0.0f, 0.0f);
    if (I) I[k] = 0.0f;
  }
}

// ---- host side ----

// Gauss-Legendre nodes and weights on [0, 1] (Newton iteration on P_n).
void gauss_legendre01(int n, double *z, double *w) {
  for (int i = 0; i < n; ++i) {
    double x = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double pn = n == 1 ? x : p1, pm = n == 1 ? 1.0 : p0;
      dp = n * (x * pn - pm) / (x * x - 1.0);
      const double dx = pn / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    z[i] = 0.5 * (1.0 - x);
    w[i] = 1.0 / ((1.0 - x * x) * dp * dp);  // 2/((1-x^2) P'^2), halved for [0, 1]
  }
}

bool smooth_length(int n) {
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

// The smallest even 2,3,5-smooth length >= at_least, or a length with a
// fixed-plan kernel up to 10 % above it.
int fft_length(int at_least) {
  int n = at_least + (at_least & 1);
  while (!smooth_length(n)) n += 2;
  for (int m = n; m <= n + n / 10; m += 2)
    if (smooth_length(m) && fixed_length(m)) return m;
  return n;
}

int radices(int n, int *r) {
  int k = 0;
  while (n % 8 == 0) r[k++] = 8, n /= 8;
  while (n % 4 == 0) r[k++] = 4, n /= 4;
  while (n % 2 == 0) r[k++] = 2, n /= 2;
  while (n % 3 == 0) r[k++] = 3, n /= 3;
  while (n % 5 == 0) r[k++] = 5, n /= 5;
  return k;
}

FastDiv fast_div(unsigned d) {
  FastDiv f;
  f.d = d;
  f.fi = d ? (float)(1.0 / (double)d) : 0.0f;
  f.fh = 0.5f * f.fi;
  return f;
}

// phi_hat(xi) = 2 int_0^1 phi(z) cos(2 pi xi z) dz by an n-node Gauss-Legendre rule, in FP64.
double es_hat_host(double xi, double beta, int n, const double *z, const double *w) {
  double sum = 0.0;
  for (int i = 0; i < n; ++i) sum += w[i] * exp(beta * (sqrt(1.0 - z[i] * z[i]) - 1.0)) * cos(2.0 * M_PI * xi * z[i]);
  return 2.0 * sum;
}

// Chebyshev interpolant of degree kCorrDeg of 2 / (W phi_hat(xi)) over u = 2 (xi / xmax)^2 - 1
// in [-1, 1] (32 Chebyshev nodes, FP64), returned as monomial coefficients in u.
void fit_correction(int W, double beta, double xmax, float *out) {
  constexpr int kNodes = 32, kQuad = 64;
  double z[kQuad], w[kQuad], g[kNodes], c[kCorrDeg + 1];
  gauss_legendre01(kQuad, z, w);
  for (int j = 0; j < kNodes; ++j) {
    const double u = cos(M_PI * (j + 0.5) / kNodes);
    const double xi = xmax * sqrt(0.5 * (u + 1.0));
    g[j] = 2.0 / (W * es_hat_host(xi, beta, kQuad, z, w));
  }
  for (int k = 0; k <= kCorrDeg; ++k) {
    double sum = 0.0;
    for (int j = 0; j < kNodes; ++j) sum += g[j] * cos(M_PI * k * (j + 0.5) / kNodes);
    c[k] = (k == 0 ? 1.0 : 2.0) * sum / kNodes;
  }
  // sum_k c_k T_k(u) -> sum_k m_k u^k, with T_{k+1} = 2 u T_k - T_{k-1}
  double t_prev[kCorrDeg + 1] = {1.0}, t_cur[kCorrDeg + 1] = {0.0, 1.0}, m[kCorrDeg + 1] = {0.0};
  m[0] += c[0];
  for (int i = 0; i <= kCorrDeg; ++i) m[i] += c[1] * t_cur[i];
  for (int k = 2; k <= kCorrDeg; ++k) {
    double t_next[kCorrDeg + 1] = {0.0};
    for (int i = 0; i <= kCorrDeg; ++i) {
      if (i > 0) t_next[i] += 2.0 * t_cur[i - 1];
      t_next[i] -= t_prev[i];
    }
    for (int i = 0; i <= kCorrDeg; ++i) {
      m[i] += c[k] * t_next[i];
      t_prev[i] = t_cur[i];
      t_cur[i] = t_next[i];
    }
  }
  for (int i = 0; i <= kCorrDeg; ++i) out[i] = (float)m[i];
}

// Window constants that depend only on the widths: the quadrature of the
// second window's transform and the first window's correction polynomial.
// Formed once per (W, W2) on first use (FP64 quadrature and fit).
struct WindowTables {
  float beta1, beta2;
  int nq;
  float qz[kMaxQuad], qc[kMaxQuad];
  float corr_xmax, corr_scale;
  float corr[kCorrDeg + 1];
};

void fill_window_tables(int W, int W2, WindowTables *t) {
  t->beta1 = (float)(2.30 * W);
  t->beta2 = W2 == W ? t->beta1 : (float)(2.65 * W2);
  t->nq = 2 * W + 6;
  double z[kMaxQuad], w[kMaxQuad];
  gauss_legendre01(t->nq, z, w);
  for (int i = 0; i < t->nq; ++i) {
    t->qz[i] = (float)z[i];
    t->qc[i] = (float)(2.0 * w[i] * exp((double)t->beta2 * (sqrt(1.0 - z[i] * z[i]) - 1.0)));
  }
  // targets have |t| <= 1/(2 sigma) per axis, so |xi| = |t| W / 2 <= W / 8 (2 % margin)
  t->corr_xmax = (float)(1.02 * W / (4.0 * kSigma));
  t->corr_scale = (float)(0.5 * W / t->corr_xmax);
  fit_correction(W, t->beta1, t->corr_xmax, t->corr);
}

// W in [kMinW, kMaxW]; W2 is W, or W - 1 for W >= 7.
const WindowTables &window_tables(int W, int W2) {
  static WindowTables tab[kMaxW + 1][2];
  static std::once_flag once;
  std::call_once(once, [] {
    for (int w = kMinW; w <= kMaxW; ++w) {
      fill_window_tables(w, w, &tab[w][0]);
      if (w >= 7) fill_window_tables(w, w - 1, &tab[w][1]);
    }
  });
  return tab[W][W2 == W ? 0 : 1];
}

int width_for(double eps) {
  if (!(eps > 0.0)) eps = 1e-6;
  int W = (int)ceil(-log10(eps)) + 1;
  return W < kMinW ? kMinW : (W > kMaxW ? kMaxW : W);
}

// Grid of one axis for atoms in [xmin, xmax] and q in [smin, smax]: spread
// window width W, FFT length >= sigma2 M, interpolation window width W2.
// quiet: report a grid that does not fit by the return code only.
int plan_axis(double xmin, double xmax, double smin, double smax, int W, double sigma2, int W2, NufftDim *D,
              bool quiet = false) {
  D->cx = 0.5 * (xmin + xmax);
  D->cs = 0.5 * (smin + smax);
  const double X = 0.5 * (xmax - xmin), S = 0.5 * (smax - smin);
  D->h = S > 0.0 ? 1.0 / (2.0 * kSigma * S) : (X > 0.0 ? 2.0 * X : 1e-9);
  const double umax = X / D->h;
  if (!(umax < 0.25 * kMaxNf))
    return quiet ? SPSIM_HIP_ERR_INVALID : fail(SPSIM_HIP_ERR_INVALID, "NUFFT grid too large (box x q-range %g)", umax);
  D->m_lo = (int)floor(-umax - 0.5 * W) - 1;
  const int m_hi = (int)ceil(umax + 0.5 * W) + 1;
  D->M = m_hi - D->m_lo + 1;
  D->nf = fft_length((int)ceil(sigma2 * D->M));
  if (D->nf > kMaxNf)
    return quiet ? SPSIM_HIP_ERR_INVALID : fail(SPSIM_HIP_ERR_INVALID, "NUFFT line of %d points exceeds %d", D->nf, kMaxNf);
  const double vmax = S * D->h * D->nf;
  D->l_lo = (int)floor(-vmax - 0.5 * W2) - 1;
  const int l_hi = (int)ceil(vmax + 0.5 * W2) + 1;
  D->L = l_hi - D->l_lo + 1;
  return SPSIM_HIP_OK;
}

// Cost of a plan in units of one window gather: the three pruned FFT passes
// (about 10 gathers per line point per pass, measured on MI355X) and the
// interpolation (W2^3 gathers per q and species).
double plan_cost(const NufftDim d[3], int W2, double nq) {
  const double fft = (double)d[0].nf * d[1].M * d[2].M + (double)d[1].nf * d[0].L * d[2].M +
                     (double)d[2].nf * d[0].L * d[1].L;
  return 10.0 * fft + nq * W2 * W2 * W2;
}

FftPass make_pass(const NufftDim &D, long long outer, long long inner) {
  FftPass p{};
  p.nlines = outer * inner;
  p.inner = inner;
  p.M = D.M;
  p.nf = D.nf;
  p.L = D.L;
  p.m_lo = D.m_lo;
  p.l_lo = D.l_lo;
  p.m0 = ((D.m_lo % D.nf) + D.nf) % D.nf;
  p.l0 = ((D.l_lo % D.nf) + D.nf) % D.nf;
  p.T = kFftLds / D.nf;
  p.nsum = 1;
  p.sum_stride = 0;
  if (p.T < 1) p.T = 1;
  if (p.T > kMaxT) p.T = kMaxT;
  p.dT = fast_div((unsigned)p.T);
  int r[kMaxStages];
  p.nstages = radices(D.nf, r);
  int span = 1;
  for (int s = 0; s < p.nstages; ++s) {
    p.st[s].R = r[s];
    p.st[s].stride = D.nf / (span * r[s]);
    p.st[s].nb = fast_div((unsigned)(D.nf / r[s]));
    p.st[s].span = fast_div((unsigned)span);
    span *= r[s];
  }
  return p;
}

// Fills the twiddle table of an nf-point line.
int launch_twiddles(int nf, float2 *tw, hipStream_t stream) {
  hipLaunchKernelGGL(twiddle_table, dim3((nf + kThreads - 1) / kThreads), dim3(kThreads), 0, stream, nf, tw);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

struct FixedKernels {
  void (*contig)(FftPass, const float2 *, float2 *, const float2 *);
  void (*strided)(FftPass, const float2 *, float2 *, const float2 *);
};
template <size_t... I>
constexpr std::array<FixedKernels, sizeof...(I)> fixed_kernel_table(std::index_sequence<I...>) {
  return {FixedKernels{fft_fixed<kFixed.n[I], kFftContig>, fft_fixed<kFixed.n[I], kFftStrided>}...};
}
constexpr std::array<FixedKernels, kNumFixed> kFixedKernels =
    fixed_kernel_table(std::make_index_sequence<kNumFixed>{});

// Position of nf in kFixed, or -1.
int fixed_index(int nf) {
  for (int i = 0; i < kNumFixed; ++i)
    if (kFixed.n[i] == nf) return i;
  return -1;
}

int launch_fft(const FftPass &pass, const float2 *in, float2 *out, const float2 *tw, hipStream_t stream) {
  FftPass p = pass;
  // every slot and line index of fft_lines stays below 2^16
  if (p.m_lo < -p.nf || p.m_lo + p.M > 2 * p.nf || p.l_lo < -p.nf || p.l_lo + p.L > 2 * p.nf || p.M > p.nf ||
      p.L >= 65536 || p.nf > kMaxNf)
    return fail(SPSIM_HIP_ERR_INVALID, "FFT line layout out of range (M %d at %d, L %d at %d, nf %d)", p.M, p.m_lo,
                p.L, p.l_lo, p.nf);
  const long long blocks = (p.nlines + p.T - 1) / p.T;
  if (blocks <= 0) return SPSIM_HIP_OK;
  if (blocks > 0x7fffffffLL) return fail(SPSIM_HIP_ERR_INVALID, "too many FFT lines");
  // the fast modes need one grid, no wrapped outputs, offsets within a line
  // (up to 2 nf inner, before the wrap) below 2^31 and a stride below 2^24
  // (24-bit multiplies)
  const bool fast = p.nsum == 1 && p.L <= p.nf && (long long)p.nf * p.inner < 0x40000000LL && p.inner < (1 << 24);
  int mode = kFftGeneral;
  if (fast) {
    const int in32 = (int)p.inner;
    p.inner32 = in32;
    p.nb_in = (int)p.st[0].nb.d * in32;
    p.m0_in = p.m0 * in32;
    p.M_in = p.M * in32;
    p.span_in = (int)p.st[p.nstages - 1].span.d * in32;
    p.l0_in = p.l0 * in32;
    p.L_in = p.L * in32;
    p.nf_in = p.nf * in32;
    mode = p.inner == 1 ? kFftContig : kFftStrided;
  }
  const bool twl = p.nf <= kFftLds;
  const size_t lds = ((size_t)p.T + (twl ? 1 : 0)) * p.nf * sizeof(float2) + 2 * (size_t)p.T * sizeof(long long);
  using Kernel = void (*)(FftPass, const float2 *, float2 *, const float2 *);
  static const Kernel kernels[2][3] = {
      {fft_lines<false, kFftGeneral>, fft_lines<false, kFftContig>, fft_lines<false, kFftStrided>},
      {fft_lines<true, kFftGeneral>, fft_lines<true, kFftContig>, fft_lines<true, kFftStrided>}};
  Kernel k = kernels[twl ? 1 : 0][mode];
  size_t lds_k = lds;
  long long nblk = blocks;
  if (mode != kFftGeneral && p.T == std::min(kFftLds / p.nf, kMaxT)) {
    const int f = fixed_index(p.nf);
    if (f >= 0) {
      // the fixed kernels take kFixedLds / nf lines per workgroup
      k = mode == kFftContig ? kFixedKernels[f].contig : kFixedKernels[f].strided;
      p.T = std::min(kFixedLds / p.nf, kMaxT);
      p.dT = fast_div((unsigned)p.T);
      nblk = (p.nlines + p.T - 1) / p.T;
      lds_k = ((size_t)p.T + 1) * p.nf * sizeof(float2) + 2 * (size_t)p.T * sizeof(long long) +
              (size_t)p.T * sizeof(float2);  // line pitch nf + 1
    }
  }
  hipLaunchKernelGGL(k, dim3((unsigned)nblk), dim3(kThreads), lds_k, stream, p, in, out, tw);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

template <int W, bool GAUSS, bool CZ>
int launch_spread(const NufftGeom &g, const SpreadAtom *rec, const SpreadBins &bins, const int *start,
                  const int *end, int bin_stride, int nsplit, size_t grid_stride, const float *dec, float2 *grid,
                  hipStream_t stream) {
  if (bins.nbins > 0x7fffffffLL) return fail(SPSIM_HIP_ERR_INVALID, "too many spread blocks");
  hipLaunchKernelGGL((spread_gather<W, GAUSS, CZ>), dim3((unsigned)bins.nbins, (unsigned)nsplit), dim3(64), 0, stream,
                     rec, start, end, bin_stride, bins, g, dec, grid, grid_stride);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

template <int W>
int launch_interp(const NufftGeom &g, int nsegs, const spsim_hip_segment *segs, const float *hkl, int npix,
                  const float2 *H, const float *ff, float b_global, const double *rots, int npat, float2 *F, float *I,
                  hipStream_t stream) {
  for (int p0 = 0; p0 < npat; p0 += 65535) {
    const int n = npat - p0 < 65535 ? npat - p0 : 65535;
    const int nt = (g.tile_n + 15) / 16;
    const unsigned gx = g.tile_n > 0 ? (unsigned)(nt * nt) : (unsigned)((npix + kThreads - 1) / kThreads);
    if (g.lds_interp)
      hipLaunchKernelGGL(nufft_interp_lds<W>, dim3(gx, n), dim3(kThreads),
                         sizeof(float2) * (kThreads / 64) * (size_t)g.box_cap, stream, hkl, npix, g, H, nsegs, segs,
                         ff, b_global, rots, p0, F, I);
    else
      hipLaunchKernelGGL(nufft_interp<W>, dim3(gx, n), dim3(kThreads), 0, stream, hkl, npix, g, H, nsegs, segs, ff,
                         b_global, rots, p0, F, I);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  return SPSIM_HIP_OK;
}

// Widths of the Gaussian spread windows (the smallest one that holds the
// widest atom's window is used).
constexpr int kGaussWidths[] = {17, 25, 33};

// gauss: Gaussian windows; cz: with complex z weights (z centred in frequency).
int dispatch_spread(int W, bool gauss, bool cz, const NufftGeom &g, const SpreadAtom *rec, const SpreadBins &bins,
                    const int *start, const int *end, int bin_stride, int nsplit, size_t grid_stride,
                    const float *dec, float2 *grid, hipStream_t s) {
#define SPSIM_SPREAD_W(w, ga, c) \
  case w:                        \
    return launch_spread<w, ga, c>(g, rec, bins, start, end, bin_stride, nsplit, grid_stride, dec, grid, s)
  if (gauss && cz) {
    switch (W) {
      SPSIM_SPREAD_W(17, true, true);
      SPSIM_SPREAD_W(25, true, true);
      SPSIM_SPREAD_W(33, true, true);
      default: return fail(SPSIM_HIP_ERR_INVALID, "Gaussian window width %d", W);
    }
  }
  if (gauss) {
    switch (W) {
      SPSIM_SPREAD_W(17, true, false);
      SPSIM_SPREAD_W(25, true, false);
      SPSIM_SPREAD_W(33, true, false);
      default: return fail(SPSIM_HIP_ERR_INVALID, "Gaussian window width %d", W);
    }
  }
  switch (W) {
    SPSIM_SPREAD_W(4, false, false);
    SPSIM_SPREAD_W(5, false, false);
    SPSIM_SPREAD_W(6, false, false);
    SPSIM_SPREAD_W(7, false, false);
    SPSIM_SPREAD_W(8, false, false);
    SPSIM_SPREAD_W(9, false, false);
    SPSIM_SPREAD_W(10, false, false);
    default: return fail(SPSIM_HIP_ERR_INVALID, "window width %d", W);
  }
#undef SPSIM_SPREAD_W
}

int dispatch_interp(int W, const NufftGeom &g, int nsegs, const spsim_hip_segment *segs, const float *hkl, int npix,
                    const float2 *H, const float *ff, float b_global, const double *rots, int npat, float2 *F,
                    float *I, hipStream_t s) {
#define SPSIM_INTERP_W(w) \
  case w:                 \
    return launch_interp<w>(g, nsegs, segs, hkl, npix, H, ff, b_global, rots, npat, F, I, s)
  switch (W) {
    SPSIM_INTERP_W(4);
    SPSIM_INTERP_W(5);
    SPSIM_INTERP_W(6);
    SPSIM_INTERP_W(7);
    SPSIM_INTERP_W(8);
    SPSIM_INTERP_W(9);
    SPSIM_INTERP_W(10);
    default: return fail(SPSIM_HIP_ERR_INVALID, "window width %d", W);
  }
#undef SPSIM_INTERP_W
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the two ping-pong buffers of the spread grid and the FFT passes:
// G (spread) and B (after y) share the first, A (after x) and H (after z) the second.
size_t grid_bytes(const NufftDim d[3], size_t nch, size_t *b1 = nullptr, size_t *b2 = nullptr) {
  const NufftDim &X = d[0], &Y = d[1], &Z = d[2];
  const size_t nG = nch * Z.M * Y.M * X.M, nA = nch * Z.M * Y.M * X.L;
  const size_t nB = nch * Z.M * Y.L * X.L, nH = nch * Z.L * Y.L * X.L;
  const size_t s1 = align256(sizeof(float2) * (nG > nB ? nG : nB)), s2 = align256(sizeof(float2) * (nA > nH ? nA : nH));
  if (b1) *b1 = s1;
  if (b2) *b2 = s2;
  return s1 + s2;
}

}  // namespace
}  // namespace spsim

namespace spsim {
namespace {

// One NUFFT evaluation. Single pattern (npat == 0): q as given, atoms rotated
// by rot9 (R^T r), grids sized to the atom box and the q box. Batch (npat >= 1):
// one set of species grids for the unrotated atoms over the ball
// |q| <= max |q| (the box of every R q), then each pattern p interpolates at
// q' = R_p q (rots: host [npat][9]) into F/I + p npix.
// hb[v] = the minimum (v < 6) or maximum of nblk workgroup records of
// bounds_partial; NaN propagates.
void fold_bounds(const double *hpart, int nblk, double *hb) {
  for (int v = 0; v < kNb; ++v) {
    double acc = v < 6 ? INFINITY : -INFINITY;
    for (int b = 0; b < nblk; ++b) {
      const double x = hpart[(size_t)b * kNb + v];
      acc = (acc != acc || x != x) ? acc + x : (v < 6 ? fmin(acc, x) : fmax(acc, x));
    }
    hb[v] = acc;
  }
}

int nufft_run(double *hb, const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
              const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
              const AtomRot &R, int npat, const double *rots, int W, double eps, float *d_F, float *d_I, int *info,
              hipStream_t s, size_t *need = nullptr, void *ws_in = nullptr);

int nufft_impl(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
               const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
               const double *rot9, int npat, const double *rots, double eps, float *d_F, float *d_I, int *info,
               hipStream_t s) {
  if (npix < 0 || natoms < 0 || nsegs < 0 || nsegs > SPSIM_HIP_NSPECIES || npat < 0)
    return fail(SPSIM_HIP_ERR_INVALID, "negative size or too many segments");
  if ((npix > 0 && !d_hkl) || (natoms > 0 && (!d_atoms || !d_segs || nsegs < 1)) || !d_ff)
    return fail(SPSIM_HIP_ERR_INVALID, "NULL input pointer");
  if (npat > 0 && !rots) return fail(SPSIM_HIP_ERR_INVALID, "NULL rotations");
  SPSIM_TRY(ensure_device(nullptr));
  const int W = width_for(eps);
  AtomRot R{};
  if (rot9) {
    R.on = 1;
    for (int i = 0; i < 9; ++i) {
      if (!std::isfinite(rot9[i])) return fail(SPSIM_HIP_ERR_INVALID, "non-finite rotation");
      R.r[i] = rot9[i];
    }
  }
  for (size_t i = 0; i < (size_t)npat * 9; ++i)
    if (!std::isfinite(rots[i])) return fail(SPSIM_HIP_ERR_INVALID, "non-finite rotation");
  if (info)
    for (int i = 0; i < 13; ++i) info[i] = 0;
  if (info) info[0] = W;
  const int nout = npat > 0 ? npat : 1;
  if (npix == 0) return SPSIM_HIP_OK;
  if (!d_F && !d_I) return fail(SPSIM_HIP_ERR_INVALID, "both outputs are NULL");
  if (natoms == 0) {
    const size_t n = (size_t)nout * npix, zb = (n + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(zero_pattern, dim3((unsigned)(zb < 4096 ? zb : 4096)), dim3(kThreads), 0, s, n, (float2 *)d_F,
                       d_I);
    SPSIM_HIP_TRY(hipGetLastError());
    return SPSIM_HIP_OK;
  }

  // 1. Bounds of the atoms and of q: per-workgroup records, folded on the host
  // after the call's one synchronisation.
  const int most = natoms > npix ? natoms : npix;
  int nblk = (most + kThreads - 1) / kThreads;
  if (nblk > kBoundBlocks) nblk = kBoundBlocks;
  void *ws = nullptr;
  SPSIM_TRY(workspace(align256(sizeof(double) * kNb * nblk), &ws));
  double *part = (double *)ws;
  hipLaunchKernelGGL(bounds_partial, dim3(nblk), dim3(kThreads), 0, s, d_atoms, d_atom_b, natoms, d_hkl, npix, R,
                     part);
  SPSIM_HIP_TRY(hipGetLastError());
  double hpart[kNb * kBoundBlocks], hb[kNb];
  SPSIM_HIP_TRY(hipMemcpyAsync(hpart, part, sizeof(double) * kNb * nblk, hipMemcpyDeviceToHost, s));
  SPSIM_HIP_TRY(hipStreamSynchronize(s));
  fold_bounds(hpart, nblk, hb);
  return nufft_run(hb, d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, R, npat, rots, W, eps,
                   d_F, d_I, info, s);
}

// Plans the grids from the folded bounds hb and enqueues every stage after the
// bounds on s (a batch synchronises at the end for its rotations' host buffer).
// With need set it only plans and stores the workspace bytes the stages take;
// with ws_in set the stages use that buffer instead of the library workspace.
int nufft_run(double *hb, const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
              const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
              const AtomRot &R, int npat, const double *rots, int W, double eps, float *d_F, float *d_I, int *info,
              hipStream_t s, size_t *need, void *ws_in) {
  const int nout = npat > 0 ? npat : 1;
  void *ws = nullptr;
  for (int v = 0; v <= 12; ++v)
    if (!std::isfinite(hb[v])) return fail(SPSIM_HIP_ERR_INVALID, "non-finite atom position or q");
  const bool gauss = d_atom_b != nullptr;
  if (gauss && !(hb[13] == hb[13] && hb[13] < INFINITY && hb[14] < INFINITY))
    return fail(SPSIM_HIP_ERR_INVALID, "non-finite B factor");
  // a Gaussian of negative width does not exist; the direct kernels take any B
  if (gauss && hb[14] > 0.0) return fail(SPSIM_HIP_ERR_INVALID, "negative B factor (%g) in the NUFFT", -hb[14]);
  if (npat > 0) {
    // every R q lies in the ball of radius max |q| ||R||_2; ||R||_2^2, the
    // largest eigenvalue of R^T R, is at most its largest absolute row sum
    // (Gershgorin): 1 for a rotation, s^2 for s times a rotation
    double g2 = 0.0;
    for (int p = 0; p < npat; ++p) {
      const double *Rp = rots + 9 * (size_t)p;
      for (int i = 0; i < 3; ++i) {
        double row = 0.0;
        for (int j = 0; j < 3; ++j)
          row += fabs(Rp[i] * Rp[j] + Rp[3 + i] * Rp[3 + j] + Rp[6 + i] * Rp[6 + j]);
        g2 = row > g2 ? row : g2;
      }
    }
    const double r = sqrt(hb[12] * g2);
    // and axis d of R q spans at most the sum over j of R_dj [q_j min, q_j max]
    // (interval arithmetic over the q box): much tighter for matrices close to
    // s times the identity (a wavelength spectrum), where the ball would
    // inflate a detector's thin q_z range to the whole sphere
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = 0; p < npat; ++p) {
      const double *Rp = rots + 9 * (size_t)p;
      for (int d = 0; d < 3; ++d) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < 3; ++j) {
          const double u = Rp[3 * d + j] * hb[3 + j], v = Rp[3 * d + j] * hb[9 + j];
          a += fmin(u, v);
          b += fmax(u, v);
        }
        lo[d] = fmin(lo[d], a);
        hi[d] = fmax(hi[d], b);
      }
    }
    for (int d = 0; d < 3; ++d) {
      hb[3 + d] = fmax(lo[d], -r);
      hb[9 + d] = fmin(hi[d], r);
    }
  }
  if (gauss) {
    // x and y not centred in frequency (cs = 0): the Gaussians' transforms at
    // q itself give the Debye-Waller factors; z stays centred (complex z
    // weights) since a detector's q_z range lies on one side of 0
    for (int d = 0; d < 2; ++d) {
      const double r = fmax(fabs(hb[3 + d]), fabs(hb[9 + d]));
      hb[3 + d] = -r;
      hb[9 + d] = r;
    }
  }
  // Spread window: the exponential of semicircle of width W, or with per-atom
  // B factors Gaussians, the widest (largest B on the finest axis) deciding
  // the kernel's width Wsp. The grid spacing h matches plan_axis.
  GaussSpread gs{0.0f, 0.0f};
  int Wsp = W;
  if (gauss) {
    const double e = eps > 0.0 ? (eps < 1e-10 ? 1e-10 : eps) : 1e-6;
    gs.tcut = (float)(-log(e));
    gs.v0 = (float)((double)gs.tcut / (M_PI * M_PI));
    const double bmax = hb[13] > 0.0 ? hb[13] : 0.0;
    int hmax = 0;
    for (int d = 0; d < 3; ++d) {
      const double S = 0.5 * (hb[9 + d] - hb[3 + d]), X = 0.5 * (hb[6 + d] - hb[d]);
      const double h = S > 0.0 ? 1.0 / (2.0 * kSigma * S) : (X > 0.0 ? 2.0 * X : 1e-9);
      const double v = (double)gs.v0 + bmax * kBToSigma2 / (h * h);
      const int hw = (int)ceil(sqrt(2.0 * v * (double)gs.tcut));
      hmax = hw > hmax ? hw : hmax;
    }
    const int need = 2 * hmax + 1;
    Wsp = 0;
    for (int w : kGaussWidths)
      if (!Wsp && w >= need) Wsp = w;
    if (!Wsp)
      return fail(SPSIM_HIP_ERR_INVALID,
                  "B factors up to %g A^2 need a %d-point Gaussian window at this grid spacing (at most %d)", bmax,
                  need, kGaussWidths[2]);
  }

  // 2. Plan the grids: the second window equal to the first, or for W >= 7
  // the narrow one (W - 1, finer frequency grid) when it costs less and fits.
  NufftGeom g{};
  const size_t nch = (size_t)nsegs;
  for (int d = 0; d < 3; ++d) SPSIM_TRY(plan_axis(hb[d], hb[6 + d], hb[3 + d], hb[9 + d], Wsp, kSigma, W, &g.d[d]));
  int W2 = W;
  if (W >= 7) {
    NufftDim fine[3];
    bool fits = true;
    for (int d = 0; d < 3 && fits; ++d)
      fits = plan_axis(hb[d], hb[6 + d], hb[3 + d], hb[9 + d], Wsp, kSigmaFine, W - 1, &fine[d], true) == SPSIM_HIP_OK;
    const double nq = (double)nout * npix;
    if (fits && grid_bytes(fine, nch) <= kMaxGridBytes && plan_cost(fine, W - 1, nq) < plan_cost(g.d, W, nq)) {
      for (int d = 0; d < 3; ++d) g.d[d] = fine[d];
      W2 = W - 1;
    }
  }
  {
    const WindowTables &t = window_tables(W, W2);
    g.beta1 = t.beta1;
    g.beta2 = t.beta2;
    g.nq = t.nq;
    for (int i = 0; i < t.nq; ++i) {
      g.qz[i] = t.qz[i];
      g.qc[i] = t.qc[i];
    }
    g.corr_xmax = t.corr_xmax;
    g.corr_scale = t.corr_scale;
    for (int i = 0; i <= kCorrDeg; ++i) g.corr[i] = t.corr[i];
    g.gauss_v0 = gauss ? gs.v0 : 0.0f;
  }
  const NufftDim &X = g.d[0], &Y = g.d[1], &Z = g.d[2];
  // Interpolation with staged boxes when a wave's q are expected to fit one.
  // For a detector-shaped list of rows of about sqrt(npix) pixels that span the
  // kept grid, one pixel moves d = (L - W2) / sqrt(npix) cells. A square list
  // (n x n) is taken in 16 x 4 pixel patches per wave: a box of about
  // (16 d + W2 + 1)(4 d + W2 + 1)(W2 + 2) points; otherwise 64 consecutive q: a
  // box of about (64 d + W2)(W2 + 1)^2.
  {
    int Lmax = X.L > Y.L ? X.L : Y.L;
    if (Z.L > Lmax) Lmax = Z.L;
    const double d = (Lmax - W2) / sqrt((double)npix);
    const int n = (int)llround(sqrt((double)npix));
    g.tile_n = (n >= 16 && (long long)n * n == npix) ? n : 0;
    const double box = g.tile_n > 0 ? (16.0 * d + W2 + 1) * (4.0 * d + W2 + 1) * (W2 + 2)
                                    : (64.0 * d + W2) * (W2 + 1) * (W2 + 1);
    g.lds_interp = box <= kBoxCap ? 1 : 0;
    // LDS per wave: half again the expected box, in 64-point steps (a wave
    // whose box does not fit gathers directly); smaller boxes leave room for
    // more workgroups per CU
    int cap = ((int)ceil(fmin(1.5 * box, (double)kBoxCap)) + 63) / 64 * 64;
    g.box_cap = cap < 256 ? 256 : (cap > kBoxCap ? kBoxCap : cap);
  }
  if (info) {
    const int v[12] = {X.M, Y.M, Z.M, X.nf, Y.nf, Z.nf, X.L, Y.L, Z.L, W2, g.lds_interp, gauss ? Wsp : 0};
    for (int i = 0; i < 12; ++i) info[1 + i] = v[i];
  }
  const size_t nG = nch * Z.M * Y.M * X.M;
  const size_t b_dec = align256(sizeof(float) * (X.M + Y.M + Z.M));
  SpreadBins bins;
  bins.nbx = (X.M + kSpreadBlock - 1) / kSpreadBlock;
  bins.nby = (Y.M + kSpreadBlock - 1) / kSpreadBlock;
  bins.nbz = (Z.M + kSpreadBlock - 1) / kSpreadBlock;
  bins.nbins = (long long)nch * bins.nbz * bins.nby * bins.nbx;
  const size_t b_rec = align256(sizeof(SpreadAtom) * (size_t)natoms);
  const size_t b_bin = align256(sizeof(int) * (size_t)bins.nbins);
  const size_t b_tw = align256(sizeof(float2) * (size_t)(X.nf + Y.nf + Z.nf));
  const size_t b_rot = align256(sizeof(double) * 9 * (size_t)npat);
  size_t b1, b2;
  const size_t b12 = grid_bytes(g.d, nch, &b1, &b2);
  if (b12 > kMaxGridBytes) return fail(SPSIM_HIP_ERR_INVALID, "NUFFT grids need %zu bytes", b12);
  // Few output blocks for the atoms (small or thin grids): split each block's
  // candidates over several one-wave workgroups (about 8192 in all, 8 waves
  // per SIMD; at most one split per 256 atoms) and sum their partial grids.
  int nsplit = 1;
  if (bins.nbins < 8192) {
    nsplit = (int)(8192 / bins.nbins);
    const int by_atoms = (natoms + 255) / 256;
    if (nsplit > by_atoms) nsplit = by_atoms;
    if (nsplit > 64) nsplit = 64;
    if (nsplit < 1) nsplit = 1;
  }
  const size_t b_part = nsplit > 1 ? align256(sizeof(float2) * nG * nsplit) : 0;
  // the binning: keys and atom indices; for few bins the (bin, run) counts,
  // their scan and the scan's chunk sums; else the sorted keys and indices,
  // the sort's scratch and the bins' [start, end)
  if (bins.nbins >= 0x7fffffffLL) return fail(SPSIM_HIP_ERR_INVALID, "too many spread blocks");
  const bool local_bins = bins.nbins <= kLocalBins;
  const long long runs = (natoms + kBinRun - 1) / kBinRun;
  const long long ncells = local_bins ? bins.nbins * runs + 1 : 0;
  const long long nchunks = (ncells + kScanChunk - 1) / kScanChunk;
  if (ncells > 0x7fffffffLL) return fail(SPSIM_HIP_ERR_INVALID, "too many atoms for the spread's bins");
  const unsigned key_bits = 32u - (unsigned)__builtin_clz((unsigned)bins.nbins);  // keys 0 .. nbins
  const size_t b_key = align256(sizeof(unsigned) * (size_t)natoms);
  size_t sort_bytes = 0;
  if (!local_bins)
    SPSIM_HIP_TRY(rocprim::radix_sort_pairs<BinSortConfig>(nullptr, sort_bytes, (const unsigned *)nullptr,
                                                           (unsigned *)nullptr, (const int *)nullptr, (int *)nullptr,
                                                           natoms, 0u, key_bits, s));
  const size_t b_key2 = local_bins ? 0 : b_key;
  const size_t b_sort = local_bins ? 0 : align256(sort_bytes);
  const size_t b_rng = local_bins ? 0 : b_bin;
  const size_t b_cell = align256(sizeof(int) * (size_t)ncells);
  const size_t b_scan = align256(sizeof(int) * (size_t)nchunks);
  const size_t b_all = b_dec + 2 * b_rec + 2 * b_rng + 2 * b_key + 2 * b_key2 + b_sort + 2 * b_cell + 2 * b_scan +
                       b_tw + b_rot + b1 + b2 + b_part;
  if (need) {
    *need = b_all;
    return SPSIM_HIP_OK;
  }
  if (ws_in)
    ws = ws_in;
  else
    SPSIM_TRY(workspace(b_all, &ws));
  char *wp = (char *)ws;
  float *dec = (float *)wp;
  wp += b_dec;
  SpreadAtom *rec = (SpreadAtom *)wp;
  wp += b_rec;
  SpreadAtom *sorted = (SpreadAtom *)wp;
  wp += b_rec;
  int *bin_end = (int *)wp;  // end and start are adjacent, cleared together
  int *bin_start = (int *)(wp + b_rng);
  wp += 2 * b_rng;
  unsigned *keys = (unsigned *)wp;
  int *index = (int *)(wp + b_key);
  wp += 2 * b_key;
  unsigned *keys_sorted = (unsigned *)wp;
  int *index_sorted = (int *)(wp + b_key2);
  void *sort_tmp = wp + 2 * b_key2;
  wp += 2 * b_key2 + b_sort;
  int *cnt = (int *)wp;
  int *pos = (int *)(wp + b_cell);
  int *csum = (int *)(wp + 2 * b_cell);
  int *coff = (int *)(wp + 2 * b_cell + b_scan);
  wp += 2 * b_cell + 2 * b_scan;
  float2 *tw = (float2 *)wp;
  wp += b_tw;
  double *d_rots = npat > 0 ? (double *)wp : nullptr;
  wp += b_rot;
  float2 *buf1 = (float2 *)wp;
  float2 *buf2 = (float2 *)(wp + b1);
  float2 *gpart = nsplit > 1 ? (float2 *)(wp + b1 + b2) : nullptr;
  if (npat > 0) {
    // pageable source: the copy has completed when the call returns
    SPSIM_HIP_TRY(hipMemcpyAsync(d_rots, rots, sizeof(double) * 9 * (size_t)npat, hipMemcpyHostToDevice, s));
  }

  // 3. Tables (deconvolution, twiddles, cleared bin ranges), atom records
  // binned by spread block (stable sort by home block), spread.
  float2 *twx = tw, *twy = tw + X.nf, *twz = tw + X.nf + Y.nf;
  {
    int most_t = X.M > X.nf ? X.M : X.nf;
    for (const NufftDim *D : {&Y, &Z}) {
      if (D->M > most_t) most_t = D->M;
      if (D->nf > most_t) most_t = D->nf;
    }
    const size_t n4 = 2 * b_rng / 16;
    size_t tb = (size_t)(most_t + kThreads - 1) / kThreads;
    const size_t zb = (n4 + kThreads - 1) / kThreads;
    if (zb > tb) tb = zb < 4096 ? zb : 4096;
    const unsigned pblocks = (unsigned)((natoms + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(tables_and_prep, dim3((unsigned)tb + pblocks), dim3(kThreads), 0, s, g, W2, dec, tw,
                       (float4 *)bin_end, n4, (int)tb, d_atoms, natoms, d_segs, nsegs, R, Wsp, d_atom_b, gs, rec,
                       bins, keys, index);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  const bool cz = gauss && g.d[2].cs != 0.0;
  float2 *spread_out = nsplit > 1 ? gpart : buf1;
  if (local_bins) {
    const int nb = (int)bins.nbins, nr = (int)runs;
    hipLaunchKernelGGL(bin_count_runs, dim3((unsigned)runs), dim3(kThreads), 0, s, keys, natoms, nb, nr, cnt);
    SPSIM_HIP_TRY(hipGetLastError());
    if (nchunks <= 1) {
      hipLaunchKernelGGL(scan_exclusive, dim3(1), dim3(kThreads), 0, s, cnt, ncells, pos);
    } else {
      hipLaunchKernelGGL(chunk_sums, dim3((unsigned)nchunks), dim3(kThreads), 0, s, cnt, ncells, csum);
      SPSIM_HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(scan_exclusive, dim3(1), dim3(kThreads), 0, s, csum, nchunks, coff);
      SPSIM_HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(scan_chunks, dim3((unsigned)nchunks), dim3(kThreads), 0, s, cnt, ncells, coff, pos);
    }
    SPSIM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(bin_place_runs, dim3((unsigned)runs), dim3(64), 0, s, rec, keys, natoms, nb, nr, pos, sorted);
    SPSIM_HIP_TRY(hipGetLastError());
    // bin h: [pos[h runs], pos[(h + 1) runs])
    SPSIM_TRY(dispatch_spread(Wsp, gauss, cz, g, sorted, bins, pos, pos + runs, nr, nsplit, nG, dec, spread_out, s));
  } else {
    const unsigned ablocks = (unsigned)((natoms + kThreads - 1) / kThreads);
    size_t sort_have = sort_bytes;
    SPSIM_HIP_TRY(rocprim::radix_sort_pairs<BinSortConfig>(sort_tmp, sort_have, keys, keys_sorted, index,
                                                           index_sorted, natoms, 0u, key_bits, s));
    hipLaunchKernelGGL(bin_gather, dim3(ablocks), dim3(kThreads), 0, s, rec, natoms, (unsigned)bins.nbins,
                       keys_sorted, index_sorted, sorted, bin_start, bin_end);
    SPSIM_HIP_TRY(hipGetLastError());
    SPSIM_TRY(dispatch_spread(Wsp, gauss, cz, g, sorted, bins, bin_start, bin_end, 1, nsplit, nG, dec, spread_out,
                              s));
  }

  // 4. Pruned FFT along x, y, z: G (buf1, or the sum of the nsplit partial
  // grids, added up as the x pass loads them) -> A (buf2) -> B (buf1) -> H (buf2).
  FftPass px = make_pass(X, (long long)nch * Z.M * Y.M, 1);
  if (nsplit > 1) {
    px.nsum = nsplit;
    px.sum_stride = (long long)nG;
  }
  SPSIM_TRY(launch_fft(px, nsplit > 1 ? gpart : buf1, buf2, twx, s));
  SPSIM_TRY(launch_fft(make_pass(Y, (long long)nch * Z.M, X.L), buf2, buf1, twy, s));
  SPSIM_TRY(launch_fft(make_pass(Z, (long long)nch, (long long)Y.L * X.L), buf1, buf2, twz, s));

  // 5. Interpolate at every q (of every pattern).
  SPSIM_TRY(dispatch_interp(W2, g, nsegs, d_segs, d_hkl, npix, buf2, d_ff, b_global, d_rots, nout, (float2 *)d_F, d_I,
                            s));
  if (npat > 0) SPSIM_HIP_TRY(hipStreamSynchronize(s));  // the rotations' host buffer may go after return
  return SPSIM_HIP_OK;
}

}  // namespace
}  // namespace spsim

extern "C" int spsim_hip_sf_nfft(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms,
                                 const float *d_atom_b, const spsim_hip_segment *d_segs, int nsegs, int natoms,
                                 const float *d_ff, float b_global, const double *rot9, double eps, float *d_F,
                                 float *d_I, int *info, void *stream) {
  return spsim::nufft_impl(d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, rot9, 0, nullptr,
                           eps, d_F, d_I, info, (hipStream_t)stream);
}

extern "C" int spsim_hip_sf_nfft_batch(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms,
                                       const float *d_atom_b, const spsim_hip_segment *d_segs, int nsegs, int natoms,
                                       const float *d_ff, float b_global, int npatterns, const double *rot9s,
                                       double eps, float *d_F, float *d_I, int *info, void *stream) {
  if (npatterns < 1) return spsim::fail(SPSIM_HIP_ERR_INVALID, "npatterns=%d", npatterns);
  return spsim::nufft_impl(d_hkl, npix, d_atoms, d_atom_b, d_segs, nsegs, natoms, d_ff, b_global, nullptr, npatterns,
                           rot9s, eps, d_F, d_I, info, (hipStream_t)stream);
}

extern "C" int spsim_hip_sf_spectrum_nfft(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms,
                                          const spsim_hip_segment *d_segs, int nsegs, int natoms,
                                          const float *d_ff, float b_global, const double *rot9, int nsamples,
                                          const double *scales, const double *weights, double eps, float *d_Fk,
                                          float *d_F, float *d_I, int *info, void *stream) {
  using namespace spsim;
  if (nsamples < 1 || nsamples > SPSIM_HIP_MAX_SPECTRUM)
    return fail(SPSIM_HIP_ERR_INVALID, "nsamples=%d outside [1, %d]", nsamples, SPSIM_HIP_MAX_SPECTRUM);
  if (!scales || !weights) return fail(SPSIM_HIP_ERR_INVALID, "NULL scales or weights");
  if (npix < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative size");
  if (npix > 0 && !d_Fk) return fail(SPSIM_HIP_ERR_INVALID, "NULL per-sample buffer");
  if (npix > 0 && !d_F && !d_I) return fail(SPSIM_HIP_ERR_INVALID, "both outputs are NULL");
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (rot9)
    for (int i = 0; i < 9; ++i) R[i] = rot9[i];
  std::vector<double> rots(9 * (size_t)nsamples);
  SpectrumWeights wt{};
  for (int k = 0; k < nsamples; ++k) {
    if (!(std::isfinite(scales[k]) && scales[k] > 0.0 && std::isfinite(weights[k])))
      return fail(SPSIM_HIP_ERR_INVALID, "sample %d: scale must be finite and > 0, weight finite", k);
    for (int i = 0; i < 9; ++i) rots[9 * (size_t)k + i] = scales[k] * R[i];
    wt.w[k] = (float)weights[k];
  }
  // per-sample sums over one set of grids (q' = s_k R q; form factors and the
  // global B at the unscaled |q|), then the weighted sum
  SPSIM_TRY(nufft_impl(d_hkl, npix, d_atoms, nullptr, d_segs, nsegs, natoms, d_ff, b_global, nullptr, nsamples,
                       rots.data(), eps, d_Fk, nullptr, info, (hipStream_t)stream));
  if (npix == 0) return SPSIM_HIP_OK;
  const size_t zb = ((size_t)npix + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(spectrum_combine, dim3((unsigned)(zb < 4096 ? zb : 4096)), dim3(kThreads), 0,
                     (hipStream_t)stream, (size_t)npix, nsamples, wt, (const float2 *)d_Fk, (float2 *)d_F, d_I);
  SPSIM_HIP_TRY(hipGetLastError());
  return SPSIM_HIP_OK;
}

extern "C" int spsim_hip_sf_nfft_molecules(const float *d_hkl, int npix, int nmol, const spsim_hip_atom *d_atoms,
                                           const float *d_atom_b, const int *atom_offsets,
                                           const spsim_hip_segment *d_segs, const int *seg_offsets,
                                           const float *d_ff, float b_global, const double *rot9s, double eps,
                                           float *d_F, float *d_I, void *stream) {
  using namespace spsim;
  hipStream_t s = (hipStream_t)stream;
  if (nmol < 0 || npix < 0) return fail(SPSIM_HIP_ERR_INVALID, "negative size");
  if (nmol == 0 || npix == 0) return SPSIM_HIP_OK;
  if (!atom_offsets || !seg_offsets || !d_hkl || !d_ff) return fail(SPSIM_HIP_ERR_INVALID, "NULL input pointer");
  if (!d_F && !d_I) return fail(SPSIM_HIP_ERR_INVALID, "both outputs are NULL");
  if (atom_offsets[0] != 0 || seg_offsets[0] != 0) return fail(SPSIM_HIP_ERR_INVALID, "offsets must start at 0");
  for (int m = 0; m < nmol; ++m) {
    const int na = atom_offsets[m + 1] - atom_offsets[m], ns = seg_offsets[m + 1] - seg_offsets[m];
    if (na < 0 || ns < 0 || ns > SPSIM_HIP_NSPECIES) return fail(SPSIM_HIP_ERR_INVALID, "molecule %d: bad offsets", m);
    if (na > 0 && (!d_atoms || !d_segs || ns < 1)) return fail(SPSIM_HIP_ERR_INVALID, "molecule %d: no atoms or segments", m);
    if (rot9s)
      for (int i = 0; i < 9; ++i)
        if (!std::isfinite(rot9s[9 * (size_t)m + i])) return fail(SPSIM_HIP_ERR_INVALID, "non-finite rotation");
  }
  SPSIM_TRY(ensure_device(nullptr));
  const int W = width_for(eps);
  // 1. Bounds: the q list's once (every molecule shares it) and every
  // molecule's atoms in one launch, read with one synchronisation.
  constexpr int kMolBlocks = 64;
  if (nmol > 65535) return fail(SPSIM_HIP_ERR_INVALID, "at most 65535 molecules per call (%d)", nmol);
  const int qblk = (npix + kThreads - 1) / kThreads < kBoundBlocks ? (npix + kThreads - 1) / kThreads : kBoundBlocks;
  const size_t b_q = align256(sizeof(double) * kNb * qblk);
  const size_t b_part = align256(sizeof(double) * kNb * kMolBlocks * (size_t)nmol);
  const size_t b_off = align256(sizeof(int) * ((size_t)nmol + 1));
  const size_t b_rot = rot9s ? align256(sizeof(double) * 9 * (size_t)nmol) : 0;
  void *ws = nullptr;
  SPSIM_TRY(workspace(b_q + b_part + b_off + b_rot, &ws));
  double *qpart = (double *)ws;
  double *part = (double *)((char *)ws + b_q);
  int *d_off = (int *)((char *)ws + b_q + b_part);
  double *d_rot = rot9s ? (double *)((char *)ws + b_q + b_part + b_off) : nullptr;
  std::vector<AtomRot> Rs(nmol);
  for (int m = 0; m < nmol; ++m) {
    Rs[m] = AtomRot{};
    if (rot9s) {
      Rs[m].on = 1;
      for (int i = 0; i < 9; ++i) Rs[m].r[i] = rot9s[9 * (size_t)m + i];
    }
  }
  hipLaunchKernelGGL(bounds_partial, dim3(qblk), dim3(kThreads), 0, s, nullptr, nullptr, 0, d_hkl, npix, AtomRot{},
                     qpart);
  SPSIM_HIP_TRY(hipGetLastError());
  if (atom_offsets[nmol] > 0) {
    // pageable sources: both copies have completed by the synchronisation below
    SPSIM_HIP_TRY(hipMemcpyAsync(d_off, atom_offsets, sizeof(int) * ((size_t)nmol + 1), hipMemcpyHostToDevice, s));
    if (rot9s)
      SPSIM_HIP_TRY(hipMemcpyAsync(d_rot, rot9s, sizeof(double) * 9 * (size_t)nmol, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bounds_molecules, dim3(kMolBlocks, (unsigned)nmol), dim3(kThreads), 0, s, d_atoms, d_atom_b,
                       d_off, d_rot, part);
    SPSIM_HIP_TRY(hipGetLastError());
  }
  std::vector<double> hq((size_t)qblk * kNb), hpart((size_t)nmol * kMolBlocks * kNb);
  SPSIM_HIP_TRY(hipMemcpyAsync(hq.data(), qpart, sizeof(double) * hq.size(), hipMemcpyDeviceToHost, s));
  if (atom_offsets[nmol] > 0)
    SPSIM_HIP_TRY(hipMemcpyAsync(hpart.data(), part, sizeof(double) * hpart.size(), hipMemcpyDeviceToHost, s));
  SPSIM_HIP_TRY(hipStreamSynchronize(s));
  double hbq[kNb];
  fold_bounds(hq.data(), qblk, hbq);
  // 2. Every molecule's grids planned on the host (workspace bytes only), then
  // the stages enqueued back to back on kWorkers streams: the kHelperStreams
  // streams forked from s, and s itself. Molecule m goes to worker
  // m mod kWorkers, each worker with a workspace slice as large as its largest
  // molecule's. Small molecules' kernels fill few CUs, so the streams overlap
  // them; s then waits for the helpers.
  std::vector<double> hbs((size_t)nmol * kNb);
  std::vector<size_t> need(nmol, 0);
  for (int m = 0; m < nmol; ++m) {
    const int na = atom_offsets[m + 1] - atom_offsets[m], ns = seg_offsets[m + 1] - seg_offsets[m];
    if (na == 0) continue;
    double *hb = hbs.data() + (size_t)m * kNb;
    fold_bounds(hpart.data() + (size_t)m * kMolBlocks * kNb, kMolBlocks, hb);
    for (int v : {3, 4, 5, 9, 10, 11, 12}) hb[v] = hbq[v];
    double hcopy[kNb];
    for (int v = 0; v < kNb; ++v) hcopy[v] = hb[v];
    SPSIM_TRY(nufft_run(hcopy, d_hkl, npix, d_atoms + atom_offsets[m], d_atom_b ? d_atom_b + atom_offsets[m] : nullptr,
                        d_segs + seg_offsets[m], ns, na, d_ff, b_global, Rs[m], 0, nullptr, W, eps, nullptr, nullptr,
                        nullptr, s, &need[m]));
  }
  constexpr int kWorkers = kHelperStreams + 1;
  size_t slice[kWorkers] = {}, off[kWorkers] = {}, total = 0;
  for (int m = 0; m < nmol; ++m)
    if (need[m] > slice[m % kWorkers]) slice[m % kWorkers] = need[m];
  for (int k = 0; k < kWorkers; ++k) {
    off[k] = total;
    total += align256(slice[k]);
  }
  SPSIM_TRY(workspace(total > 0 ? total : 256, &ws));  // the bounds' records are on the host already
  hipStream_t hs[kWorkers];
  hipEvent_t ev[kHelperStreams + 1];
  SPSIM_TRY(helper_streams(hs, ev));
  hs[kHelperStreams] = s;
  SPSIM_HIP_TRY(hipEventRecord(ev[kHelperStreams], s));
  for (int k = 0; k < kHelperStreams; ++k) SPSIM_HIP_TRY(hipStreamWaitEvent(hs[k], ev[kHelperStreams], 0));
  int rc = SPSIM_HIP_OK;
  for (int m = 0; m < nmol && rc == SPSIM_HIP_OK; ++m) {
    const int na = atom_offsets[m + 1] - atom_offsets[m], ns = seg_offsets[m + 1] - seg_offsets[m];
    const int k = m % kWorkers;
    float *F = d_F ? d_F + (size_t)m * npix * 2 : nullptr;
    float *I = d_I ? d_I + (size_t)m * npix : nullptr;
    if (na == 0) {
      const size_t zb = ((size_t)npix + kThreads - 1) / kThreads;
      hipLaunchKernelGGL(zero_pattern, dim3((unsigned)(zb < 4096 ? zb : 4096)), dim3(kThreads), 0, hs[k],
                         (size_t)npix, (float2 *)F, I);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) rc = hip_fail(e, "zero_pattern");
      continue;
    }
    rc = nufft_run(hbs.data() + (size_t)m * kNb, d_hkl, npix, d_atoms + atom_offsets[m],
                   d_atom_b ? d_atom_b + atom_offsets[m] : nullptr, d_segs + seg_offsets[m], ns, na, d_ff, b_global,
                   Rs[m], 0, nullptr, W, eps, F, I, nullptr, hs[k], nullptr, (char *)ws + off[k]);
  }
  // join even after a failure, so that s orders after whatever was enqueued
  for (int k = 0; k < kHelperStreams; ++k) {
    SPSIM_HIP_TRY(hipEventRecord(ev[k], hs[k]));
    SPSIM_HIP_TRY(hipStreamWaitEvent(s, ev[k], 0));
  }
  return rc;
}

extern "C" int spsim_hip_fft_pruned(const float *d_in, long long outer, long long inner, int M, int m_lo, int nf,
                                    int L, int l_lo, float *d_out, void *stream) {
  using namespace spsim;
  if (outer < 0 || inner < 1 || M < 1 || L < 1 || !d_in || !d_out)
    return fail(SPSIM_HIP_ERR_INVALID, "bad FFT shape or NULL pointer");
  if (nf < 2 || nf > kMaxNf || !smooth_length(nf) || M > nf)
    return fail(SPSIM_HIP_ERR_INVALID, "FFT length %d must be a 2,3,5-smooth number in [2, %d] and >= M", nf, kMaxNf);
  if (m_lo < -nf || m_lo + M > 2 * nf || l_lo < -nf || l_lo + L > 2 * nf)
    return fail(SPSIM_HIP_ERR_INVALID, "first point or output outside [-nf, nf], or last one beyond 2 nf");
  SPSIM_TRY(ensure_device(nullptr));
  NufftDim D{};
  D.M = M;
  D.m_lo = m_lo;
  D.nf = nf;
  D.L = L;
  D.l_lo = l_lo;
  void *ws = nullptr;
  SPSIM_TRY(workspace(sizeof(float2) * (size_t)nf, &ws));
  SPSIM_TRY(launch_twiddles(nf, (float2 *)ws, (hipStream_t)stream));
  return launch_fft(make_pass(D, outer, inner), (const float2 *)d_in, (float2 *)d_out, (const float2 *)ws,
                    (hipStream_t)stream);
}
