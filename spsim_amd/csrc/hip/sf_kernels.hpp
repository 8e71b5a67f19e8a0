// Internal launch interface of the structure-factor kernels (sf_kernels.hip).
#pragma once

#include "common.hpp"

namespace spsim {

// Per-molecule description inside a batch: packed atoms [atom_begin, atom_end)
// and species segments [seg_begin, seg_begin + nseg) of the shared arrays; the
// segments' starts are relative to atom_begin.
struct MolInfo {
  int atom_begin, atom_end, seg_begin, nseg;
};

// Row-major rotation R applied to q (q' = R q). Kernels apply R^T to the atoms
// instead, in FP64, which gives the same phase q'.r = q.(R^T r).
struct Rot {
  double r[9];
};

// Degree of the spectrum factor's polynomials (in v^2) and the reach of the
// fit: |2 pi e_k t| <= kSpectrumReach radians for every sample k.
constexpr int kSpectrumDegree = 4;
constexpr double kSpectrumReach = 1.5;

// Wavelength samples of a spectrum pattern: phase scales and normalised
// weights, and the polynomial form of their sum about the centre scale sc:
// with e_k = s_k / sc - 1 and v = t / tlim (t = sc q.r in turns),
//   sum_k W_k cos(2 pi e_k t) ~ P(v^2),  sum_k W_k sin(2 pi e_k t) ~ v Q(v^2)
// for |t| <= tlim. tlim < 0 turns the polynomial off.
struct Spectrum {
  int n;
  float w[SPSIM_HIP_MAX_SPECTRUM];
  double s[SPSIM_HIP_MAX_SPECTRUM];
  double sc;
  float tlim, inv_tlim;
  float p[kSpectrumDegree + 1], q[kSpectrumDegree + 1];
};

// Fills sp.sc, sp.tlim, sp.inv_tlim, sp.p and sp.q from sp.n, sp.s and sp.w
// (host code). The polynomial is used from 4 samples on.
void spectrum_fit(Spectrum &sp);

struct PatternLaunch {
  const float *hkl;
  int npix;
  const spsim_hip_atom *atoms;
  const float *atom_b;  // may be null
  const spsim_hip_segment *segs;
  int nsegs;
  int natoms;
  const float *ff;
  float b_global;
  float2 *F;  // may be null
  float *I;   // may be null
  int has_rot;
  Rot rot;
};

// One work item of the batched kernels: the atoms [a_begin, a_end) (relative
// to the molecule's first packed atom) of one pattern on one tile. When the
// pattern's atoms are split over nchunks > 1 items, the tile's items share the
// counter counters[group] and the partial slots slot .. slot + nchunks - 1
// (slot_entries(tile, grid) double2 entries each), this item's being slot + chunk.
struct BatchItem {
  int pattern, tile, a_begin, a_end;
  int group, chunk, nchunks, slot;
};

struct BatchLaunch {
  const float *hkl;
  int npix;
  int nx, ny;  // detector shape (grid work items)
  const spsim_hip_atom *atoms;
  const float *atom_b;  // may be null
  const spsim_hip_segment *segs;
  const MolInfo *mols;
  const int *pattern_mol;
  const double *rot;  // [npatterns][9]
  const BatchItem *items;  // cost-sorted
  int nitems;
  unsigned *queue;       // work counter, reset by launch_batch
  unsigned *counters;    // per split tile, zero between runs
  double2 *partials;     // partial slots of split tiles
  const float *ff;
  float b_global;
  float2 *F;  // [npatterns][npix] or null
  float *I;   // [npatterns][npix]
};

// Valid tiles: per-pixel kernels take pixels_per_lane in {1,2,4,8}; the
// mirror-symmetric (grid) kernels take quads per lane in {1,2}, or
// kOctantLane for one octant (8 pixels) per lane on a square detector;
// waves == 4; chunk_atoms in {64,128,256}.
constexpr int kOctantLane = 8;
bool tile_supported(const spsim_hip_tile &t, bool grid);
// Pixels (per-pixel kernels) or quads (grid kernels) covered by one tile.
inline int tile_units(const spsim_hip_tile &t) { return t.pixels_per_lane * t.waves * 64; }
__host__ __device__ inline int grid_quads(int nx, int ny) { return ((nx + 1) / 2) * ((ny + 1) / 2); }
// Octants x <= y of the lower quadrant of an n x n detector.
__host__ __device__ inline int grid_octants(int n) {
  const int h = (n + 1) / 2;
  return h * (h + 1) / 2;
}
// Work units of a grid tile (quads, or octants for kOctantLane) and their count.
inline int grid_units_per_tile(const spsim_hip_tile &t) {
  return t.pixels_per_lane == kOctantLane ? t.waves * 64 : tile_units(t);
}
inline int grid_units(const spsim_hip_tile &t, int nx, int ny) {
  return t.pixels_per_lane == kOctantLane ? grid_octants(nx) : grid_quads(nx, ny);
}
// double2 entries of one partial slot of the batched kernels: each lane's
// pixels of a tile (K per lane, 4 K for quads, 8 for an octant).
inline int slot_entries(const spsim_hip_tile &t, bool grid) {
  const int per_lane = !grid ? t.pixels_per_lane : (t.pixels_per_lane == kOctantLane ? 8 : 4 * t.pixels_per_lane);
  return per_lane * t.waves * 64;
}
// Pixels per grid work unit.
inline int grid_unit_pixels(const spsim_hip_tile &t) { return t.pixels_per_lane == kOctantLane ? 8 : 4; }

// One pattern. splits >= 1; with splits > 1 the library workspace holds the partial sums.
int launch_pattern(const PatternLaunch &a, const spsim_hip_tile &t, int splits, hipStream_t s);
// Same on an nx x ny detector list, exploiting its mirror symmetry where present.
int launch_grid(const PatternLaunch &a, int nx, int ny, const spsim_hip_tile &t, int splits, hipStream_t s);
// Same with a wavelength spectrum (per-pixel kernel; pixels_per_lane 1 or 2).
int launch_spectrum(const PatternLaunch &a, const Spectrum &sp, const spsim_hip_tile &t, int splits, hipStream_t s);
// Persistent kernel: `blocks` workgroups pull work items from the queue. With
// grid != 0 the items are quad tiles of the a.nx x a.ny detector.
int launch_batch(const BatchLaunch &a, const spsim_hip_tile &t, bool grid, int blocks, hipStream_t s);
// Workgroups of the batch kernel that fit on one CU at once for this tile.
int batch_blocks_per_cu(const spsim_hip_tile &t, bool has_b, bool grid, int *blocks);
// Default tile for a single pattern of npix pixels and natoms atoms.
spsim_hip_tile default_tile(int natoms, int npix, bool grid);
// Atom ranges per pixel tile of one pattern (at most max_splits, each at
// least one LDS chunk): the count that minimises the launch's length in rounds
// of `slots` resident workgroups.
int auto_splits(int tiles, int natoms, int chunk_atoms, int slots, int max_splits);
// Workgroups of `kernel` resident on the device at once (cus CUs).
int resident_slots(const void *kernel, int cus);
// Largest automatic partial-sum buffer of a single pattern (splits x npix double2
// entries in the library workspace): 512 MiB, 32 atom ranges of a 1024 x 1024 pattern.
constexpr size_t kWorkspacePartials = (size_t)1 << 25;
// True when the [ny][nx][3] host list is exactly mirror-symmetric in x and y.
bool host_mirror_symmetric(const float *hkl, int nx, int ny);
// True when the n x n host list is also transpose-symmetric: q(y, x) = (qy, qx, qz)(x, y).
bool host_transpose_symmetric(const float *hkl, int n);

}  // namespace spsim
