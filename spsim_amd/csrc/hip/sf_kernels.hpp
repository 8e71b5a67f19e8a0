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
};

struct BatchLaunch {
  const float *hkl;
  int npix;
  const spsim_hip_atom *atoms;
  const float *atom_b;  // may be null
  const spsim_hip_segment *segs;
  const MolInfo *mols;
  const int *pattern_mol;
  const double *rot;  // [npatterns][9]
  const int2 *items;  // {pattern, tile}, cost-sorted
  int nitems;
  unsigned *queue;    // work counter, reset by launch_batch
  const float *ff;
  float b_global;
  float2 *F;  // [npatterns][npix] or null
  float *I;   // [npatterns][npix]
};

// Valid tile parameters: pixels_per_lane in {1,2,4,8}, waves == 4, chunk in {64,128,256}.
bool tile_supported(const spsim_hip_tile &t);
inline int tile_pixels(const spsim_hip_tile &t) { return t.pixels_per_lane * t.waves * 64; }

// One pattern. splits >= 1; with splits > 1 the library workspace holds the partial sums.
int launch_pattern(const PatternLaunch &a, const spsim_hip_tile &t, int splits, hipStream_t s);
// Persistent kernel: grid of `blocks` workgroups pulls work items from the queue.
int launch_batch(const BatchLaunch &a, const spsim_hip_tile &t, int blocks, hipStream_t s);
// Workgroups of the batch kernel that fit on one CU at once for this tile.
int batch_blocks_per_cu(const spsim_hip_tile &t, bool has_b, int *blocks);
// Default tile and split count for a single pattern of npix pixels and natoms atoms.
spsim_hip_tile default_tile(int natoms, int npix, bool has_b);
int auto_splits(const spsim_hip_tile &t, int natoms, int npix, int cus);

}  // namespace spsim
