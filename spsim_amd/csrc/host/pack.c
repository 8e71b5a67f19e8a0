/*
 * Atom packing: turns the reference's Molecule arrays into the species-sorted
 * records the GPU kernels stream (spsim_hip_pack_atoms in spsim_hip.h).
 *
 * Species grouping follows the reference's two observations: the CPU path
 * caches f_Z(|q|) per element present (src/diffraction.c:824-828) and the
 * CUDA path sorts atoms by Z to evaluate f only on a change of Z
 * (src/diffraction.cu:254-271). Here the sort is a stable counting sort, so
 * atoms keep their input order inside each species.
 */
#include <stdlib.h>

#include "spsim_hip.h"

static int species_of(int z, int j, int natoms, int delta_atoms, int mode) {
  if (mode == SPSIM_HIP_PACK_VECTOR) {
    /* src/diffraction.c:903-947: the 4-wide loop honours delta_atoms, the
       scalar tail (:949-956) does not; Z == 0 is not skipped. */
    if (delta_atoms && j < 4 * (natoms / 4)) return SPSIM_HIP_UNIT_SPECIES;
    return z;
  }
  if (mode == SPSIM_HIP_PACK_NFFT) {
    /* src/diffraction.c:556-563, 599-624: every element present gets its own
       transform, Z == 0 included; delta_atoms skips the form factor
       (:659-661), so every atom scatters with f = 1. */
    return delta_atoms ? SPSIM_HIP_UNIT_SPECIES : z;
  }
  /* src/diffraction.c:830-843 */
  if (z == 0) return -1;
  if (delta_atoms) return SPSIM_HIP_UNIT_SPECIES;
  return z;
}

int spsim_hip_pack_atoms(const int *Z, const float *pos, const float *amp, const float *b, int natoms, int delta_atoms,
                         int mode, spsim_hip_atom *out_atoms, float *out_b, int *out_perm, spsim_hip_segment *out_segs,
                         int *out_natoms, int *out_nsegs) {
  int count[SPSIM_HIP_NSPECIES];
  int offset[SPSIM_HIP_NSPECIES];
  int j, s, n, nseg;
  if (natoms < 0 || !out_atoms || !out_segs || !out_natoms || !out_nsegs) return SPSIM_HIP_ERR_INVALID;
  if (natoms > 0 && (!Z || !pos)) return SPSIM_HIP_ERR_INVALID;
  if (mode != SPSIM_HIP_PACK_PLAIN && mode != SPSIM_HIP_PACK_VECTOR && mode != SPSIM_HIP_PACK_NFFT)
    return SPSIM_HIP_ERR_INVALID;
  for (s = 0; s < SPSIM_HIP_NSPECIES; ++s) count[s] = 0;
  for (j = 0; j < natoms; ++j) {
    /* Z indexes a 100-row table; anything outside it is a caller error (the
       reference would read out of bounds). */
    if (Z[j] < 0 || Z[j] >= SPSIM_HIP_UNIT_SPECIES) return SPSIM_HIP_ERR_INVALID;
    s = species_of(Z[j], j, natoms, delta_atoms, mode);
    if (s >= 0) ++count[s];
  }
  n = 0;
  nseg = 0;
  for (s = 0; s < SPSIM_HIP_NSPECIES; ++s) {
    offset[s] = n;
    if (count[s] > 0) {
      out_segs[nseg].start = n;
      out_segs[nseg].count = count[s];
      out_segs[nseg].species = s;
      out_segs[nseg].reserved = 0;
      ++nseg;
    }
    n += count[s];
  }
  for (j = 0; j < natoms; ++j) {
    int k;
    s = species_of(Z[j], j, natoms, delta_atoms, mode);
    if (s < 0) continue;
    k = offset[s]++;
    out_atoms[k].x = pos[3 * j + 0];
    out_atoms[k].y = pos[3 * j + 1];
    out_atoms[k].z = pos[3 * j + 2];
    /* Delta atoms ignore the illumination weight (src/diffraction.c:840-842);
       the NFFT path has unit strengths (:619-620). */
    out_atoms[k].w = (s == SPSIM_HIP_UNIT_SPECIES || !amp || mode == SPSIM_HIP_PACK_NFFT) ? 1.0f : amp[j];
    if (out_b) out_b[k] = b ? b[j] : 0.0f;
    if (out_perm) out_perm[k] = j;
  }
  *out_natoms = n;
  *out_nsegs = nseg;
  return SPSIM_HIP_OK;
}
