/*
 * Molecule constructors (reference src/molecule.c:33-63, 394-464). Memory
 * layout and ownership follow the reference: atomic_number and pos are
 * realloc'd arrays owned by the Molecule and released by free_mol.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "molecule.h"
#include "spsim_internal.h"

Molecule *alloc_mol(void) {
  Molecule *mol = (Molecule *)malloc(sizeof(Molecule));
  if (!mol) return NULL;
  mol->natoms = 0;
  mol->atomic_number = NULL;
  mol->pos = NULL;
  mol->delta_atoms = 0;
  return mol;
}

/* Grows the arrays geometrically (the reference reallocs per atom); the
   capacity is not part of the struct, so it is recomputed from natoms. */
static int capacity_for(int n) {
  int c = 1;
  while (c < n) c <<= 1;
  return c;
}

void add_atom_to_mol(Molecule *mol, int atomic_number, float x, float y, float z) {
  if (!mol) return;
  const int i = mol->natoms;
  if (i == 0 || capacity_for(i) == i) {
    const int cap = capacity_for(i + 1);
    int *zn = (int *)realloc(mol->atomic_number, sizeof(int) * cap);
    if (!zn) return;
    mol->atomic_number = zn;
    float *pn = (float *)realloc(mol->pos, sizeof(float) * 3 * cap);
    if (!pn) return;
    mol->pos = pn;
  }
  mol->atomic_number[i] = atomic_number;
  mol->pos[3 * i + 0] = x;
  mol->pos[3 * i + 1] = y;
  mol->pos[3 * i + 2] = z;
  mol->natoms = i + 1;
}

void free_mol(Molecule *mol) {
  if (!mol) return;
  free(mol->atomic_number);
  free(mol->pos);
  free(mol);
}

Molecule *make_mol(Image *atomic_number, Image *pos) {
  if (!atomic_number || !pos || !atomic_number->image || !pos->image) return NULL;
  if (pos->image->x != 3) {
    fprintf(stderr, "Array of positions must have 3 times N dimensions, where N denotes the number of atoms.\n");
    return NULL;
  }
  if (atomic_number->image->x != pos->image->y) {
    fprintf(stderr, "Array of atomic numbers and atomic positions do not have compatible shapes.\n");
    return NULL;
  }
  const int N = (int)atomic_number->image->x;
  Molecule *mol = alloc_mol();
  if (!mol) return NULL;
  mol->atomic_number = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
  mol->pos = (float *)malloc(sizeof(float) * 3 * (N > 0 ? N : 1));
  if (!mol->atomic_number || !mol->pos) {
    free_mol(mol);
    return NULL;
  }
  mol->natoms = N;
  for (int i = 0; i < N; ++i) {
    mol->atomic_number[i] = (int)atomic_number->image->data[i].re;
    for (int d = 0; d < 3; ++d) mol->pos[3 * i + d] = pos->image->data[3 * i + d].re;
  }
  return mol;
}

void origin_to_center_of_mass(Molecule *mol) {
  /* electron-weighted centre (weights = Z), accumulated in float as the reference does */
  float com[3] = {0.f, 0.f, 0.f};
  int n = 0;
  if (!mol || mol->natoms <= 0) return;
  for (int i = 0; i < mol->natoms; ++i) {
    n += mol->atomic_number[i];
    for (int d = 0; d < 3; ++d) com[d] += mol->pos[3 * i + d] * mol->atomic_number[i];
  }
  if (n == 0) return;
  for (int d = 0; d < 3; ++d) com[d] /= n;
  for (int i = 0; i < mol->natoms; ++i)
    for (int d = 0; d < 3; ++d) mol->pos[3 * i + d] -= com[d];
}

Molecule *get_Molecule_from_formula(Chem_Formula *form, Options *opts) {
  if (!form || !opts) return NULL;
  Molecule *res = alloc_mol();
  if (!res) return NULL;
  int total = 0;
  for (int i = 0; i < form->n_elements; ++i) total += form->quantity[i];
  res->atomic_number = (int *)malloc(sizeof(int) * (total > 0 ? total : 1));
  res->pos = (float *)malloc(sizeof(float) * 3 * (total > 0 ? total : 1));
  if (!res->atomic_number || !res->pos) {
    free_mol(res);
    return NULL;
  }
  int j = 0;
  const float box = opts->box_dimension;
  for (int i = 0; i < form->n_elements; ++i) {
    for (int q = 0; q < form->quantity[i]; ++q, ++j) {
      res->atomic_number[j] = form->atomic_number[i];
      float *p = &res->pos[3 * j];
      if (opts->box_type == BOX_SPHERICAL) {
        /* rejection sampling of a ball of diameter box_dimension */
        float r;
        do {
          r = 0;
          for (int d = 0; d < 3; ++d) {
            p[d] = (float)((p_drand48() - 0.5) * box);
            r += p[d] * p[d];
          }
          r = sqrtf(r);
        } while (r > box / 2.0);
      } else if (opts->box_type == BOX_PARALLEL) {
        for (int d = 0; d < 3; ++d) p[d] = (float)(p_drand48() * box);
      } else {
        p[0] = p[1] = p[2] = 0.0f;
      }
    }
  }
  res->natoms = total;
  return res;
}
