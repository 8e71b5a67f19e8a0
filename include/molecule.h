/*
 * spsim-mi355x drop-in: molecule constructors (reference include/molecule.h:22-31).
 */
#ifndef _MOLECULE_H_
#define _MOLECULE_H_ 1

#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/molecule.c:394-400 */
Molecule *alloc_mol(void);
/* src/molecule.c:402-411 */
void add_atom_to_mol(Molecule *mol, int atomic_number, float x, float y, float z);
/* src/molecule.c:413-417 */
void free_mol(Molecule *mol);
/* src/molecule.c:419-442: atomic_number is a 1 x N image, pos a 3 x N image. */
Molecule *make_mol(Image *atomic_number, Image *pos);
/* src/molecule.c:444-464: shifts the origin to the electron-weighted centre. */
void origin_to_center_of_mass(Molecule *mol);
/* src/molecule.c:33-63: random atoms for a chemical formula inside box_dimension. */
Molecule *get_Molecule_from_formula(Chem_Formula *form, Options *opts);

#ifdef __cplusplus
}
#endif

#endif
