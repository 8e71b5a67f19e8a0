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
/* src/molecule.c:65-298: ATOM/HETATM records plus the REMARK 350 BIOMT
 * assembly operators; positions in metres. Aborts when the file cannot be
 * opened, as the reference does. */
Molecule *get_Molecule_from_pdb(char *filename);
/* src/molecule.c:367-392: one ATOM line per atom (Angstrom) and END. */
void write_pdb_from_mol(char *filename, Molecule *mol);

/* No counterpart in the reference: get_Molecule_from_pdb with a status
 * (0; -1 file not readable; -2 out of memory) instead of abort. */
int spsim_read_pdb(const char *filename, Molecule **out, char *err, int errlen);
/* No counterpart in the reference: write_pdb_from_mol with a status (0 or -1). */
int spsim_write_pdb(const char *filename, const Molecule *mol);
/* No counterpart in the reference: atomic number of a two-character PDB
 * element field (src/molecule.c:305-343), 0 when it names no element. */
int spsim_element_from_symbol(const char field[2]);

#ifdef __cplusplus
}
#endif

#endif
