/*
 * PDB molecule input and output (reference src/molecule.c:65-298 reader,
 * :305-364 element symbols, :367-392 writer).
 *
 * Reader. ATOM and HETATM records give one atom each: x, y and z come from
 * the fixed columns [30,38), [38,46) and [46,54) (0-based, end exclusive),
 * parsed as a double, rounded to float, then scaled by 1e-10 in double and
 * rounded to float again (metres), as the reference does. The element comes
 * from columns [76,78): a leading blank is dropped, letters are upper-cased,
 * and the pair is looked up in the periodic table; a pair that is not a
 * symbol falls back to its first letter alone. "REMARK 350   BIOMT" records
 * (column 18: row 1..3; columns [20,24): operator number; [24,33), [34,43),
 * [44,53): the row of the rotation; [59,68): the translation in Angstrom)
 * define the biological assembly: after reading, operator 0 is applied to
 * every atom in place, then operators 1..n-1 to the result, giving
 * natoms * n atoms ordered by operator, each coordinate computed as
 * ((r0 x + r1 y) + r2 z) + t in float. Operator 0 is applied one coordinate
 * at a time in place, so a first operator that is not the identity sees its
 * own updated x (and y), exactly as in the reference (src/molecule.c:254-275;
 * PDB files list the identity first).
 *
 * Differences from the reference, all in cases where it reads or writes
 * memory it does not own:
 *   - an atom whose element cannot be identified (blank, or a line shorter
 *     than 77 characters) is skipped with a warning; the reference's skip
 *     sets the previous atom's atomic number to 0 instead
 *     (src/molecule.c:196-201), and it reads stale buffer bytes for lines
 *     shorter than 76 characters;
 *   - a BIOMT row number outside 1..3 or an operator number below 1 is
 *     ignored with a warning, and operators that are never defined are zero
 *     matrices (the reference leaves them uninitialised);
 *   - a missing file is reported through the status of spsim_read_pdb
 *     (get_Molecule_from_pdb still aborts, as the reference does).
 *
 * Writer. One "ATOM" line per atom in the reference's layout (serial modulo
 * 99999, the two-character symbol, coordinates in Angstrom as %8.3f) and a
 * closing "END". It also fills the standard element columns 77-78, which
 * the reference leaves out, so that the reader above can read the file back.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "molecule.h"

/* Element symbols for Z = 1..105, two characters, left-justified, upper case
 * (105 is the historical "HA"). */
static const char kSymbols[105][3] = {
    "H ", "HE", "LI", "BE", "B ", "C ", "N ", "O ", "F ", "NE", "NA", "MG", "AL", "SI", "P ", "S ", "CL", "AR",
    "K ", "CA", "SC", "TI", "V ", "CR", "MN", "FE", "CO", "NI", "CU", "ZN", "GA", "GE", "AS", "SE", "BR", "KR",
    "RB", "SR", "Y ", "ZR", "NB", "MO", "TC", "RU", "RH", "PD", "AG", "CD", "IN", "SN", "SB", "TE", "I ", "XE",
    "CS", "BA", "LA", "CE", "PR", "ND", "PM", "SM", "EU", "GD", "TB", "DY", "HO", "ER", "TM", "YB", "LU", "HF",
    "TA", "W ", "RE", "OS", "IR", "PT", "AU", "HG", "TL", "PB", "BI", "PO", "AT", "RN", "FR", "RA", "AC", "TH",
    "PA", "U ", "NP", "PU", "AM", "CM", "BK", "CF", "ES", "FM", "MD", "NO", "LR", "RF", "HA"};

static int find_symbol(char a, char b) {
  for (int z = 0; z < 105; z++)
    if (kSymbols[z][0] == a && kSymbols[z][1] == b) return z + 1;
  return 0;
}

/* Atomic number of a two-character element field, 0 if none
 * (src/molecule.c:305-343). */
int spsim_element_from_symbol(const char field[2]) {
  char a = field[0], b = field[1];
  if (isspace((unsigned char)a)) { /* left-justify, as the reader does */
    a = b;
    b = ' ';
  }
  if (isalpha((unsigned char)a)) {
    a = (char)toupper((unsigned char)a);
  } else if (isalpha((unsigned char)b)) {
    a = (char)toupper((unsigned char)b);
    b = ' ';
  } else {
    return 0;
  }
  b = isalpha((unsigned char)b) ? (char)toupper((unsigned char)b) : ' ';
  int z = find_symbol(a, b);
  if (!z) z = find_symbol(a, ' ');
  return z;
}

/* atoi of line[a, min(b, len)). */
static int field_int(const char *line, size_t len, size_t a, size_t b) {
  char tmp[64];
  if (a >= len) return 0;
  if (b > len) b = len;
  size_t n = b - a;
  if (n >= sizeof tmp) n = sizeof tmp - 1;
  memcpy(tmp, line + a, n);
  tmp[n] = 0;
  return atoi(tmp);
}

/* atof of line[a, min(b, len)). */
static double field_value(const char *line, size_t len, size_t a, size_t b) {
  char tmp[64];
  if (a >= len) return 0.0;
  if (b > len) b = len;
  size_t n = b - a;
  if (n >= sizeof tmp) n = sizeof tmp - 1;
  memcpy(tmp, line + a, n);
  tmp[n] = 0;
  return atof(tmp);
}

typedef struct {
  float *rot;   /* [n][9] */
  float *trans; /* [n][3], metres */
  int n, cap;
} biomt;

static int biomt_reserve(biomt *s, int n) {
  if (n <= s->cap) return 0;
  int cap = s->cap ? s->cap : 1;
  while (cap < n) cap *= 2;
  float *r = (float *)realloc(s->rot, sizeof(float) * 9 * (size_t)cap);
  if (!r) return -1;
  s->rot = r;
  float *t = (float *)realloc(s->trans, sizeof(float) * 3 * (size_t)cap);
  if (!t) return -1;
  s->trans = t;
  memset(s->rot + 9 * (size_t)s->cap, 0, sizeof(float) * 9 * (size_t)(cap - s->cap));
  memset(s->trans + 3 * (size_t)s->cap, 0, sizeof(float) * 3 * (size_t)(cap - s->cap));
  s->cap = cap;
  return 0;
}

/* Reads one line of any length; returns its length without the newline,
 * -1 at end of file, -2 when out of memory. */
static long read_line(FILE *f, char **buf, size_t *cap) {
  size_t n = 0;
  int c;
  while ((c = fgetc(f)) != EOF) {
    if (n + 2 > *cap) {
      size_t nc = *cap ? 2 * *cap : 256;
      char *nb = (char *)realloc(*buf, nc);
      if (!nb) return -2;
      *buf = nb;
      *cap = nc;
    }
    if (c == '\n') break;
    (*buf)[n++] = (char)c;
  }
  if (c == EOF && n == 0) return -1;
  (*buf)[n] = 0;
  if (n && (*buf)[n - 1] == '\r') (*buf)[--n] = 0;
  return (long)n;
}

/* Status-returning reader: 0 on success (*out is a new Molecule), -1 when
 * the file cannot be read, -2 when out of memory. err (may be NULL) receives
 * a message. */
int spsim_read_pdb(const char *filename, Molecule **out, char *err, int errlen) {
  char dummy[8];
  if (!err || errlen <= 0) {
    err = dummy;
    errlen = (int)sizeof dummy;
  }
  err[0] = 0;
  if (!out) return -1;
  *out = NULL;
  FILE *fp = filename ? fopen(filename, "r") : NULL;
  if (!fp) {
    snprintf(err, (size_t)errlen, "cannot open PDB file %s", filename ? filename : "(null)");
    return -1;
  }
  Molecule *mol = alloc_mol();
  biomt sym = {0};
  char *line = NULL;
  size_t cap = 0;
  int cap_atoms = 0, electrons = 0, lineno = 0, rc = 0;
  long len;
  if (!mol) rc = -2;
  while (rc == 0 && (len = read_line(fp, &line, &cap)) >= 0) {
    lineno++;
    size_t L = (size_t)len;
    if (strncmp(line, "ATOM", 4) == 0 || strncmp(line, "HETATM", 6) == 0) {
      int Z = 0;
      if (L >= 77) {
        const char f[2] = {line[76], L >= 78 ? line[77] : ' '};
        Z = spsim_element_from_symbol(f);
      }
      if (!Z) {
        fprintf(stderr, "WARNING: Null atom at line %d '%s'. Skipping\n", lineno, line);
        continue;
      }
      if (mol->natoms >= cap_atoms) {
        int nc = cap_atoms ? 2 * cap_atoms : 1024;
        int *zn = (int *)realloc(mol->atomic_number, sizeof(int) * (size_t)nc);
        if (!zn) { rc = -2; break; }
        mol->atomic_number = zn;
        float *pn = (float *)realloc(mol->pos, sizeof(float) * 3 * (size_t)nc);
        if (!pn) { rc = -2; break; }
        mol->pos = pn;
        cap_atoms = nc;
      }
      const int i = mol->natoms;
      for (int d = 0; d < 3; d++) {
        float v = (float)field_value(line, L, 30 + 8 * (size_t)d, 38 + 8 * (size_t)d);
        mol->pos[3 * i + d] = (float)((double)v * 1e-10);
      }
      mol->atomic_number[i] = Z;
      mol->natoms = i + 1;
      electrons += Z;
    } else if (strncmp(line, "REMARK 350   BIOMT", 18) == 0) {
      int row = (L > 18 && isdigit((unsigned char)line[18])) ? line[18] - '0' - 1 : -1;
      int op = field_int(line, L, 20, 24) - 1;
      if (row < 0 || row > 2 || op < 0) {
        fprintf(stderr, "WARNING: unusable BIOMT record at line %d. Skipping\n", lineno);
        continue;
      }
      if (biomt_reserve(&sym, op + 1) != 0) { rc = -2; break; }
      if (op + 1 > sym.n) sym.n = op + 1;
      sym.rot[9 * op + 3 * row + 0] = (float)field_value(line, L, 24, 33);
      sym.rot[9 * op + 3 * row + 1] = (float)field_value(line, L, 34, 43);
      sym.rot[9 * op + 3 * row + 2] = (float)field_value(line, L, 44, 53);
      sym.trans[3 * op + row] = (float)(field_value(line, L, 59, 68) * 1e-10);
    }
  }
  if (len == -2) rc = -2;
  fclose(fp);
  free(line);
  if (rc == 0) {
    fprintf(stderr, "Read %d atoms with %d electrons\n", mol->natoms, electrons);
    fprintf(stderr, "Read %d symmetry transformations\n", sym.n);
  }
  if (rc == 0 && sym.n > 0 && mol->natoms > 0) {
    const int n = mol->natoms;
    const size_t total = (size_t)n * (size_t)sym.n;
    int *zn = (int *)realloc(mol->atomic_number, sizeof(int) * total);
    float *pn = zn ? (float *)realloc(mol->pos, sizeof(float) * 3 * total) : NULL;
    if (zn) mol->atomic_number = zn;
    if (pn) mol->pos = pn;
    if (!zn || !pn) {
      rc = -2;
    } else {
      for (int e = 0; e < sym.n; e++) {
        const float *R = sym.rot + 9 * e;
        const float *t = sym.trans + 3 * e;
        for (int i = 0; i < n; i++) {
          /* the source row is read live: for operator 0 it is also the
             destination, so rows 2 and 3 see the updated x (and y) */
          const float *p = mol->pos + 3 * (size_t)i;
          float *o = mol->pos + 3 * ((size_t)e * n + i);
          mol->atomic_number[(size_t)e * n + i] = mol->atomic_number[i];
          for (int r = 0; r < 3; r++) {
            float acc = p[0] * R[3 * r + 0];
            acc = acc + p[1] * R[3 * r + 1];
            acc = acc + p[2] * R[3 * r + 2];
            o[r] = acc + t[r];
          }
        }
      }
      mol->natoms = n * sym.n;
    }
  }
  if (rc == 0 && sym.n > 0)
    fprintf(stderr, "%d atoms with %d electrons after symmetrization\n", mol->natoms, electrons * sym.n);
  free(sym.rot);
  free(sym.trans);
  if (rc != 0) {
    free_mol(mol);
    snprintf(err, (size_t)errlen, "out of memory reading %s", filename);
    return rc;
  }
  *out = mol;
  return 0;
}

Molecule *get_Molecule_from_pdb(char *filename) {
  Molecule *mol = NULL;
  char err[1200];
  int rc = spsim_read_pdb(filename, &mol, err, (int)sizeof err);
  if (rc == -1) {
    fprintf(stderr, "Error reading PDB: %s\n", err);
    abort();
  }
  if (rc != 0) {
    fprintf(stderr, "%s\n", err);
    return NULL;
  }
  return mol;
}

/* Status variant of write_pdb_from_mol: 0, or -1 when the file cannot be written. */
int spsim_write_pdb(const char *filename, const Molecule *mol) {
  if (!filename || !mol) return -1;
  FILE *f = fopen(filename, "w");
  if (!f) return -1;
  for (int i = 0; i < mol->natoms; i++) {
    const int Z = mol->atomic_number[i];
    const char *sym = (Z >= 1 && Z <= 105) ? kSymbols[Z - 1] : "  ";
    char elem[3] = {sym[1] == ' ' ? ' ' : sym[0], sym[1] == ' ' ? sym[0] : sym[1], 0};
    fprintf(f, "ATOM  %5d  %.2s      A   1    %8.3f%8.3f%8.3f%22s%s\n", i % 99999, sym,
            (double)mol->pos[3 * i + 0] * 1e10, (double)mol->pos[3 * i + 1] * 1e10,
            (double)mol->pos[3 * i + 2] * 1e10, "", elem);
  }
  fprintf(f, "END\n");
  int bad = ferror(f);
  if (fclose(f) != 0) bad = 1;
  return bad ? -1 : 0;
}

void write_pdb_from_mol(char *filename, Molecule *mol) {
  if (spsim_write_pdb(filename, mol) != 0) {
    fprintf(stderr, "Cannot open file %s\n", filename ? filename : "(null)");
    abort();
  }
}
