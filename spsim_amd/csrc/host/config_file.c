/*
 * spsim.conf reader and writer (reference src/config.c:89-426 and :428-642).
 *
 * The file format is libconfig's: settings `name = value;` (`:` works as well
 * as `=`, and `;` or `,` may end a setting), groups `{ ... }`, arrays
 * `[ ... ]`, lists `( ... )`, comments `#`, `//` and C-style block comments,
 * adjacent string literals concatenated, integers in decimal or hex with an
 * optional L suffix, floats, booleans and `@include "file"`. libconfig itself
 * is not a dependency: the parser below implements that grammar.
 *
 * The Options fields are filled exactly as the reference does it (same keys,
 * same enum strings, `detector_binning` before the per-axis keys, the
 * chemical formula read from `chemical_formula/atomZ` for Z = 0..99 with zero
 * quantities skipped, then the derived fields of src/config.c:386-421).
 * Values keep libconfig's typing without auto-conversion, as the reference's
 * manual documents (doc/UserManual.tex:91-93): a float key written without a
 * decimal point or exponent (`detector_readout_noise = 10;`) reads as 0, and
 * an integer key given a float or a boolean reads as 0. Each such case is
 * reported on stderr, since the silent 0 is a common mistake. One deliberate
 * difference: the writer spells the real-space blurring key
 * `real_space_blurring`, the key the reader looks up (the reference writes
 * `real_sapce_blurring`, src/config.c:545, so the value was lost on a round
 * trip).
 * use_cuda is kept as read: this library has no CPU path, so the reference's
 * "no CUDA device" fallback (src/config.c:335-356) does not apply.
 */
#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include "config.h"

/* ------------------------------------------------------------------ tree */

typedef enum { CFG_GROUP, CFG_INT, CFG_INT64, CFG_FLOAT, CFG_STRING, CFG_BOOL, CFG_ARRAY, CFG_LIST } cfg_type;

typedef struct cfg_node {
  char *name; /* NULL for elements of arrays and lists */
  cfg_type type;
  long long ival;
  double fval;
  char *sval;
  struct cfg_node *child, *last, *next;
} cfg_node;

static void node_free(cfg_node *n) {
  while (n) {
    cfg_node *next = n->next;
    node_free(n->child);
    free(n->name);
    free(n->sval);
    free(n);
    n = next;
  }
}

static cfg_node *node_new(cfg_node *parent, char *name, cfg_type type) {
  cfg_node *n = (cfg_node *)calloc(1, sizeof(cfg_node));
  if (!n) return NULL;
  n->name = name;
  n->type = type;
  if (parent->last)
    parent->last->next = n;
  else
    parent->child = n;
  parent->last = n;
  return n;
}

static cfg_node *member(const cfg_node *group, const char *name, size_t len) {
  if (!group || group->type != CFG_GROUP) return NULL;
  for (cfg_node *c = group->child; c; c = c->next)
    if (c->name && strlen(c->name) == len && strncmp(c->name, name, len) == 0) return c;
  return NULL;
}

/* Path lookup with libconfig's separators '.', '/' and ':'. */
static const cfg_node *lookup(const cfg_node *root, const char *path) {
  const cfg_node *n = root;
  while (*path) {
    size_t len = strcspn(path, "./:");
    if (len) {
      n = member(n, path, len);
      if (!n) return NULL;
    }
    path += len;
    if (*path) path++;
  }
  return n == root ? NULL : n;
}

/* ---------------------------------------------------------------- parser */

#define CFG_MAX_INCLUDE_DEPTH 10

typedef struct {
  const char *p, *end;
  const char *file;
  int line;
  int depth;
  char *err;
  int errlen;
  int err_line;
} parser;

static int fail_at(parser *P, const char *msg) {
  if (P->err && P->errlen > 0 && P->err_line == 0) {
    snprintf(P->err, (size_t)P->errlen, "%s", msg);
  }
  if (P->err_line == 0) P->err_line = P->line > 0 ? P->line : 1;
  return -1;
}

static int skip_space(parser *P) {
  while (P->p < P->end) {
    char c = *P->p;
    if (c == '\n') {
      P->line++;
      P->p++;
    } else if (isspace((unsigned char)c)) {
      P->p++;
    } else if (c == '#' || (c == '/' && P->p + 1 < P->end && P->p[1] == '/')) {
      while (P->p < P->end && *P->p != '\n') P->p++;
    } else if (c == '/' && P->p + 1 < P->end && P->p[1] == '*') {
      P->p += 2;
      for (;;) {
        if (P->p + 1 >= P->end) {
          if (P->p < P->end && *P->p == '\n') P->line++;
          P->p = P->end;
          return fail_at(P, "unterminated comment");
        }
        if (*P->p == '\n') P->line++;
        if (P->p[0] == '*' && P->p[1] == '/') {
          P->p += 2;
          break;
        }
        P->p++;
      }
    } else {
      break;
    }
  }
  return 0;
}

static int is_name_start(char c) { return isalpha((unsigned char)c) || c == '*'; }
static int is_name_char(char c) { return isalnum((unsigned char)c) || c == '*' || c == '_' || c == '-'; }

typedef struct {
  char *s;
  size_t len, cap;
} strbuf;

static int sb_put(strbuf *b, char c) {
  if (b->len + 2 > b->cap) {
    size_t cap = b->cap ? 2 * b->cap : 64;
    char *s = (char *)realloc(b->s, cap);
    if (!s) return -1;
    b->s = s;
    b->cap = cap;
  }
  b->s[b->len++] = c;
  b->s[b->len] = 0;
  return 0;
}

static int hexval(char c) {
  if (c >= '0' && c <= '9') return c - '0';
  if (c >= 'a' && c <= 'f') return c - 'a' + 10;
  if (c >= 'A' && c <= 'F') return c - 'A' + 10;
  return -1;
}

/* One or more adjacent string literals, concatenated. */
static int parse_string(parser *P, char **out) {
  strbuf b = {0};
  if (sb_put(&b, 0) != 0) return fail_at(P, "out of memory");
  b.len = 0;
  while (P->p < P->end && *P->p == '"') {
    P->p++;
    for (;;) {
      if (P->p >= P->end || *P->p == '\n') {
        free(b.s);
        return fail_at(P, "unterminated string");
      }
      char c = *P->p++;
      if (c == '"') break;
      if (c == '\\') {
        if (P->p >= P->end) break;
        char e = *P->p++;
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'f': c = '\f'; break;
          case '\\': c = '\\'; break;
          case '"': c = '"'; break;
          case 'x': {
            int h1 = P->p < P->end ? hexval(P->p[0]) : -1;
            int h2 = P->p + 1 < P->end ? hexval(P->p[1]) : -1;
            if (h1 < 0 || h2 < 0) {
              free(b.s);
              return fail_at(P, "invalid \\x escape in string");
            }
            c = (char)(h1 * 16 + h2);
            P->p += 2;
            break;
          }
          default:
            free(b.s);
            return fail_at(P, "invalid escape sequence in string");
        }
      }
      if (sb_put(&b, c) != 0) {
        free(b.s);
        return fail_at(P, "out of memory");
      }
    }
    if (skip_space(P) != 0) {
      free(b.s);
      return -1;
    }
  }
  *out = b.s;
  return 0;
}

static int parse_scalar_token(parser *P, cfg_node *n) {
  const char *start = P->p;
  while (P->p < P->end && (isalnum((unsigned char)*P->p) || *P->p == '.' || *P->p == '+' || *P->p == '-')) P->p++;
  size_t len = (size_t)(P->p - start);
  if (len == 0 || len > 127) return fail_at(P, "syntax error");
  char tok[128];
  memcpy(tok, start, len);
  tok[len] = 0;
  if (strcasecmp(tok, "true") == 0 || strcasecmp(tok, "false") == 0) {
    n->type = CFG_BOOL;
    n->ival = tolower((unsigned char)tok[0]) == 't';
    return 0;
  }
  const char *digits = tok + ((tok[0] == '+' || tok[0] == '-') ? 1 : 0);
  int hex = digits[0] == '0' && (digits[1] == 'x' || digits[1] == 'X');
  int is_float = !hex && strpbrk(tok, ".eE") != NULL;
  if (!isdigit((unsigned char)digits[0]) && !(is_float && digits[0] == '.')) return fail_at(P, "syntax error");
  char *endp = NULL;
  errno = 0;
  if (is_float) {
    n->type = CFG_FLOAT;
    n->fval = strtod(tok, &endp);
    if (*endp) return fail_at(P, "invalid floating point value");
    return 0;
  }
  size_t tl = strlen(tok);
  int suffix = 0;
  while (tl > 0 && (tok[tl - 1] == 'L' || tok[tl - 1] == 'l') && suffix < 2) {
    tok[--tl] = 0;
    suffix++;
  }
  long long v;
  if (hex) {
    unsigned long long u = strtoull(tok + ((tok[0] == '+' || tok[0] == '-') ? 1 : 0), &endp, 16);
    v = (long long)u;
    if (tok[0] == '-') v = -v;
  } else {
    v = strtoll(tok, &endp, 10);
  }
  if (*endp || errno == ERANGE) return fail_at(P, "invalid integer value");
  n->ival = v;
  n->type = (suffix || v > INT_MAX || v < INT_MIN) ? CFG_INT64 : CFG_INT;
  return 0;
}

static int parse_settings(parser *P, cfg_node *group, char close);

static int parse_value(parser *P, cfg_node *n) {
  if (P->p >= P->end) return fail_at(P, "unexpected end of file");
  char c = *P->p;
  if (c == '{') {
    P->p++;
    n->type = CFG_GROUP;
    return parse_settings(P, n, '}');
  }
  if (c == '[' || c == '(') {
    char close = c == '[' ? ']' : ')';
    n->type = c == '[' ? CFG_ARRAY : CFG_LIST;
    P->p++;
    if (skip_space(P) != 0) return -1;
    if (P->p < P->end && *P->p == close) {
      P->p++;
      return 0;
    }
    for (;;) {
      cfg_node *e = node_new(n, NULL, CFG_INT);
      if (!e) return fail_at(P, "out of memory");
      if (skip_space(P) != 0 || parse_value(P, e) != 0) return -1;
      if (n->type == CFG_ARRAY && (e->type == CFG_GROUP || e->type == CFG_ARRAY || e->type == CFG_LIST))
        return fail_at(P, "arrays hold scalar values only");
      if (skip_space(P) != 0) return -1;
      if (P->p < P->end && *P->p == ',') {
        P->p++;
        if (skip_space(P) != 0) return -1;
        continue;
      }
      if (P->p < P->end && *P->p == close) {
        P->p++;
        return 0;
      }
      return fail_at(P, c == '[' ? "expected ',' or ']'" : "expected ',' or ')'");
    }
  }
  if (c == '"') {
    n->type = CFG_STRING;
    return parse_string(P, &n->sval);
  }
  return parse_scalar_token(P, n);
}

static int read_text(const char *filename, char **text, size_t *len);

static int parse_include(parser *P, cfg_node *group) {
  static const char kw[] = "@include";
  if ((size_t)(P->end - P->p) < sizeof(kw) - 1 || strncmp(P->p, kw, sizeof(kw) - 1) != 0)
    return fail_at(P, "syntax error");
  P->p += sizeof(kw) - 1;
  if (skip_space(P) != 0) return -1;
  if (P->p >= P->end || *P->p != '"') return fail_at(P, "expected a file name after @include");
  char *path = NULL;
  if (parse_string(P, &path) != 0) return -1;
  if (P->depth >= CFG_MAX_INCLUDE_DEPTH) {
    free(path);
    return fail_at(P, "include nesting too deep");
  }
  char *text = NULL;
  size_t len = 0;
  if (read_text(path, &text, &len) != 0) {
    free(path);
    return fail_at(P, "cannot read included file");
  }
  int include_line = P->line;
  parser Q = *P;
  Q.p = text;
  Q.end = text + len;
  Q.file = path;
  Q.line = 1;
  Q.depth = P->depth + 1;
  int rc = parse_settings(&Q, group, 0);
  if (rc != 0) {
    P->err_line = include_line;
    if (P->err && P->errlen > 0) {
      size_t used = strlen(P->err);
      snprintf(P->err + used, (size_t)P->errlen - used, " (in %s line %d)", path, Q.err_line);
    }
  }
  free(text);
  free(path);
  return rc;
}

/* Settings until `close` (0: end of input). */
static int parse_settings(parser *P, cfg_node *group, char close) {
  for (;;) {
    if (skip_space(P) != 0) return -1;
    if (P->p >= P->end) {
      if (close) return fail_at(P, "unexpected end of file");
      return 0;
    }
    if (close && *P->p == close) {
      P->p++;
      return 0;
    }
    if (*P->p == '@') {
      if (parse_include(P, group) != 0) return -1;
      continue;
    }
    if (!is_name_start(*P->p)) return fail_at(P, "syntax error");
    const char *start = P->p;
    while (P->p < P->end && is_name_char(*P->p)) P->p++;
    size_t len = (size_t)(P->p - start);
    if (member(group, start, len)) return fail_at(P, "duplicate setting name");
    if (skip_space(P) != 0) return -1;
    if (P->p >= P->end || (*P->p != '=' && *P->p != ':')) return fail_at(P, "expected '=' or ':'");
    P->p++;
    if (skip_space(P) != 0) return -1;
    char *name = (char *)malloc(len + 1);
    if (!name) return fail_at(P, "out of memory");
    memcpy(name, start, len);
    name[len] = 0;
    cfg_node *n = node_new(group, name, CFG_INT);
    if (!n) {
      free(name);
      return fail_at(P, "out of memory");
    }
    if (parse_value(P, n) != 0) return -1;
    if (skip_space(P) != 0) return -1;
    if (P->p < P->end && (*P->p == ';' || *P->p == ',')) P->p++;
  }
}

static int read_text(const char *filename, char **text, size_t *len) {
  FILE *f = fopen(filename, "rb");
  if (!f) return -1;
  strbuf b = {0};
  char chunk[4096];
  size_t got;
  int rc = 0;
  while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) {
    for (size_t i = 0; i < got && rc == 0; i++) rc = sb_put(&b, chunk[i]);
    if (rc) break;
  }
  if (ferror(f)) rc = -1;
  fclose(f);
  if (rc) {
    free(b.s);
    return -1;
  }
  if (!b.s) {
    b.s = (char *)calloc(1, 1);
    if (!b.s) return -1;
  }
  *text = b.s;
  *len = b.len;
  return 0;
}

/* ---------------------------------------------------- typed lookups */

/* libconfig's float lookup: only a float setting has a value. */
static double as_float(const cfg_node *n, const char *key) {
  if (n->type == CFG_FLOAT) return n->fval;
  if (n->type == CFG_INT || n->type == CFG_INT64)
    fprintf(stderr, "Warning: %s has no decimal point and reads as 0 (write %lld.0)\n", key, n->ival);
  else
    fprintf(stderr, "Warning: %s is not a float and reads as 0\n", key);
  return 0.0;
}

/* libconfig's int lookup: int settings, and 64-bit ones that fit in an int. */
static int as_int(const cfg_node *n, const char *key) {
  if (n->type == CFG_INT) return (int)n->ival;
  if (n->type == CFG_INT64 && n->ival <= INT_MAX && n->ival >= INT_MIN) return (int)n->ival;
  fprintf(stderr, "Warning: %s is not an integer (or does not fit in one) and reads as 0\n", key);
  return 0;
}

static const char *as_string(const cfg_node *n) { return n && n->type == CFG_STRING ? n->sval : NULL; }

/* ------------------------------------------------------- key table */

enum { IN_OPTS, IN_DET, IN_EXP };
enum { K_FLOAT, K_INT };

typedef struct {
  const char *key;
  int kind;
  int where;
  size_t off;
} option_key;

#define OPT(k, kind, field) {k, kind, IN_OPTS, offsetof(Options, field)}
#define DET(k, kind, field) {k, kind, IN_DET, offsetof(CCD, field)}
#define EXP(k, kind, field) {k, kind, IN_EXP, offsetof(Experiment, field)}

/* Numeric keys in the order the writer emits them (src/config.c:470-631).
 * The string-valued keys (input_type, box_type, experiment_polarization,
 * pdb_filename, precalculated_sf, hkl_grid) and detector_binning are handled
 * separately. */
static const option_key numeric_keys[] = {
    DET("detector_distance", K_FLOAT, distance),
    DET("detector_width", K_FLOAT, width),
    DET("detector_height", K_FLOAT, height),
    DET("detector_depth", K_FLOAT, depth),
    DET("detector_center_x", K_FLOAT, center_x),
    DET("detector_center_y", K_FLOAT, center_y),
    DET("detector_center_z", K_FLOAT, center_z),
    DET("detector_pixel_width", K_FLOAT, pixel_width),
    DET("detector_pixel_height", K_FLOAT, pixel_height),
    DET("detector_pixel_depth", K_FLOAT, pixel_depth),
    DET("detector_quantum_efficiency", K_FLOAT, quantum_efficiency),
    DET("detector_electron_hole_production_energy", K_FLOAT, electron_hole_production_energy),
    DET("detector_readout_noise", K_FLOAT, readout_noise),
    DET("detector_dark_current", K_FLOAT, dark_current),
    DET("detector_linear_full_well", K_FLOAT, linear_full_well),
    DET("detector_binning_x", K_INT, binning_x),
    DET("detector_binning_y", K_INT, binning_y),
    DET("detector_binning_z", K_INT, binning_z),
    DET("detector_maximum_value", K_FLOAT, maximum_value),
    EXP("experiment_wavelength", K_FLOAT, wavelength),
    EXP("experiment_bandwidth", K_FLOAT, bandwidth),
    EXP("experiment_exposure_time", K_FLOAT, exposure_time),
    EXP("experiment_beam_intensity", K_FLOAT, beam_intensity),
    EXP("experiment_beam_center_x", K_FLOAT, beam_center_x),
    EXP("experiment_beam_center_y", K_FLOAT, beam_center_y),
    EXP("experiment_beam_fwhm", K_FLOAT, beam_fwhm),
    EXP("experiment_beam_energy", K_FLOAT, beam_energy),
    EXP("experiment_photon_energy", K_FLOAT, photon_energy),
    EXP("experiment_focal_diameter", K_FLOAT, focal_diameter),
    OPT("box_dimension", K_FLOAT, box_dimension),
    DET("detector_spherical", K_INT, spherical),
    DET("detector_gaussian_blurring", K_FLOAT, gaussian_blurring),
    DET("real_space_blurring", K_FLOAT, real_space_blurring),
    OPT("use_fft_for_sf", K_INT, use_fft_for_sf),
    OPT("use_nfft_for_sf", K_INT, use_nfft_for_sf),
    OPT("b_factor", K_FLOAT, b_factor),
    OPT("origin_to_com", K_INT, origin_to_com),
    OPT("phi", K_FLOAT, euler_orientation[0]),
    OPT("theta", K_FLOAT, euler_orientation[1]),
    OPT("psi", K_FLOAT, euler_orientation[2]),
    OPT("random_orientation", K_INT, random_orientation),
    OPT("number_of_patterns", K_INT, n_patterns),
    OPT("vectorize", K_INT, vectorize),
    OPT("delta_atoms", K_INT, delta_atoms),
    OPT("fast_exit", K_INT, fast_exit),
    OPT("crystal_size_a", K_INT, crystal_size[0]),
    OPT("crystal_size_b", K_INT, crystal_size[1]),
    OPT("crystal_size_c", K_INT, crystal_size[2]),
    OPT("crystal_cell_a", K_FLOAT, crystal_cell[0]),
    OPT("crystal_cell_b", K_FLOAT, crystal_cell[1]),
    OPT("crystal_cell_c", K_FLOAT, crystal_cell[2]),
    OPT("crystal_cell_alpha", K_FLOAT, crystal_cell[3]),
    OPT("crystal_cell_beta", K_FLOAT, crystal_cell[4]),
    OPT("crystal_cell_gamma", K_FLOAT, crystal_cell[5]),
    /* the writer emits experiment_polarization before use_cuda */
    OPT("use_cuda", K_INT, use_cuda),
    OPT("wavelength_samples", K_INT, wavelength_samples),
    OPT("random_seed", K_INT, random_seed),
    OPT("output_photons", K_INT, output_photons),
    OPT("output_noiseless_photons", K_INT, output_noiseless_photons),
    OPT("output_count", K_INT, output_count),
    OPT("output_noiseless_count", K_INT, output_noiseless_count),
    OPT("output_solid_angles", K_INT, output_solid_angles),
    OPT("output_scattering_factors", K_INT, output_scattering_factors),
    OPT("output_real_space", K_INT, output_real_space),
    OPT("verbosity_level", K_INT, verbosity_level),
};
#define N_NUMERIC_KEYS (sizeof numeric_keys / sizeof numeric_keys[0])

static void *field_ptr(const option_key *k, Options *o) {
  char *base = k->where == IN_OPTS ? (char *)o : k->where == IN_DET ? (char *)o->detector : (char *)o->experiment;
  return base + k->off;
}

static const void *field_cptr(const option_key *k, const Options *o) {
  const char *base = k->where == IN_OPTS  ? (const char *)o
                     : k->where == IN_DET ? (const char *)o->detector
                                          : (const char *)o->experiment;
  return base + k->off;
}

static const char *const polarization_names[] = {"ignore", "vertical", "horizontal", "unpolarized"};

/* ---------------------------------------------------------- reading */

static int copy_path(char *dst, const char *src) {
  size_t n = strlen(src);
  if (n >= 1024) return -1;
  memcpy(dst, src, n + 1);
  return 0;
}

static int read_chem_formula(const cfg_node *root, Chem_Formula *cf) {
  int count = 0;
  int Z[100], qty[100];
  char path[64];
  for (int i = 0; i < 100; i++) {
    snprintf(path, sizeof path, "chemical_formula/atom%d", i);
    const cfg_node *n = lookup(root, path);
    int q = n ? as_int(n, path) : 0;
    if (q) {
      Z[count] = i;
      qty[count] = q;
      count++;
    }
  }
  free(cf->atomic_number);
  free(cf->quantity);
  cf->atomic_number = NULL;
  cf->quantity = NULL;
  cf->n_elements = 0;
  if (!count) return 0;
  cf->atomic_number = (int *)malloc(sizeof(int) * (size_t)count);
  cf->quantity = (int *)malloc(sizeof(int) * (size_t)count);
  if (!cf->atomic_number || !cf->quantity) {
    free(cf->atomic_number);
    free(cf->quantity);
    cf->atomic_number = cf->quantity = NULL;
    return -1;
  }
  memcpy(cf->atomic_number, Z, sizeof(int) * (size_t)count);
  memcpy(cf->quantity, qty, sizeof(int) * (size_t)count);
  cf->n_elements = count;
  return 0;
}

/* rint of the float quotient, as the reference computes nx/ny/nz; a quotient
 * that is not a finite int (a zero pixel size) gives 0 instead of the
 * reference's undefined conversion. */
static int pixel_count(float extent, float pixel) {
  float r = extent / pixel;
  double v = rint((double)r);
  return (isfinite(v) && v <= (double)INT_MAX && v >= (double)INT_MIN) ? (int)v : 0;
}

/* Derived fields, src/config.c:386-421, with the reference's float arithmetic. */
static void derive_fields(Options *res) {
  CCD *det = res->detector;
  Experiment *exp = res->experiment;
  det->nx = pixel_count(det->width, det->pixel_width);
  det->ny = pixel_count(det->height, det->pixel_height);
  if (det->pixel_depth) {
    det->nz = pixel_count(det->depth, det->pixel_depth);
  } else {
    det->nz = 1;
    det->binning_z = 1;
  }
  if (det->nz > 1 && exp->polarization != POLARIZATION_IGNORE) {
    fprintf(stderr,
            "Warning: For a 3D detector polarization effects (Thomson correction) have to be neglected "
            "(polarization = \"ignore\")!\n");
    exp->polarization = POLARIZATION_IGNORE;
  }
  if (exp->photon_energy) {
    float lambda = (float)(1.240e-6 / exp->photon_energy);
    if (exp->wavelength) {
      if ((lambda - exp->wavelength) / (lambda + exp->wavelength) > 0.01)
        fprintf(stderr, "Warning: experiment_wavelength does not agree with experiment_photon_energy!\n");
    } else {
      exp->wavelength = lambda;
    }
  } else if (exp->wavelength) {
    exp->photon_energy = (float)(1.240e-6 / exp->wavelength);
  } else {
    fprintf(stderr, "Warning: you need to specify the experiment_wavelength!\n");
  }
  if (!exp->beam_intensity && exp->focal_diameter && exp->beam_energy) {
    float eV = (float)(1.240e-6 / exp->wavelength);
    float nphotons = (float)(exp->beam_energy * 6.24150974e18 / eV);
    float area = exp->focal_diameter * exp->focal_diameter;
    exp->beam_intensity = nphotons / area;
  }
}

static int apply_tree(const cfg_node *root, Options *res, char *err, int errlen) {
  const char *s;
  const cfg_node *n;
  if ((s = as_string(lookup(root, "box_type")))) {
    if (strcmp(s, "spherical") == 0)
      res->box_type = BOX_SPHERICAL;
    else if (strcmp(s, "parallelepipedic") == 0)
      res->box_type = BOX_PARALLEL;
  }
  if ((s = as_string(lookup(root, "input_type")))) {
    if (strcmp(s, "chemical_formula") == 0) {
      res->input_type = CHEM_FORMULA;
      if (read_chem_formula(root, res->chem_formula) != 0) {
        snprintf(err, (size_t)errlen, "out of memory");
        return -1;
      }
    } else if (strcmp(s, "pdb") == 0) {
      res->input_type = PDB;
      const char *f = as_string(lookup(root, "pdb_filename"));
      if (!f) {
        snprintf(err, (size_t)errlen,
                 "Error: You need to specify \"pdb_filename\" when you use input_type = \"pdb\"");
        return -2;
      }
      if (copy_path(res->pdb_filename, f) != 0) {
        snprintf(err, (size_t)errlen, "pdb_filename longer than 1023 characters");
        return -1;
      }
    }
  }
  if ((n = lookup(root, "detector_binning"))) {
    res->detector->binning_x = as_int(n, "detector_binning");
    res->detector->binning_y = res->detector->binning_x;
    res->detector->binning_z = 1;
  }
  for (size_t i = 0; i < N_NUMERIC_KEYS; i++) {
    const option_key *k = &numeric_keys[i];
    if (!(n = lookup(root, k->key))) continue;
    if (k->kind == K_FLOAT)
      *(float *)field_ptr(k, res) = (float)as_float(n, k->key);
    else
      *(int *)field_ptr(k, res) = as_int(n, k->key);
  }
  if ((s = as_string(lookup(root, "experiment_polarization")))) {
    int found = 0;
    for (int p = 0; p < 4; p++)
      if (strcmp(s, polarization_names[p]) == 0) {
        res->experiment->polarization = p;
        found = 1;
      }
    if (!found) fprintf(stderr, "Warning: polarization is set to an undefined option!\n");
  }
  if ((s = as_string(lookup(root, "precalculated_sf"))) && copy_path(res->sf_filename, s) != 0) {
    snprintf(err, (size_t)errlen, "precalculated_sf longer than 1023 characters");
    return -1;
  }
  if ((s = as_string(lookup(root, "hkl_grid"))) && copy_path(res->hkl_grid_filename, s) != 0) {
    snprintf(err, (size_t)errlen, "hkl_grid longer than 1023 characters");
    return -1;
  }
  derive_fields(res);
  return 0;
}

/* Status-returning reader. 0: read; -1: file or syntax error, or a value the
 * Options fields cannot hold; -2: input_type "pdb" without pdb_filename.
 * err (may be NULL) receives "<file>:<line>: <message>" on failure. */
int spsim_read_options_file(const char *filename, Options *opt, char *err, int errlen) {
  char dummy[8];
  if (!err || errlen <= 0) {
    err = dummy;
    errlen = (int)sizeof dummy;
  }
  err[0] = 0;
  if (!filename || !opt || !opt->detector || !opt->experiment || !opt->chem_formula) {
    snprintf(err, (size_t)errlen, "null argument");
    return -1;
  }
  char *text = NULL;
  size_t len = 0;
  if (read_text(filename, &text, &len) != 0) {
    snprintf(err, (size_t)errlen, "%s:0: file I/O error", filename);
    return -1;
  }
  char msg[512] = {0};
  parser P = {text, text + len, filename, 1, 0, msg, (int)sizeof msg, 0};
  cfg_node root = {0};
  root.type = CFG_GROUP;
  int rc = parse_settings(&P, &root, 0);
  free(text);
  if (rc != 0) {
    snprintf(err, (size_t)errlen, "%s:%d: %s", filename, P.err_line, msg);
    node_free(root.child);
    return -1;
  }
  rc = apply_tree(&root, opt, msg, (int)sizeof msg);
  node_free(root.child);
  if (rc != 0) snprintf(err, (size_t)errlen, "%s", msg);
  return rc;
}

/* src/config.c:89-426: exits on a syntax error (status 1) and on input_type
 * "pdb" without pdb_filename (status 0), as the reference does. */
void read_options_file(char *filename, Options *res) {
  char err[640];
  int rc = spsim_read_options_file(filename, res, err, (int)sizeof err);
  if (rc == -2) {
    fprintf(stderr, "%s\n", err);
    exit(0);
  }
  if (rc != 0) {
    fprintf(stderr, "Error parsing %s: %s\n", filename ? filename : "(null)", err);
    exit(1);
  }
}

/* ---------------------------------------------------------- writing */

static void write_string(FILE *f, const char *key, const char *s) {
  fprintf(f, "%s = \"", key);
  for (; *s; s++) {
    unsigned char c = (unsigned char)*s;
    if (c == '"' || c == '\\')
      fprintf(f, "\\%c", c);
    else if (c == '\n')
      fputs("\\n", f);
    else if (c == '\t')
      fputs("\\t", f);
    else if (c == '\r')
      fputs("\\r", f);
    else if (c == '\f')
      fputs("\\f", f);
    else if (c < 0x20)
      fprintf(f, "\\x%02x", c);
    else
      fputc(c, f);
  }
  fputs("\";\n", f);
}

/* Floats as %.8e: nine significant digits restore every float exactly, and
 * the exponent keeps the value a float literal for libconfig. */
static void write_key(FILE *f, const option_key *k, const Options *o) {
  if (k->kind == K_FLOAT)
    fprintf(f, "%s = %.8e;\n", k->key, (double)*(const float *)field_cptr(k, o));
  else
    fprintf(f, "%s = %d;\n", k->key, *(const int *)field_cptr(k, o));
}

/* Writes the keys of src/config.c:428-642 in the same order. 0 on success,
 * -1 when the file cannot be written. */
int spsim_write_options_file(const char *filename, const Options *res) {
  if (!filename || !res || !res->detector || !res->experiment) return -1;
  FILE *f = fopen(filename, "w");
  if (!f) return -1;
  if (res->input_type == CHEM_FORMULA) {
    write_string(f, "input_type", "chemical_formula");
    fputs("chemical_formula = {\n", f);
    const Chem_Formula *cf = res->chem_formula;
    for (int i = 0; cf && i < cf->n_elements; i++)
      fprintf(f, "  atom%d = %d;\n", cf->atomic_number[i], cf->quantity[i]);
    fputs("};\n", f);
  } else if (res->input_type == PDB) {
    write_string(f, "input_type", "pdb");
    write_string(f, "pdb_filename", res->pdb_filename);
  } else if (!res->input_type) {
    write_string(f, "input_type", "none");
  }
  if (res->box_type == BOX_SPHERICAL)
    write_string(f, "box_type", "spherical");
  else if (res->box_type == BOX_PARALLEL)
    write_string(f, "box_type", "parallelepipedic");
  const int pol = res->experiment->polarization;
  for (size_t i = 0; i < N_NUMERIC_KEYS; i++) {
    if (strcmp(numeric_keys[i].key, "use_cuda") == 0 && pol >= 0 && pol < 4)
      write_string(f, "experiment_polarization", polarization_names[pol]);
    write_key(f, &numeric_keys[i], res);
  }
  if (res->sf_filename[0]) write_string(f, "precalculated_sf", res->sf_filename);
  int bad = ferror(f);
  if (fclose(f) != 0) bad = 1;
  return bad ? -1 : 0;
}

/* src/config.c:428-642 (the reference ignores write failures; this reports
 * them on stderr). */
void write_options_file(char *filename, Options *res) {
  if (spsim_write_options_file(filename, res) != 0)
    fprintf(stderr, "spsim: could not write %s\n", filename ? filename : "(null)");
}
