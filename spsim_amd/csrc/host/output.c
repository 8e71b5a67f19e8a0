/*
 * Pattern output files (reference src/spsim.c:146-279 and src/io.c:63-93).
 *
 * output_array / output_carray write one image per file, "<basename>-%05d.cxi",
 * as the reference does through libspimage's sp_image_write. libspimage is not
 * available, so the exact layout it writes cannot be reproduced; the files
 * follow the CXI 1.5 conventions instead:
 *   /cxi_version                                  int 150
 *   /entry_1/image_1/data                         float [ny][nx] ([nz][ny][nx] for a
 *                                                 volume); complex images as the
 *                                                 compound {float r, i}
 *   /entry_1/image_1/mask                         int, 1 everywhere (as make_image sets it)
 *   /entry_1/image_1/orientation                  double [3][3], when the pattern has a rotation
 *   /entry_1/instrument_1/detector_1/distance     double, m
 *   /entry_1/instrument_1/detector_1/{x,y}_pixel_size   double, m (pixel size times
 *                                                 binning, as make_image records it;
 *                                                 z_pixel_size for a volume)
 *   /entry_1/instrument_1/source_1/{wavelength, energy} double, m and J
 *   /entry_1/image_1/detector_1, /entry_1/image_1/source_1  soft links to the above
 * HDF5 is loaded when the first file is written (the library named by
 * SPSIM_HDF5_LIBRARY, else libhdf5.so from the loader's search path or
 * /opt/conda/lib), so the library has no link-time dependency on it; without
 * HDF5 the writers report the fact and write nothing.
 *
 * output_files writes the arrays the Options flags select, with the
 * reference's base names. The binned outputs (counts_out, noiseless_counts_out)
 * are written with their binned dimensions; the reference wrote them as
 * nx x ny x nz images, reading past the binned arrays whenever binning > 1.
 * The real-space outputs need an FFT of the structure factor and are not
 * written (a message says so), nor is VTK output (output_sf_vtk): the
 * reference's VTK writer (src/io.c) is outside this library's scope.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spsim.h"

/* ---------------------------------------------------------------- HDF5 */

/* The HDF5 1.10 C ABI: identifiers are int64_t, sizes unsigned long long. */
typedef int64_t h5id;
typedef unsigned long long h5size;
#define H5_DEFAULT ((h5id)0) /* H5P_DEFAULT, H5S_ALL, H5E_DEFAULT */
#define H5_ACC_TRUNC 0x0002u
#define H5_SCALAR 0
#define H5_COMPOUND 6

typedef struct {
  int (*open)(void);
  int (*eset_auto)(h5id, void *, void *);
  h5id (*fcreate)(const char *, unsigned, h5id, h5id);
  int (*fclose)(h5id);
  h5id (*gcreate)(h5id, const char *, h5id, h5id, h5id);
  int (*gclose)(h5id);
  h5id (*screate_simple)(int, const h5size *, const h5size *);
  h5id (*screate)(int);
  int (*sclose)(h5id);
  h5id (*dcreate)(h5id, const char *, h5id, h5id, h5id, h5id, h5id);
  int (*dwrite)(h5id, h5id, h5id, h5id, h5id, const void *);
  int (*dclose)(h5id);
  h5id (*tcreate)(int, size_t);
  int (*tinsert)(h5id, const char *, size_t, h5id);
  int (*tclose)(h5id);
  int (*lsoft)(const char *, h5id, const char *, h5id, h5id);
  h5id native_float, native_double, native_int;
} h5api;

static h5api g_h5;
static int g_h5_state = 0; /* 0 untried, 1 loaded, -1 unavailable */

static int load_hdf5(void) {
  if (g_h5_state) return g_h5_state > 0;
  /* an explicit SPSIM_HDF5_LIBRARY is the only library tried */
  const char *env = getenv("SPSIM_HDF5_LIBRARY");
  const char *names[] = {"libhdf5.so", "libhdf5_serial.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
  void *h = NULL;
  if (env && env[0]) {
    h = dlopen(env, RTLD_NOW | RTLD_LOCAL);
  } else {
    for (size_t i = 0; i < sizeof names / sizeof names[0] && !h; ++i) h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
  }
  g_h5_state = -1;
  if (!h) {
    fprintf(stderr, "spsim: HDF5 (libhdf5.so) not found; set SPSIM_HDF5_LIBRARY to write .cxi files\n");
    return 0;
  }
#define SYM(field, name)                                              \
  do {                                                                \
    *(void **)(&g_h5.field) = dlsym(h, name);                         \
    if (!g_h5.field) {                                                \
      fprintf(stderr, "spsim: HDF5 library lacks %s\n", name);        \
      return 0;                                                       \
    }                                                                 \
  } while (0)
  SYM(open, "H5open");
  SYM(eset_auto, "H5Eset_auto2");
  SYM(fcreate, "H5Fcreate");
  SYM(fclose, "H5Fclose");
  SYM(gcreate, "H5Gcreate2");
  SYM(gclose, "H5Gclose");
  SYM(screate_simple, "H5Screate_simple");
  SYM(screate, "H5Screate");
  SYM(sclose, "H5Sclose");
  SYM(dcreate, "H5Dcreate2");
  SYM(dwrite, "H5Dwrite");
  SYM(dclose, "H5Dclose");
  SYM(tcreate, "H5Tcreate");
  SYM(tinsert, "H5Tinsert");
  SYM(tclose, "H5Tclose");
  SYM(lsoft, "H5Lcreate_soft");
#undef SYM
  if (g_h5.open() < 0) return 0;
  const h5id *f = (const h5id *)dlsym(h, "H5T_NATIVE_FLOAT_g");
  const h5id *d = (const h5id *)dlsym(h, "H5T_NATIVE_DOUBLE_g");
  const h5id *n = (const h5id *)dlsym(h, "H5T_NATIVE_INT_g");
  if (!f || !d || !n) {
    fprintf(stderr, "spsim: HDF5 library lacks the native type globals\n");
    return 0;
  }
  g_h5.native_float = *f;
  g_h5.native_double = *d;
  g_h5.native_int = *n;
  g_h5.eset_auto(H5_DEFAULT, NULL, NULL); /* errors are reported here, not on HDF5's stack printer */
  g_h5_state = 1;
  return 1;
}

static int write_dataset(h5id loc, const char *name, h5id type, int rank, const h5size *dims, const void *data) {
  h5id space = rank ? g_h5.screate_simple(rank, dims, NULL) : g_h5.screate(H5_SCALAR);
  if (space < 0) return -1;
  h5id ds = g_h5.dcreate(loc, name, type, space, H5_DEFAULT, H5_DEFAULT, H5_DEFAULT);
  int rc = -1;
  if (ds >= 0) {
    rc = g_h5.dwrite(ds, type, H5_DEFAULT, H5_DEFAULT, H5_DEFAULT, data) < 0 ? -1 : 0;
    g_h5.dclose(ds);
  }
  g_h5.sclose(space);
  return rc;
}

static int write_double(h5id loc, const char *name, double v) {
  return write_dataset(loc, name, g_h5.native_double, 0, NULL, &v);
}

typedef struct {
  int nx, ny, nz;
  double distance, pixel[3], wavelength;
  const SpRotation *rot;
} image_meta;

/* One CXI file holding a float (cplx == 0) or complex image of nx*ny*nz values. */
static int write_cxi(const char *filename, const void *values, int cplx, const image_meta *m) {
  if (!load_hdf5()) return -1;
  const size_t n = (size_t)m->nx * m->ny * m->nz;
  int *mask = (int *)malloc(sizeof(int) * (n ? n : 1));
  if (!mask) return -1;
  for (size_t i = 0; i < n; ++i) mask[i] = 1;
  int rc = -1;
  h5id file = g_h5.fcreate(filename, H5_ACC_TRUNC, H5_DEFAULT, H5_DEFAULT);
  if (file < 0) {
    free(mask);
    fprintf(stderr, "spsim: cannot create %s\n", filename);
    return -1;
  }
  h5id entry = -1, image = -1, instr = -1, det = -1, src = -1, ctype = -1;
  const int version = 150;
  h5size dims[3];
  int rank;
  if (m->nz > 1) {
    dims[0] = (h5size)m->nz;
    dims[1] = (h5size)m->ny;
    dims[2] = (h5size)m->nx;
    rank = 3;
  } else {
    dims[0] = (h5size)m->ny;
    dims[1] = (h5size)m->nx;
    rank = 2;
  }
  if (write_dataset(file, "cxi_version", g_h5.native_int, 0, NULL, &version) < 0) goto done;
  if ((entry = g_h5.gcreate(file, "entry_1", H5_DEFAULT, H5_DEFAULT, H5_DEFAULT)) < 0) goto done;
  if ((image = g_h5.gcreate(entry, "image_1", H5_DEFAULT, H5_DEFAULT, H5_DEFAULT)) < 0) goto done;
  if (cplx) {
    if ((ctype = g_h5.tcreate(H5_COMPOUND, 2 * sizeof(float))) < 0) goto done;
    if (g_h5.tinsert(ctype, "r", 0, g_h5.native_float) < 0 ||
        g_h5.tinsert(ctype, "i", sizeof(float), g_h5.native_float) < 0)
      goto done;
    if (write_dataset(image, "data", ctype, rank, dims, values) < 0) goto done;
  } else if (write_dataset(image, "data", g_h5.native_float, rank, dims, values) < 0) {
    goto done;
  }
  if (write_dataset(image, "mask", g_h5.native_int, rank, dims, mask) < 0) goto done;
  if (m->rot) {
    double R[9];
    const h5size rd[2] = {3, 3};
    for (int i = 0; i < 9; ++i) R[i] = m->rot->data[i];
    if (write_dataset(image, "orientation", g_h5.native_double, 2, rd, R) < 0) goto done;
  }
  if ((instr = g_h5.gcreate(entry, "instrument_1", H5_DEFAULT, H5_DEFAULT, H5_DEFAULT)) < 0) goto done;
  if ((det = g_h5.gcreate(instr, "detector_1", H5_DEFAULT, H5_DEFAULT, H5_DEFAULT)) < 0) goto done;
  if (write_double(det, "distance", m->distance) < 0 || write_double(det, "x_pixel_size", m->pixel[0]) < 0 ||
      write_double(det, "y_pixel_size", m->pixel[1]) < 0)
    goto done;
  if (m->nz > 1 && write_double(det, "z_pixel_size", m->pixel[2]) < 0) goto done;
  if ((src = g_h5.gcreate(instr, "source_1", H5_DEFAULT, H5_DEFAULT, H5_DEFAULT)) < 0) goto done;
  if (write_double(src, "wavelength", m->wavelength) < 0) goto done;
  if (m->wavelength > 0 && write_double(src, "energy", 1.98644521e-25 / m->wavelength) < 0) goto done;
  if (g_h5.lsoft("/entry_1/instrument_1/detector_1", image, "detector_1", H5_DEFAULT, H5_DEFAULT) < 0 ||
      g_h5.lsoft("/entry_1/instrument_1/source_1", image, "source_1", H5_DEFAULT, H5_DEFAULT) < 0)
    goto done;
  rc = 0;
done:
  if (ctype >= 0) g_h5.tclose(ctype);
  if (src >= 0) g_h5.gclose(src);
  if (det >= 0) g_h5.gclose(det);
  if (instr >= 0) g_h5.gclose(instr);
  if (image >= 0) g_h5.gclose(image);
  if (entry >= 0) g_h5.gclose(entry);
  if (g_h5.fclose(file) < 0) rc = -1;
  free(mask);
  if (rc) fprintf(stderr, "spsim: writing %s failed\n", filename);
  return rc;
}

/* Image metadata as make_image records it (src/spsim.c:166-174). */
static image_meta meta_for(const Options *opts, const SpRotation *rot, int binned) {
  const CCD *d = opts->detector;
  const int bx = d->binning_x > 0 ? d->binning_x : 1, by = d->binning_y > 0 ? d->binning_y : 1;
  const int bz = d->binning_z > 0 ? d->binning_z : 1, nz = d->nz > 0 ? d->nz : 1;
  image_meta m;
  m.nx = binned ? d->nx / bx : d->nx;
  m.ny = binned ? d->ny / by : d->ny;
  m.nz = binned ? nz / bz : nz;
  m.distance = d->distance;
  m.pixel[0] = (double)(d->pixel_width * d->binning_x);
  m.pixel[1] = (double)(d->pixel_height * d->binning_y);
  m.pixel[2] = (double)(d->pixel_depth * d->binning_z);
  m.wavelength = opts->experiment ? opts->experiment->wavelength : 0.0;
  m.rot = rot;
  return m;
}

static void write_named(const char *basename, int index, const void *array, int cplx, const SpRotation *rot,
                        const Options *opts, int binned) {
  char name[1100];
  if (!basename || !array || !opts || !opts->detector) return;
  snprintf(name, sizeof name, "%s-%05d.cxi", basename, index);
  const image_meta m = meta_for(opts, rot, binned);
  write_cxi(name, array, cplx, &m);
}

void output_array(char *basename, int index, float *array, SpRotation *rot, Options *opts) {
  write_named(basename, index, array, 0, rot, opts, 0);
}

void output_carray(char *basename, int index, Complex *array, SpRotation *rot, Options *opts) {
  write_named(basename, index, array, 1, rot, opts, 0);
}

void output_files(Diffraction_Pattern *pattern, Options *opts, int index) {
  if (!pattern || !opts || !opts->detector) return;
  CCD *d = opts->detector;
  const SpRotation *rot = pattern->rot;
  if (opts->output_scatt_int || opts->output_intensities)
    write_named("scattering_int", index, pattern->ints, 0, rot, opts, 0);
  if (opts->output_sf_vtk) fprintf(stderr, "spsim: VTK output (src/io.c) is not built; nothing written\n");
  if (opts->output_noiseless_photons)
    write_named("noiseless_photon_out", index, d->photons_per_pixel, 0, rot, opts, 0);
  if (opts->output_scattering_factors) write_named("scattering_factors", index, pattern->F, 1, rot, opts, 0);
  if (opts->output_solid_angles) write_named("solid_angles", index, d->solid_angle, 0, rot, opts, 0);
  if (opts->output_real_space || opts->output_realspace_histogram)
    fprintf(stderr, "spsim: real-space outputs need the FFT path and are not written\n");
  if (opts->output_photons) write_named("photon_out", index, d->photon_count, 0, rot, opts, 0);
  if (opts->output_count) write_named("counts_out", index, d->real_output, 0, rot, opts, 1);
  if (opts->output_noiseless_count) write_named("noiseless_counts_out", index, d->noiseless_output, 0, rot, opts, 1);
}
