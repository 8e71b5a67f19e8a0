/*
 * spsim-mi355x C-ABI: device-resident entry points of the structure-factor hot
 * path on AMD MI355X (gfx950).
 *
 * These have no counterpart in the reference's public headers; each one
 * replaces the inner part of a reference function so that callers that keep
 * data in HBM (PyTorch, the batched pipeline, the benchmark) can skip the
 * host round trip the reference API implies:
 *
 *   spsim_hip_hkl_detector      <- get_HKL_list_for_detector   src/diffraction.c:237-297
 *   spsim_hip_rotate_hkl        <- apply_orientation_to_HKL_list src/diffraction.c:300-323
 *   spsim_hip_sf_pattern        <- compute_pattern_on_list      src/diffraction.c:770-853
 *                                  (and cuda_compute_pattern_on_list2, src/diffraction.cu:235-339)
 *   spsim_hip_batch_*           <- the per-shot loop of main/simulate_shot, src/spsim.c:392-396,
 *                                  run as one persistent launch over many molecules/orientations
 *   spsim_hip_detector_chain    <- calculate_thomson_correction .. calculate_noiseless_detector_output
 *                                  src/diffraction.c:1109-1204, src/noise.c:32-51,
 *                                  src/amplification.c:54-153 (fused)
 *   spsim_hip_detector_chain_binned <- the same chain when only the binned outputs
 *                                  (src/amplification.c:77-153) are kept
 *   spsim_hip_sf_nfft           <- compute_pattern_on_list_by_nfft src/diffraction.c:528-675 (NNFFT path)
 *   spsim_hip_sf_adjoint        <- (no reference counterpart) gradients of a Poisson
 *                                  negative log-likelihood w.r.t. atom positions and B factors
 *
 * Conventions: all pointers named d_* are device pointers (hipMalloc or
 * torch.cuda memory) on the current HIP device; `stream` is a hipStream_t
 * passed as void* (NULL = the default stream). Every function returns
 * SPSIM_HIP_OK or a negative error code and never exits the process;
 * spsim_hip_last_error() describes the last failure on the calling thread.
 * Calls are stream-ordered and asynchronous unless stated otherwise. Entry
 * points that use "the library workspace" share one device buffer per GPU:
 * such calls on different streams must not overlap (issue them on one stream,
 * or synchronise between them); growing the buffer synchronises the device.
 */
#ifndef SPSIM_HIP_H
#define SPSIM_HIP_H 1

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPSIM_HIP_OK 0
#define SPSIM_HIP_ERR_INVALID (-1)  /* bad argument (null pointer, negative size, bad shape) */
#define SPSIM_HIP_ERR_HIP (-2)      /* a HIP runtime call failed */
#define SPSIM_HIP_ERR_NOMEM (-3)    /* host or device allocation failed */
#define SPSIM_HIP_ERR_NODEVICE (-4) /* no usable gfx950 device */

/* Species index of a packed atom: 0..99 = Cromer-Mann row (atomic number Z),
   100 = unit point scatterer (f = 1, no B factor), used for delta_atoms. */
#define SPSIM_HIP_NSPECIES 101
#define SPSIM_HIP_UNIT_SPECIES 100
#define SPSIM_HIP_FF_STRIDE 9 /* a1..a4, b1..b4, c  (b in A^2) */

/* How raw (Z, pos) atoms are turned into packed records. */
#define SPSIM_HIP_PACK_PLAIN 0  /* compute_pattern_on_list: skip Z==0, delta -> unit species */
#define SPSIM_HIP_PACK_VECTOR 1 /* vector_compute_pattern_on_list quirks (see diffraction.h) */
#define SPSIM_HIP_PACK_NFFT 2   /* compute_pattern_on_list_by_nfft: Z==0 kept (f = row 0), delta -> unit
                                   species for every atom, amplitudes ignored */

/* Packed atom record in HBM (16 bytes): position in metres and the real
   amplitude weight (sqrt of the beam illumination at the atom, 1 for a flat beam). */
typedef struct {
  float x, y, z, w;
} spsim_hip_atom;

/* A run of packed atoms of one species: atoms [start, start+count). */
typedef struct {
  int start, count, species, reserved;
} spsim_hip_segment;

/* Kernel tiling: pixels per lane, waves per workgroup, atoms staged per LDS chunk. */
typedef struct {
  int pixels_per_lane; /* 1, 2, 4 or 8 */
  int waves;           /* 4 */
  int chunk_atoms;     /* 64, 128 or 256 */
  int atom_splits;     /* >= 1: atom ranges per pixel tile (single pattern); 0 = automatic */
} spsim_hip_tile;

/* Library / device */
const char *spsim_hip_last_error(void);
int spsim_hip_device_count(int *count);
/* Number of compute units of the current device. */
int spsim_hip_cu_count(int *cus);
int spsim_hip_synchronize(void *stream);

/* Default Cromer-Mann table [101][9] (row 100 = unit scatterer) into host memory.
   The values equal the reference's include/atomsf.cdata rows; this function does
   not apply the ./atomsf.lib override (diffraction.c does, like the reference). */
int spsim_hip_default_form_factors(float *out_101x9);

/* Host-side packing (no device work): sorts atoms by species, drops/relabels
   per `mode`, and writes packed records, per-atom B (if b != NULL), the
   permutation perm[packed] = original index and the species segments.
   Capacities: out_atoms/out_b/out_perm >= natoms, out_segs >= SPSIM_HIP_NSPECIES.
   amp may be NULL (all 1). */
int spsim_hip_pack_atoms(const int *Z, const float *pos, const float *amp, const float *b, int natoms,
                         int delta_atoms, int mode, spsim_hip_atom *out_atoms, float *out_b, int *out_perm,
                         spsim_hip_segment *out_segs, int *out_natoms, int *out_nsegs);

/* Scattering vectors of a flat (spherical = 0) or spherical-grid detector, exactly
   as get_HKL_list_for_detector: FP64 geometry, FP32 [ny][nx][3] store (m^-1). */
int spsim_hip_hkl_detector(int nx, int ny, float width, float height, float distance, float center_x,
                           float center_y, int spherical, float wavelength, float *d_hkl, void *stream);

/* d_out[i] = R * d_in[i] for a row-major 3x3 R, formed in FP64 and rounded to FP32. */
int spsim_hip_rotate_hkl(const float *d_in, int npix, const double *rot9, float *d_out, void *stream);

/* <- get_HKL_list_for_3d_detector, src/diffraction.c:325-369: the uniform
   reciprocal-space grid of an nx x ny x nz detector volume, [nz][ny][nx][3],
   q = p / (distance * wavelength) with p = ((i - (n-1)/2)/n) * extent - center
   per axis (FP64, stored FP32, bit-identical to the host formula). */
int spsim_hip_hkl_detector_3d(int nx, int ny, int nz, float width, float height, float depth, float center_x,
                              float center_y, float center_z, float distance, float wavelength, float *d_hkl,
                              void *stream);

/* One pattern: F[i] = sum_j f_s(|q'_i|) exp(-B_eff s^2) w_j e^{-B_j s^2} exp(-2 pi i q'_i.r_j),
   I[i] = |F[i]|^2, with q'_i = R q_i, s = |q| / 2 in A^-1, B_eff = b_global (0 for the
   unit species). rot9 (host, row-major 3x3) may be NULL (R = identity); the
   kernels apply it exactly, in FP64, as R^T to the atoms. d_atom_b may be NULL
   (no per-atom Debye-Waller factor). d_F ([npix][2]) or d_I may be NULL. The
   nsegs segments must tile the packed atoms [0, natoms) in order (as
   spsim_hip_pack_atoms writes them). tile may be NULL (default tiling). Uses a
   library-owned device workspace when atom_splits > 1 (grown on demand, kept
   until spsim_hip_release). */
int spsim_hip_sf_pattern(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                         const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff,
                         float b_global, const double *rot9, float *d_F, float *d_I, const spsim_hip_tile *tile,
                         void *stream);

/* One pattern on a full nx x ny detector whose q list is [ny][nx][3] (as
   spsim_hip_hkl_detector writes it); same result as spsim_hip_sf_pattern.
   When the list is mirror-symmetric (a detector centred on the beam: the pixel
   mirrored in x has q = (-qx, qy, qz), in y (qx, -qy, qz), bit for bit), each
   lane evaluates the four mirror pixels together and shares their sines,
   cosines, form factor and Debye-Waller factor; workgroups whose pixels are not
   symmetric fall back to the per-pixel sum. On a square detector (nx == ny)
   whose list is also transpose-symmetric (q(y, x) = (qy, qx, qz)(x, y)), one
   lane evaluates an octant, a quad and its transpose (8 pixels sharing qz r_z);
   workgroups without that symmetry fall back to two quads. tile->pixels_per_lane
   is the quads per lane here (1 or 2), or 8 for one octant per lane (nx == ny
   only; the default for square detectors, with atom_splits chosen to fill the
   GPU). tile may be NULL. */
int spsim_hip_sf_grid(const float *d_hkl, int nx, int ny, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                      const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
                      const double *rot9, float *d_F, float *d_I, const spsim_hip_tile *tile, void *stream);

/* spsim_hip_sf_grid without any allocation. No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc). Same
   arguments and kernels; the atom-range partial sums use a fixed 32 MiB device
   array (2^21 double2 entries, held by each device the library runs on until
   the process ends) instead of the library workspace, so splits x nx x ny may
   not exceed 2^21. With tile->atom_splits > 0 a larger product is rejected
   (SPSIM_HIP_ERR_INVALID, nothing launched); with an automatic split count the
   count is lowered to what fits (1 range needs no array at all, so every
   detector size is accepted). Every call on one device shares that device's
   array: calls that overlap on one device (two streams, two threads) race. With
   the same tile and split count the result is bitwise that of
   spsim_hip_sf_grid. */
int spsim_hip_sf_grid_noalloc(const float *d_hkl, int nx, int ny, const spsim_hip_atom *d_atoms,
                              const float *d_atom_b, const spsim_hip_segment *d_segs, int nsegs, int natoms,
                              const float *d_ff, float b_global, const double *rot9, float *d_F, float *d_I,
                              const spsim_hip_tile *tile, void *stream);

/* ---- Wavelength spectrum (replaces the spectrum branch of the reference's GPU
   path, cuda_compute_pattern_on_list2 src/diffraction.cu:300-318 with the kernel
   CUDA_spectrum_scattering_from_all_atoms :416-476) ---- */

#define SPSIM_HIP_MAX_SPECTRUM 65 /* wavelength samples per pattern */

/* Samples of a Gaussian spectrum of relative width `bandwidth` (the full
   range lambda (1 +- bandwidth/2) is +-2 sigma), as the reference draws them
   (src/diffraction.cu:302-303, 431-460): an even count above 1 is raised by
   one; lambda_k = lambda (1 - bandwidth/2) + k/(S-1) lambda bandwidth; weights
   exp(-x^2/2), x = (lambda_k - lambda)/sigma, normalised to sum 1; phase scales
   s_k = lambda / lambda_k. bandwidth == 0 or nsamples <= 1 gives one sample
   {1, 1}. Writes at most SPSIM_HIP_MAX_SPECTRUM entries of scales and weights
   and returns their count, or SPSIM_HIP_ERR_INVALID when more are needed. */
int spsim_hip_spectrum_samples(double wavelength, double bandwidth, int nsamples, double *scales, double *weights);

/* One pattern under a spectrum: F[i] = sum_j f_s(|q'_i|) w_j e^{-B_j s^2}
   sum_k weights[k] exp(-2 pi i scales[k] q'_i.r_j), I = |F|^2 (an amplitude
   average over wavelengths with the form factor at the unscaled |q|, as the
   reference's spectrum kernel computes it). scales and weights are host arrays
   of nsamples <= SPSIM_HIP_MAX_SPECTRUM entries. Other arguments as
   spsim_hip_sf_pattern; tile->pixels_per_lane is 1 or 2. */
int spsim_hip_sf_spectrum(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                          const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff,
                          float b_global, const double *rot9, int nsamples, const double *scales,
                          const double *weights, float *d_F, float *d_I, const spsim_hip_tile *tile, void *stream);

/* ---- Non-uniform FFT structure factor (replaces compute_pattern_on_list_by_nfft,
   src/diffraction.c:528-675, the NNFFT path behind the use_nfft_for_sf option) ---- */

/* F[i] = sum_s f_s(|q_i|) sum_{j in s} w_j exp(-2 pi i q_i.r_j), I = |F|^2, over the
   species segments of the packed atoms (f_s with the global B, 1 for the unit
   species) and q_i rotated as q' = R q (rot9: host row-major 3x3, may be NULL;
   applied exactly in FP64 as R^T to the atoms, like spsim_hip_sf_pattern),
   computed by a type-3 non-uniform FFT instead of the direct sum:
   atoms are spread onto per-species grids with an exponential-of-semicircle
   window, the grids go through a pruned 3-D FFT, and each q interpolates from
   them. eps sets the window width W = ceil(-log10 eps) + 1 in [4, 10] (eps <= 0:
   1e-6); the error of F is about eps times sum_j |w_j| f_s.
   d_atom_b (packed order, may be NULL) gives every atom its own B factor
   (f_s exp(-B_j s^2) with s = |q| / 2 in 1/A, as in spsim_hip_sf_pattern):
   each atom is then spread as a Gaussian whose width carries its B (17, 25
   or 33 points, chosen from the largest B; larger needs and negative B are
   rejected); the x and y grids are then not centred in frequency, and z
   keeps its centring through complex z weights.
   The grids scale with (atom box x q range) per axis; lines longer than 3840
   points or grids above 64 GiB are rejected. Each q interpolates with a second
   window: the spread window itself (FFT lines >= 2 M), or, for W >= 7 when that
   costs less (many q per grid), a window of width W - 1 on lines >= 4 M, which
   keeps the error. info (host, may be NULL) receives 13 ints
   {W, Mx, My, Mz, nfx, nfy, nfz, Lx, Ly, Lz, W2, staged, Wg}: window width,
   spread grid, FFT and kept-output sizes per axis, interpolation window width,
   1 when the interpolation staged each wave's grid box in LDS (lists whose
   consecutive q are close, such as detector rows), and the Gaussian spread
   window's width (0 without per-atom B). Uses
   the library workspace; synchronises `stream` once (to read the bounds that
   size the grids), the rest is stream-ordered. */
int spsim_hip_sf_nfft(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                      const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
                      const double *rot9, double eps, float *d_F, float *d_I, int *info, void *stream);

/* Many orientations of one molecule through the NUFFT: pattern p is
   spsim_hip_sf_nfft's result for q rotated as q' = R_p q (rot9s: host
   [npatterns][9], row-major, applied exactly in FP64), written to
   d_F + p*npix*2 and d_I + p*npix. One set of species grids is built for the
   unrotated atoms over the ball |q| <= max |q| max_p ||R_p||_2 (which holds
   every R q; the norm is bounded by Gershgorin's bound on R^T R, so any R_p
   is covered: 1 for a rotation, s for s times a rotation), narrowed per axis
   to the interval-arithmetic range of R_p over the q box, and every pattern
   only interpolates from it, so the spread and FFT run once.
   The grids are larger than a single pattern's when q does not fill that ball
   (a flat detector's Ewald cap). Returns after the last pattern is written
   (it synchronises `stream`). Other arguments as spsim_hip_sf_nfft. */
int spsim_hip_sf_nfft_batch(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                            const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff, float b_global,
                            int npatterns, const double *rot9s, double eps, float *d_F, float *d_I, int *info,
                            void *stream);

/* A spectrum (as spsim_hip_sf_spectrum) through the NUFFT: the sums at
   q'_k = scales[k] R q of every sample come from one set of species grids, as
   spsim_hip_sf_nfft_batch computes them for the matrices scales[k] R (R =
   rot9, or the identity when NULL), into d_Fk [nsamples][npix][2] (a device
   buffer the caller provides); then F = sum_k weights[k] F_k and I = |F|^2
   (d_F and d_I may each be NULL). Form factors and the global B are taken at
   the unscaled |q|, as in spsim_hip_sf_spectrum. Per-atom B factors are not
   taken: a Gaussian-spread atom's Debye-Waller factor would be evaluated at the
   scaled q. nsamples <= SPSIM_HIP_MAX_SPECTRUM; scales finite and > 0. No
   counterpart in the reference (its NFFT path has no spectrum); it computes
   the same sums as the spectrum branch of src/diffraction.cu:300-318 to the
   NUFFT's accuracy. Synchronises `stream` once, before the weighted sum. */
int spsim_hip_sf_spectrum_nfft(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms,
                               const spsim_hip_segment *d_segs, int nsegs, int natoms, const float *d_ff,
                               float b_global, const double *rot9, int nsamples, const double *scales,
                               const double *weights, double eps, float *d_Fk, float *d_F, float *d_I, int *info,
                               void *stream);

/* Many molecules through the NUFFT, one pattern each: pattern m is
   spsim_hip_sf_nfft's result for molecule m at q rotated by rot9s[m] (host
   [nmol][9], row-major; NULL: no rotation), written to d_F + m*npix*2 and
   d_I + m*npix. Molecule m owns the packed atoms (and B factors, when
   d_atom_b is given) [atom_offsets[m], atom_offsets[m+1]) and the species
   segments [seg_offsets[m], seg_offsets[m+1]), whose starts count from the
   molecule's first atom (host offset arrays of nmol + 1 entries, from 0). The
   bounds of every molecule are read with one synchronisation; the molecules'
   stages are then enqueued back to back, so the host does not wait between
   molecules. They run on `stream` and on 31 library-owned streams forked
   from it and joined back into it (work enqueued on `stream` afterwards runs
   after every pattern), each stream with its own slice of the library
   workspace. Calls
   that overlap on one device race, as for every entry point sharing the
   workspace. Same grids and accuracy as spsim_hip_sf_nfft. */
int spsim_hip_sf_nfft_molecules(const float *d_hkl, int npix, int nmol, const spsim_hip_atom *d_atoms,
                                const float *d_atom_b, const int *atom_offsets, const spsim_hip_segment *d_segs,
                                const int *seg_offsets, const float *d_ff, float b_global, const double *rot9s,
                                double eps, float *d_F, float *d_I, void *stream);

/* The pruned FFT stage of spsim_hip_sf_nfft on its own: every line of the
   complex input view [outer][M][inner] is placed at points m = m_lo + i
   (i < M) of a zero line of nf points (slot m mod nf), transformed forward
   (exp(-2 pi i l m / nf)), and the values at l = l_lo + i (i < L) are written
   to the view [outer][L][inner]. nf must be 2,3,5-smooth, in [2, 3840], and >= M;
   m_lo and l_lo lie in [-nf, nf], m_lo + M and l_lo + L are at most 2 nf. Uses
   the library workspace (the twiddle table). */
int spsim_hip_fft_pruned(const float *d_in, long long outer, long long inner, int M, int m_lo, int nf, int L,
                         int l_lo, float *d_out, void *stream);

/* Free the library's cached device workspace and graphs on the current device. */
int spsim_hip_release(void);

/* ---- Batched persistent launch over many molecules and orientations ---- */

typedef struct spsim_hip_batch spsim_hip_batch;

/* Molecules are given as host arrays, concatenated: molecule m owns raw atoms
   [mol_offsets[m], mol_offsets[m+1]). atom_b may be NULL. Pattern p shows
   molecule pattern_mol[p] with q rotated by rot9[p] (row-major 3x3, q' = R q,
   applied exactly in FP64, as in spsim_hip_sf_pattern). hkl is the shared
   [ny][nx][3] host q list. When that list is mirror-symmetric the plan runs the
   quad kernel (tile->pixels_per_lane = quads per lane), or, when it is also
   square and transpose-symmetric, the octant kernel (pixels_per_lane = 8, the
   default there); otherwise the per-pixel kernel. The plan copies everything
   to the device, builds the cost-sorted (pattern, tile, atom range) work list
   and owns its device memory until spsim_hip_batch_destroy. tile may be NULL
   (default tiling). A pattern whose molecule has more atoms than the plan's
   item bound (spsim_hip_batch_set_item_atoms) is split into near-equal atom
   ranges, one work item per range and tile; the tile's last item to finish
   adds the ranges' FP64 partial sums in range order, so results do not depend
   on the order in which items finish. */
int spsim_hip_batch_create(const float *hkl, int nx, int ny, int nmol, const int *mol_offsets, const int *Z,
                           const float *pos, const float *amp, const float *atom_b, int delta_atoms,
                           int npatterns, const int *pattern_mol, const double *rot9, const float *ff,
                           float b_global, const spsim_hip_tile *tile, spsim_hip_batch **out);
/* Runs the persistent kernel: d_I [npatterns][npix] (required), d_F [npatterns][npix][2]
   (may be NULL). use_graph != 0 replays a captured hipGraph (captured on first use
   for this stream and these output pointers; the stream must not be NULL). */
int spsim_hip_batch_run(spsim_hip_batch *b, float *d_F, float *d_I, int use_graph, void *stream);
/* Plain path for comparison: one single-pattern launch per pattern with a host
   synchronisation after each, no work list and no graph. */
int spsim_hip_batch_run_plain(spsim_hip_batch *b, float *d_F, float *d_I, void *stream);
/* 1 when the plan runs the mirror-symmetric quad kernel, 0 for the per-pixel kernel. */
int spsim_hip_batch_uses_grid(const spsim_hip_batch *b);
/* Tile of the persistent kernel; setting it rebuilds the work list. */
int spsim_hip_batch_set_tile(spsim_hip_batch *b, const spsim_hip_tile *tile);
int spsim_hip_batch_get_tile(const spsim_hip_batch *b, spsim_hip_tile *tile, int *blocks);
/* Bound on the atoms of one work item; rebuilds the work list. 0 (the default):
   automatic, about an eighth of one resident workgroup's share of the whole
   run and at least 4 LDS chunks; > 0: that many atoms; < 0: never split a
   pattern's atoms. */
int spsim_hip_batch_set_item_atoms(spsim_hip_batch *b, int item_atoms);
/* Work items of one run, the item bound in effect (0: no pattern is split) and
   the number of (pattern, tile) pairs split over several items. Any pointer
   may be NULL. */
int spsim_hip_batch_describe(const spsim_hip_batch *b, int *nitems, int *item_atoms, int *split_tiles);
/* Autotuner: times every supported tile (pixels per lane x LDS chunk) on a
   slice of the plan's own work (patterns spread over its atom-count range),
   keeps the kernel family (pixels or quads per lane, octants) with the
   highest pixel-atom terms/s and, within it, the LDS chunk whose sample runs
   fastest, and returns it (best_rate: its terms/s; may be NULL). Synchronises
   the stream. */
int spsim_hip_batch_autotune(spsim_hip_batch *b, void *stream, spsim_hip_tile *best, double *best_rate);
/* Total pixel-atom terms of one run (sum over patterns of natoms * npix). */
int spsim_hip_batch_terms(const spsim_hip_batch *b, double *terms);
void spsim_hip_batch_destroy(spsim_hip_batch *b);

/* ---- Detector chain (fused epilogue) ---- */

typedef struct {
  int nx, ny;             /* raw pixels */
  int bin_x, bin_y;       /* binning; binned image is (ny/bin_y) x (nx/bin_x) */
  float width, height;    /* m */
  float distance;         /* m */
  float pixel_width, pixel_height;
  float center_x, center_y;
  int spherical;
  int polarization;       /* POLARIZATION_* */
  float beam_intensity;   /* photons / m^2 */
  float quantum_efficiency;
  float wavelength;       /* m */
  float electron_hole_production_energy; /* J; 0 = one electron per photon (noisy branch) */
  float dark_current;     /* e- / (pixel s) */
  float exposure_time;    /* s */
  float readout_noise;    /* e- rms per binned pixel */
  float linear_full_well; /* e- */
  float maximum_value;    /* ADC counts */
  int nz, bin_z;          /* 3-D detector volume: raw voxels nx*ny*nz, binned (nz/bin_z) deep;
                             0 means 1 (a 2-D detector) */
} spsim_hip_detector;

/* For each of npatterns intensity images d_I[p*npix .. ] (npix = nx*ny*nz) runs
   Thomson x solid angle x beam -> photons, Poisson(photons*QE) -> count,
   electrons, binned readout with dark current + Gaussian readout noise clamped
   at 0 and the ADC factor, and the noiseless binned output. Randomness comes
   from Philox-4x32-10 keyed by seed, with counters (pattern_base + p, pixel),
   so results do not depend on launch shape. Any output pointer may be NULL.
   Raw outputs are [npatterns][npix], binned ones [npatterns][nbinned] with
   nbinned = (nx/bin_x)*(ny/bin_y)*(nz/bin_z), z slowest. Thomson and solid
   angle depend on (x, y) only, as in the reference's loops. */
int spsim_hip_detector_chain(const spsim_hip_detector *det, const float *d_I, int npatterns, uint64_t seed,
                             uint64_t pattern_base, float *d_thomson, float *d_solid_angle, float *d_photons,
                             float *d_count, float *d_electrons, float *d_real_output,
                             float *d_noiseless_output, void *stream);

/* The binned outputs of spsim_hip_detector_chain only (real and noiseless,
   either may be NULL), with one photon count per bin instead of one per raw
   pixel: the pixels of a bin whose mean photons x QE is below 1000 have
   independent Poisson counts, and their sum is one Poisson variable with the
   summed mean, which is drawn once per bin (counters (pattern_base + p, bin));
   pixels on the reference's Gaussian branch (mean >= 1000, src/noise.c:44-46)
   keep their own draw. The binned real output therefore has the distribution
   of the per-pixel chain's, not its values; the noiseless output is bitwise
   the per-pixel chain's. With d_real_output, uses the library workspace (one
   bit per bin and pattern marks the bins its second pass recomputes). */
int spsim_hip_detector_chain_binned(const spsim_hip_detector *det, const float *d_I, int npatterns, uint64_t seed,
                                    uint64_t pattern_base, float *d_real_output, float *d_noiseless_output,
                                    void *stream);

/* ---- Adjoint ---- */

/* Poisson NLL of expected counts lambda_i = K_i * I_i against counts y_i:
   loss = sum_i (lambda_i - y_i log lambda_i); d_gI[i] = dloss/dI_i. d_K may be
   NULL (K = 1). d_loss (1 double) may be NULL. */
int spsim_hip_poisson_nll(const float *d_I, const float *d_y, const float *d_K, int npix, double *d_loss,
                          float *d_gI, void *stream);

/* Given the forward F of the same inputs (d_F [npix][2]) and g_i = dL/dI_i (d_gI),
   writes d_grad_pos [natoms][3] = dL/dr_j (per packed atom, in units of L per
   metre) and d_grad_b [natoms] = dL/dB_j (units of L per A^2), recomputing the
   phases per pixel chunk instead of storing atom x pixel terms. h_segs is the
   HOST copy of the segments (as spsim_hip_pack_atoms returns them); d_grad_b
   may be NULL. Uses the library workspace and synchronises the stream. */
int spsim_hip_sf_adjoint(const float *d_hkl, int npix, const spsim_hip_atom *d_atoms, const float *d_atom_b,
                         const spsim_hip_segment *h_segs, int nsegs, int natoms, const float *d_ff,
                         float b_global, const float *d_F, const float *d_gI, float *d_grad_pos,
                         float *d_grad_b, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SPSIM_HIP_H */
