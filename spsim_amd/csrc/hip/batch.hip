// Batched structure factors: many molecules and orientations in one persistent
// launch over a cost-sorted (pattern, pixel-tile, atom-range) work list,
// optionally replayed from a captured hipGraph, plus the plain
// one-launch-per-pattern path and the tile autotuner.
//
// Atom ranges: a pattern whose molecule has more than item_atoms atoms is
// split into ceil(N / item_atoms) near-equal ranges, one work item each per
// tile, whose FP64 partial sums the tile's last item adds in range order. By
// default item_atoms bounds the largest item to about an eighth of one
// workgroup slot's share of the whole run (total atom-tiles / (8 x resident
// workgroups)), and at least 4 LDS chunks, so a few very large molecules
// among many small ones no longer leave most of the GPU idle at the end of
// the run, while runs whose molecules are all small against that bound keep
// one item per (pattern, tile) and write their outputs directly.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "sf_kernels.hpp"

using spsim::BatchItem;
using spsim::MolInfo;

struct spsim_hip_batch {
  int npix = 0, nx = 0, ny = 0, npatterns = 0, nmol = 0;
  bool grid = false;    // mirror-symmetric detector: quad work items
  bool square = false;  // also n x n and transpose-symmetric: octant work items allowed
  spsim_hip_tile tile{};
  int blocks = 0;
  int cus = 0;
  bool has_b = false;
  float b_global = 0.0f;
  int nitems = 0;
  int item_atoms_req = 0;  // 0: automatic, > 0: fixed bound, < 0: never split
  int item_atoms = 0;      // in effect (0: no pattern split)
  int split_tiles = 0;     // (pattern, tile) pairs with more than one item
  double terms = 0.0;

  // device memory
  float *hkl = nullptr;
  spsim_hip_atom *atoms = nullptr;
  float *atom_b = nullptr;
  spsim_hip_segment *segs = nullptr;  // per molecule, starts relative to the molecule's first atom
  MolInfo *mols = nullptr;
  int *pattern_mol = nullptr;
  double *rot = nullptr;
  BatchItem *items = nullptr;
  unsigned *queue = nullptr;
  unsigned *counters = nullptr;  // one per split (pattern, tile), zero between runs
  double2 *partials = nullptr;   // partial slots of split tiles
  float *ff = nullptr;

  // host copies
  std::vector<MolInfo> h_mols;
  std::vector<int> h_pattern_mol;
  std::vector<double> h_rot;

  // captured graph
  hipGraphExec_t exec = nullptr;
  hipStream_t g_stream = nullptr;
  float *g_F = nullptr, *g_I = nullptr;
};

namespace spsim {
namespace {

template <class T>
int dmalloc(T **p, size_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return fail(SPSIM_HIP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", n * sizeof(T), hipGetErrorString(e));
  return SPSIM_HIP_OK;
}

template <class T>
int dupload(T **p, const T *h, size_t n) {
  SPSIM_TRY(dmalloc(p, n));
  if (n) SPSIM_HIP_TRY(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return SPSIM_HIP_OK;
}

void free_batch(spsim_hip_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  void *ptrs[] = {b->hkl,  b->atoms, b->atom_b,   b->segs,     b->mols, b->pattern_mol,
                  b->rot,  b->items, b->counters, b->partials, b->queue, b->ff};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete b;
}

int natoms_of(const spsim_hip_batch &b, int p) {
  const MolInfo &m = b.h_mols[b.h_pattern_mol[p]];
  return m.atom_end - m.atom_begin;
}

int tiles_of(const spsim_hip_batch &b, const spsim_hip_tile &t) {
  const int units = b.grid ? grid_units(t, b.nx, b.ny) : b.npix;
  const int per_tile = b.grid ? grid_units_per_tile(t) : tile_units(t);
  return (units + per_tile - 1) / per_tile;
}

// Largest atom range per work item for tile t with `blocks` resident
// workgroups (0: no pattern is split).
int item_atoms_for(const spsim_hip_batch &b, const spsim_hip_tile &t, int blocks) {
  if (b.item_atoms_req < 0) return 0;
  int bound = b.item_atoms_req;
  if (bound == 0) {
    double atom_tiles = 0.0;
    for (int p = 0; p < b.npatterns; ++p) atom_tiles += (double)natoms_of(b, p);
    atom_tiles *= tiles_of(b, t);
    const double share = atom_tiles / (8.0 * std::max(blocks, 1));
    bound = (int)std::min<double>(std::max<double>(std::ceil(share), 4.0 * t.chunk_atoms), 1 << 30);
  }
  int biggest = 0;
  for (int p = 0; p < b.npatterns; ++p) biggest = std::max(biggest, natoms_of(b, p));
  return biggest > bound ? bound : 0;
}

struct ItemList {
  std::vector<BatchItem> items;
  int groups = 0;       // split (pattern, tile) pairs
  long long slots = 0;  // partial slots
};

// Work items for tile t, atoms split into ranges of at most item_atoms (0: no
// split), heaviest first so the persistent workgroups finish together
// (longest-processing-time order). only_patterns / max_tiles restrict the
// list to a sample (autotuner).
ItemList build_items(const spsim_hip_batch &b, const spsim_hip_tile &t, int item_atoms,
                     const std::vector<int> *only_patterns = nullptr, int max_tiles = -1) {
  std::vector<int> pats;
  if (only_patterns) {
    pats = *only_patterns;
  } else {
    pats.resize(b.npatterns);
    std::iota(pats.begin(), pats.end(), 0);
  }
  std::stable_sort(pats.begin(), pats.end(), [&](int a, int c) { return natoms_of(b, a) > natoms_of(b, c); });
  int tiles = tiles_of(b, t);
  if (max_tiles > 0) tiles = std::min(tiles, max_tiles);
  ItemList L;
  for (int p : pats) {
    const int n = natoms_of(b, p);
    const int nch = (item_atoms > 0 && n > item_atoms) ? (n + item_atoms - 1) / item_atoms : 1;
    for (int tl = 0; tl < tiles; ++tl) {
      const int group = nch > 1 ? L.groups++ : -1;
      const int slot = nch > 1 ? (int)L.slots : -1;
      if (nch > 1) L.slots += nch;
      for (int c = 0; c < nch; ++c) {
        BatchItem it;
        it.pattern = p;
        it.tile = tl;
        it.a_begin = (int)((long long)n * c / nch);
        it.a_end = (int)((long long)n * (c + 1) / nch);
        it.group = group;
        it.chunk = c;
        it.nchunks = nch;
        it.slot = slot;
        L.items.push_back(it);
      }
    }
  }
  std::stable_sort(L.items.begin(), L.items.end(), [](const BatchItem &x, const BatchItem &y) {
    return x.a_end - x.a_begin > y.a_end - y.a_begin;
  });
  return L;
}

// Device counters (zeroed) and partial slots for an item list.
int alloc_split_state(const ItemList &L, const spsim_hip_tile &t, bool grid, unsigned **counters,
                      double2 **partials) {
  *counters = nullptr;
  *partials = nullptr;
  if (L.groups == 0) return SPSIM_HIP_OK;
  SPSIM_TRY(dmalloc(counters, (size_t)L.groups));
  SPSIM_HIP_TRY(hipMemsetAsync(*counters, 0, sizeof(unsigned) * (size_t)L.groups, nullptr));
  SPSIM_HIP_TRY(hipStreamSynchronize(nullptr));
  return dmalloc(partials, (size_t)L.slots * slot_entries(t, grid));
}

BatchLaunch launch_args(const spsim_hip_batch *b, const BatchItem *items, int nitems, unsigned *queue,
                        unsigned *counters, double2 *partials, float *F, float *I) {
  BatchLaunch a;
  a.hkl = b->hkl;
  a.npix = b->npix;
  a.nx = b->nx;
  a.ny = b->ny;
  a.atoms = b->atoms;
  a.atom_b = b->has_b ? b->atom_b : nullptr;
  a.segs = b->segs;
  a.mols = b->mols;
  a.pattern_mol = b->pattern_mol;
  a.rot = b->rot;
  a.items = items;
  a.nitems = nitems;
  a.queue = queue;
  a.counters = counters;
  a.partials = partials;
  a.ff = b->ff;
  a.b_global = b->b_global;
  a.F = (float2 *)F;
  a.I = I;
  return a;
}

int set_tile(spsim_hip_batch *b, const spsim_hip_tile &t) {
  if (!tile_supported(t, b->grid) || (b->grid && t.pixels_per_lane == kOctantLane && !b->square))
    return fail(SPSIM_HIP_ERR_INVALID, "unsupported tile (octants need a square, transpose-symmetric q list)");
  int per_cu = 1;
  SPSIM_TRY(batch_blocks_per_cu(t, b->has_b, b->grid, &per_cu));
  const int blocks = b->cus * per_cu;
  const int item_atoms = item_atoms_for(*b, t, blocks);
  ItemList L = build_items(*b, t, item_atoms);
  if (L.slots > 0x7fffffffLL || L.items.size() > 0x7fffffffULL)
    return fail(SPSIM_HIP_ERR_INVALID, "work list too large (%zu items, %lld partial slots); raise item_atoms",
                L.items.size(), L.slots);
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  void *old[] = {b->items, b->counters, b->partials};
  for (void *p : old)
    if (p) SPSIM_HIP_TRY(hipFree(p));
  b->items = nullptr;
  b->counters = nullptr;
  b->partials = nullptr;
  SPSIM_TRY(dupload(&b->items, L.items.data(), L.items.size()));
  SPSIM_TRY(alloc_split_state(L, t, b->grid, &b->counters, &b->partials));
  b->nitems = (int)L.items.size();
  b->item_atoms = item_atoms;
  b->split_tiles = L.groups;
  b->tile = t;
  b->blocks = blocks;
  return SPSIM_HIP_OK;
}

}  // namespace
}  // namespace spsim

extern "C" {

int spsim_hip_batch_create(const float *hkl, int nx, int ny, int nmol, const int *mol_offsets, const int *Z,
                           const float *pos, const float *amp, const float *atom_b, int delta_atoms, int npatterns,
                           const int *pattern_mol, const double *rot9, const float *ff, float b_global,
                           const spsim_hip_tile *tile, spsim_hip_batch **out) {
  using namespace spsim;
  if (!out) return fail(SPSIM_HIP_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (nx <= 0 || ny <= 0 || nmol <= 0 || npatterns <= 0)
    return fail(SPSIM_HIP_ERR_INVALID, "nx, ny, nmol and npatterns must be > 0");
  if ((long long)nx * ny > (1 << 28)) return fail(SPSIM_HIP_ERR_INVALID, "detector too large");
  const int npix = nx * ny;
  if (!hkl || !mol_offsets || !pattern_mol || !rot9) return fail(SPSIM_HIP_ERR_INVALID, "NULL input array");
  if (mol_offsets[0] != 0) return fail(SPSIM_HIP_ERR_INVALID, "mol_offsets[0] must be 0");
  for (int m = 0; m < nmol; ++m)
    if (mol_offsets[m + 1] < mol_offsets[m]) return fail(SPSIM_HIP_ERR_INVALID, "mol_offsets must be non-decreasing");
  const int total = mol_offsets[nmol];
  if (total > 0 && (!Z || !pos)) return fail(SPSIM_HIP_ERR_INVALID, "Z/pos is NULL");
  for (int p = 0; p < npatterns; ++p)
    if (pattern_mol[p] < 0 || pattern_mol[p] >= nmol) return fail(SPSIM_HIP_ERR_INVALID, "pattern_mol[%d] out of range", p);
  int cus = 0;
  SPSIM_TRY(ensure_device(&cus));

  spsim_hip_batch *b = new spsim_hip_batch();
  b->npix = npix;
  b->nx = nx;
  b->ny = ny;
  b->grid = host_mirror_symmetric(hkl, nx, ny);
  b->square = b->grid && nx == ny && host_transpose_symmetric(hkl, nx);
  b->npatterns = npatterns;
  b->nmol = nmol;
  b->cus = cus;
  b->has_b = atom_b != nullptr;
  b->b_global = b_global;

  // Pack every molecule: species-sorted atoms with segments relative to the molecule.
  std::vector<spsim_hip_atom> h_atoms(std::max(total, 1));
  std::vector<float> h_b(std::max(total, 1), 0.0f);
  std::vector<spsim_hip_segment> h_segs;
  std::vector<spsim_hip_segment> tmp(SPSIM_HIP_NSPECIES);
  int packed_total = 0;
  b->h_mols.resize(nmol);
  for (int m = 0; m < nmol; ++m) {
    const int o = mol_offsets[m], n = mol_offsets[m + 1] - o;
    int nat = 0, nseg = 0;
    int rc = spsim_hip_pack_atoms(Z + o, pos + 3 * (size_t)o, amp ? amp + o : nullptr, atom_b ? atom_b + o : nullptr, n,
                                  delta_atoms, SPSIM_HIP_PACK_PLAIN, h_atoms.data() + packed_total,
                                  h_b.data() + packed_total, nullptr, tmp.data(), &nat, &nseg);
    if (rc != SPSIM_HIP_OK) {
      free_batch(b);
      return fail(rc, "molecule %d: atomic numbers must lie in [0, 99]", m);
    }
    b->h_mols[m] = MolInfo{packed_total, packed_total + nat, (int)h_segs.size(), nseg};
    h_segs.insert(h_segs.end(), tmp.begin(), tmp.begin() + nseg);
    packed_total += nat;
  }
  b->h_pattern_mol.assign(pattern_mol, pattern_mol + npatterns);
  b->h_rot.assign(rot9, rot9 + 9 * (size_t)npatterns);
  for (int p = 0; p < npatterns; ++p) {
    const MolInfo &m = b->h_mols[pattern_mol[p]];
    b->terms += (double)(m.atom_end - m.atom_begin) * npix;
  }

  float ffh[SPSIM_HIP_NSPECIES * SPSIM_HIP_FF_STRIDE];
  if (ff)
    std::memcpy(ffh, ff, sizeof ffh);
  else
    spsim_hip_default_form_factors(ffh);

  int rc = SPSIM_HIP_OK;
  if ((rc = dupload(&b->hkl, hkl, 3 * (size_t)npix)) ||
      (rc = dupload(&b->atoms, h_atoms.data(), (size_t)std::max(packed_total, 1))) ||
      (b->has_b && (rc = dupload(&b->atom_b, h_b.data(), (size_t)std::max(packed_total, 1)))) ||
      (rc = dupload(&b->segs, h_segs.data(), h_segs.size())) || (rc = dupload(&b->mols, b->h_mols.data(), (size_t)nmol)) ||
      (rc = dupload(&b->pattern_mol, pattern_mol, (size_t)npatterns)) ||
      (rc = dupload(&b->rot, rot9, 9 * (size_t)npatterns)) || (rc = dmalloc(&b->queue, 1)) ||
      (rc = dupload(&b->ff, ffh, (size_t)SPSIM_HIP_NSPECIES * SPSIM_HIP_FF_STRIDE))) {
    free_batch(b);
    return rc;
  }
  const int avg_atoms = packed_total / std::max(nmol, 1);
  spsim_hip_tile t = tile ? *tile : default_tile(avg_atoms, npix, b->grid);
  if (!tile && b->square) t.pixels_per_lane = kOctantLane;
  if ((rc = set_tile(b, t)) != SPSIM_HIP_OK) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return SPSIM_HIP_OK;
}

int spsim_hip_batch_set_tile(spsim_hip_batch *b, const spsim_hip_tile *tile) {
  using namespace spsim;
  if (!b || !tile) return fail(SPSIM_HIP_ERR_INVALID, "NULL argument");
  return set_tile(b, *tile);
}

int spsim_hip_batch_get_tile(const spsim_hip_batch *b, spsim_hip_tile *tile, int *blocks) {
  using namespace spsim;
  if (!b || !tile) return fail(SPSIM_HIP_ERR_INVALID, "NULL argument");
  *tile = b->tile;
  if (blocks) *blocks = b->blocks;
  return SPSIM_HIP_OK;
}

int spsim_hip_batch_set_item_atoms(spsim_hip_batch *b, int item_atoms) {
  using namespace spsim;
  if (!b) return fail(SPSIM_HIP_ERR_INVALID, "NULL plan");
  const int old = b->item_atoms_req;
  b->item_atoms_req = item_atoms;
  const int rc = set_tile(b, b->tile);
  if (rc != SPSIM_HIP_OK) b->item_atoms_req = old;
  return rc;
}

int spsim_hip_batch_describe(const spsim_hip_batch *b, int *nitems, int *item_atoms, int *split_tiles) {
  using namespace spsim;
  if (!b) return fail(SPSIM_HIP_ERR_INVALID, "NULL plan");
  if (nitems) *nitems = b->nitems;
  if (item_atoms) *item_atoms = b->item_atoms;
  if (split_tiles) *split_tiles = b->split_tiles;
  return SPSIM_HIP_OK;
}

int spsim_hip_batch_run(spsim_hip_batch *b, float *d_F, float *d_I, int use_graph, void *stream) {
  using namespace spsim;
  if (!b || !d_I) return fail(SPSIM_HIP_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  BatchLaunch a = launch_args(b, b->items, b->nitems, b->queue, b->counters, b->partials, d_F, d_I);
  if (!use_graph) return launch_batch(a, b->tile, b->grid, b->blocks, s);
  if (!s) return fail(SPSIM_HIP_ERR_INVALID, "graph replay needs a non-default stream");
  if (!b->exec || b->g_stream != s || b->g_F != d_F || b->g_I != d_I) {
    if (b->exec) {
      SPSIM_HIP_TRY(hipGraphExecDestroy(b->exec));
      b->exec = nullptr;
    }
    hipGraph_t g = nullptr;
    SPSIM_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = launch_batch(a, b->tile, b->grid, b->blocks, s);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != SPSIM_HIP_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    b->g_stream = s;
    b->g_F = d_F;
    b->g_I = d_I;
  }
  SPSIM_HIP_TRY(hipGraphLaunch(b->exec, s));
  return SPSIM_HIP_OK;
}

int spsim_hip_batch_run_plain(spsim_hip_batch *b, float *d_F, float *d_I, void *stream) {
  using namespace spsim;
  if (!b || !d_I) return fail(SPSIM_HIP_ERR_INVALID, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  for (int p = 0; p < b->npatterns; ++p) {
    const MolInfo &m = b->h_mols[b->h_pattern_mol[p]];
    const spsim_hip_atom *atoms = b->atoms + m.atom_begin;
    const float *ab = b->has_b ? b->atom_b + m.atom_begin : nullptr;
    const double *R = &b->h_rot[9 * (size_t)p];
    float *F = d_F ? d_F + 2 * (size_t)p * b->npix : nullptr;
    float *I = d_I + (size_t)p * b->npix;
    const int nat = m.atom_end - m.atom_begin;
    if (b->grid)
      SPSIM_TRY(spsim_hip_sf_grid(b->hkl, b->nx, b->ny, atoms, ab, b->segs + m.seg_begin, m.nseg, nat, b->ff,
                                  b->b_global, R, F, I, nullptr, s));
    else
      SPSIM_TRY(spsim_hip_sf_pattern(b->hkl, b->npix, atoms, ab, b->segs + m.seg_begin, m.nseg, nat, b->ff,
                                     b->b_global, R, F, I, nullptr, s));
    SPSIM_HIP_TRY(hipStreamSynchronize(s));
  }
  return SPSIM_HIP_OK;
}

int spsim_hip_batch_uses_grid(const spsim_hip_batch *b) { return b && b->grid ? 1 : 0; }

int spsim_hip_batch_terms(const spsim_hip_batch *b, double *terms) {
  using namespace spsim;
  if (!b || !terms) return fail(SPSIM_HIP_ERR_INVALID, "NULL argument");
  *terms = b->terms;
  return SPSIM_HIP_OK;
}

// Times each supported tile of the plan's kernel on representative slices of
// the work (patterns spread over the cost range, all of their tiles and atom
// ranges as the whole plan would split them). Picks the kernel family with the
// highest marginal rate and, within it, the LDS chunk by the larger sample's
// time. Leaves the plan on the chosen tile.
int spsim_hip_batch_autotune(spsim_hip_batch *b, void *stream, spsim_hip_tile *best_out, double *best_rate) {
  using namespace spsim;
  if (!b) return fail(SPSIM_HIP_ERR_INVALID, "NULL plan");
  hipStream_t s = (hipStream_t)stream;
  // Two samples of patterns at cost quantiles: `big` (up to 48) and `small`
  // (every third of them). Each candidate runs both in full; the marginal
  // rate (extra terms over extra time) cancels the run's tail, whose length
  // depends on the tile's item size and would otherwise decide the choice on
  // samples much smaller than the plan.
  std::vector<int> order(b->npatterns);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return natoms_of(*b, x) > natoms_of(*b, y); });
  std::vector<int> big, small;
  const int want = std::min(b->npatterns, 48);
  double big_terms = 0.0, small_terms = 0.0;
  for (int i = 0; i < want; ++i) {
    const int p = order[(size_t)i * b->npatterns / want];
    big.push_back(p);
    big_terms += (double)natoms_of(*b, p) * b->npix;
    if (i % 3 == 1) {
      small.push_back(p);
      small_terms += (double)natoms_of(*b, p) * b->npix;
    }
  }
  // quads per lane {1, 2} (+ octants on a square list) or pixels per lane {1, 2, 4, 8}
  const int ks_grid[3] = {1, 2, kOctantLane}, ks_pix[4] = {1, 2, 4, 8};
  const int *ks = b->grid ? ks_grid : ks_pix;
  const int nks = b->grid ? (b->square ? 3 : 2) : 4;
  const int cs[3] = {64, 128, 256};
  float *dI = nullptr;
  unsigned *q = nullptr;
  SPSIM_TRY(dmalloc(&dI, (size_t)b->npatterns * b->npix));
  if (dmalloc(&q, 1) != SPSIM_HIP_OK) {
    (void)hipFree(dI);
    return SPSIM_HIP_ERR_NOMEM;
  }
  hipEvent_t e0, e1;
  SPSIM_HIP_TRY(hipEventCreate(&e0));
  SPSIM_HIP_TRY(hipEventCreate(&e1));
  // Runs the given patterns in full (every tile, split as the whole plan
  // would split them) with tile t; *ms is the faster of two timed runs after a
  // warm-up (the marginal rate is a difference of two times, so a single
  // run's noise could swap tiles that differ by 1 %).
  auto time_sample = [&](const spsim_hip_tile &t, int blocks, const std::vector<int> &pats, float *ms) -> int {
    *ms = 0.0f;
    ItemList L = build_items(*b, t, item_atoms_for(*b, t, blocks), &pats);
    BatchItem *ditems = nullptr;
    unsigned *dcount = nullptr;
    double2 *dpart = nullptr;
    SPSIM_TRY(dupload(&ditems, L.items.data(), L.items.size()));
    int r = alloc_split_state(L, t, b->grid, &dcount, &dpart);
    if (r == SPSIM_HIP_OK) {
      BatchLaunch a = launch_args(b, ditems, (int)L.items.size(), q, dcount, dpart, nullptr, dI);
      r = launch_batch(a, t, b->grid, blocks, s);  // warm-up
      for (int rep = 0; rep < 2 && r == SPSIM_HIP_OK; ++rep) {
        hipEventRecord(e0, s);
        r = launch_batch(a, t, b->grid, blocks, s);
        hipEventRecord(e1, s);
        const hipError_t e = hipEventSynchronize(e1);
        float m = 0.0f;
        if (e != hipSuccess) r = hip_fail(e, "hipEventSynchronize");
        else hipEventElapsedTime(&m, e0, e1);
        if (r == SPSIM_HIP_OK && (rep == 0 || m < *ms)) *ms = m;
      }
    }
    if (dcount) (void)hipFree(dcount);
    if (dpart) (void)hipFree(dpart);
    (void)hipFree(ditems);
    return r;
  };
  spsim_hip_tile best = b->tile;
  double best_r = -1.0;
  struct Cand {
    spsim_hip_tile t;
    double r;
    float ms_big;
  };
  std::vector<Cand> cands;
  int rc = SPSIM_HIP_OK;
  for (int ki = 0; ki < nks && rc == SPSIM_HIP_OK; ++ki) {
    for (int ci = 0; ci < 3 && rc == SPSIM_HIP_OK; ++ci) {
      spsim_hip_tile t{ks[ki], 4, cs[ci], 0};
      int per_cu = 1;
      if ((rc = batch_blocks_per_cu(t, b->has_b, b->grid, &per_cu)) != SPSIM_HIP_OK) break;
      const int blocks = b->cus * per_cu;
      float ms_big = 0.0f, ms_small = 0.0f;
      if ((rc = time_sample(t, blocks, big, &ms_big)) != SPSIM_HIP_OK) break;
      if (!small.empty() && (rc = time_sample(t, blocks, small, &ms_small)) != SPSIM_HIP_OK) break;
      double r = -1.0;
      if (!small.empty() && ms_big > ms_small)
        r = (big_terms - small_terms) / ((ms_big - ms_small) * 1e-3);
      else if (ms_big > 0.0f)
        r = big_terms / (ms_big * 1e-3);
      cands.push_back(Cand{t, r, ms_big});
    }
  }
  // The kernel family (pixels or quads per lane, or octants) with the highest
  // marginal rate; within it, the LDS chunk with the fastest big sample. The
  // chunks of one family differ by 1-2 %, less than the marginal rate's spread
  // between runs (the small sample's time varies most, and the rate divides by
  // a difference of times), while the big sample's time ranks them as whole
  // runs do. Families differ by 10 % and more, and their tails differ.
  if (rc == SPSIM_HIP_OK && !cands.empty()) {
    int fam = -1;
    for (const Cand &c : cands)
      if (c.r > best_r) {
        best_r = c.r;
        fam = c.t.pixels_per_lane;
      }
    float best_ms = 0.0f;
    bool found = false;
    for (const Cand &c : cands)
      if (c.t.pixels_per_lane == fam && c.ms_big > 0.0f && (!found || c.ms_big < best_ms)) {
        best_ms = c.ms_big;
        best = c.t;
        best_r = c.r;
        found = true;
      }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dI);
  (void)hipFree(q);
  if (rc != SPSIM_HIP_OK) return rc;
  SPSIM_TRY(set_tile(b, best));
  if (best_out) *best_out = best;
  if (best_rate) *best_rate = best_r;
  return SPSIM_HIP_OK;
}

void spsim_hip_batch_destroy(spsim_hip_batch *b) { spsim::free_batch(b); }

}  // extern "C"
