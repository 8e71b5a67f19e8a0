"""spsim_hip_sf_grid_noalloc, the structure factor on a detector grid with its
atom-range partial sums in the library's fixed device array: called through
the C-ABI twice in a row (the second call smaller, so it reuses the array
that the first left holding its partials) and checked bit for bit against
spsim_hip_sf_grid with the same tile and split count, and against the FP64
oracle. An explicit split count that does not fit the array is rejected
before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import _lib
from spsim_amd import device as dv
from spsim_amd import workloads as wl
from tests.conftest import oracle_threads, rel_l2

pytestmark = pytest.mark.gpu

SCRATCH_PARTIALS = 1 << 21  # double2 entries of the library's no-allocation array


def _call(fn_name, q, n, atoms, ff, F, I, tile, B=0.0):
    L = _lib.load()
    fn = getattr(L, fn_name)
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    return fn(vp(q), n, n, vp(atoms.atoms), vp(atoms.b), vp(atoms.segs), atoms.nsegs, atoms.natoms, vp(ff),
              C.c_float(B), None, vp(F), vp(I), C.byref(tile) if tile is not None else None, None)


def _case(gpu, natoms, n, seed, tile):
    import torch

    rng = np.random.default_rng(seed)
    Z, pos, B = wl.random_molecule(natoms, 3e-9, rng, atom_b_range=(10.0, 40.0))
    det = wl.Detector(n, 75e-6, 0.1, 0.15e-9)
    q = dv.hkl_detector(n, n, det.width, det.width, det.distance, det.wavelength, device=gpu)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=B), gpu)
    ff = dv.form_factor_table(device=gpu)
    out = {}
    for name in ("spsim_hip_sf_grid", "spsim_hip_sf_grid_noalloc"):
        F = torch.full((n * n, 2), float("nan"), dtype=torch.float32, device=gpu)
        I = torch.full((n * n,), float("nan"), dtype=torch.float32, device=gpu)
        _lib.check(_call(name, q, n, atoms, ff, F, I, tile, B=2.0), name)
        torch.cuda.synchronize()
        out[name] = (F.cpu().numpy(), I.cpu().numpy())
    Fr, Ir = out["spsim_hip_sf_grid"]
    Fn, In = out["spsim_hip_sf_grid_noalloc"]
    assert np.array_equal(Fn.view(np.uint32), Fr.view(np.uint32))
    assert np.array_equal(In.view(np.uint32), Ir.view(np.uint32))
    sub = np.sort(rng.choice(n * n, 1024, replace=False))
    _, ref = oracle.sf_f64(q.cpu().numpy()[sub], Z, pos, atom_b=B, B=2.0, threads=oracle_threads())
    err = rel_l2(In[sub], ref)
    print(f"no-alloc grid: {natoms} atoms on {n}x{n}, tile {tile and (tile.pixels_per_lane, tile.chunk_atoms, tile.atom_splits)}, "
          f"rel err vs FP64 {err:.2e}", flush=True)
    assert err < 1e-5


def test_noalloc_matches_grid_two_calls(gpu):
    # first call: octants, 4 explicit atom ranges on 256 x 256 (4 x 65536 partials)
    _case(gpu, 3000, 256, 71, _lib.Tile(8, 4, 256, 4))
    # second call, smaller: automatic tile and split count on 96 x 96
    _case(gpu, 1200, 96, 72, None)


def test_noalloc_rejects_oversized_split(gpu):
    import torch

    n = 1024
    splits = SCRATCH_PARTIALS // (n * n) + 1  # one atom range more than the array holds at this size
    rng = np.random.default_rng(73)
    Z, pos, _ = wl.random_molecule(600, 3e-9, rng)
    q = dv.hkl_detector(n, n, n * 75e-6, n * 75e-6, 0.1, 0.15e-9, device=gpu)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos), gpu)
    ff = dv.form_factor_table(device=gpu)
    I = torch.zeros((n * n,), dtype=torch.float32, device=gpu)
    rc = _call("spsim_hip_sf_grid_noalloc", q, n, atoms, ff, None, I, _lib.Tile(8, 4, 64, splits))
    assert rc != 0
    assert "no-allocation" in _lib.last_error()
    torch.cuda.synchronize()
    assert float(I.abs().sum()) == 0.0  # nothing was launched
