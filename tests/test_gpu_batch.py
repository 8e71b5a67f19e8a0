"""Batched persistent launch (spsim_hip_batch_*) against the oracle, against
the plain per-pattern path, and under repeated hipGraph replays."""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import _lib
from spsim_amd import workloads as wl
from tests.conftest import oracle_threads, rel_l2

pytestmark = pytest.mark.gpu


def vp(a):
    return a.ctypes.data_as(C.c_void_p)


def make_plan(mols, rots, pattern_mol, q, nx, ny, atom_b=None, tile=None, b_global=0.0):
    offs = np.zeros(len(mols) + 1, np.int32)
    offs[1:] = np.cumsum([len(m[0]) for m in mols])
    Z = np.ascontiguousarray(np.concatenate([m[0] for m in mols]), np.int32)
    P = np.ascontiguousarray(np.concatenate([m[1] for m in mols]), np.float32).reshape(-1, 3)
    R = np.ascontiguousarray(np.asarray(rots).reshape(-1, 9), np.float64)
    pm = np.ascontiguousarray(pattern_mol, np.int32)
    plan = C.c_void_p()
    B = None if atom_b is None else np.ascontiguousarray(atom_b, np.float32)
    _lib.check(_lib.load().spsim_hip_batch_create(vp(q), nx, ny, len(mols), vp(offs), vp(Z), vp(P), None,
                                                  None if B is None else vp(B), 0, len(pm), vp(pm), vp(R), None,
                                                  float(b_global), C.byref(tile) if tile else None, C.byref(plan)),
               "batch_create")
    return plan, offs, Z, P


@pytest.fixture(scope="module", params=["centred", "off-centre"])
def setup(request):
    """Five molecules (one empty) shown in nine orientations on a 40 x 40
    detector; centred (mirror-symmetric: quad kernel) or off-centre (per-pixel)."""
    rng = np.random.default_rng(21)
    sizes = [1, 0, 333, 2000, 5000]
    mols = [wl.random_molecule(n, 2e-9 + 1e-9 * (n / 5000.0) ** (1 / 3), rng)[:2] if n else
            (np.zeros(0, np.int32), np.zeros((0, 3), np.float32)) for n in sizes]
    npat = 9
    pattern_mol = np.array([4, 3, 2, 1, 0, 4, 2, 3, 4], np.int32)
    rots = wl.uniform_rotations(npat, rng)
    det = wl.Detector(40, 5e-4, 0.1, 0.2e-9)
    cx = 0.0 if request.param == "centred" else 1.5e-4
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, cx, 0, 0, det.wavelength)
    return mols, rots, pattern_mol, q, det.n, request.param == "centred"


def reference(setup):
    mols, rots, pattern_mol, q, n, _ = setup
    out = []
    for p in range(len(pattern_mol)):
        Z, pos = mols[pattern_mol[p]]
        # batched patterns rotate q exactly (q' = R q in FP64)
        out.append(oracle.sf_f64(q, Z, pos, R=rots[p], threads=oracle_threads())[1] if len(Z) else np.zeros(q.shape[0]))
    return np.stack(out)


def test_batched_paths_match_oracle_and_each_other(gpu, setup):
    import torch

    mols, rots, pattern_mol, q, n, centred = setup
    ref = reference(setup)
    L = _lib.load()
    plan, *_ = make_plan(mols, rots, pattern_mol, q, n, n)
    try:
        assert L.spsim_hip_batch_uses_grid(plan) == int(centred)
        npat, npix = len(pattern_mol), q.shape[0]
        stream = torch.cuda.Stream()
        sp = C.c_void_p(stream.cuda_stream)
        I = torch.zeros((npat, npix), dtype=torch.float32, device=gpu)
        F = torch.zeros((npat, npix, 2), dtype=torch.float32, device=gpu)
        _lib.check(L.spsim_hip_batch_run(plan, C.c_void_p(F.data_ptr()), C.c_void_p(I.data_ptr()), 0, sp), "run")
        stream.synchronize()
        Ib = I.cpu().numpy()
        for p in range(npat):
            if ref[p].max() == 0:
                assert np.all(Ib[p] == 0)
            else:
                assert rel_l2(Ib[p], ref[p]) < 1e-5, p
        Fb = F.cpu().numpy()
        np.testing.assert_allclose(Fb[..., 0] ** 2 + Fb[..., 1] ** 2, Ib, rtol=1e-5, atol=1e-6 * Ib.max())
        # plain path: one single-pattern launch per pattern
        I2 = torch.zeros_like(I)
        _lib.check(L.spsim_hip_batch_run_plain(plan, None, C.c_void_p(I2.data_ptr()), sp), "plain")
        stream.synchronize()
        assert rel_l2(I2.cpu().numpy(), Ib) < 1e-6
        # graph replays, with eager work and synchronisation in between
        for rep in range(4):
            I.zero_()
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                _lib.check(L.spsim_hip_batch_run(plan, C.c_void_p(F.data_ptr()), C.c_void_p(I.data_ptr()), 1, sp),
                           "graph")
            stream.synchronize()
            assert np.array_equal(I.cpu().numpy(), Ib), rep
        terms = C.c_double()
        L.spsim_hip_batch_terms(plan, C.byref(terms))
        assert terms.value == sum(len(mols[m][0]) for m in pattern_mol) * npix
    finally:
        L.spsim_hip_batch_destroy(plan)


def test_autotune_and_every_tile(gpu, setup):
    import torch

    mols, rots, pattern_mol, q, n, centred = setup
    ref = reference(setup)
    L = _lib.load()
    plan, *_ = make_plan(mols, rots, pattern_mol, q, n, n)
    try:
        stream = torch.cuda.Stream()
        sp = C.c_void_p(stream.cuda_stream)
        best = _lib.Tile()
        rate = C.c_double()
        _lib.check(L.spsim_hip_batch_autotune(plan, sp, C.byref(best), C.byref(rate)), "autotune")
        # centred square detector: quads per lane 1, 2 or one octant (8) per lane
        assert best.pixels_per_lane in ((1, 2, 8) if centred else (1, 2, 4, 8)) and best.chunk_atoms in (64, 128, 256)
        assert rate.value > 0
        I = torch.zeros((len(pattern_mol), q.shape[0]), dtype=torch.float32, device=gpu)
        for k in ((1, 2, 8) if centred else (1, 2, 4, 8)):
            for c in (64, 128, 256):
                t = _lib.Tile(k, 4, c, 0)
                _lib.check(L.spsim_hip_batch_set_tile(plan, C.byref(t)), "set_tile")
                _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I.data_ptr()), 0, sp), "run")
                stream.synchronize()
                Ib = I.cpu().numpy()
                for p in range(len(pattern_mol)):
                    if ref[p].max() > 0:
                        assert rel_l2(Ib[p], ref[p]) < 1e-5, (k, c, p)
    finally:
        L.spsim_hip_batch_destroy(plan)


def test_batched_per_atom_b(gpu):
    import torch

    rng = np.random.default_rng(4)
    mols, Bs = [], []
    for n in (700, 1500):
        Z, pos, B = wl.random_molecule(n, 3e-9, rng, atom_b_range=(10.0, 40.0))
        mols.append((Z, pos))
        Bs.append(B)
    rots = wl.uniform_rotations(3, rng)
    pm = np.array([1, 0, 1], np.int32)
    det = wl.Detector(32, 1e-3, 0.1, 0.15e-9)
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    L = _lib.load()
    plan, *_ = make_plan(mols, rots, pm, q, det.n, det.n, atom_b=np.concatenate(Bs), b_global=4.0)
    try:
        I = torch.zeros((3, q.shape[0]), dtype=torch.float32, device=gpu)
        _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I.data_ptr()), 0, None), "run")
        torch.cuda.synchronize()
        for p in range(3):
            Z, pos = mols[pm[p]]
            _, ref = oracle.sf_f64(q, Z, pos, atom_b=Bs[pm[p]], B=4.0, R=rots[p])
            assert rel_l2(I.cpu().numpy()[p], ref) < 1e-5
    finally:
        L.spsim_hip_batch_destroy(plan)


def test_invalid_arguments_are_rejected(gpu):
    L = _lib.load()
    plan = C.c_void_p()
    q = np.zeros((4, 3), np.float32)
    offs = np.array([0, 1], np.int32)
    Z = np.array([6], np.int32)
    P = np.zeros(3, np.float32)
    R = np.eye(3).ravel()
    bad_pm = np.array([3], np.int32)
    rc = L.spsim_hip_batch_create(vp(q), 2, 2, 1, vp(offs), vp(Z), vp(P), None, None, 0, 1, vp(bad_pm), vp(R), None,
                                  0.0, None, C.byref(plan))
    assert rc == -1 and "out of range" in _lib.last_error()
    badZ = np.array([150], np.int32)
    pm = np.array([0], np.int32)
    rc = L.spsim_hip_batch_create(vp(q), 2, 2, 1, vp(offs), vp(badZ), vp(P), None, None, 0, 1, vp(pm), vp(R), None,
                                  0.0, None, C.byref(plan))
    assert rc == -1


def test_config3_full_workload(gpu):
    """Batched uneven instances at full size: 512 molecules with log-uniform
    atom counts in [2e3, 2e5] (radius 5 nm (N/5000)^(1/3)) at uniform random
    orientations on a 512 x 512 detector (75 um at 0.1 m, 0.2 nm). The
    autotuned persistent kernel, replayed from a hipGraph, against the FP64
    oracle on 2048 random pixels of 4 patterns (smallest, largest and two
    between), against the plain per-pattern path on every pattern, and
    identical over graph replays."""
    import torch

    det, mols, rots = wl.config3(512)
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength)
    npat = len(mols)
    L = _lib.load()
    plan, *_ = make_plan(mols, rots, np.arange(npat), q, det.n, det.n)
    try:
        stream = torch.cuda.Stream()
        sp = C.c_void_p(stream.cuda_stream)
        best = _lib.Tile()
        _lib.check(L.spsim_hip_batch_autotune(plan, sp, C.byref(best), None), "autotune")
        I = torch.zeros((npat, q.shape[0]), dtype=torch.float32, device=gpu)
        with torch.cuda.stream(stream):
            _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I.data_ptr()), 1, sp), "graph")
        stream.synchronize()
        Ib = I.cpu().numpy()
        sizes = np.array([len(m[0]) for m in mols])
        rng = np.random.default_rng(33)
        errs = []
        for p in (int(sizes.argmin()), int(sizes.argmax()), 100, 400):
            sub = np.sort(rng.choice(q.shape[0], 2048, replace=False))
            Z, pos = mols[p]
            _, ref = oracle.sf_f64(q[sub], Z, pos, R=rots[p], threads=oracle_threads())
            errs.append(rel_l2(Ib[p, sub], ref))
        print(f"config 3: tile {best.pixels_per_lane}/{best.chunk_atoms}, rel err vs FP64 (4 patterns x 2048 px) "
              f"max {max(errs):.2e}")
        assert max(errs) < 1e-5
        for rep in range(2):
            I.zero_()
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I.data_ptr()), 1, sp), "graph")
            stream.synchronize()
            assert np.array_equal(I.cpu().numpy(), Ib), rep
        I2 = torch.zeros_like(I)
        _lib.check(L.spsim_hip_batch_run_plain(plan, None, C.c_void_p(I2.data_ptr()), sp), "plain")
        stream.synchronize()
        I2h = I2.cpu().numpy()
        worst = max(rel_l2(I2h[p], Ib[p]) for p in range(npat))
        assert worst < 1e-6, worst
    finally:
        L.spsim_hip_batch_destroy(plan)


@pytest.mark.parametrize("shape", ["octants", "quads", "pixels"])
def test_atom_ranges_match_unsplit_plan_and_oracle(gpu, shape):
    """Patterns split into atom ranges (item bound 700 atoms: the 2000- and
    5000-atom molecules become 3 and 8 ranges per tile) against the same plan
    with no split, against the FP64 oracle, and over hipGraph replays (the
    tiles' counters clear themselves, so every replay equals the first)."""
    import torch

    rng = np.random.default_rng(31)
    sizes = [0, 1, 333, 2000, 5000]
    mols = [wl.random_molecule(n, 2e-9 + 1e-9 * (n / 5000.0) ** (1 / 3), rng)[:2] if n else
            (np.zeros(0, np.int32), np.zeros((0, 3), np.float32)) for n in sizes]
    pattern_mol = np.array([4, 3, 2, 1, 0, 4, 3], np.int32)
    npat = len(pattern_mol)
    rots = wl.uniform_rotations(npat, rng)
    nx, ny, cx = {"octants": (40, 40, 0.0), "quads": (48, 40, 0.0), "pixels": (40, 40, 1.5e-4)}[shape]
    q = oracle.hkl_detector(nx, ny, nx * 5e-4 / 40, ny * 5e-4 / 40, 0.1, cx, 0, 0, 0.2e-9)
    L = _lib.load()
    plan, *_ = make_plan(mols, rots, pattern_mol, q, nx, ny)
    try:
        stream = torch.cuda.Stream()
        sp = C.c_void_p(stream.cuda_stream)
        npix = q.shape[0]
        _lib.check(L.spsim_hip_batch_set_item_atoms(plan, -1), "no split")
        nitems, bound, split = C.c_int(), C.c_int(), C.c_int()
        L.spsim_hip_batch_describe(plan, C.byref(nitems), C.byref(bound), C.byref(split))
        assert bound.value == 0 and split.value == 0
        I0 = torch.zeros((npat, npix), dtype=torch.float32, device=gpu)
        _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I0.data_ptr()), 0, sp), "run unsplit")
        _lib.check(L.spsim_hip_batch_set_item_atoms(plan, 700), "split")
        L.spsim_hip_batch_describe(plan, C.byref(nitems), C.byref(bound), C.byref(split))
        assert bound.value == 700 and split.value > 0
        tile, blocks = _lib.Tile(), C.c_int()
        L.spsim_hip_batch_get_tile(plan, C.byref(tile), C.byref(blocks))
        print(f"{shape}: tile {tile.pixels_per_lane}/{tile.chunk_atoms}, {nitems.value} items, "
              f"{split.value} split tiles", flush=True)
        I1 = torch.full((npat, npix), float("nan"), dtype=torch.float32, device=gpu)
        _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I1.data_ptr()), 0, sp), "run split")
        stream.synchronize()
        a, b = I0.cpu().numpy(), I1.cpu().numpy()
        for p in range(npat):
            Z, pos = mols[pattern_mol[p]]
            if len(Z) == 0:
                assert np.all(b[p] == 0)
                continue
            assert rel_l2(b[p], a[p]) < 1e-6, p
            ref = oracle.sf_f64(q, Z, pos, R=rots[p], threads=oracle_threads())[1]
            assert rel_l2(b[p], ref) < 1e-5, p
        # graph replays with eager work on the device between them
        I2 = torch.full((npat, npix), float("nan"), dtype=torch.float32, device=gpu)
        first = None
        for rep in range(3):
            with torch.cuda.stream(stream):
                _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I2.data_ptr()), 1, sp), "graph")
            stream.synchronize()
            got = I2.cpu().numpy().copy()
            if first is None:
                first = got
                assert np.array_equal(got.view(np.uint32), b.view(np.uint32))
            else:
                assert np.array_equal(got.view(np.uint32), first.view(np.uint32)), rep
            I2.fill_(float("nan"))
            torch.cuda.synchronize()
    finally:
        L.spsim_hip_batch_destroy(plan)


def test_automatic_atom_ranges_on_a_skewed_batch(gpu):
    """One 60,000-atom molecule among fifteen 400-atom ones on a 128 x 128
    detector: the automatic item bound splits the large molecule's tiles (so
    its work spreads over the GPU) and leaves the small ones whole; every
    pattern against the FP64 oracle on a pixel sample, and the split plan
    against the unsplit one."""
    import torch

    rng = np.random.default_rng(32)
    big = wl.random_molecule(60000, 6e-9, rng)[:2]
    small = [wl.random_molecule(400, 1.5e-9, rng)[:2] for _ in range(15)]
    mols = [big] + small
    npat = len(mols)
    pattern_mol = np.arange(npat, dtype=np.int32)
    rots = wl.uniform_rotations(npat, rng)
    n = 128
    q = oracle.hkl_detector(n, n, n * 75e-6, n * 75e-6, 0.1, 0, 0, 0, 0.2e-9)
    L = _lib.load()
    plan, *_ = make_plan(mols, rots, pattern_mol, q, n, n)
    try:
        stream = torch.cuda.Stream()
        sp = C.c_void_p(stream.cuda_stream)
        nitems, bound, split = C.c_int(), C.c_int(), C.c_int()
        L.spsim_hip_batch_describe(plan, C.byref(nitems), C.byref(bound), C.byref(split))
        print(f"skewed batch: item bound {bound.value} atoms, {nitems.value} items, {split.value} split tiles",
              flush=True)
        assert 0 < bound.value < 60000 and split.value > 0
        I = torch.zeros((npat, n * n), dtype=torch.float32, device=gpu)
        with torch.cuda.stream(stream):
            _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I.data_ptr()), 1, sp), "graph")
        stream.synchronize()
        Ib = I.cpu().numpy()
        sub = np.sort(rng.choice(n * n, 1024, replace=False))
        for p in range(npat):
            Z, pos = mols[p]
            ref = oracle.sf_f64(q[sub], Z, pos, R=rots[p], threads=oracle_threads())[1]
            assert rel_l2(Ib[p, sub], ref) < 1e-5, p
        _lib.check(L.spsim_hip_batch_set_item_atoms(plan, -1), "no split")
        I0 = torch.zeros_like(I)
        _lib.check(L.spsim_hip_batch_run(plan, None, C.c_void_p(I0.data_ptr()), 0, sp), "unsplit")
        stream.synchronize()
        assert rel_l2(Ib, I0.cpu().numpy()) < 1e-6
    finally:
        L.spsim_hip_batch_destroy(plan)
