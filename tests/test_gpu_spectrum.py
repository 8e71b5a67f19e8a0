"""Wavelength-spectrum structure factor on the GPU (spsim_hip_sf_spectrum and
the drop-in dispatch of compute_sf) against the FP64 spectrum oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import _lib
from spsim_amd import device as dv
from spsim_amd import spsim
from spsim_amd import workloads as wl
from tests.conftest import oracle_threads, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-5


def setup_case(n=40, natoms=700, seed=0, lam=0.5e-9, with_b=True):
    rng = np.random.default_rng(seed)
    Z, pos, B = wl.random_molecule(natoms, 3e-9, rng, atom_b_range=(10.0, 40.0) if with_b else None)
    Z[:4] = 0  # skipped atoms
    q = oracle.hkl_detector(n, n, n * 2e-4, n * 2e-4, 0.05, 0.7e-4, -0.3e-4, 0, lam)
    return Z, pos, B, q


def run(gpu, q, Z, pos, sc, w, B=None, b_global=0.0, tile=None, R=None):
    import torch

    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=B), gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    F, I = dv.structure_factor_spectrum(hkl, atoms, ff, sc, w, B=b_global, tile=tile, R=R)
    torch.cuda.synchronize()
    Fh = F.cpu().numpy()
    return Fh[:, 0] + 1j * Fh[:, 1], I.cpu().numpy()


@pytest.mark.parametrize("nsamples,bw", [(5, 0.02), (9, 0.1), (1, 0.02), (4, 0.005)])
def test_spectrum_kernel_vs_f64(gpu, nsamples, bw):
    Z, pos, B, q = setup_case()
    sc, w = dv.spectrum_samples(0.5e-9, bw, nsamples)
    assert len(sc) == (nsamples + 1 if nsamples > 1 and nsamples % 2 == 0 else nsamples)
    F_ref, I_ref = oracle.sf_spectrum_f64(q, Z, pos, sc, w, atom_b=B, B=3.0, threads=oracle_threads())
    for tile in (None, dv.make_tile(1, 64, 1), dv.make_tile(2, 128, 3), dv.make_tile(2, 256, 2)):
        F, I = run(gpu, q, Z, pos, sc, w, B=B, b_global=3.0, tile=tile)
        assert rel_l2(I, I_ref) < TOL, tile
        assert rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag]) < TOL, tile


def test_spectrum_without_b_and_with_rotation(gpu):
    Z, pos, _, q = setup_case(seed=3, with_b=False)
    sc, w = dv.spectrum_samples(0.5e-9, 0.05, 7)
    R = wl.uniform_rotations(1, np.random.default_rng(8))[0]
    _, I_ref = oracle.sf_spectrum_f64(q, Z, pos, sc, w, R=R, threads=oracle_threads())
    _, I = run(gpu, q, Z, pos, sc, w, R=R)
    assert rel_l2(I, I_ref) < TOL


@pytest.mark.parametrize("nsamples,bw", [(7, 0.05), (33, 0.3), (65, 0.005)])
def test_spectrum_large_phases(gpu, nsamples, bw):
    """Phases |s_c q.r| up to about 60 turns (10 nm molecule, wide angles at
    0.15 nm). From 4 samples on the kernel takes the sample sum as
    e^{-2 pi i t} G(t) with G a fitted polynomial for |t| <= tlim =
    1.5 / (2 pi max|s_k/s_c - 1|), s_c the midpoint of the scales, and the
    per-sample FP64 sum beyond it: the three spectra put tlim at 9.5, 1.6 and 95
    turns, so both forms meet inside waves, and at 95 turns (beyond every phase)
    only the polynomial runs."""
    rng = np.random.default_rng(21)
    Z, pos, B = wl.random_molecule(1500, 10e-9, rng, atom_b_range=(5.0, 20.0))
    n = 40
    q = oracle.hkl_detector(n, n, n * 2.4e-3, n * 2.4e-3, 0.05, 1.1e-3, -0.7e-3, 0, 0.15e-9)
    sc, w = dv.spectrum_samples(0.15e-9, bw, nsamples)
    s_c = 0.5 * (np.min(sc) + np.max(sc))
    tlim = 1.5 / (2 * np.pi * np.max(np.abs(np.asarray(sc) / s_c - 1)))
    t = np.abs(s_c * (q.astype(np.float64) @ pos.T.astype(np.float64)))
    frac_far = np.mean(t > tlim)
    assert t.max() > 40
    if tlim < t.max():
        assert 0.05 < frac_far < 0.99, frac_far
    else:
        assert frac_far == 0.0
    F_ref, I_ref = oracle.sf_spectrum_f64(q, Z, pos, sc, w, atom_b=B, threads=oracle_threads())
    for tile in (None, dv.make_tile(2, 128, 1)):
        F, I = run(gpu, q, Z, pos, sc, w, B=B, tile=tile)
        assert rel_l2(I, I_ref) < TOL, tile
        assert rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag]) < TOL, tile


@pytest.mark.parametrize("nsamples,lo,hi,seed", [(4, 0.97, 1.05, 1), (6, 0.8, 1.25, 2), (17, 1.6, 1.7, 3)])
def test_spectrum_arbitrary_samples(gpu, nsamples, lo, hi, seed):
    """Caller-given scales and weights (spsim_hip_sf_spectrum takes any): unevenly
    spaced, asymmetric about 1 or far from it, with unnormalised random weights, so
    the fitted sample factor is centred on the scales' midpoint and spans an
    uneven set. Against the FP64 spectrum oracle given the same samples."""
    rng = np.random.default_rng(30 + seed)
    Z, pos, B, q = setup_case(seed=seed)
    sc = np.sort(rng.uniform(lo, hi, nsamples))
    w = rng.uniform(0.1, 1.0, nsamples)
    F_ref, I_ref = oracle.sf_spectrum_f64(q, Z, pos, sc, w, atom_b=B, threads=oracle_threads())
    for tile in (None, dv.make_tile(1, 128, 2)):
        F, I = run(gpu, q, Z, pos, sc, w, B=B, tile=tile)
        assert rel_l2(I, I_ref) < TOL, tile
        assert rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag]) < TOL, tile


def test_single_sample_matches_plain_kernel(gpu):
    import torch

    Z, pos, B, q = setup_case(seed=4)
    _, I_s = run(gpu, q, Z, pos, [1.0], [1.0], B=B)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=B), gpu)
    _, I_p = dv.structure_factor(torch.from_numpy(q).to(gpu), atoms, dv.form_factor_table(device=gpu))
    assert rel_l2(I_s, I_p.cpu().numpy()) < 1e-6


def test_invalid_sample_counts(gpu):
    import torch

    Z, pos, B, q = setup_case(n=4, natoms=10)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos), gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    I = torch.empty(q.shape[0], device=gpu)
    L = _lib.load()
    sc = np.ones(70)
    for n in (0, _lib.MAX_SPECTRUM + 1):
        rc = L.spsim_hip_sf_spectrum(C.c_void_p(hkl.data_ptr()), q.shape[0], C.c_void_p(atoms.atoms.data_ptr()), None,
                                     C.c_void_p(atoms.segs.data_ptr()), atoms.nsegs, atoms.natoms,
                                     C.c_void_p(ff.data_ptr()), 0.0, None, n, sc.ctypes.data_as(C.c_void_p),
                                     sc.ctypes.data_as(C.c_void_p), None, C.c_void_p(I.data_ptr()), None, None)
        assert rc == -1 and "nsamples" in _lib.last_error()


def test_compute_sf_takes_the_spectrum_path(gpu):
    """use_cuda with a bandwidth and several wavelength samples: the reference's
    GPU spectrum branch (src/diffraction.cu:313-318); an even sample count is
    raised in the caller's options."""
    opts = spsim.set_defaults()
    d = opts.detector
    d.nx = d.ny = 24
    d.nz = 1
    d.pixel_width = d.pixel_height = 1e-3
    d.width = d.height = float(np.float32(24 * 1e-3))
    d.distance = 0.05
    opts.experiment.wavelength = 0.2e-9
    opts.experiment.bandwidth = 0.05
    opts.wavelength_samples = 4
    opts.use_cuda = 1
    opts.verbosity_level = 0
    rng = np.random.default_rng(6)
    Z, pos, _ = wl.random_molecule(300, 2.5e-9, rng)
    mol = spsim.get_molecule_from_atoms(Z, pos)
    n = C.c_int(0)
    ptr = spsim.get_HKL_list_for_detector(opts.detector, opts.experiment, C.byref(n))
    q = np.ctypeslib.as_array(ptr, shape=(n.value * 3,)).reshape(-1, 3).copy()
    pat = spsim.compute_sf(mol, ptr, n.value, opts)
    assert opts.wavelength_samples == 5
    sc, w = oracle.spectrum_samples(float(np.float32(0.2e-9)), float(np.float32(0.05)), 5)
    _, I_ref = oracle.sf_spectrum_f64(q, Z, pos, sc, w)
    assert rel_l2(spsim.float_array(pat.ints, n.value), I_ref) < TOL
    _, I_plain = oracle.sf_f64(q, Z, pos)
    assert rel_l2(I_ref, I_plain) > 1e-4  # the bandwidth matters at this q range
    spsim.free_diffraction_pattern(pat)
    # bandwidth 0: the plain path
    opts.experiment.bandwidth = 0.0
    pat = spsim.compute_sf(mol, ptr, n.value, opts)
    assert rel_l2(spsim.float_array(pat.ints, n.value), I_plain) < TOL
    spsim.free_diffraction_pattern(pat)


@pytest.mark.parametrize("nsamples,bw,rotate", [(5, 0.02, False), (9, 0.1, True), (1, 0.02, False)])
def test_spectrum_nfft_vs_f64(gpu, nsamples, bw, rotate):
    """The spectrum through the NUFFT (spsim_hip_sf_spectrum_nfft): every sample's
    sums from one set of grids at q'_k = s_k R q, then the weighted sum. Against
    the FP64 spectrum oracle, with the global B and skipped Z = 0 atoms."""
    import torch

    Z, pos, _, q = setup_case(seed=5, with_b=False)
    sc, w = dv.spectrum_samples(0.5e-9, bw, nsamples)
    R = wl.uniform_rotations(1, np.random.default_rng(9))[0] if rotate else None
    F_ref, I_ref = oracle.sf_spectrum_f64(q, Z, pos, sc, w, B=3.0, R=R, threads=oracle_threads())
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos), gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    F, I, info = dv.structure_factor_spectrum_nfft(hkl, atoms, ff, sc, w, B=3.0, R=R)
    torch.cuda.synchronize()
    Fh = F.cpu().numpy()
    assert rel_l2(I.cpu().numpy(), I_ref) < TOL, info
    assert rel_l2(Fh, np.c_[F_ref.real, F_ref.imag]) < TOL, info
    _, I_only, _ = dv.structure_factor_spectrum_nfft(hkl, atoms, ff, sc, w, B=3.0, R=R, want_F=False)
    torch.cuda.synchronize()
    # the spread bins atoms with global atomics, so calls agree to rounding, not bitwise
    assert rel_l2(I_only.cpu().numpy(), I.cpu().numpy()) < 1e-6
    with pytest.raises(ValueError, match="per-atom B"):
        dv.structure_factor_spectrum_nfft(hkl, dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=np.ones(len(Z), np.float32)),
                                                              gpu), ff, sc, w)
    with pytest.raises(RuntimeError, match="scale"):
        dv.structure_factor_spectrum_nfft(hkl, atoms, ff, [1.0, -1.0], [0.5, 0.5])


def test_nfft_batch_scaled_rotations(gpu):
    """spsim_hip_sf_nfft_batch with matrices s_p R_p (s_p from 0.6 to 1.4): the
    grids cover the ball of radius max |q| max_p s_p, so every pattern matches
    the FP64 oracle's sum at q' = s_p R_p q (form factors at the unscaled |q|,
    the spectrum oracle with one sample)."""
    import torch

    Z, pos, _, q = setup_case(seed=6, with_b=False)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos), gpu)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    Rs = wl.uniform_rotations(3, np.random.default_rng(10))
    scales = (1.4, 0.6, 1.0)
    F, _, info = dv.structure_factor_nfft_batch(hkl, atoms, ff, [s * R for s, R in zip(scales, Rs)], B=1.0)
    torch.cuda.synchronize()
    Fh = F.cpu().numpy()
    for p, (s, R) in enumerate(zip(scales, Rs)):
        F_ref, _ = oracle.sf_spectrum_f64(q, Z, pos, [s], [1.0], B=1.0, R=R, threads=oracle_threads())
        assert rel_l2(Fh[p], np.c_[F_ref.real, F_ref.imag]) < TOL, (p, s, info)
