"""The non-uniform FFT structure factor on the GPU (spsim_hip_sf_nfft,
spsim_hip_fft_pruned and the compute_pattern_on_list_by_nfft drop-in) against
numpy's FFT and the FP64 oracle of the sums the reference's NNFFT path
approximates (oracle.sf_nfft_f64)."""
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


def fft_reference(x, m_lo, nf, L, l_lo):
    outer, M, inner = x.shape
    line = np.zeros((outer, nf, inner), np.complex128)
    line[:, (m_lo + np.arange(M)) % nf, :] = x
    f = np.fft.fft(line, axis=1)
    return f[:, (l_lo + np.arange(L)) % nf, :]


@pytest.mark.parametrize("nf,M,m_lo,L,l_lo,outer,inner", [
    (2, 1, 0, 2, 0, 3, 1), (3, 2, -1, 3, -1, 2, 5), (4, 4, -2, 4, 0, 4, 1), (5, 3, -1, 5, -2, 3, 2),
    (6, 3, -1, 7, -3, 2, 1), (12, 6, -3, 8, -4, 5, 3), (30, 15, -7, 20, -10, 4, 1), (60, 30, -15, 35, -17, 3, 17),
    (96, 48, -24, 60, -30, 6, 1), (100, 50, -25, 60, -30, 2, 9), (125, 60, -30, 70, -35, 3, 1),
    (480, 240, -120, 250, -125, 4, 1), (480, 240, -120, 250, -125, 2, 33), (1000, 500, -250, 520, -260, 2, 1),
    (2048, 1000, -500, 1030, -515, 2, 3), (3840, 1900, -950, 1930, -965, 1, 2),
    (64, 10, 60, 12, 110, 3, 2), (8, 5, -2, 20, -8, 3, 4), (16, 16, 16, 40, -16, 2, 1)])
def test_fft_pruned_matches_numpy(gpu, nf, M, m_lo, L, l_lo, outer, inner):
    import torch

    rng = np.random.default_rng(nf * 7 + inner)
    x = (rng.normal(size=(outer, M, inner)) + 1j * rng.normal(size=(outer, M, inner))).astype(np.complex64)
    out = dv.fft_pruned(torch.from_numpy(x).to(gpu), M, m_lo, nf, L, l_lo).cpu().numpy()
    ref = fft_reference(x.astype(np.complex128), m_lo, nf, L, l_lo)
    assert rel_l2(np.c_[out.real.ravel(), out.imag.ravel()], np.c_[ref.real.ravel(), ref.imag.ravel()]) < 2e-6


def smooth_even_lengths(lo, hi):
    out = []
    for n in range(lo, hi + 1, 2):
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1:
            out.append(n)
    return out


@pytest.mark.parametrize("inner", [1, 6])
def test_fft_pruned_every_smooth_length(gpu, inner):
    """Every even 2,3,5-smooth line length from 16 to 1920 (each has a kernel of
    its own when its plan takes at most 3 stages of radices up to 16, the
    general kernel otherwise), contiguous and strided, on half-filled lines
    with about half the outputs kept, against numpy's FFT."""
    import torch

    rng = np.random.default_rng(inner)
    worst = 0.0
    for nf in smooth_even_lengths(16, 1920):
        M, L = nf // 2 - 1, nf // 2 + 3
        m_lo, l_lo = -(M // 2), -(L // 2) + 1
        outer = 3
        x = (rng.normal(size=(outer, M, inner)) + 1j * rng.normal(size=(outer, M, inner))).astype(np.complex64)
        out = dv.fft_pruned(torch.from_numpy(x).to(gpu), M, m_lo, nf, L, l_lo).cpu().numpy()
        ref = fft_reference(x.astype(np.complex128), m_lo, nf, L, l_lo)
        err = rel_l2(np.c_[out.real.ravel(), out.imag.ravel()], np.c_[ref.real.ravel(), ref.imag.ravel()])
        assert err < 2e-6, (nf, err)
        worst = max(worst, err)
    print("worst relative L2 error", worst)


def test_fft_pruned_rejects_bad_lengths(gpu):
    import torch

    x = torch.zeros((1, 4, 1), dtype=torch.complex64, device=gpu)
    for nf in (7, 4, 4096):  # not smooth / shorter than M / longer than 3840
        M = 4 if nf != 4 else 5
        xx = x if M == 4 else torch.zeros((1, 5, 1), dtype=torch.complex64, device=gpu)
        with pytest.raises(RuntimeError):
            dv.fft_pruned(xx, M, 0, nf, 4, 0)
    # points or outputs beyond the wrapped range [-nf, 2 nf)
    with pytest.raises(RuntimeError, match="outside"):
        dv.fft_pruned(x, 4, -100, 64, 4, 0)
    with pytest.raises(RuntimeError, match="outside"):
        dv.fft_pruned(x, 4, 0, 64, 4, 126)


def molecule(n, radius, seed, z0=3):
    rng = np.random.default_rng(seed)
    Z, pos, _ = wl.random_molecule(n, radius, rng)
    Z[:z0] = 0
    return Z, pos


def detector_q(n=40, lam=0.3e-9, pix=1e-4, dist=0.05, cx=0.0, cy=0.0):
    return oracle.hkl_detector(n, n, n * pix, n * pix, dist, cx, cy, 0, lam)


def run_nfft(gpu, q, Z, pos, B=0.0, eps=1e-6, delta=False, R=None):
    import torch

    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, delta_atoms=delta, mode=_lib.PACK_NFFT), gpu)
    F, I, info = dv.structure_factor_nfft(torch.from_numpy(np.ascontiguousarray(q, np.float32)).to(gpu), atoms,
                                          dv.form_factor_table(device=gpu), B=B, eps=eps, R=R)
    torch.cuda.synchronize()
    Fh = F.cpu().numpy()
    return Fh[:, 0] + 1j * Fh[:, 1], I.cpu().numpy(), info


def cplx_err(F, F_ref):
    return rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag])


@pytest.mark.parametrize("B", [0.0, 12.0])
def test_nfft_vs_exact_detector(gpu, B):
    Z, pos = molecule(900, 4e-9, seed=1)
    q = detector_q(48, cx=1e-4, cy=-2e-4)
    F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=B, threads=oracle_threads())
    F, I, info = run_nfft(gpu, q, Z, pos, B=B)
    # one pattern of few q: the FFT dominates, so q interpolates with the spread window
    assert info["W"] == 7 and info["W2"] == 7 and all(n >= 2 * m for n, m in zip(info["nf"], info["M"]))
    assert rel_l2(I, I_ref) < 1e-5, info
    assert cplx_err(F, F_ref) < 1e-5


def test_nfft_error_follows_eps(gpu):
    Z, pos = molecule(600, 3e-9, seed=2)
    q = detector_q(40, lam=0.2e-9)
    F_ref, _ = oracle.sf_nfft_f64(q, Z, pos, threads=oracle_threads())
    errs = []
    for eps in (1e-2, 1e-3, 1e-4, 1e-5):
        F, _, info = run_nfft(gpu, q, Z, pos, eps=eps)
        errs.append(cplx_err(F, F_ref))
        assert errs[-1] < 3 * eps, (eps, info, errs)
    assert errs[0] > errs[-1] * 10


def test_nfft_three_d_volume_and_rotation(gpu):
    Z, pos = molecule(500, 3e-9, seed=3)
    q = oracle.hkl_3d(16, 16, 16, 16e-4, 16e-4, 16e-4, 0.05, 0.3e-9)
    R = wl.uniform_rotations(1, np.random.default_rng(4))[0]
    q = oracle.rotate_hkl(q, R)
    F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=4.0, threads=oracle_threads())
    F, I, info = run_nfft(gpu, q, Z, pos, B=4.0)
    assert rel_l2(I, I_ref) < 1e-5, info
    assert cplx_err(F, F_ref) < 1e-5


def test_nfft_exact_rotation(gpu):
    """rot9 rotates q exactly (applied as R^T to the atoms in FP64), as in the
    direct entry points; the result matches the oracle's unrounded R q."""
    Z, pos = molecule(700, 3.5e-9, seed=12)
    q = detector_q(40, lam=0.2e-9)
    rng = np.random.default_rng(13)
    for R in wl.uniform_rotations(3, rng):
        F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=5.0, R=R, threads=oracle_threads())
        F, I, info = run_nfft(gpu, q, Z, pos, B=5.0, R=R)
        assert rel_l2(I, I_ref) < 1e-5, info
        assert cplx_err(F, F_ref) < 1e-5
    # the identity rotation is the unrotated result
    F0, _, _ = run_nfft(gpu, q, Z, pos, B=5.0)
    F1, _, _ = run_nfft(gpu, q, Z, pos, B=5.0, R=np.eye(3))
    assert cplx_err(F1, F0) < 1e-6
    with pytest.raises(RuntimeError, match="rotation"):
        run_nfft(gpu, q, Z, pos, R=np.full((3, 3), np.nan))


def test_nfft_delta_atoms(gpu):
    Z, pos = molecule(400, 3e-9, seed=5, z0=10)
    q = detector_q(32)
    F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, delta=True, B=20.0, threads=oracle_threads())
    F, I, _ = run_nfft(gpu, q, Z, pos, B=20.0, delta=True)
    assert rel_l2(I, I_ref) < 1e-5
    assert cplx_err(F, F_ref) < 1e-5


def test_nfft_degenerate_geometries(gpu):
    q = detector_q(24)
    # planar molecule: every atom at z = 0 (zero-width box along z)
    Z, pos = molecule(300, 3e-9, seed=7, z0=0)
    pos[:, 2] = 0.0
    F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos)
    _, I, info = run_nfft(gpu, q, Z, pos)
    assert rel_l2(I, I_ref) < 1e-5, info
    # one atom: F = f(|q|) exp(-2 pi i q.r)
    p1 = np.array([[1.5e-9, -0.5e-9, 2e-9]], np.float32)
    F_ref, _ = oracle.sf_nfft_f64(q, [8], p1, B=3.0)
    F, _, _ = run_nfft(gpu, q, np.array([8], np.int32), p1, B=3.0)
    assert cplx_err(F, F_ref) < 1e-5
    # every q the same point
    qs = np.repeat(q[100:101], 50, axis=0)
    F_ref, _ = oracle.sf_nfft_f64(qs, Z, pos)
    F, _, _ = run_nfft(gpu, qs, Z, pos)
    assert cplx_err(F, F_ref) < 1e-5
    # q = 0 only: F = sum of f(0); the error bound is about eps sum |f| (eps = 1e-6)
    q0 = np.zeros((8, 3), np.float32)
    F_ref, _ = oracle.sf_nfft_f64(q0, Z, pos)
    F, _, _ = run_nfft(gpu, q0, Z, pos)
    assert cplx_err(F, F_ref) < 3e-6


def test_nfft_empty_inputs(gpu):
    import torch

    q = detector_q(8)
    F, I, _ = run_nfft(gpu, q, np.zeros(0, np.int32), np.zeros((0, 3), np.float32))
    assert np.all(F == 0) and np.all(I == 0)
    Z, pos = molecule(20, 1e-9, seed=8)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT), gpu)
    hkl = torch.zeros((0, 3), dtype=torch.float32, device=gpu)
    F, I, _ = dv.structure_factor_nfft(hkl, atoms, dv.form_factor_table(device=gpu))
    assert F.shape == (0, 2) and I.shape == (0,)


def test_nfft_rejects_bad_inputs(gpu):
    import torch

    Z, pos = molecule(50, 2e-9, seed=9)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT), gpu)
    ff = dv.form_factor_table(device=gpu)
    q = detector_q(8).copy()
    q[5, 1] = np.nan
    with pytest.raises(RuntimeError, match="non-finite"):
        dv.structure_factor_nfft(torch.from_numpy(q).to(gpu), atoms, ff)
    # a box x q-range product beyond the longest FFT line
    big = (detector_q(8) * 1e4).astype(np.float32)
    with pytest.raises(RuntimeError, match="too large|exceeds"):
        dv.structure_factor_nfft(torch.from_numpy(big).to(gpu), atoms, ff)
    # per-atom B factors must be finite
    bad_b = np.ones(50, np.float32)
    bad_b[7] = np.nan
    with pytest.raises(RuntimeError, match="non-finite B"):
        dv.structure_factor_nfft(torch.from_numpy(detector_q(8)).to(gpu),
                                 dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=bad_b), gpu), ff)


def test_compute_sf_takes_the_nfft_path(gpu):
    """use_nfft_for_sf selects compute_pattern_on_list_by_nfft (src/spsim.c:119-122),
    which keeps the Z == 0 atoms and ignores the illumination profile."""
    opts = spsim.set_defaults()
    d = opts.detector
    d.nx = d.ny = 32
    d.nz = 1
    d.pixel_width = d.pixel_height = 1e-4
    d.width = d.height = float(np.float32(32 * 1e-4))
    d.distance = 0.05
    opts.experiment.wavelength = 0.25e-9
    opts.use_nfft_for_sf = 1
    opts.b_factor = 6.0
    opts.verbosity_level = 0
    Z, pos = molecule(700, 3.5e-9, seed=10, z0=6)
    mol = spsim.get_molecule_from_atoms(Z, pos)
    n = C.c_int(0)
    ptr = spsim.get_HKL_list_for_detector(opts.detector, opts.experiment, C.byref(n))
    q = np.ctypeslib.as_array(ptr, shape=(n.value * 3,)).reshape(-1, 3).copy()
    pat = spsim.compute_sf(mol, ptr, n.value, opts)
    assert pat
    I = spsim.float_array(pat.ints, n.value)
    _, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=6.0)
    assert rel_l2(I, I_ref) < 1e-5
    np.testing.assert_array_equal(np.ctypeslib.as_array(pat.HKL_list, shape=(n.value * 3,)), q.ravel())
    spsim.free_diffraction_pattern(pat)
    # the direct path skips the Z == 0 atoms, so the two differ
    _, I_plain = oracle.sf_f64(q, Z, pos, B=6.0)
    assert rel_l2(I_plain, I_ref) > 1e-3


def test_nfft_long_lines_many_species_and_widest_window(gpu):
    """A 200 nm needle along x gives FFT lines of more than 1920 points (the
    twiddles then come from global memory); 12 elements give 12 species grids;
    eps = 1e-9 selects the widest window (W = 10), where FP32 sets the floor."""
    import torch

    rng = np.random.default_rng(11)
    n = 1500
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = rng.uniform(-100e-9, 100e-9, n)
    pos[:, 1:] = rng.uniform(-0.4e-9, 0.4e-9, (n, 2))
    Z = rng.choice([1, 6, 7, 8, 11, 12, 15, 16, 17, 19, 20, 26], n).astype(np.int32)
    q = detector_q(40, lam=0.1e-9, pix=4e-4)  # |q| up to 1.6e9 per m: x lines of 2500 points
    F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=2.0, threads=oracle_threads())
    F, I, info = run_nfft(gpu, q, Z, pos, B=2.0)
    assert info["nf"][0] > 1920, info
    assert rel_l2(I, I_ref) < 1e-5, info
    assert cplx_err(F, F_ref) < 1e-5
    F, I, info = run_nfft(gpu, q, Z, pos, B=2.0, eps=1e-9)
    assert info["W"] == 10
    assert cplx_err(F, F_ref) < 3e-6, info
    # intensities only
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT), gpu)
    F0, I0, _ = dv.structure_factor_nfft(torch.from_numpy(q).to(gpu), atoms, dv.form_factor_table(device=gpu),
                                         B=2.0, want_F=False)
    assert F0 is None
    assert rel_l2(I0.cpu().numpy(), I_ref) < 1e-5


def test_nfft_batch_of_orientations(gpu):
    """spsim_hip_sf_nfft_batch: one species grid set over the ball |q| <= max|q|,
    interpolated at R_p q for every pattern; each pattern matches the FP64
    oracle with that rotation and the single-pattern call with rot9."""
    import torch

    Z, pos = molecule(800, 3.5e-9, seed=14)
    q = detector_q(36, lam=0.25e-9)
    rots = wl.uniform_rotations(4, np.random.default_rng(15))
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT), gpu)
    ff = dv.form_factor_table(device=gpu)
    hkl = torch.from_numpy(q).to(gpu)
    F, I, info = dv.structure_factor_nfft_batch(hkl, atoms, ff, rots, B=3.0)
    torch.cuda.synchronize()
    Fh, Ih = F.cpu().numpy(), I.cpu().numpy()
    assert Fh.shape == (4, q.shape[0], 2) and info["W"] == 7
    for p, R in enumerate(rots):
        F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=3.0, R=R, threads=oracle_threads())
        Fp = Fh[p, :, 0] + 1j * Fh[p, :, 1]
        assert rel_l2(Ih[p], I_ref) < 1e-5, (p, info)
        assert cplx_err(Fp, F_ref) < 1e-5
        F1, _, _ = run_nfft(gpu, q, Z, pos, B=3.0, R=R)
        assert cplx_err(Fp, F1) < 3e-6
    # intensities only, and bad arguments
    _, I2, _ = dv.structure_factor_nfft_batch(hkl, atoms, ff, rots[:2], B=3.0, want_F=False)
    assert rel_l2(I2.cpu().numpy(), Ih[:2]) < 1e-6
    with pytest.raises(RuntimeError, match="rotation"):
        dv.structure_factor_nfft_batch(hkl, atoms, ff, np.full((1, 3, 3), np.inf))
    with pytest.raises(RuntimeError, match="npatterns"):
        dv.structure_factor_nfft_batch(hkl, atoms, ff, np.zeros((0, 9)))


def test_nfft_batch_narrow_interpolation_window(gpu):
    """Many orientations over one grid: the interpolation dominates the cost, so
    the plan takes the narrow second window (width W - 1 on FFT lines of at least
    4 M points), with the error of the default one. At W = 6 (eps 1e-5) the
    window stays."""
    import torch

    Z, pos = molecule(600, 3e-9, seed=21)
    q = detector_q(48, lam=0.25e-9)
    rots = wl.uniform_rotations(24, np.random.default_rng(22))
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT), gpu)
    ff = dv.form_factor_table(device=gpu)
    hkl = torch.from_numpy(q).to(gpu)
    F, I, info = dv.structure_factor_nfft_batch(hkl, atoms, ff, rots, B=2.0)
    torch.cuda.synchronize()
    assert info["W"] == 7 and info["W2"] == 6, info
    assert all(n >= 4 * m for n, m in zip(info["nf"], info["M"])), info
    Fh, Ih = F.cpu().numpy(), I.cpu().numpy()
    for p in (0, 11, 23):
        F_ref, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=2.0, R=rots[p], threads=oracle_threads())
        assert rel_l2(Ih[p], I_ref) < 1e-5, (p, info)
        assert cplx_err(Fh[p, :, 0] + 1j * Fh[p, :, 1], F_ref) < 1e-5, (p, info)
    F6, _, info6 = dv.structure_factor_nfft_batch(hkl, atoms, ff, rots[:3], B=2.0, eps=1e-5)
    torch.cuda.synchronize()
    assert info6["W"] == 6 and info6["W2"] == 6, info6
    F6h = F6.cpu().numpy()
    F_ref, _ = oracle.sf_nfft_f64(q, Z, pos, B=2.0, R=rots[2], threads=oracle_threads())
    assert cplx_err(F6h[2, :, 0] + 1j * F6h[2, :, 1], F_ref) < 3e-5


def test_nfft_staged_interpolation_on_a_large_detector(gpu):
    """A 512 x 512 detector over a small molecule: 64 consecutive pixels of a row
    share most of their windows, so the interpolation stages each wave's grid box
    in LDS. One pattern and a batch of orientations match the FP64 oracle on a
    pixel sample, and the batch agrees with one call per orientation."""
    import torch

    Z, pos = molecule(500, 3e-9, seed=31)
    q = detector_q(512, lam=0.25e-9, pix=2e-5)
    rng = np.random.default_rng(32)
    sample = np.sort(rng.choice(q.shape[0], 1536, replace=False))
    F, I, info = run_nfft(gpu, q, Z, pos, B=2.0)
    assert info["staged"] == 1, info
    F_ref, I_ref = oracle.sf_nfft_f64(q[sample], Z, pos, B=2.0, threads=oracle_threads())
    assert rel_l2(I[sample], I_ref) < 1e-5, info
    assert cplx_err(F[sample], F_ref) < 1e-5, info
    rots = wl.uniform_rotations(3, np.random.default_rng(33))
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT), gpu)
    ff = dv.form_factor_table(device=gpu)
    Fb, _, infob = dv.structure_factor_nfft_batch(torch.from_numpy(q).to(gpu), atoms, ff, rots, B=2.0)
    torch.cuda.synchronize()
    assert infob["staged"] == 1, infob
    Fbh = Fb.cpu().numpy()
    for p, R in enumerate(rots):
        Fp = Fbh[p, :, 0] + 1j * Fbh[p, :, 1]
        F_ref, _ = oracle.sf_nfft_f64(q[sample], Z, pos, B=2.0, R=R, threads=oracle_threads())
        assert cplx_err(Fp[sample], F_ref) < 1e-5, (p, infob)
        F1, _, _ = run_nfft(gpu, q, Z, pos, B=2.0, R=R)
        assert cplx_err(Fp, F1) < 3e-6


def test_nfft_per_atom_b_factors(gpu):
    """Per-atom B factors through the NUFFT: every atom is spread as a Gaussian
    whose transform carries its Debye-Waller factor exp(-B_j s^2). With plain
    packing (Z = 0 skipped, amplitudes kept) the result is the direct sum's:
    one off-centre pattern and a batch of orientations match the FP64 oracle
    with the same per-atom B; a looser eps stays within its error; B factors
    too wide for the grid are rejected."""
    import torch

    Z, pos = molecule(900, 4e-9, seed=41)
    rng = np.random.default_rng(42)
    B = rng.uniform(10, 40, len(Z)).astype(np.float32)
    q = detector_q(48, lam=0.15e-9, pix=2e-4, cx=1e-4, cy=-5e-5)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=B), gpu)
    ff = dv.form_factor_table(device=gpu)
    hkl = torch.from_numpy(q).to(gpu)
    F, I, info = dv.structure_factor_nfft(hkl, atoms, ff, B=2.0)
    torch.cuda.synchronize()
    assert info["Wg"] in (17, 25, 33), info
    Fh = F.cpu().numpy()
    F_ref, I_ref = oracle.sf_f64(q, Z, pos, atom_b=B, B=2.0, threads=oracle_threads())
    assert rel_l2(I.cpu().numpy(), I_ref) < 1e-5, info
    assert cplx_err(Fh[:, 0] + 1j * Fh[:, 1], F_ref) < 1e-5, info
    # a looser accuracy target: narrower Gaussians, error within it
    F4, _, info4 = dv.structure_factor_nfft(hkl, atoms, ff, B=2.0, eps=1e-4)
    torch.cuda.synchronize()
    assert info4["Wg"] <= info["Wg"], (info4, info)
    F4h = F4.cpu().numpy()
    assert cplx_err(F4h[:, 0] + 1j * F4h[:, 1], F_ref) < 3e-4
    # many orientations over one grid
    rots = wl.uniform_rotations(3, np.random.default_rng(43))
    Fb, _, infob = dv.structure_factor_nfft_batch(hkl, atoms, ff, rots, B=2.0)
    torch.cuda.synchronize()
    assert infob["Wg"] in (17, 25, 33), infob
    Fbh = Fb.cpu().numpy()
    for p, R in enumerate(rots):
        F_ref, _ = oracle.sf_f64(q, Z, pos, atom_b=B, B=2.0, R=R, threads=oracle_threads())
        assert cplx_err(Fbh[p, :, 0] + 1j * Fbh[p, :, 1], F_ref) < 1e-5, (p, infob)
    # B = 0 for every atom: the exact Gaussian spread of point atoms
    atoms0 = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=np.zeros(len(Z), np.float32)), gpu)
    F0, _, info0 = dv.structure_factor_nfft(hkl, atoms0, ff)
    torch.cuda.synchronize()
    F_ref0, _ = oracle.sf_f64(q, Z, pos, threads=oracle_threads())
    F0h = F0.cpu().numpy()
    assert info0["Wg"] == 17 and cplx_err(F0h[:, 0] + 1j * F0h[:, 1], F_ref0) < 1e-5, info0
    # B far too wide for the grid spacing
    big = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=np.full(len(Z), 5e4, np.float32)), gpu)
    with pytest.raises(RuntimeError, match="Gaussian window"):
        dv.structure_factor_nfft(hkl, big, ff)
    neg = B.copy()
    neg[3] = -1.0
    with pytest.raises(RuntimeError, match="negative B"):
        dv.structure_factor_nfft(hkl, dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=neg), gpu), ff)


def test_nfft_per_atom_b_rotation_volume_and_high_b(gpu):
    """Per-atom B with an exact rotation (rot9), on a rotated 3-D q volume (every
    axis off-centre in frequency), and with large B factors that need the
    25-point Gaussian window: each against the FP64 direct oracle."""
    import torch

    ff = dv.form_factor_table(device=gpu)
    Z, pos = molecule(700, 3.5e-9, seed=51)
    rng = np.random.default_rng(52)
    B = rng.uniform(0, 60, len(Z)).astype(np.float32)
    atoms = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=B), gpu)
    # exact rotation of a detector's q list
    q = detector_q(40, lam=0.2e-9, pix=2e-4)
    R = wl.uniform_rotations(1, np.random.default_rng(53))[0]
    F, _, info = dv.structure_factor_nfft(torch.from_numpy(q).to(gpu), atoms, ff, B=1.0, R=R)
    torch.cuda.synchronize()
    F_ref, _ = oracle.sf_f64(q, Z, pos, atom_b=B, B=1.0, R=R, threads=oracle_threads())
    Fh = F.cpu().numpy()
    assert cplx_err(Fh[:, 0] + 1j * Fh[:, 1], F_ref) < 1e-5, info
    # a rotated 3-D volume of q
    qv = oracle.rotate_hkl(oracle.hkl_3d(16, 16, 16, 16e-4, 16e-4, 16e-4, 0.05, 0.3e-9),
                           wl.uniform_rotations(1, np.random.default_rng(54))[0])
    F, _, info = dv.structure_factor_nfft(torch.from_numpy(np.ascontiguousarray(qv, np.float32)).to(gpu), atoms, ff)
    torch.cuda.synchronize()
    F_ref, _ = oracle.sf_f64(qv, Z, pos, atom_b=B, threads=oracle_threads())
    Fh = F.cpu().numpy()
    assert cplx_err(Fh[:, 0] + 1j * Fh[:, 1], F_ref) < 1e-5, info
    # B factors wide enough for the 25-point window at this resolution
    q2 = detector_q(48, lam=0.1e-9, pix=4e-4)
    Bh = rng.uniform(100, 400, len(Z)).astype(np.float32)
    atoms_h = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=Bh), gpu)
    F, _, info = dv.structure_factor_nfft(torch.from_numpy(q2).to(gpu), atoms_h, ff)
    torch.cuda.synchronize()
    assert info["Wg"] == 25, info
    F_ref, _ = oracle.sf_f64(q2, Z, pos, atom_b=Bh, threads=oracle_threads())
    Fh = F.cpu().numpy()
    assert cplx_err(Fh[:, 0] + 1j * Fh[:, 1], F_ref) < 1e-5, info


def test_nfft_results_repeat_bit_for_bit(gpu):
    """The NUFFT's spread binning is a stable sort and its candidate lists keep
    lane order, so repeated calls on the same input give the same bits: a large
    grid (more than 8192 spread blocks), a small one (each block's atoms split
    over several workgroups whose partial grids the first FFT pass adds), per-atom
    B factors (Gaussian spreading) and the many-molecule call."""
    import torch

    ff = dv.form_factor_table(device=gpu)
    rng = np.random.default_rng(71)
    Zl, posl, _ = wl.random_molecule(30000, 12e-9, rng)
    ql = dv.hkl_detector(1024, 1024, 1024 * 75e-6, 1024 * 75e-6, 0.1, 0.15e-9, device=gpu)
    Zs, poss = molecule(2000, 3e-9, seed=72, z0=0)
    qs = torch.from_numpy(np.ascontiguousarray(detector_q(40), np.float32)).to(gpu)
    Bs = rng.uniform(10, 40, len(Zs)).astype(np.float32)
    cases = [("large grid", ql, dv.DeviceAtoms(dv.pack_atoms(Zl, posl, mode=_lib.PACK_NFFT), gpu)),
             ("small grid", qs, dv.DeviceAtoms(dv.pack_atoms(Zs, poss, mode=_lib.PACK_NFFT), gpu)),
             ("per-atom B", qs, dv.DeviceAtoms(dv.pack_atoms(Zs, poss, atom_b=Bs), gpu))]
    for name, q, atoms in cases:
        runs = []
        for _ in range(3):
            F, _, info = dv.structure_factor_nfft(q, atoms, ff)
            torch.cuda.synchronize()
            runs.append(F.cpu().numpy().copy())
        if name == "large grid":
            blocks = np.prod([(m + 7) // 8 for m in info["M"]])
            assert blocks * 4 > 8192, info
        assert all(np.array_equal(runs[0], r) for r in runs[1:]), name
    mols = [molecule(n, r, seed, z0=0) for n, r, seed in ((500, 2e-9, 73), (1500, 4e-9, 74), (60, 1e-9, 75))]
    dm = dv.DeviceMolecules([dv.pack_atoms(Z, pos) for Z, pos in mols], gpu)
    rots = wl.uniform_rotations(len(mols), np.random.default_rng(76))
    runs = []
    for _ in range(3):
        F, _ = dv.structure_factor_nfft_molecules(qs, dm, ff, rots)
        torch.cuda.synchronize()
        runs.append(F.cpu().numpy().copy())
    assert all(np.array_equal(runs[0], r) for r in runs[1:]), "molecules"


def test_nfft_many_molecules_in_one_call(gpu):
    """spsim_hip_sf_nfft_molecules: molecules of uneven sizes (one empty), each at
    its own rotation, with one synchronisation for all their bounds. There are
    more molecules than the call's streams, so streams run several in turn over
    their workspace slice. Every pattern matches the FP64 oracle and the
    one-molecule call; with per-atom B factors too."""
    import torch

    q = detector_q(40, lam=0.2e-9, pix=2e-4)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    spec = [(300, 2e-9, 61), (0, 0.0, 0), (900, 4e-9, 62), (60, 1e-9, 63), (1500, 5e-9, 66), (120, 1.5e-9, 67),
            (40, 0.8e-9, 68), (700, 3e-9, 69), (200, 2.5e-9, 70)]
    mols = []
    for n, r, seed in spec:
        if n == 0:
            mols.append((np.zeros(0, np.int32), np.zeros((0, 3), np.float32)))
        else:
            mols.append(molecule(n, r, seed, z0=0))
    rots = wl.uniform_rotations(len(mols), np.random.default_rng(64))
    rng = np.random.default_rng(65)
    for with_b in (False, True):
        Bs = [rng.uniform(5, 30, len(Z)).astype(np.float32) if with_b else None for Z, _ in mols]
        packed = [dv.pack_atoms(Z, pos, atom_b=B) for (Z, pos), B in zip(mols, Bs)]
        dm = dv.DeviceMolecules(packed, gpu)
        F, I = dv.structure_factor_nfft_molecules(hkl, dm, ff, rots, B=1.5)
        torch.cuda.synchronize()
        Fh = F.cpu().numpy()
        for m, ((Z, pos), B) in enumerate(zip(mols, Bs)):
            Fm = Fh[m, :, 0] + 1j * Fh[m, :, 1]
            if len(Z) == 0:
                assert not np.any(Fm), m
                continue
            F_ref, _ = oracle.sf_f64(q, Z, pos, atom_b=B, B=1.5, R=rots[m], threads=oracle_threads())
            assert cplx_err(Fm, F_ref) < 1e-5, (m, with_b)
            one = dv.DeviceAtoms(packed[m], gpu)
            F1, _, _ = dv.structure_factor_nfft(hkl, one, ff, B=1.5, R=rots[m])
            torch.cuda.synchronize()
            F1h = F1.cpu().numpy()
            assert cplx_err(Fm, F1h[:, 0] + 1j * F1h[:, 1]) < 1e-6, (m, with_b)
    with pytest.raises(RuntimeError, match="offsets"):
        bad = dv.DeviceMolecules([dv.pack_atoms(*mols[0])], gpu)
        bad.atom_offsets[1] = -1
        dv.structure_factor_nfft_molecules(hkl, bad, ff)


def test_nfft_molecules_intensity_only_mixed_windows(gpu):
    """spsim_hip_sf_nfft_molecules without rotations and with only the intensity
    output, over molecules whose per-atom B factors need different Gaussian
    windows (17 and 25 points at this resolution): each pattern matches the FP64
    oracle and the one-molecule call."""
    import torch

    q = detector_q(48, lam=0.1e-9, pix=4e-4)
    hkl = torch.from_numpy(q).to(gpu)
    ff = dv.form_factor_table(device=gpu)
    rng = np.random.default_rng(81)
    spec = [(500, 3e-9, 82, (5, 30)), (700, 3.5e-9, 83, (100, 400)), (300, 2e-9, 84, (10, 40)),
            (400, 2.5e-9, 85, (150, 300)), (200, 1.5e-9, 86, (0, 5)), (600, 3e-9, 87, (20, 60))]
    mols, Bs = [], []
    for n, r, seed, (b0, b1) in spec:
        mols.append(molecule(n, r, seed, z0=0))
        Bs.append(rng.uniform(b0, b1, n).astype(np.float32))
    packed = [dv.pack_atoms(Z, pos, atom_b=B) for (Z, pos), B in zip(mols, Bs)]
    dm = dv.DeviceMolecules(packed, gpu)
    F, I = dv.structure_factor_nfft_molecules(hkl, dm, ff, None, B=0.5, want_F=False)
    torch.cuda.synchronize()
    assert F is None
    Ih = I.cpu().numpy()
    widths = set()
    for m, ((Z, pos), B) in enumerate(zip(mols, Bs)):
        _, I_ref = oracle.sf_f64(q, Z, pos, atom_b=B, B=0.5, threads=oracle_threads())
        assert rel_l2(Ih[m], I_ref) < 1e-5, m
        _, I1, info = dv.structure_factor_nfft(hkl, dv.DeviceAtoms(packed[m], gpu), ff, B=0.5)
        torch.cuda.synchronize()
        widths.add(info["Wg"])
        assert rel_l2(Ih[m], I1.cpu().numpy()) < 1e-6, (m, info)
    assert widths == {17, 25}, widths


def test_nfft_molecules_full_batched_workload_against_direct(gpu):
    """The full batched workload (512 molecules of 2e3-2e5 atoms at random
    orientations, 512 x 512 detector) in one spsim_hip_sf_nfft_molecules call:
    every pattern's intensities within 1e-5 (normwise) of the direct kernel's,
    which the other tests pin to the FP64 oracle at 3e-7. Both W2 plans occur."""
    import torch

    det, mols, rots = wl.config3(512)
    q = dv.hkl_detector(det.n, det.n, det.width, det.width, det.distance, det.wavelength, device=gpu)
    packed = [dv.pack_atoms(Z, pos, mode=_lib.PACK_NFFT) for Z, pos in mols]
    dm = dv.DeviceMolecules(packed, gpu)
    ff = dv.form_factor_table(device=gpu)
    _, I = dv.structure_factor_nfft_molecules(q, dm, ff, rots, eps=1e-6, want_F=False)
    torch.cuda.synchronize()
    worst, plans = 0.0, set()
    for m in range(len(mols)):
        a = dv.DeviceAtoms(packed[m], gpu)
        _, Id = dv.structure_factor(q, a, ff, want_F=False, R=rots[m])
        e = (torch.linalg.vector_norm(I[m].double() - Id.double()) / torch.linalg.vector_norm(Id.double())).item()
        assert e < 1e-5, (m, a.natoms, e)
        worst = max(worst, e)
        _, _, info = dv.structure_factor_nfft(q, a, ff, eps=1e-6, want_F=False, R=rots[m])
        plans.add(info["W2"])
    assert plans == {6, 7}, plans
    print(f"worst of 512 patterns vs the direct kernel: {worst:.3g}")
