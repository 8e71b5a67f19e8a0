"""The reference's C API (include/*.h) and Python binding names, run on the GPU,
against the oracle: compute_pattern_on_list, vector_compute_pattern_on_list,
compute_sf, get_HKL_list_for_detector, apply_orientation_to_HKL_list and the
detector stages / simulate_shot."""
import ctypes as C

import numpy as np
import pytest

import oracle
import spsim  # the reference's module name (src/spsim.py), top-level spsim.py
from spsim_amd import workloads as wl
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def make_opts(n=48, pw=2e-4, dist=0.08, lam=0.2e-9, verbosity=0):
    opts = spsim.set_defaults()
    d = opts.detector
    d.nx = d.ny = n
    d.nz = 1
    d.pixel_width = d.pixel_height = pw
    d.width = d.height = float(np.float32(n * pw))
    d.distance = dist
    opts.experiment.wavelength = lam
    opts.verbosity_level = verbosity
    return opts


def hkl_of(opts):
    n = C.c_int(0)
    ptr = spsim.get_HKL_list_for_detector(opts.detector, opts.experiment, C.byref(n))
    assert ptr is not None
    return ptr, n.value, np.ctypeslib.as_array(ptr, shape=(n.value * 3,)).reshape(-1, 3).copy()


def molecule(natoms=400, radius=3e-9, seed=0):
    Z, pos, _ = wl.random_molecule(natoms, radius, np.random.default_rng(seed))
    Z[:3] = 0
    return Z, pos, spsim.get_molecule_from_atoms(Z, pos)


def test_get_hkl_list_matches_oracle(gpu):
    opts = make_opts(37)
    opts.detector.center_x = 1e-4
    ptr, n, q = hkl_of(opts)
    d = opts.detector
    ref = oracle.hkl_detector(d.nx, d.ny, d.width, d.height, d.distance, d.center_x, d.center_y, 0,
                              opts.experiment.wavelength)
    assert n == 37 * 37
    assert np.array_equal(q.view(np.uint32), ref.view(np.uint32))


def test_compute_pattern_on_list(gpu):
    opts = make_opts()
    Z, pos, mol = molecule()
    ptr, n, q = hkl_of(opts)
    pat = spsim.compute_pattern_on_list(mol, ptr, n, 7.5, opts.experiment, opts)
    assert pat is not None and pat.HKL_list_size == n
    I = spsim.float_array(pat.ints, n)
    F = spsim.complex_array(pat.F, n)
    np.testing.assert_array_equal(spsim.float_array(pat.HKL_list, 3 * n), q.ravel())
    F_ref, I_ref = oracle.sf_f64(q, Z, pos, B=7.5)
    assert rel_l2(I, I_ref) < 1e-5
    assert rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag]) < 1e-5
    _, I_sem = oracle.sf_refsem(q, Z, pos, B=7.5)
    assert rel_l2(I, I_sem) < 1e-5
    np.testing.assert_allclose(I, np.abs(F) ** 2, rtol=1e-5, atol=1e-6 * I.max())
    spsim.free_diffraction_pattern(pat)


def test_illumination_delta_and_vector_path(gpu):
    opts = make_opts(32)
    opts.experiment.beam_fwhm = 4e-9
    opts.experiment.beam_center_x = 1e-9
    Z, pos, mol = molecule(203, seed=1)
    ptr, n, q = hkl_of(opts)
    sig = np.float32(4e-9) / 2.355
    il = np.exp(-(((pos[:, 0].astype(np.float64) - np.float32(1e-9)) ** 2 + pos[:, 1].astype(np.float64) ** 2)
                  / (2 * float(sig) ** 2))).astype(np.float32)
    for delta in (0, 1):
        opts.delta_atoms = delta
        pat = spsim.compute_pattern_on_list(mol, ptr, n, 0.0, opts.experiment, opts)
        _, I_ref = oracle.sf_f64(q, Z, pos, amp=np.sqrt(il), delta=bool(delta))
        assert rel_l2(spsim.float_array(pat.ints, n), I_ref) < 1e-5
        _, I_sem = oracle.sf_refsem(q, Z, pos, illum=il, delta=bool(delta))
        assert rel_l2(spsim.float_array(pat.ints, n), I_sem) < 2e-5
        spsim.free_diffraction_pattern(pat)
        # vector path: amplitude illum^(1/4), Z = 0 kept, delta only on the first 4*floor(N/4) atoms
        pat = spsim.vector_compute_pattern_on_list(mol, ptr, n, 0.0, opts.experiment, opts)
        _, I_vec = oracle.sf_f64(q, Z, pos, amp=np.sqrt(np.sqrt(il)), delta=bool(delta), vector_mode=True)
        assert rel_l2(spsim.float_array(pat.ints, n), I_vec) < 1e-5
        spsim.free_diffraction_pattern(pat)


def test_compute_sf_dispatch(gpu):
    opts = make_opts(24)
    Z, pos, mol = molecule(120, seed=2)
    ptr, n, q = hkl_of(opts)
    opts.use_cuda, opts.vectorize = 1, 1
    pat = spsim.compute_sf(mol, ptr, n, opts)
    assert pat.rot is None
    assert rel_l2(spsim.float_array(pat.ints, n), oracle.sf_f64(q, Z, pos)[1]) < 1e-5
    opts.use_cuda = 0
    pat2 = spsim.compute_sf(mol, ptr, n, opts)
    assert rel_l2(spsim.float_array(pat2.ints, n), oracle.sf_f64(q, Z, pos, vector_mode=True)[1]) < 1e-5
    spsim.free_diffraction_pattern(pat)
    spsim.free_diffraction_pattern(pat2)


def test_apply_orientation(gpu):
    opts = make_opts(20)
    ptr, n, q = hkl_of(opts)
    pp = C.pointer(ptr)
    nn = C.c_int(n)
    rot = spsim.apply_orientation_to_HKL_list(pp, C.byref(nn), opts)  # Euler (0,0,0): identity
    assert [rot.data[i] for i in range(9)] == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    np.testing.assert_array_equal(np.ctypeslib.as_array(ptr, shape=(3 * n,)).reshape(-1, 3), q)
    opts.random_orientation = 1
    spsim.sp_srand(5)
    rot = spsim.apply_orientation_to_HKL_list(pp, C.byref(nn), opts)
    R = np.array([rot.data[i] for i in range(9)], np.float64).reshape(3, 3)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-6)
    got = np.ctypeslib.as_array(ptr, shape=(3 * n,)).reshape(-1, 3)
    assert np.array_equal(got.view(np.uint32), oracle.rotate_hkl(q, R).view(np.uint32))


def test_bad_atomic_number_returns_null(gpu):
    opts = make_opts(8)
    mol = spsim.get_molecule_from_atoms([6, 120], [[0, 0, 0], [1e-10, 0, 0]])
    ptr, n, _ = hkl_of(opts)
    assert spsim.compute_pattern_on_list(mol, ptr, n, 0.0, opts.experiment, opts) is None


def detector_opts(n=64, bins=4):
    opts = make_opts(n, pw=2e-5, dist=0.05, lam=1.3e-9)
    d = opts.detector
    d.binning_x = d.binning_y = bins
    d.quantum_efficiency = 0.15
    d.electron_hole_production_energy = 5.8e-19
    d.readout_noise = 10.0
    d.dark_current = 0.1
    d.linear_full_well = 2e5
    d.maximum_value = 65535.0
    opts.experiment.exposure_time = 1e-13
    opts.experiment.beam_intensity = 1e32  # bright enough for all noise branches
    return opts


def ccd_array(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def test_simulate_shot_vs_staged_and_oracle(gpu):
    opts = detector_opts()
    Z, pos, mol = molecule(600, seed=3)
    spsim.sp_srand(11)
    pat = spsim.simulate_shot(mol, opts)
    assert pat is not None
    d, e = opts.detector, opts.experiment
    n, nb = d.nx * d.ny, (d.nx // 4) * (d.ny // 4)
    shot = {k: ccd_array(getattr(d, k), n) for k in ("thomson_correction", "solid_angle", "photons_per_pixel",
                                                     "photon_count", "electrons_per_pixel")}
    shot["real_output"] = ccd_array(d.real_output, nb)
    shot["noiseless_output"] = ccd_array(d.noiseless_output, nb)
    I = spsim.float_array(pat.ints, n)
    # the same stages called one by one, with the same seed, give identical arrays
    spsim.sp_srand(11)
    spsim.calculate_thomson_correction(d, e)
    spsim.calculate_pixel_solid_angle(d)
    d.photons_per_pixel = None
    spsim.calculate_photons_per_pixel(pat, opts)
    spsim.generate_poisson_noise(d)
    spsim.calculate_electrons_per_pixel(d, e)
    spsim.calculate_real_detector_output(d, e)
    spsim.calculate_noiseless_detector_output(d, e)
    for k in ("thomson_correction", "solid_angle", "photons_per_pixel", "photon_count", "electrons_per_pixel"):
        assert np.array_equal(ccd_array(getattr(d, k), n), shot[k]), k
    for k in ("real_output", "noiseless_output"):
        assert np.array_equal(ccd_array(getattr(d, k), nb), shot[k]), k
    # against the oracle
    th = oracle.thomson(d.nx, d.ny, d.width, d.height, d.distance, 0, 0, 0)
    om = oracle.solid_angle(d.nx, d.ny, d.width, d.height, d.distance, d.pixel_width, d.pixel_height)
    assert np.array_equal(shot["thomson_correction"], th)
    np.testing.assert_array_max_ulp(shot["solid_angle"], om, maxulp=1)
    ph = oracle.photons(shot["thomson_correction"], shot["solid_angle"], I, e.beam_intensity)
    assert np.array_equal(shot["photons_per_pixel"], ph)
    epp = oracle.electrons_per_photon(e.wavelength, d.electron_hole_production_energy)
    np.testing.assert_array_equal(shot["electrons_per_pixel"], shot["photon_count"] * np.float32(epp))
    epp_nl = oracle.electrons_per_photon(e.wavelength, d.electron_hole_production_energy, noiseless=True)
    nl = oracle.noiseless_output(ph, d.nx, d.ny, 4, 4, d.quantum_efficiency, epp_nl, d.linear_full_well,
                                 d.maximum_value)
    assert np.array_equal(shot["noiseless_output"], nl)
    spsim.free_stuff(pat, opts)


def test_simulate_shot_through_the_nfft_path(gpu):
    """simulate_shot with use_nfft_for_sf: the pattern comes from
    compute_pattern_on_list_by_nfft (Z == 0 atoms included) and the detector
    stages run on its intensities."""
    opts = detector_opts()
    opts.use_nfft_for_sf = 1
    Z, pos, mol = molecule(600, seed=4)
    _, _, q = hkl_of(opts)
    spsim.sp_srand(5)
    pat = spsim.simulate_shot(mol, opts)
    assert pat is not None
    d, e = opts.detector, opts.experiment
    n, nb = d.nx * d.ny, (d.nx // 4) * (d.ny // 4)
    I = spsim.float_array(pat.ints, n)
    _, I_ref = oracle.sf_nfft_f64(q, Z, pos, B=opts.b_factor)
    assert rel_l2(I, I_ref) < 1e-5
    _, I_plain = oracle.sf_f64(q, Z, pos, B=opts.b_factor)
    assert rel_l2(I_plain, I_ref) > 1e-4  # the three Z == 0 atoms count on this path
    th = ccd_array(d.thomson_correction, n)
    om = ccd_array(d.solid_angle, n)
    ph = oracle.photons(th, om, I, e.beam_intensity)
    assert np.array_equal(ccd_array(d.photons_per_pixel, n), ph)
    epp_nl = oracle.electrons_per_photon(e.wavelength, d.electron_hole_production_energy, noiseless=True)
    nl = oracle.noiseless_output(ph, d.nx, d.ny, 4, 4, d.quantum_efficiency, epp_nl, d.linear_full_well,
                                 d.maximum_value)
    assert np.array_equal(ccd_array(d.noiseless_output, nb), nl)
    spsim.free_stuff(pat, opts)


def test_atomsf_lib_override_intensity(gpu, tmp_path):
    """compute_pattern_on_list with an ./atomsf.lib in the working directory
    (src/diffraction.c:118-177): the table is read once per process, so a
    child process runs the drop-in call there; its intensities match the FP64
    oracle fed the same overridden rows (S and Fe replaced, a1 = -1 in every
    other row)."""
    import subprocess
    import sys
    import textwrap
    from pathlib import Path

    from tests.test_oracle_formfactor import atomsf_lib_text, override_table

    root = Path(__file__).resolve().parents[1]
    (tmp_path / "atomsf.lib").write_text(atomsf_lib_text())
    rng = np.random.default_rng(17)
    Z = rng.choice(np.array([6, 16, 26], np.int32), 300).astype(np.int32)
    pos = (rng.uniform(-1, 1, (300, 3)) * 2e-9).astype(np.float32)
    np.save(tmp_path / "Z.npy", Z)
    np.save(tmp_path / "pos.npy", pos)
    code = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {str(root)!r})
        import torch  # noqa: F401  (the library binds to torch's HIP runtime)
        import ctypes as C
        import numpy as np
        import spsim
        from tests.test_gpu_dropin import make_opts, hkl_of
        opts = make_opts(24)
        mol = spsim.get_molecule_from_atoms(np.load("Z.npy"), np.load("pos.npy"))
        ptr, n, q = hkl_of(opts)
        print("atomsf.lib child: compute_pattern_on_list 300 atoms x", n, flush=True)
        pat = spsim.compute_pattern_on_list(mol, ptr, n, 2.5, opts.experiment, opts)
        assert pat is not None, spsim.last_error()
        np.save("I.npy", spsim.float_array(pat.ints, n))
        np.save("q.npy", q)
    """)
    r = subprocess.run([sys.executable, "-c", code], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    I, q = np.load(tmp_path / "I.npy"), np.load(tmp_path / "q.npy")
    ff = override_table(oracle.default_form_factors())
    _, I_ref = oracle.sf_f64(q, Z, pos, B=2.5, ff=ff)
    assert rel_l2(I, I_ref) < 1e-5
    _, I_builtin = oracle.sf_f64(q, Z, pos, B=2.5)
    assert rel_l2(I_builtin, I_ref) > 1e-2  # the override changes the pattern
