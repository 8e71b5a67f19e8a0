"""GPU parity of the structure-factor path against the oracle (FP64 and the
reference's own FP32 semantics), through the C-ABI (spsim_hip_* entry points)."""
import numpy as np
import pytest

import oracle
from spsim_amd import device as dv
from spsim_amd import workloads as wl
from tests.conftest import oracle_threads, rel_l2

pytestmark = pytest.mark.gpu

# Tolerance of the FP32-sum / FP64-phase kernel against the FP64 oracle on the
# same FP32 inputs (normwise relative error of I and of F).
TOL_F64 = 1e-5


def _det_hkl(det: wl.Detector):
    return oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0.0, 0.0, 0, det.wavelength)


def _gpu_sf(gpu, hkl, Z, pos, amp=None, B=0.0, atom_b=None, delta=False, mode=0, tile=None):
    import torch

    packed = dv.pack_atoms(Z, pos, amp=amp, atom_b=atom_b, delta_atoms=delta, mode=mode)
    da = dv.DeviceAtoms(packed, gpu)
    ff = dv.form_factor_table(device=gpu)
    q = torch.from_numpy(np.ascontiguousarray(hkl, np.float32)).to(gpu)
    F, I = dv.structure_factor(q, da, ff, B=B, tile=tile)
    torch.cuda.synchronize()
    F = F.cpu().numpy()
    return F[:, 0] + 1j * F[:, 1], I.cpu().numpy()


def test_hkl_detector_bit_exact(gpu):
    import torch

    for det, spherical, cx, cy in [(wl.CFG1_DET, 0, 0.0, 0.0), (wl.CFG5_DET, 0, 1e-4, -2e-4),
                                   (wl.Detector(37, 1e-4, 0.05, 1e-9), 1, 0.0, 0.0)]:
        ref = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, cx, cy, spherical, det.wavelength)
        got = dv.hkl_detector(det.n, det.n, det.width, det.width, det.distance, det.wavelength, cx, cy, spherical,
                              device=gpu)
        torch.cuda.synchronize()
        got = got.cpu().numpy()
        if spherical:
            # the spherical grid goes through sin/atan on the host in both paths
            np.testing.assert_array_equal(got, ref)
        else:
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
                f"max |diff| {np.abs(got - ref).max()}"


def test_rotate_hkl_bit_exact(gpu):
    import torch

    rng = np.random.default_rng(11)
    hkl = _det_hkl(wl.CFG1_DET)
    R = wl.uniform_rotations(1, rng)[0]
    ref = oracle.rotate_hkl(hkl, R)
    got = dv.rotate_hkl(torch.from_numpy(hkl).to(gpu), R)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_sf_small_molecule_vs_f64(gpu):
    rng = np.random.default_rng(5)
    Z, pos, _ = wl.random_molecule(300, 3e-9, rng)
    Z[:5] = 0          # skipped atoms (src/diffraction.c:830-832)
    Z[5:15] = 1        # hydrogen
    Z[15:20] = 79      # gold
    hkl = _det_hkl(wl.Detector(64, 3e-4, 0.1, 0.2e-9))
    F_ref, I_ref = oracle.sf_f64(hkl, Z, pos, threads=oracle_threads())
    F, I = _gpu_sf(gpu, hkl, Z, pos)
    assert rel_l2(I, I_ref) < TOL_F64
    assert rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag]) < TOL_F64


def test_sf_config1_vs_f64_and_reference_semantics(gpu):
    det, Z, pos, _ = wl.config1()
    hkl = _det_hkl(det)
    F_ref, I_ref = oracle.sf_f64(hkl, Z, pos, threads=oracle_threads())
    F, I = _gpu_sf(gpu, hkl, Z, pos)
    err = rel_l2(I, I_ref)
    assert err < TOL_F64, err
    # The reference's own FP32 evaluation differs from FP64 by its rounding; the
    # GPU result must agree with it at least as well as the reference agrees
    # with FP64 (plus the GPU's own error).
    _, I_sem = oracle.sf_refsem(hkl, Z, pos, threads=oracle_threads())
    ref_err = rel_l2(I_sem, I_ref)
    assert rel_l2(I, I_sem) <= ref_err + 2 * err + 1e-7


def test_sf_global_b_and_illumination(gpu):
    rng = np.random.default_rng(7)
    Z, pos, _ = wl.random_molecule(1000, 4e-9, rng)
    amp = rng.uniform(0.2, 1.0, size=Z.shape[0]).astype(np.float32)
    hkl = _det_hkl(wl.Detector(48, 4e-4, 0.08, 0.15e-9))
    F_ref, I_ref = oracle.sf_f64(hkl, Z, pos, amp=amp, B=25.0, threads=oracle_threads())
    F, I = _gpu_sf(gpu, hkl, Z, pos, amp=amp, B=25.0)
    assert rel_l2(I, I_ref) < TOL_F64


def test_sf_per_atom_b(gpu):
    rng = np.random.default_rng(8)
    Z, pos, B = wl.random_molecule(2000, 6e-9, rng, atom_b_range=(10.0, 40.0))
    hkl = _det_hkl(wl.Detector(64, 1.2e-3, 0.1, 0.15e-9))
    F_ref, I_ref = oracle.sf_f64(hkl, Z, pos, atom_b=B, threads=oracle_threads())
    F, I = _gpu_sf(gpu, hkl, Z, pos, atom_b=B)
    assert rel_l2(I, I_ref) < TOL_F64


def test_sf_config2_subset_vs_f64(gpu):
    """Full config 2 on the GPU (100k atoms, per-atom B, 1024^2 pixels); a random
    4096-pixel subset is checked against the FP64 oracle."""
    det, Z, pos, B = wl.config2()
    hkl = _det_hkl(det)
    F, I = _gpu_sf(gpu, hkl, Z, pos, atom_b=B)
    rng = np.random.default_rng(22)
    sub = np.sort(rng.choice(hkl.shape[0], 4096, replace=False))
    F_ref, I_ref = oracle.sf_f64(hkl[sub], Z, pos, atom_b=B, threads=oracle_threads())
    err = rel_l2(I[sub], I_ref)
    print(f"config2 subset rel err {err:.3e}")
    assert err < TOL_F64


@pytest.mark.parametrize("K", [1, 2, 4, 8])
@pytest.mark.parametrize("chunk", [64, 128, 256])
def test_sf_tile_shapes_agree(gpu, K, chunk):
    rng = np.random.default_rng(9)
    Z, pos, B = wl.random_molecule(700, 3e-9, rng, atom_b_range=(5.0, 30.0))
    hkl = _det_hkl(wl.Detector(45, 4e-4, 0.1, 0.2e-9))  # 2025 pixels: ragged tiles
    _, I_ref = oracle.sf_f64(hkl, Z, pos, atom_b=B, threads=oracle_threads())
    for splits in (1, 3):
        _, I = _gpu_sf(gpu, hkl, Z, pos, atom_b=B, tile=dv.make_tile(K, chunk, splits))
        assert rel_l2(I, I_ref) < TOL_F64, (K, chunk, splits)


def test_sf_deterministic(gpu):
    det, Z, pos, _ = wl.config1()
    hkl = _det_hkl(det)
    _, I1 = _gpu_sf(gpu, hkl, Z, pos)
    _, I2 = _gpu_sf(gpu, hkl, Z, pos)
    assert np.array_equal(I1, I2)


def test_sf_edge_cases(gpu):
    hkl = _det_hkl(wl.Detector(7, 1e-3, 0.1, 0.2e-9))
    # no atoms at all
    F, I = _gpu_sf(gpu, hkl, np.zeros(0, np.int32), np.zeros((0, 3), np.float32))
    assert np.all(I == 0) and np.all(F == 0)
    # only Z == 0 atoms: skipped by the plain path
    F, I = _gpu_sf(gpu, hkl, np.zeros(3, np.int32), np.ones((3, 3), np.float32) * 1e-9)
    assert np.all(I == 0)
    # one atom: I = f^2 at every pixel
    Z = np.array([6], np.int32)
    pos = np.array([[1e-9, -2e-9, 3e-9]], np.float32)
    _, I = _gpu_sf(gpu, hkl, Z, pos)
    _, I_ref = oracle.sf_f64(hkl, Z, pos)
    np.testing.assert_allclose(I, I_ref, rtol=2e-6)
    # a single pixel
    _, I = _gpu_sf(gpu, hkl[:1], Z, pos)
    np.testing.assert_allclose(I, I_ref[:1], rtol=2e-6)


def test_sf_delta_atoms_and_vector_mode(gpu):
    rng = np.random.default_rng(10)
    Z, pos, _ = wl.random_molecule(203, 3e-9, rng)  # not a multiple of 4
    Z[:7] = 0
    amp = rng.uniform(0.5, 1.0, size=Z.shape[0]).astype(np.float32)
    hkl = _det_hkl(wl.Detector(32, 5e-4, 0.1, 0.2e-9))
    for delta in (False, True):
        for mode in (0, 1):
            _, I_ref = oracle.sf_f64(hkl, Z, pos, amp=amp, delta=delta, vector_mode=bool(mode), B=10.0)
            _, I = _gpu_sf(gpu, hkl, Z, pos, amp=amp, delta=delta, mode=mode, B=10.0)
            assert rel_l2(I, I_ref) < TOL_F64, (delta, mode)


def _gpu_grid(gpu, hkl, nx, ny, Z, pos, atom_b=None, B=0.0, tile=None):
    import torch

    da = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=atom_b), gpu)
    ff = dv.form_factor_table(device=gpu)
    q = torch.from_numpy(np.ascontiguousarray(hkl, np.float32)).to(gpu)
    F, I = dv.structure_factor_grid(q, nx, ny, da, ff, B=B, tile=tile)
    torch.cuda.synchronize()
    F = F.cpu().numpy()
    return F[:, 0] + 1j * F[:, 1], I.cpu().numpy()


@pytest.mark.parametrize("nx,ny", [(64, 64), (45, 45), (50, 33), (1, 7), (2, 2)])
def test_sf_grid_mirror_kernel_vs_f64(gpu, nx, ny):
    rng = np.random.default_rng(30)
    Z, pos, B = wl.random_molecule(900, 4e-9, rng, atom_b_range=(10.0, 40.0))
    hkl = oracle.hkl_detector(nx, ny, np.float32(nx * 3e-4).item(), np.float32(ny * 3e-4).item(), 0.05, 0.0, 0.0, 0,
                              0.15e-9)
    tiles = [None, dv.make_tile(1, 64, 1), dv.make_tile(2, 256, 3)]
    if nx == ny:  # square: octant tiles (8 pixels per lane), also with atom splits
        tiles += [dv.make_tile(8, 128, 1), dv.make_tile(8, 64, 3)]
    for has_b in (False, True):
        F_ref, I_ref = oracle.sf_f64(hkl, Z, pos, atom_b=B if has_b else None, B=3.0, threads=oracle_threads())
        for tile in tiles:
            F, I = _gpu_grid(gpu, hkl, nx, ny, Z, pos, atom_b=B if has_b else None, B=3.0, tile=tile)
            assert rel_l2(I, I_ref) < TOL_F64, (has_b, tile)
            assert rel_l2(np.c_[F.real, F.imag], np.c_[F_ref.real, F_ref.imag]) < TOL_F64


def test_sf_grid_falls_back_without_symmetry(gpu):
    rng = np.random.default_rng(31)
    Z, pos, _ = wl.random_molecule(500, 3e-9, rng)
    n = 40
    # off-centre detector: no mirror symmetry anywhere
    hkl = oracle.hkl_detector(n, n, np.float32(n * 5e-4).item(), np.float32(n * 5e-4).item(), 0.1, 1e-4, -2e-4, 0,
                              0.2e-9)
    _, I_ref = oracle.sf_f64(hkl, Z, pos)
    _, I = _gpu_grid(gpu, hkl, n, n, Z, pos)
    assert rel_l2(I, I_ref) < TOL_F64
    # a symmetric detector with one pixel disturbed: only that workgroup falls back
    hkl2 = _det_hkl(wl.Detector(n, 5e-4, 0.1, 0.2e-9)).copy()
    hkl2[7, 0] *= np.float32(1.001)
    _, I_ref2 = oracle.sf_f64(hkl2, Z, pos)
    _, I2 = _gpu_grid(gpu, hkl2, n, n, Z, pos)
    assert rel_l2(I2, I_ref2) < TOL_F64
    # an arbitrary (non-detector) list of n*n vectors
    hkl3 = rng.normal(0, 2e9, size=(n * n, 3)).astype(np.float32)
    _, I_ref3 = oracle.sf_f64(hkl3, Z, pos)
    _, I3 = _gpu_grid(gpu, hkl3, n, n, Z, pos)
    assert rel_l2(I3, I_ref3) < TOL_F64
    # square and centred but with rectangular pixels: mirror-symmetric, not
    # transpose-symmetric, so octant workgroups fall back to two quads
    hkl4 = oracle.hkl_detector(n, n, np.float32(n * 5e-4).item(), np.float32(n * 4e-4).item(), 0.1, 0.0, 0.0, 0,
                               0.2e-9)
    _, I_ref4 = oracle.sf_f64(hkl4, Z, pos)
    for tile in (None, dv.make_tile(8, 128, 1)):
        _, I4 = _gpu_grid(gpu, hkl4, n, n, Z, pos, tile=tile)
        assert rel_l2(I4, I_ref4) < TOL_F64
    # one transposed pixel disturbed
    hkl5 = _det_hkl(wl.Detector(n, 5e-4, 0.1, 0.2e-9)).copy()
    hkl5[3 * n + 11, 1] *= np.float32(1.0001)
    _, I_ref5 = oracle.sf_f64(hkl5, Z, pos)
    _, I5 = _gpu_grid(gpu, hkl5, n, n, Z, pos, tile=dv.make_tile(8, 256, 1))
    assert rel_l2(I5, I_ref5) < TOL_F64


def test_sf_grid_octant_tile_needs_square(gpu):
    import torch

    from spsim_amd import _lib

    rng = np.random.default_rng(32)
    Z, pos, _ = wl.random_molecule(50, 2e-9, rng)
    hkl = oracle.hkl_detector(8, 6, 8 * 5e-4, 6 * 5e-4, 0.1, 0.0, 0.0, 0, 0.2e-9)
    with pytest.raises(RuntimeError, match="octants need nx == ny"):
        _gpu_grid(gpu, hkl, 8, 6, Z, pos, tile=dv.make_tile(8, 128, 1))
    assert "octants" in _lib.last_error()


def test_sf_grid_config2_subset_vs_f64(gpu):
    det, Z, pos, B = wl.config2()
    hkl = _det_hkl(det)
    F, I = _gpu_grid(gpu, hkl, det.n, det.n, Z, pos, atom_b=B)
    rng = np.random.default_rng(23)
    sub = np.sort(rng.choice(hkl.shape[0], 4096, replace=False))
    _, I_ref = oracle.sf_f64(hkl[sub], Z, pos, atom_b=B, threads=oracle_threads())
    err = rel_l2(I[sub], I_ref)
    print(f"config2 (mirror kernel) subset rel err {err:.3e}")
    assert err < TOL_F64


@pytest.mark.parametrize("grid", [False, True])
def test_sf_exact_rotation(gpu, grid):
    """q' = R q applied exactly (atoms rotated by R^T in FP64) in both kernels."""
    import torch

    rng = np.random.default_rng(32)
    Z, pos, B = wl.random_molecule(1200, 5e-9, rng, atom_b_range=(10.0, 40.0))
    R = wl.uniform_rotations(1, rng)[0]
    n = 36
    hkl = _det_hkl(wl.Detector(n, 8e-4, 0.1, 0.15e-9))
    _, I_ref = oracle.sf_f64(hkl, Z, pos, atom_b=B, R=R, threads=oracle_threads())
    da = dv.DeviceAtoms(dv.pack_atoms(Z, pos, atom_b=B), gpu)
    ff = dv.form_factor_table(device=gpu)
    q = torch.from_numpy(hkl).to(gpu)
    if grid:
        _, I = dv.structure_factor_grid(q, n, n, da, ff, R=R)
    else:
        _, I = dv.structure_factor(q, da, ff, R=R)
    torch.cuda.synchronize()
    assert rel_l2(I.cpu().numpy(), I_ref) < TOL_F64
