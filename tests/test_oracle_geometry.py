"""Pins the oracle's detector q-vectors (src/diffraction.c:237-297) against the
Ewald-sphere formulas of the user manual (doc/UserManual.tex:159-179), which
are written in a different form, and against symmetry."""
import math

import mpmath as mp
import numpy as np

import oracle
from spsim_amd import workloads as wl


def manual_hkl(x: float, y: float, d: float, lam: float):
    """h, k, l of doc/UserManual.tex:159-179 in high precision. The manual's l
    is 1/lambda - 1/(lambda^2 chi); the code stores q_z = -l."""
    mp.mp.dps = 40
    x, y, d, lam = (mp.mpf(float(v)) for v in (x, y, d, lam))
    chi = mp.sqrt((x / (lam * d)) ** 2 + (y / (lam * d)) ** 2 + (1 / lam) ** 2)
    h = x / (lam ** 2 * d * chi)
    k = y / (lam ** 2 * d * chi)
    l = 1 / lam - 1 / (lam ** 2 * chi)
    return float(h), float(k), float(-l)


def test_matches_manual_formula():
    nx, ny, pw, dist, lam = 64, 48, 1e-4, 0.05, 1.5e-10
    width, height = np.float32(nx * pw).item(), np.float32(ny * pw).item()
    cx, cy = 2e-4, -1e-4
    q = oracle.hkl_detector(nx, ny, width, height, dist, cx, cy, 0, lam)
    rng = np.random.default_rng(0)
    for _ in range(40):
        x, y = int(rng.integers(nx)), int(rng.integers(ny))
        f32 = lambda v: float(np.float32(v))
        px = ((x - (nx - 1) / 2) / nx) * f32(width) - f32(cx)
        py = ((y - (ny - 1) / 2) / ny) * f32(height) - f32(cy)
        h, k, l = manual_hkl(px, py, f32(dist), f32(lam))
        np.testing.assert_allclose(q[y * nx + x], [h, k, l], rtol=2e-7, atol=1e-7 * abs(h) + 1.0)


def test_ewald_sphere_and_bragg_law():
    det = wl.CFG2_DET
    q = oracle.hkl_detector(det.n, det.n, det.width, det.width, det.distance, 0, 0, 0, det.wavelength).astype(np.float64)
    lam = float(np.float32(det.wavelength))
    # q lies on the Ewald sphere through the origin: |q + k0| = 1/lambda with k0 = (0,0,1/lambda)
    r = np.linalg.norm(q + np.array([0, 0, 1 / lam]), axis=1)
    np.testing.assert_allclose(r, 1 / lam, rtol=1e-6)
    assert np.all(q[:, 2] <= 0)
    # |q| = 2 sin(theta) / lambda with tan(2 theta) = rho / d
    n = det.n
    x = np.arange(n)
    pxv = ((x - (n - 1) / 2) / n) * float(np.float32(det.width))
    PX, PY = np.meshgrid(pxv, pxv)
    two_theta = np.arctan(np.hypot(PX, PY).ravel() / float(np.float32(det.distance)))
    np.testing.assert_allclose(np.linalg.norm(q, axis=1), 2 * np.sin(two_theta / 2) / lam, rtol=1e-6)


def test_centrosymmetric_grid():
    # centred detector: pixel (x, y) and (n-1-x, n-1-y) have opposite (qx, qy), equal qz
    n = 33
    q = oracle.hkl_detector(n, n, np.float32(n * 1e-4).item(), np.float32(n * 1e-4).item(), 0.1, 0, 0, 0, 2e-10)
    q = q.reshape(n, n, 3)
    np.testing.assert_array_equal(q[:, :, 0], -q[::-1, ::-1, 0])
    np.testing.assert_array_equal(q[:, :, 1], -q[::-1, ::-1, 1])
    np.testing.assert_array_equal(q[:, :, 2], q[::-1, ::-1, 2])
    # odd n: the middle pixel looks straight down the beam, q = 0
    assert np.all(q[n // 2, n // 2] == 0)


def test_spherical_flat_grid():
    n, pw, dist, lam = 21, 2e-4, 0.1, 1e-9
    w = np.float32(n * pw).item()
    q = oracle.hkl_detector(n, n, w, w, dist, 0, 0, 1, lam).reshape(n, n, 3)
    assert np.all(q[:, :, 2] == 0)
    # uniform spacing 2 sin(atan(p/d)/2)/lambda between neighbours
    step = 2 * math.sin(math.atan(float(np.float32(w / n)) / dist) / 2) / float(np.float32(lam))
    np.testing.assert_allclose(np.diff(q[0, :, 0]), step, rtol=1e-5)


def test_rotation_quarter_turns_are_exact():
    q = oracle.hkl_detector(16, 16, 1.6e-3, 1.6e-3, 0.1, 0, 0, 0, 2e-10)
    Rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float64)
    out = oracle.rotate_hkl(q, Rz)
    np.testing.assert_array_equal(out[:, 0], -q[:, 1])
    np.testing.assert_array_equal(out[:, 1], q[:, 0])
    np.testing.assert_array_equal(out[:, 2], q[:, 2])
    R = wl.uniform_rotations(3, np.random.default_rng(1))
    for Ri in R:
        np.testing.assert_allclose(Ri @ Ri.T, np.eye(3), atol=1e-12)
        out = oracle.rotate_hkl(q, Ri)
        np.testing.assert_allclose(np.linalg.norm(out, axis=1), np.linalg.norm(q, axis=1), rtol=1e-6)
        np.testing.assert_allclose(out, (Ri @ q.astype(np.float64).T).T, rtol=1e-6, atol=1e-3)


# ---------------------------------------------------------------- 3-D detector grid

def test_hkl_3d_closed_form():
    """q = p / (d lambda) on a uniform grid (src/diffraction.c:325-369), with
    d lambda rounded to float as the reference forms it."""
    import mpmath as mp

    mp.mp.dps = 40
    nx, ny, nz = 7, 6, 5
    W, H, D = (float(np.float32(v)) for v in (2e-3, 1.5e-3, 1e-3))
    dist, lam = float(np.float32(0.1)), float(np.float32(1e-9))
    cx, cy, cz = (float(np.float32(v)) for v in (1e-5, -2e-5, 0.0))
    q = oracle.hkl_3d(nx, ny, nz, W, H, D, dist, lam, cx, cy, cz).reshape(nz, ny, nx, 3)
    dl = mp.mpf(float(np.float32(np.float32(dist) * np.float32(lam))))
    for (z, y, x) in [(0, 0, 0), (4, 5, 6), (2, 3, 3), (1, 2, 5)]:
        p = [((mp.mpf(i) - (mp.mpf(n) - 1) / 2) / n) * e - c
             for i, n, e, c in ((x, nx, W, cx), (y, ny, H, cy), (z, nz, D, cz))]
        want = [float(np.float32(float(v / dl))) for v in p]
        assert list(q[z, y, x]) == [np.float32(v) for v in want], (z, y, x)
    # uniform spacing extent / (n d lambda) along each axis
    sx = np.diff(q[..., 0].astype(np.float64), axis=2)
    np.testing.assert_allclose(sx, W / nx / float(dl), rtol=1e-6)
    sz = np.diff(q[..., 2].astype(np.float64), axis=0)
    np.testing.assert_allclose(sz, D / nz / float(dl), rtol=1e-6)


def test_hkl_3d_centred_is_antisymmetric():
    q = oracle.hkl_3d(6, 5, 4, 1e-3, 1e-3, 1e-3, 0.05, 2e-10)
    np.testing.assert_array_equal(q[::-1], -q)
    q = oracle.hkl_3d(5, 5, 5, 1e-3, 1e-3, 1e-3, 0.05, 2e-10)
    assert np.all(q[q.shape[0] // 2] == 0)
