"""Pins the oracle's detector chain (src/diffraction.c:1109-1204,
src/noise.c:32-51, src/amplification.c:54-153) with closed forms, the Philox
known-answer vectors and distribution tests."""
import math

import numpy as np
import pytest

import oracle

R0 = 2.81794e-15
HC = 1.98644521e-25


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert oracle.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def geometry(nx=9, ny=7, pw=1e-3, dist=0.05, cx=2e-4, cy=-3e-4):
    w, h = np.float32(nx * pw).item(), np.float32(ny * pw).item()
    x, y = np.meshgrid(np.arange(nx), np.arange(ny))
    f = lambda v: float(np.float32(v))
    px = ((x - (nx - 1) / 2) / nx) * f(w) - f(cx)
    py = ((y - (ny - 1) / 2) / ny) * f(h) - f(cy)
    return nx, ny, w, h, dist, cx, cy, px.ravel(), py.ravel()


def test_thomson_polarisation_closed_forms():
    nx, ny, w, h, dist, cx, cy, px, py = geometry()
    d = float(np.float32(dist))
    r2 = px ** 2 + py ** 2 + d * d
    np.testing.assert_allclose(oracle.thomson(nx, ny, w, h, dist, cx, cy, 0), R0 ** 2, rtol=1e-7)
    # sin(chi) = py / r  ->  cos^2 chi = 1 - py^2 / r^2
    np.testing.assert_allclose(oracle.thomson(nx, ny, w, h, dist, cx, cy, 1), R0 ** 2 * (1 - py ** 2 / r2), rtol=1e-6)
    np.testing.assert_allclose(oracle.thomson(nx, ny, w, h, dist, cx, cy, 2), R0 ** 2 * (1 - px ** 2 / r2), rtol=1e-6)
    np.testing.assert_allclose(oracle.thomson(nx, ny, w, h, dist, cx, cy, 3), R0 ** 2 * (1 + d * d / r2) / 2,
                               rtol=1e-6)


def test_solid_angle_closed_forms():
    nx, ny, w, h, dist, cx, cy, px, py = geometry()
    pw = ph = 1e-3
    A = float(np.float32(pw)) * float(np.float32(ph))
    d = float(np.float32(dist))
    r = np.sqrt(px ** 2 + py ** 2 + d * d)
    # A/d^2 cos^3(atan(rho/d)) = A d / r^3
    np.testing.assert_allclose(oracle.solid_angle(nx, ny, w, h, dist, pw, ph, cx, cy, 0), A * d / r ** 3, rtol=1e-6)
    np.testing.assert_allclose(oracle.solid_angle(nx, ny, w, h, dist, pw, ph, cx, cy, 1), A / d ** 2, rtol=1e-6)


def test_photons_and_electrons():
    rng = np.random.default_rng(0)
    th, om, I = (rng.uniform(0.5, 2, 50).astype(np.float32) for _ in range(3))
    np.testing.assert_allclose(oracle.photons(th, om, I, 3.0), th.astype(np.float64) * om * I * 3.0, rtol=1e-6)
    lam, eh = 1.3e-9, 5.8e-19
    assert oracle.electrons_per_photon(lam, eh) == pytest.approx(HC / lam / eh, rel=1e-6)
    assert oracle.electrons_per_photon(lam, 0.0) == 1.0
    assert math.isinf(oracle.electrons_per_photon(lam, 0.0, noiseless=True))


def test_binned_outputs_by_hand():
    nx, ny, bx, by = 4, 4, 2, 2
    ph = np.arange(16, dtype=np.float32)
    qe, epp, fw, mx = 0.5, 2.0, 100.0, 50.0
    out = oracle.noiseless_output(ph, nx, ny, bx, by, qe, epp, fw, mx)
    # bins: {0,1,4,5}, {2,3,6,7}, {8,9,12,13}, {10,11,14,15}; each x qe*epp*adc = 0.5
    np.testing.assert_allclose(out, np.array([10, 18, 42, 50]) * 0.5)
    el = np.full(16, 3.0, np.float32)
    noise = np.array([1.0, -100.0, 0.0, 2.5], np.float32)
    out = oracle.real_output(el, nx, ny, bx, by, 0.25, 2.0, noise, fw, mx)
    # each bin: 4 * (3 + 0.5) + noise, clamped at 0, times adc = 0.5
    np.testing.assert_allclose(out, np.array([15.0, 0.0, 14.0, 16.5]) * 0.5)
    # no ADC scaling when full well or maximum value is zero
    out = oracle.real_output(el, nx, ny, bx, by, 0.0, 0.0, noise * 0, 0.0, mx)
    np.testing.assert_allclose(out, 12.0)


@pytest.mark.parametrize("mean", [0.0, 0.3, 4.0, 25.0, 300.0, 999.0, 5000.0, 2.5e5])
def test_poisson_moments(mean):
    n = 40000
    x = np.array([oracle.poisson(mean, 7, i, 3) for i in range(n)])
    assert np.all(x == np.floor(x)) and np.all(x >= 0)
    if mean == 0:
        assert np.all(x == 0)
        return
    if mean < 1000:
        m, v = mean, mean
    else:
        # (int)(mean + sqrt(mean) N(0,1)) truncates: mean - 1/2, variance mean + 1/12
        m, v = mean - 0.5, mean + 1.0 / 12
    se_m = math.sqrt(v / n)
    se_v = v * math.sqrt(2.0 / n) * (1 + 1.0 / max(mean, 1))
    assert abs(x.mean() - m) < 5 * se_m + 1e-12
    assert abs(x.var() - v) < 5 * se_v


def test_poisson_small_mean_pmf():
    mean, n = 1.7, 60000
    x = np.array([oracle.poisson(mean, 11, i, 0) for i in range(n)]).astype(int)
    for k in range(6):
        p = math.exp(-mean) * mean ** k / math.factorial(k)
        assert abs(np.mean(x == k) - p) < 5 * math.sqrt(p * (1 - p) / n)


def test_poisson_streams_are_independent():
    a = np.array([oracle.poisson(50.0, 1, i, 0) for i in range(5000)])
    b = np.array([oracle.poisson(50.0, 1, i, 1) for i in range(5000)])
    c = np.array([oracle.poisson(50.0, 2, i, 0) for i in range(5000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.06 and abs(np.corrcoef(a, c)[0, 1]) < 0.06


def test_std_normal_moments():
    n = 40000
    x = np.array([oracle.std_normal(5, i, 0, 2) for i in range(n)])
    assert abs(x.mean()) < 5 / math.sqrt(n)
    assert abs(x.var() - 1) < 5 * math.sqrt(2 / n)
    assert abs(np.mean(x < -1.0) - 0.158655) < 0.01


def test_three_d_binning_hand_sums():
    """Binned 3-D outputs (src/amplification.c:88-110, 133-150): each output voxel
    is the float sum over its bx x by x bz block in (z, y, x) order."""
    nx, ny, nz, bx, by, bz = 5, 4, 6, 2, 2, 3
    rng = np.random.default_rng(3)
    ph = rng.uniform(0, 10, nx * ny * nz).astype(np.float32)
    out = oracle.noiseless_output(ph, nx, ny, bx, by, 1.0, 1.0, 0.0, 0.0, nz=nz, bz=bz)
    vol = ph.reshape(nz, ny, nx)
    want = []
    for z in range(nz // bz):
        for y in range(ny // by):
            for x in range(nx // bx):
                acc = np.float32(0)
                for zi in range(bz):
                    for yi in range(by):
                        for xi in range(bx):
                            acc = np.float32(acc + vol[z * bz + zi, y * by + yi, x * bx + xi])
                want.append(acc)
    assert np.array_equal(out, np.array(want, np.float32))
    # real output: + dark * exposure per voxel, + noise per bin, clamp, ADC; a 2-D call is nz = bz = 1
    el = rng.uniform(0, 5, nx * ny * nz).astype(np.float32)
    noise = rng.normal(0, 3, (nx // bx) * (ny // by) * (nz // bz)).astype(np.float32)
    ro = oracle.real_output(el, nx, ny, bx, by, 0.5, 2.0, noise, 100.0, 50.0, nz=nz, bz=bz)
    ev = el.reshape(nz, ny, nx)[: nz // bz * bz, : ny // by * by, : nx // bx * bx]  # whole bins only
    blocks = ev.reshape(nz // bz, bz, ny // by, by, nx // bx, bx).transpose(0, 2, 4, 1, 3, 5).reshape(len(noise), -1)
    np.testing.assert_allclose(ro, np.maximum(blocks.sum(1) + 1.0 * bx * by * bz + noise, 0) * 0.5, rtol=1e-5)
    # the first z slab of a bz = 1 volume is the 2-D result of the first slice
    two_d = oracle.noiseless_output(ph[: nx * ny], nx, ny, bx, by, 1.0, 1.0, 0.0, 0.0)
    slab = oracle.noiseless_output(ph, nx, ny, bx, by, 1.0, 1.0, 0.0, 0.0, nz=nz, bz=1)
    assert np.array_equal(slab[: two_d.size], two_d)


def _bin_counts(ph, nx, ny, b, qe, epp, seed, pattern):
    """Joint counts per bin recovered from oracle.real_output_bin_draws with no
    dark current, no readout noise and unit ADC (acc = count x epp exactly for
    integer epp)."""
    out = oracle.real_output_bin_draws(ph, nx, ny, b, b, qe, epp, 0.0, 0.0, 0.0, 0.0, 0.0, seed, pattern)
    return out.astype(np.float64) / epp


@pytest.mark.parametrize("bin_mean", [0.5, 3.2, 40.0, 14400.0])
def test_bin_draws_poisson_moments(bin_mean):
    """One draw per bin over 4 x 4 bins of pixels below the Gaussian branch:
    the bin count is Poisson with the summed mean (inversion below 10, PTRS
    above, up to 16 x 900 photons)."""
    nx = ny = 256
    ph = np.full(nx * ny, bin_mean / 16.0, np.float32)  # QE 1
    lam = float(np.float32(bin_mean / 16.0)) * 16
    x = _bin_counts(ph, nx, ny, 4, 1.0, 1.0, 5, 3)
    n = x.size
    z = (x - lam).sum() / math.sqrt(lam * n)
    assert abs(z) < 5, z
    vr = ((x - x.mean()) ** 2).sum() / (n - 1) / lam
    assert abs(vr - 1) < 5 * math.sqrt(2.0 / n), vr
    assert np.all(x == np.rint(x)) and x.min() >= 0


def test_bin_draws_small_mean_pmf():
    ph = np.full(512 * 512, 0.2 / 16, np.float32)
    x = _bin_counts(ph, 512, 512, 4, 1.0, 1.0, 9, 0)
    lam = float(np.float32(0.2 / 16)) * 16
    for k in range(4):
        p = math.exp(-lam) * lam ** k / math.factorial(k)
        assert abs(np.mean(x == k) - p) < 5 * math.sqrt(p * (1 - p) / x.size) + 1e-6, k


def test_bin_draws_with_gaussian_pixels():
    """One pixel of each bin at mean 5000 (the reference's truncated Gaussian:
    mean 4999.5, variance 5000 + 1/12) and fifteen at 0.2: the bin count has
    the summed moments."""
    nx = ny = 256
    ph = np.full((ny, nx), 0.2, np.float32)
    ph[::4, ::4] = 5000.0
    x = _bin_counts(ph.ravel(), nx, ny, 4, 1.0, 1.0, 6, 1)
    m = 4999.5 + 15 * float(np.float32(0.2))
    v = 5000.0 + 1.0 / 12 + 15 * float(np.float32(0.2))
    n = x.size
    assert abs((x - m).sum() / math.sqrt(v * n)) < 5
    assert abs(((x - x.mean()) ** 2).sum() / (n - 1) / v - 1) < 5 * math.sqrt(2.0 / n)


def test_bin_draws_all_gaussian_equals_per_pixel_chain():
    """When every pixel takes the Gaussian branch, the joint Poisson count is
    0 and, without readout noise, the binned real output is bitwise the
    per-pixel chain's with the same draws."""
    nx, ny, b = 32, 24, 4
    rng = np.random.default_rng(4)
    ph = rng.uniform(2000, 9000, nx * ny).astype(np.float32)
    qe, epp, dark, t, fw, mv, seed, pat = 0.5, 263.5, 0.1, 1e-3, 2e5, 65535.0, 12, 7
    got = oracle.real_output_bin_draws(ph, nx, ny, b, b, qe, epp, dark, t, 0.0, fw, mv, seed, pat)
    lam = (ph * np.float32(qe)).astype(np.float64)
    cnt = np.array([oracle.poisson(l, seed, i, pat) for i, l in enumerate(lam)], np.float32)
    el = cnt * np.float32(epp)
    nb = (nx // b) * (ny // b)
    ref = oracle.real_output(el, nx, ny, b, b, dark, t, np.zeros(nb, np.float32), fw, mv)
    assert np.array_equal(got, ref)


def test_bin_draws_readout_noise_moments():
    """No photons: the real output is max(N(0, sigma), 0) x ADC per bin, from
    the Box-Muller pair of each bin's block."""
    nx = ny = 512
    s = 10.0
    out = oracle.real_output_bin_draws(np.zeros(nx * ny, np.float32), nx, ny, 4, 4, 1.0, 263.5, 0.0, 0.0, s, 0.0,
                                       0.0, 3, 1).astype(np.float64)
    n = out.size
    assert abs(np.mean(out == 0) - 0.5) < 5 * math.sqrt(0.25 / n)
    assert abs(out.mean() - s / math.sqrt(2 * math.pi)) < 5 * s * math.sqrt(0.5 / n)
    assert abs((out ** 2).mean() - s * s / 2) < 5 * s * s * math.sqrt(2.0 / n)
