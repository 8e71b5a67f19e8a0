"""Wavelength-spectrum oracle (oracle_spectrum_samples / oracle_sf_spectrum_f64,
the spectrum branch of the reference's GPU path, src/diffraction.cu:300-318 and
416-476) and the library's own sample generator, on the CPU.

What pins them: closed forms of the samples (wavelengths evenly spread over
lambda (1 +- bw/2), Gaussian weights exp(-2 (2k/(S-1) - 1)^2) normalised to one,
phase scales lambda / lambda_k, even counts raised by one), the plain FP64
oracle for a single sample, atoms at the origin, Friedel symmetry, and
linearity over samples for point scatterers: with f = 1 the spectrum pattern
is the weighted sum of plain amplitudes at q scaled exactly by s_k (the plain
oracle's exact q rotation with R = s_k 1).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import _lib
from spsim_amd import workloads as wl


def samples_closed_form(lam, bw, S):
    if S > 1 and S % 2 == 0:
        S += 1
    if bw == 0 or S <= 1:
        return np.ones(1), np.ones(1)
    k = np.arange(S)
    w = np.exp(-2.0 * (2.0 * k / (S - 1) - 1.0) ** 2)
    return 1.0 / (1.0 - bw / 2 + k * bw / (S - 1)), w / w.sum()


@pytest.mark.parametrize("lam,bw,S", [(1e-9, 0.01, 5), (0.2e-9, 0.05, 4), (1.3e-9, 0.002, 65), (1e-10, 0.3, 3),
                                      (1e-9, 0.0, 5), (1e-9, 0.01, 1), (1e-9, 0.01, 2)])
def test_samples_closed_form(lam, bw, S):
    sc, w = oracle.spectrum_samples(lam, bw, S)
    sc_ref, w_ref = samples_closed_form(lam, bw, S)
    np.testing.assert_allclose(sc, sc_ref, rtol=1e-14)
    # (lambda_k - lambda) cancels about log10(1/bw) digits before the Gaussian
    np.testing.assert_allclose(w, w_ref, rtol=1e-11)
    assert abs(w.sum() - 1) < 1e-15
    assert np.allclose(w, w[::-1], rtol=1e-11)                      # symmetric spectrum
    # the library's own generator (host C) gives the same samples
    n = _lib.MAX_SPECTRUM
    a, b = np.zeros(n), np.zeros(n)
    k = _lib.load().spsim_hip_spectrum_samples(lam, bw, S, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p))
    assert k == sc.size
    np.testing.assert_allclose(a[:k], sc, rtol=1e-15)
    np.testing.assert_allclose(b[:k], w, rtol=1e-12)


def test_library_sample_limits():
    L = _lib.load()
    a, b = np.zeros(_lib.MAX_SPECTRUM), np.zeros(_lib.MAX_SPECTRUM)
    vp = lambda x: x.ctypes.data_as(C.c_void_p)
    assert L.spsim_hip_spectrum_samples(1e-9, 0.01, 64, vp(a), vp(b)) == 65   # raised to odd, at the limit
    assert L.spsim_hip_spectrum_samples(1e-9, 0.01, 66, vp(a), vp(b)) == -1   # 67 samples exceed the limit
    assert L.spsim_hip_spectrum_samples(-1e-9, 0.01, 5, vp(a), vp(b)) < 0


def molecule(n=60, seed=0):
    rng = np.random.default_rng(seed)
    Z, pos, B = wl.random_molecule(n, 2e-9, rng, atom_b_range=(10.0, 30.0))
    return Z, pos, B


def qlist(n=12, lam=1e-9):
    return oracle.hkl_detector(n, n, n * 3e-4, n * 3e-4, 0.05, 0.0, 0.0, 0, lam)


def test_single_sample_is_the_plain_oracle():
    Z, pos, B = molecule()
    q = qlist()
    F1, I1 = oracle.sf_spectrum_f64(q, Z, pos, [1.0], [1.0], atom_b=B, B=2.0)
    F0, I0 = oracle.sf_f64(q, Z, pos, atom_b=B, B=2.0)
    np.testing.assert_allclose(I1, I0, rtol=1e-13)
    np.testing.assert_allclose(F1, F0, rtol=1e-13, atol=1e-13 * np.abs(F0).max())


def test_atoms_at_origin_and_friedel():
    q = qlist(9)
    sc, w = oracle.spectrum_samples(1e-9, 0.2, 7)
    Z = np.array([6, 8, 8, 16], np.int32)
    F, I = oracle.sf_spectrum_f64(q, Z, np.zeros((4, 3), np.float32), sc, w)
    ff = oracle.default_form_factors()
    d = np.linalg.norm(q.astype(np.float64), axis=1)
    fsum = sum(np.array([oracle.form_factor(x, z, ff=ff) for x in d]) for z in Z)
    np.testing.assert_allclose(I, fsum ** 2, rtol=1e-13)
    Zm, pos, _ = molecule(40, 3)
    _, Ip = oracle.sf_spectrum_f64(q, Zm, pos, sc, w)
    _, Im = oracle.sf_spectrum_f64(-q, Zm, pos, sc, w)
    np.testing.assert_allclose(Ip, Im, rtol=1e-12)


def test_linearity_over_samples_for_point_scatterers():
    """f = 1 for every species: F_spectrum(q) = sum_k w_k F_plain(s_k q) exactly."""
    unit = np.zeros((101, 9), np.float32)
    unit[:, 8] = 1.0
    Z, pos, _ = molecule(80, 5)
    q = qlist(10)
    sc, w = oracle.spectrum_samples(1e-9, 0.08, 9)
    F, I = oracle.sf_spectrum_f64(q, Z, pos, sc, w, ff=unit)
    acc = np.zeros(q.shape[0], np.complex128)
    for s_k, w_k in zip(sc, w):
        Fk, _ = oracle.sf_f64(q, Z, pos, ff=unit, R=np.eye(3) * s_k)
        acc += w_k * Fk
    np.testing.assert_allclose(F, acc, rtol=1e-12, atol=1e-12 * np.abs(acc).max())
    np.testing.assert_allclose(I, np.abs(acc) ** 2, rtol=1e-11)


def test_narrow_band_tends_to_plain():
    Z, pos, _ = molecule(50, 7)
    q = qlist(8)
    _, I0 = oracle.sf_f64(q, Z, pos)
    prev = None
    for bw in (1e-2, 1e-3, 1e-4):
        sc, w = oracle.spectrum_samples(1e-9, bw, 5)
        _, I = oracle.sf_spectrum_f64(q, Z, pos, sc, w)
        err = np.linalg.norm(I - I0) / np.linalg.norm(I0)
        if prev is not None:
            assert err < prev / 20          # quadratic in the bandwidth
        prev = err
    assert prev < 1e-6
