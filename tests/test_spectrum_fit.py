"""The host fit of the wavelength-spectrum kernel's sample factor (spectrum_fit,
a host function of the library: no GPU needed), against exact FP64 sample sums.

With sc the midpoint of the phase scales s_k, e_k = s_k / sc - 1 and t = sc q.r in
turns, the kernel takes sum_k W_k e^{-2 pi i s_k q.r} = e^{-2 pi i t} G(t), with
    sum_k W_k cos(2 pi e_k t) = P(v^2),  sum_k W_k sin(2 pi e_k t) = v Q(v^2),
v = t / tlim, tlim = 1.5 / (2 pi max|e_k|), P and Q of degree 4 evaluated in FP32.
Checked here: the FP32 polynomials against the FP64 sums over |t| <= tlim,
P(0) = sum W_k (u = 0 is one of the interpolation points), the range, and the
cases that switch the polynomial off (fewer than 4 samples) or make G constant
(all scales equal)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import _lib

DEG = 4  # degree of P and Q in v^2
REACH = 1.5  # max |2 pi e_k t| over the fitted range, radians
NMAX = _lib.MAX_SPECTRUM


class Spectrum(C.Structure):
    """The library's launch record of a spectrum (sf_kernels.hpp)."""
    _fields_ = [("n", C.c_int), ("w", C.c_float * NMAX), ("s", C.c_double * NMAX), ("sc", C.c_double),
                ("tlim", C.c_float), ("inv_tlim", C.c_float), ("p", C.c_float * (DEG + 1)),
                ("q", C.c_float * (DEG + 1))]


def fit(scales, weights) -> Spectrum:
    f = getattr(_lib.load(), "_ZN5spsim12spectrum_fitERNS_8SpectrumE")
    f.restype = None
    f.argtypes = [C.POINTER(Spectrum)]
    sp = Spectrum()
    sp.n = len(scales)
    for k, (s, w) in enumerate(zip(scales, weights)):
        sp.s[k] = float(s)
        sp.w[k] = float(w)
    f(C.byref(sp))
    return sp


def horner32(c, u):
    acc = np.full_like(u, np.float32(c[-1]))
    for a in c[-2::-1]:
        acc = (acc * u + np.float32(a)).astype(np.float32)
    return acc


def max_error(sp, scales, weights, npts=20001):
    """Largest deviation of the FP32 polynomials from the FP64 sample sums over |t| <= tlim."""
    s = np.asarray(scales, np.float64)
    w = np.asarray(weights, np.float32).astype(np.float64)  # the kernel's weights are FP32
    e = s / sp.sc - 1.0
    t = np.linspace(-sp.tlim, sp.tlim, npts)
    v = (t.astype(np.float32) * np.float32(sp.inv_tlim)).astype(np.float32)
    u = (v * v).astype(np.float32)
    p = np.array(sp.p[:], np.float32)
    q = np.array(sp.q[:], np.float32)
    gc = horner32(p, u)
    gs = (v * horner32(q, u)).astype(np.float32)
    ph = 2.0 * np.pi * e[None, :] * t[:, None]
    ec = (w * np.cos(ph)).sum(1)
    es = (w * np.sin(ph)).sum(1)
    return max(np.abs(gc - ec).max(), np.abs(gs - es).max())


@pytest.mark.parametrize("bw,n", [(0.02, 5), (0.1, 9), (0.3, 33), (0.005, 65), (0.05, 7)])
def test_fit_of_gaussian_spectra(bw, n):
    sc, w = oracle.spectrum_samples(1.5e-10, bw, n)
    sp = fit(sc, w)
    mid = 0.5 * (min(sc) + max(sc))
    assert sp.sc == pytest.approx(mid, rel=1e-15)
    emax = max(abs(s / mid - 1.0) for s in sc)
    assert sp.tlim == pytest.approx(REACH / (2 * np.pi * emax), rel=1e-6)
    assert sp.p[0] == pytest.approx(float(np.sum(np.asarray(w, np.float32))), rel=2e-7)
    assert sp.q[0] == pytest.approx(2 * np.pi * sp.tlim * float(np.sum(np.asarray(w, np.float32) * (np.asarray(sc) / mid - 1))),
                                    abs=1e-6)
    assert max_error(sp, sc, w) < 1.2e-7  # Gaussian spectra weigh the extreme samples little


@pytest.mark.parametrize("seed,n,lo,hi", [(1, 4, 0.97, 1.05), (2, 6, 0.8, 1.25), (3, 17, 1.6, 1.7), (4, 65, 0.5, 0.52)])
def test_fit_of_uneven_samples(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    sc = np.sort(rng.uniform(lo, hi, n))
    w = rng.uniform(0.1, 1.0, n)
    sp = fit(sc, w)
    assert sp.tlim > 0
    assert max_error(sp, sc, w) < 3e-7 * float(np.sum(np.abs(w)))


def test_few_samples_take_the_per_sample_sum():
    for n in (1, 2, 3):
        sp = fit(np.linspace(0.99, 1.01, n), np.ones(n) / n)
        assert sp.tlim < 0  # every lane takes the per-sample FP64 sum
        assert sp.sc == 1.0


def test_equal_scales_give_a_constant_factor():
    w = [0.1, 0.2, 0.3, 0.4]
    sp = fit([1.25] * 4, w)
    assert sp.sc == 1.25
    assert sp.tlim >= 1e29  # G(t) = sum W_k for every t
    assert sp.p[0] == pytest.approx(1.0, rel=1e-7)
    assert all(abs(x) < 1e-7 for x in sp.p[1:]) and all(abs(x) < 1e-7 for x in sp.q[:])
