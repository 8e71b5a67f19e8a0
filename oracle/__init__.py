"""ORACLE — test infrastructure only.

NumPy/ctypes front end of oracle/oracle.c, the CPU restatement of spsim's
structure-factor hot path (see the header of oracle.c for what each function
restates and how it is pinned). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, always as the checker, never as the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            from . import build as _b

            _b.build()
        # PyTorch bundles its own OpenMP runtime under the SONAME libgomp.so.1;
        # importing it first lets this library share it instead of loading
        # the system copy beside it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(path))
        L.oracle_form_factor.restype = C.c_double
        L.oracle_form_factor.argtypes = [C.c_double, C.c_int, C.c_double, _f32p]
        L.oracle_hkl_detector.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_int, C.c_float, _f32p]
        L.oracle_rotate_hkl.argtypes = [_f32p, C.c_int, _f64p, _f32p]
        L.oracle_hkl_3d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_float, _f32p]
        L.oracle_sf_f64.argtypes = [_f32p, C.c_int, _i32p, _f32p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                    _f32p, C.c_float, C.c_void_p, _f64p, C.c_int, C.c_void_p]
        L.oracle_spectrum_samples.restype = C.c_int
        L.oracle_spectrum_samples.argtypes = [C.c_double, C.c_double, C.c_int, _f64p, _f64p]
        L.oracle_sf_spectrum_f64.argtypes = [_f32p, C.c_int, _i32p, _f32p, C.c_void_p, C.c_void_p, C.c_int, _f32p,
                                             C.c_float, C.c_int, _f64p, _f64p, C.c_void_p, _f64p, C.c_int,
                                             C.c_void_p]
        L.oracle_sf_refsem.argtypes = [_f32p, C.c_int, _i32p, _f32p, C.c_void_p, C.c_int, C.c_int, _f32p, C.c_float,
                                       C.c_void_p, _f32p, C.c_int]
        L.oracle_sf_grad_f64.argtypes = [_f32p, C.c_int, _i32p, _f32p, C.c_void_p, C.c_void_p, C.c_int, _f32p,
                                         C.c_float, _f64p, _f64p, C.c_void_p, C.c_int]
        L.oracle_sf_grad_given_f_f64.argtypes = [_f32p, C.c_int, _f64p, _i32p, _f32p, C.c_void_p, C.c_void_p,
                                                 _f32p, C.c_float, _f64p, C.c_int, _i32p, _f64p, C.c_void_p,
                                                 C.c_int]
        L.oracle_thomson.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                     _f32p]
        L.oracle_solid_angle.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, C.c_int, _f32p]
        L.oracle_photons.argtypes = [_f32p, _f32p, _f32p, C.c_float, C.c_size_t, _f32p]
        L.oracle_electrons_per_photon.restype = C.c_float
        L.oracle_electrons_per_photon.argtypes = [C.c_float, C.c_float, C.c_int]
        L.oracle_noiseless_output.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                              C.c_float, C.c_float, C.c_float, C.c_float, _f32p]
        L.oracle_real_output.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                         C.c_float, _f32p, C.c_float, C.c_float, _f32p]
        L.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.oracle_std_normal.restype = C.c_double
        L.oracle_std_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        L.oracle_poisson.restype = C.c_double
        L.oracle_poisson.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32]
        L.oracle_poisson_exact.restype = C.c_double
        L.oracle_poisson_exact.argtypes = [C.c_double, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
        L.oracle_real_output_bin_draws.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                                   C.c_float, C.c_float, C.c_uint64, C.c_uint32, _f32p]
        _LIB = L
    return _LIB


def _opt(a, dtype):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr, arr.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------- form factors

ATOMSF_FIXTURE = HERE.parent / "tests" / "golden" / "atomsf_table.npy"


def reference_atomsf() -> np.ndarray:
    """The reference's compiled-in Cromer-Mann table atomsf[100][9]
    (include/atomsf.cdata, src/diffraction.c:58-60) as float32, read from the
    golden fixture that tests/golden/make_atomsf_fixture.py wrote from that
    file. It does not come from the product's own copy of the table."""
    vals = np.load(ATOMSF_FIXTURE, allow_pickle=False)
    assert vals.shape == (100, 9) and vals.dtype == np.float32, (vals.shape, vals.dtype)
    return vals


def default_form_factors() -> np.ndarray:
    """[101][9] table: rows 0..99 the reference's atomsf table (reference_atomsf),
    row 100 the unit scatterer (f = c = 1) that stands for delta atoms."""
    unit = np.zeros((1, 9), np.float32)
    unit[0, 8] = 1.0
    return np.ascontiguousarray(np.vstack([reference_atomsf(), unit]))


def form_factor(d: float, Z: int, B: float = 0.0, ff: np.ndarray | None = None) -> float:
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32)
    return lib().oracle_form_factor(float(d), int(Z), float(B), ff.ravel())


# --------------------------------------------------------------- geometry

def hkl_detector(nx, ny, width, height, distance, center_x, center_y, spherical, wavelength) -> np.ndarray:
    out = np.zeros(nx * ny * 3, np.float32)
    lib().oracle_hkl_detector(nx, ny, width, height, distance, center_x, center_y, spherical, wavelength, out)
    return out.reshape(ny * nx, 3)


def hkl_3d(nx, ny, nz, width, height, depth, distance, wavelength, center_x=0.0, center_y=0.0, center_z=0.0):
    out = np.zeros(nx * ny * nz * 3, np.float32)
    lib().oracle_hkl_3d(nx, ny, nz, width, height, depth, center_x, center_y, center_z, distance, wavelength, out)
    return out.reshape(-1, 3)


def rotate_hkl(hkl: np.ndarray, R: np.ndarray) -> np.ndarray:
    hkl = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3)
    out = np.empty_like(hkl)
    lib().oracle_rotate_hkl(hkl.ravel(), hkl.shape[0], np.ascontiguousarray(R, np.float64).ravel(), out.ravel())
    return out


# --------------------------------------------------------------- structure factor

def sf_f64(hkl, Z, pos, amp=None, atom_b=None, delta=False, vector_mode=False, B=0.0, ff=None, threads=0, R=None):
    """FP64 oracle. R (3x3) rotates q exactly (q' = R q) first.
    Returns (F complex128 [npix], I float64 [npix])."""
    hkl = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3)
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32)
    amp_a, amp_p = _opt(amp, np.float32)
    b_a, b_p = _opt(atom_b, np.float32)
    n = hkl.shape[0]
    F = np.zeros(2 * n, np.float64)
    I = np.zeros(n, np.float64)
    R_a, R_p = _opt(None if R is None else np.asarray(R, np.float64).reshape(9), np.float64)
    lib().oracle_sf_f64(hkl.ravel(), n, Z, pos.ravel(), amp_p, b_p, Z.shape[0], int(delta), int(vector_mode),
                        ff.ravel(), float(B), F.ctypes.data_as(C.c_void_p), I, int(threads), R_p)
    return F[0::2] + 1j * F[1::2], I


def sf_nfft_f64(hkl, Z, pos, delta=False, B=0.0, ff=None, threads=0, R=None):
    """FP64 value of the sums compute_pattern_on_list_by_nfft approximates
    (src/diffraction.c:528-675): over every element present, Z == 0 included
    (:556-563, :599-606), f_Z(|q|) with the global B
    (multiply_pattern_on_list_with_scattering_factor :392-400; f = 1 for every
    atom under delta_atoms, :659-661) times sum_j exp(-2 pi i q.r_j) with unit
    strengths (:619-620). The NNFFT's own approximation error is not modelled.
    Built from sf_f64: the Z > 0 atoms exactly as compute_pattern_on_list sums
    them (amplitude 1), plus the Z == 0 atoms as unit scatterers times row 0 of
    the form-factor table. R (3x3) rotates q exactly first (q' = R q, |q| kept).
    Returns (F complex128, I float64)."""
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    if delta:
        # delta_atoms ignores the element, so any non-zero row stands for Z == 0
        return sf_f64(hkl, np.where(Z == 0, 1, Z).astype(np.int32), pos, delta=True, threads=threads, R=R)
    F, _ = sf_f64(hkl, Z, pos, B=B, ff=ff, threads=threads, R=R)
    z0 = Z == 0
    if z0.any():
        Fu, _ = sf_f64(hkl, np.ones(int(z0.sum()), np.int32), pos[z0], delta=True, threads=threads, R=R)
        q = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3).astype(np.float64)
        d = np.sqrt((q * q).sum(1))
        F = F + np.array([form_factor(x, 0, B, ff) for x in d]) * Fu
    return F, (F * F.conj()).real


def spectrum_samples(wavelength, bandwidth, nsamples):
    """(scales, weights) of the spectrum path; an even count above 1 is raised by one."""
    n = max(1, int(nsamples) + 1)
    sc = np.zeros(n, np.float64)
    w = np.zeros(n, np.float64)
    k = lib().oracle_spectrum_samples(float(wavelength), float(bandwidth), int(nsamples), sc, w)
    return sc[:k].copy(), w[:k].copy()


def sf_spectrum_f64(hkl, Z, pos, scales, weights, amp=None, atom_b=None, B=0.0, ff=None, threads=0, R=None):
    """FP64 spectrum structure factor. Returns (F complex128 [npix], I float64 [npix])."""
    hkl = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3)
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32)
    sc = np.ascontiguousarray(scales, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    amp_a, amp_p = _opt(amp, np.float32)
    b_a, b_p = _opt(atom_b, np.float32)
    R_a, R_p = _opt(None if R is None else np.asarray(R, np.float64).reshape(9), np.float64)
    n = hkl.shape[0]
    F = np.zeros(2 * n, np.float64)
    I = np.zeros(n, np.float64)
    lib().oracle_sf_spectrum_f64(hkl.ravel(), n, Z, pos.ravel(), amp_p, b_p, Z.shape[0], ff.ravel(), float(B),
                                 sc.shape[0], sc, w, F.ctypes.data_as(C.c_void_p), I, int(threads), R_p)
    return F[0::2] + 1j * F[1::2], I


def sf_refsem(hkl, Z, pos, illum=None, delta=False, B=0.0, ff=None, threads=0):
    """compute_pattern_on_list in the reference's own precision. Returns (F complex64, I float32)."""
    hkl = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3)
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32)
    il_a, il_p = _opt(illum, np.float32)
    n = hkl.shape[0]
    F = np.zeros(2 * n, np.float32)
    I = np.zeros(n, np.float32)
    lib().oracle_sf_refsem(hkl.ravel(), n, Z, pos.ravel(), il_p, Z.shape[0], int(delta), ff.ravel(), float(B),
                           F.ctypes.data_as(C.c_void_p), I, int(threads))
    return (F[0::2] + 1j * F[1::2]).astype(np.complex64), I


def sf_grad_f64(hkl, Z, pos, g, amp=None, atom_b=None, B=0.0, ff=None, threads=0):
    """Gradient of sum_p g_p I_p w.r.t. positions [N,3] and B factors [N] (FP64)."""
    hkl = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3)
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32)
    amp_a, amp_p = _opt(amp, np.float32)
    b_a, b_p = _opt(atom_b, np.float32)
    g = np.ascontiguousarray(g, np.float64)
    N = Z.shape[0]
    gp = np.zeros(3 * N, np.float64)
    gb = np.zeros(N, np.float64)
    lib().oracle_sf_grad_f64(hkl.ravel(), hkl.shape[0], Z, pos.ravel(), amp_p, b_p, N, ff.ravel(), float(B), g, gp,
                             gb.ctypes.data_as(C.c_void_p), int(threads))
    return gp.reshape(N, 3), gb


def sf_grad_given_f_f64(hkl, F, Z, pos, g, atoms, amp=None, atom_b=None, B=0.0, ff=None, threads=0):
    """The gradient of sum_p g_p I_p for the listed atoms only, given the FP64
    forward F (complex [npix], e.g. from sf_f64 over every atom). Returns
    (grad_pos [len(atoms), 3], grad_b [len(atoms)]) in the order of `atoms`."""
    hkl = np.ascontiguousarray(hkl, np.float32).reshape(-1, 3)
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32)
    Fr = np.empty(2 * hkl.shape[0], np.float64)
    Fr[0::2] = np.real(F)
    Fr[1::2] = np.imag(F)
    amp_a, amp_p = _opt(amp, np.float32)
    b_a, b_p = _opt(atom_b, np.float32)
    g = np.ascontiguousarray(g, np.float64)
    sub = np.ascontiguousarray(atoms, np.int32)
    assert sub.size == 0 or (sub.min() >= 0 and sub.max() < Z.shape[0])
    gp = np.zeros(3 * sub.size, np.float64)
    gb = np.zeros(sub.size, np.float64)
    lib().oracle_sf_grad_given_f_f64(hkl.ravel(), hkl.shape[0], Fr, Z, pos.ravel(), amp_p, b_p, ff.ravel(),
                                     float(B), g, int(sub.size), sub, gp, gb.ctypes.data_as(C.c_void_p),
                                     int(threads))
    return gp.reshape(-1, 3), gb


# --------------------------------------------------------------- detector chain

def thomson(nx, ny, width, height, distance, cx=0.0, cy=0.0, polarization=0):
    out = np.zeros(nx * ny, np.float32)
    lib().oracle_thomson(nx, ny, width, height, distance, cx, cy, polarization, out)
    return out


def solid_angle(nx, ny, width, height, distance, pixel_width, pixel_height, cx=0.0, cy=0.0, spherical=0):
    out = np.zeros(nx * ny, np.float32)
    lib().oracle_solid_angle(nx, ny, width, height, distance, pixel_width, pixel_height, cx, cy, spherical, out)
    return out


def photons(th, om, I, beam):
    th, om, I = (np.ascontiguousarray(a, np.float32) for a in (th, om, I))
    out = np.zeros_like(I)
    lib().oracle_photons(th, om, I, float(beam), I.size, out)
    return out


def electrons_per_photon(wavelength, e_h, noiseless=False):
    return lib().oracle_electrons_per_photon(wavelength, e_h, int(noiseless))


def noiseless_output(ph, nx, ny, bx, by, qe, epp, full_well, max_value, nz=1, bz=1):
    ph = np.ascontiguousarray(ph, np.float32)
    out = np.zeros((nx // bx) * (ny // by) * (nz // bz), np.float32)
    lib().oracle_noiseless_output(ph, nx, ny, nz, bx, by, bz, qe, epp, full_well, max_value, out)
    return out


def real_output(el, nx, ny, bx, by, dark, exposure, noise, full_well, max_value, nz=1, bz=1):
    el = np.ascontiguousarray(el, np.float32)
    noise = np.ascontiguousarray(noise, np.float32)
    out = np.zeros((nx // bx) * (ny // by) * (nz // bz), np.float32)
    lib().oracle_real_output(el, nx, ny, nz, bx, by, bz, dark, exposure, noise, full_well, max_value, out)
    return out


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return list(o)


def std_normal(seed, idx, pattern, purpose):
    return lib().oracle_std_normal(seed, idx, pattern, purpose)


def poisson(mean, seed, idx, pattern):
    return lib().oracle_poisson(float(mean), int(seed), int(idx), int(pattern))


def poisson_exact(mean, seed, idx, pattern, purpose):
    return lib().oracle_poisson_exact(float(mean), int(seed), int(idx), int(pattern), int(purpose))


def real_output_bin_draws(ph, nx, ny, bx, by, qe, epp, dark, exposure, readout_noise, full_well, max_value, seed,
                          pattern, nz=1, bz=1):
    """Binned real output with one Poisson draw per bin (see oracle.c)."""
    ph = np.ascontiguousarray(ph, np.float32)
    out = np.zeros((nx // bx) * (ny // by) * (nz // bz), np.float32)
    lib().oracle_real_output_bin_draws(ph, nx, ny, nz, bx, by, bz, qe, epp, dark, exposure, readout_noise, full_well,
                                       max_value, int(seed), int(pattern), out)
    return out
