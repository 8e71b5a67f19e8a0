"""Device-resident Python front end of the C-ABI (include/spsim_hip.h).

PyTorch provides device memory and the stream; every computation is a call
into libspsim_amd.so with raw device pointers. Inputs are NumPy arrays or
CUDA (HIP) tensors; outputs are torch tensors on the current device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import Atom, DetectorParams, Segment, Tile, check


def _ptr(t: torch.Tensor | None) -> C.c_void_p | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _stream(stream: torch.cuda.Stream | None = None) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def default_form_factors() -> np.ndarray:
    out = np.zeros((_lib.NSPECIES, _lib.FF_STRIDE), np.float32)
    check(_lib.load().spsim_hip_default_form_factors(out.ctypes.data_as(C.c_void_p)), "spsim_hip_default_form_factors")
    return out


@dataclass
class PackedAtoms:
    """Species-sorted atoms in host memory (see spsim_hip_pack_atoms)."""

    atoms: np.ndarray  # [n, 4] float32: x, y, z, amplitude
    b: np.ndarray | None  # [n] float32 per-atom B (A^2) or None
    perm: np.ndarray  # [n] int32: packed index -> original index
    segs: np.ndarray  # [nsegs, 4] int32: start, count, species, 0

    @property
    def natoms(self) -> int:
        return int(self.atoms.shape[0])


def pack_atoms(Z, pos, amp=None, atom_b=None, delta_atoms=False, mode=_lib.PACK_PLAIN) -> PackedAtoms:
    Z = np.ascontiguousarray(Z, np.int32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = Z.shape[0]
    if pos.shape[0] != n:
        raise ValueError("Z and pos have different atom counts")
    amp_a = None if amp is None else np.ascontiguousarray(amp, np.float32)
    b_a = None if atom_b is None else np.ascontiguousarray(atom_b, np.float32)
    atoms = np.zeros((max(n, 1), 4), np.float32)
    outb = np.zeros(max(n, 1), np.float32)
    perm = np.zeros(max(n, 1), np.int32)
    segs = np.zeros((_lib.NSPECIES, 4), np.int32)
    nat, nseg = C.c_int(0), C.c_int(0)
    vp = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    check(_lib.load().spsim_hip_pack_atoms(vp(Z), vp(pos), vp(amp_a), vp(b_a), n, int(bool(delta_atoms)), int(mode),
                                           vp(atoms), vp(outb), vp(perm), vp(segs), C.byref(nat), C.byref(nseg)),
          "spsim_hip_pack_atoms")
    k = nat.value
    return PackedAtoms(atoms[:k].copy(), outb[:k].copy() if b_a is not None else None, perm[:k].copy(),
                       segs[:nseg.value].copy())


class DeviceAtoms:
    """PackedAtoms uploaded to HBM."""

    def __init__(self, packed: PackedAtoms, device=None):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.packed = packed
        self.natoms = packed.natoms
        self.nsegs = int(packed.segs.shape[0])
        self.atoms = torch.from_numpy(packed.atoms.reshape(-1, 4) if packed.natoms else np.zeros((1, 4), np.float32)).to(dev)
        self.b = None if packed.b is None else torch.from_numpy(packed.b if packed.natoms else np.zeros(1, np.float32)).to(dev)
        self.segs = torch.from_numpy(packed.segs if self.nsegs else np.zeros((1, 4), np.int32)).to(dev)
        self.perm = torch.from_numpy(packed.perm.astype(np.int64)).to(dev)


def form_factor_table(ff: np.ndarray | None = None, device=None) -> torch.Tensor:
    ff = default_form_factors() if ff is None else np.ascontiguousarray(ff, np.float32).reshape(_lib.NSPECIES, 9)
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    return torch.from_numpy(ff).to(dev)


def make_tile(pixels_per_lane=4, chunk_atoms=256, atom_splits=0, waves=4) -> Tile:
    return Tile(pixels_per_lane, waves, chunk_atoms, atom_splits)


def _rot(R):
    if R is None:
        return None, None
    a = np.ascontiguousarray(R, np.float64).reshape(9)
    return a, a.ctypes.data_as(C.c_void_p)


def structure_factor(hkl: torch.Tensor, atoms: DeviceAtoms, ff: torch.Tensor, B: float = 0.0,
                     F: torch.Tensor | None = None, I: torch.Tensor | None = None, tile: Tile | None = None,
                     stream=None, want_F: bool = True, R=None):
    """One pattern on the q list hkl ([npix, 3] float32 device tensor), with q
    rotated exactly by R (3x3) when given. Returns (F [npix, 2] float32 or
    None, I [npix] float32)."""
    npix = int(hkl.shape[0])
    if want_F and F is None:
        F = torch.empty((npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((npix,), dtype=torch.float32, device=hkl.device)
    Ra, Rp = _rot(R)
    rc = _lib.load().spsim_hip_sf_pattern(_ptr(hkl), npix, _ptr(atoms.atoms), _ptr(atoms.b), _ptr(atoms.segs),
                                          atoms.nsegs, atoms.natoms, _ptr(ff), float(B), Rp,
                                          _ptr(F if want_F else None), _ptr(I),
                                          C.byref(tile) if tile is not None else None, _stream(stream))
    check(rc, "spsim_hip_sf_pattern")
    return (F if want_F else None), I


def structure_factor_grid(hkl: torch.Tensor, nx: int, ny: int, atoms: DeviceAtoms, ff: torch.Tensor,
                          B: float = 0.0, F: torch.Tensor | None = None, I: torch.Tensor | None = None,
                          tile: Tile | None = None, stream=None, want_F: bool = True, R=None):
    """One pattern on a full nx x ny detector list ([ny*nx, 3]), q rotated
    exactly by R when given; uses the mirror-symmetric kernel where the list is
    symmetric (tile.pixels_per_lane = quads per lane, 1 or 2). Returns (F or None, I)."""
    npix = nx * ny
    if int(hkl.shape[0]) != npix:
        raise ValueError("hkl must have nx*ny rows")
    if want_F and F is None:
        F = torch.empty((npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((npix,), dtype=torch.float32, device=hkl.device)
    Ra, Rp = _rot(R)
    rc = _lib.load().spsim_hip_sf_grid(_ptr(hkl), nx, ny, _ptr(atoms.atoms), _ptr(atoms.b), _ptr(atoms.segs),
                                       atoms.nsegs, atoms.natoms, _ptr(ff), float(B), Rp, _ptr(F if want_F else None),
                                       _ptr(I), C.byref(tile) if tile is not None else None, _stream(stream))
    check(rc, "spsim_hip_sf_grid")
    return (F if want_F else None), I


def spectrum_samples(wavelength: float, bandwidth: float, nsamples: int):
    """(scales, weights) of a Gaussian spectrum, as spsim_hip_spectrum_samples draws them."""
    sc = np.zeros(_lib.MAX_SPECTRUM, np.float64)
    w = np.zeros(_lib.MAX_SPECTRUM, np.float64)
    n = _lib.load().spsim_hip_spectrum_samples(float(wavelength), float(bandwidth), int(nsamples),
                                               sc.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    if n < 1:
        raise ValueError(f"spectrum of {nsamples} samples is not supported (at most {_lib.MAX_SPECTRUM})")
    return sc[:n].copy(), w[:n].copy()


def structure_factor_spectrum(hkl: torch.Tensor, atoms: DeviceAtoms, ff: torch.Tensor, scales, weights,
                              B: float = 0.0, F: torch.Tensor | None = None, I: torch.Tensor | None = None,
                              tile: Tile | None = None, stream=None, want_F: bool = True, R=None):
    """One pattern averaged over wavelength samples (phase scales, weights); returns (F or None, I)."""
    npix = int(hkl.shape[0])
    sc = np.ascontiguousarray(scales, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    if sc.shape != w.shape:
        raise ValueError("scales and weights differ in length")
    if want_F and F is None:
        F = torch.empty((npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((npix,), dtype=torch.float32, device=hkl.device)
    Ra, Rp = _rot(R)
    rc = _lib.load().spsim_hip_sf_spectrum(_ptr(hkl), npix, _ptr(atoms.atoms), _ptr(atoms.b), _ptr(atoms.segs),
                                           atoms.nsegs, atoms.natoms, _ptr(ff), float(B), Rp, int(sc.shape[0]),
                                           sc.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                                           _ptr(F if want_F else None), _ptr(I),
                                           C.byref(tile) if tile is not None else None, _stream(stream))
    check(rc, "spsim_hip_sf_spectrum")
    return (F if want_F else None), I


def structure_factor_nfft(hkl: torch.Tensor, atoms: DeviceAtoms, ff: torch.Tensor, B: float = 0.0,
                          eps: float = 1e-6, F: torch.Tensor | None = None, I: torch.Tensor | None = None,
                          stream=None, want_F: bool = True, R=None):
    """One pattern by the type-3 non-uniform FFT (spsim_hip_sf_nfft), with q
    rotated exactly by R (3x3) when given; atoms packed with PACK_NFFT give the
    reference's NFFT semantics, atoms packed with per-atom B factors are spread
    as Gaussians that carry them. Returns (F or None, I, info) with
    info = {W, M, nf, L, W2, staged, Wg} (per-axis sizes as tuples; W2 the
    interpolation window's width; staged 1 when the interpolation staged grid
    boxes in LDS; Wg the Gaussian spread width, 0 without per-atom B)."""
    npix = int(hkl.shape[0])
    if want_F and F is None:
        F = torch.empty((npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((npix,), dtype=torch.float32, device=hkl.device)
    info = (C.c_int * 13)()
    Ra, Rp = _rot(R)
    rc = _lib.load().spsim_hip_sf_nfft(_ptr(hkl), npix, _ptr(atoms.atoms), _ptr(atoms.b), _ptr(atoms.segs),
                                       atoms.nsegs, atoms.natoms, _ptr(ff), float(B), Rp, float(eps),
                                       _ptr(F if want_F else None), _ptr(I), C.cast(info, C.c_void_p),
                                       _stream(stream))
    check(rc, "spsim_hip_sf_nfft")
    v = list(info)
    return (F if want_F else None), I, {"W": v[0], "M": tuple(v[1:4]), "nf": tuple(v[4:7]), "L": tuple(v[7:10]), "W2": v[10], "staged": v[11], "Wg": v[12]}


def structure_factor_spectrum_nfft(hkl: torch.Tensor, atoms: DeviceAtoms, ff: torch.Tensor, scales, weights,
                                   B: float = 0.0, eps: float = 1e-6, R=None, F: torch.Tensor | None = None,
                                   I: torch.Tensor | None = None, stream=None, want_F: bool = True):
    """One pattern averaged over wavelength samples through the NUFFT
    (spsim_hip_sf_spectrum_nfft; no per-atom B). Returns (F or None, I, info)."""
    if atoms.b is not None:
        raise ValueError("the spectrum NUFFT takes no per-atom B factors")
    npix = int(hkl.shape[0])
    sc = np.ascontiguousarray(scales, np.float64)
    w = np.ascontiguousarray(weights, np.float64)
    if sc.shape != w.shape:
        raise ValueError("scales and weights differ in length")
    if want_F and F is None:
        F = torch.empty((npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((npix,), dtype=torch.float32, device=hkl.device)
    Fk = torch.empty((int(sc.shape[0]), npix, 2), dtype=torch.float32, device=hkl.device)
    info = (C.c_int * 13)()
    Ra, Rp = _rot(R)  # Ra keeps the matrix alive through the call
    rc = _lib.load().spsim_hip_sf_spectrum_nfft(_ptr(hkl), npix, _ptr(atoms.atoms), _ptr(atoms.segs), atoms.nsegs,
                                                atoms.natoms, _ptr(ff), float(B), Rp, int(sc.shape[0]),
                                                sc.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                                                float(eps), _ptr(Fk), _ptr(F if want_F else None), _ptr(I),
                                                C.cast(info, C.c_void_p), _stream(stream))
    check(rc, "spsim_hip_sf_spectrum_nfft")
    v = list(info)
    return (F if want_F else None), I, {"W": v[0], "M": tuple(v[1:4]), "nf": tuple(v[4:7]), "L": tuple(v[7:10]),
                                        "W2": v[10], "staged": v[11], "Wg": v[12]}


def structure_factor_nfft_batch(hkl: torch.Tensor, atoms: DeviceAtoms, ff: torch.Tensor, rots, B: float = 0.0,
                                eps: float = 1e-6, F: torch.Tensor | None = None, I: torch.Tensor | None = None,
                                stream=None, want_F: bool = True):
    """Many orientations of one molecule through the NUFFT
    (spsim_hip_sf_nfft_batch): pattern p at q rotated exactly by rots[p].
    Returns (F [npat, npix, 2] or None, I [npat, npix], info)."""
    R = np.ascontiguousarray(rots, np.float64).reshape(-1, 9)
    npat, npix = R.shape[0], int(hkl.shape[0])
    if want_F and F is None:
        F = torch.empty((npat, npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((npat, npix), dtype=torch.float32, device=hkl.device)
    info = (C.c_int * 13)()
    rc = _lib.load().spsim_hip_sf_nfft_batch(_ptr(hkl), npix, _ptr(atoms.atoms), _ptr(atoms.b), _ptr(atoms.segs),
                                             atoms.nsegs, atoms.natoms, _ptr(ff), float(B), npat,
                                             R.ctypes.data_as(C.c_void_p),
                                             float(eps), _ptr(F if want_F else None), _ptr(I),
                                             C.cast(info, C.c_void_p), _stream(stream))
    check(rc, "spsim_hip_sf_nfft_batch")
    v = list(info)
    return (F if want_F else None), I, {"W": v[0], "M": tuple(v[1:4]), "nf": tuple(v[4:7]), "L": tuple(v[7:10]), "W2": v[10], "staged": v[11], "Wg": v[12]}


class DeviceMolecules:
    """Several molecules packed one after another in HBM for
    spsim_hip_sf_nfft_molecules: atoms [n, 4] (and B factors), segments whose
    starts count from each molecule's first atom, and the host offset arrays."""

    def __init__(self, packed_list, device=None):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.nmol = len(packed_list)
        self.atom_offsets = np.zeros(self.nmol + 1, np.int32)
        self.seg_offsets = np.zeros(self.nmol + 1, np.int32)
        for m, p in enumerate(packed_list):
            self.atom_offsets[m + 1] = self.atom_offsets[m] + p.natoms
            self.seg_offsets[m + 1] = self.seg_offsets[m] + p.segs.shape[0]
        atoms = [p.atoms.reshape(-1, 4)[:p.natoms] for p in packed_list]
        segs = [p.segs for p in packed_list]
        has_b = any(p.b is not None for p in packed_list)
        if has_b and not all(p.b is not None for p in packed_list):
            raise ValueError("either every molecule has per-atom B factors or none has")
        self.atoms = torch.from_numpy(np.ascontiguousarray(np.concatenate(atoms + [np.zeros((1, 4), np.float32)]))).to(dev)
        self.segs = torch.from_numpy(np.ascontiguousarray(np.concatenate(segs + [np.zeros((1, 4), np.int32)]))).to(dev)
        self.b = None
        if has_b:
            self.b = torch.from_numpy(np.ascontiguousarray(np.concatenate([p.b[:p.natoms] for p in packed_list]
                                                                          + [np.zeros(1, np.float32)]))).to(dev)


def structure_factor_nfft_molecules(hkl: torch.Tensor, mols: DeviceMolecules, ff: torch.Tensor, rots=None,
                                    B: float = 0.0, eps: float = 1e-6, F: torch.Tensor | None = None,
                                    I: torch.Tensor | None = None, stream=None, want_F: bool = True):
    """One pattern per molecule through the NUFFT (spsim_hip_sf_nfft_molecules),
    molecule m at q rotated exactly by rots[m] (None: no rotation). Returns
    (F [nmol, npix, 2] or None, I [nmol, npix])."""
    npix = int(hkl.shape[0])
    if want_F and F is None:
        F = torch.empty((mols.nmol, npix, 2), dtype=torch.float32, device=hkl.device)
    if I is None:
        I = torch.empty((mols.nmol, npix), dtype=torch.float32, device=hkl.device)
    R = None if rots is None else np.ascontiguousarray(rots, np.float64).reshape(mols.nmol, 9)
    rc = _lib.load().spsim_hip_sf_nfft_molecules(
        _ptr(hkl), npix, mols.nmol, _ptr(mols.atoms), _ptr(mols.b), mols.atom_offsets.ctypes.data_as(C.c_void_p),
        _ptr(mols.segs), mols.seg_offsets.ctypes.data_as(C.c_void_p), _ptr(ff), float(B),
        None if R is None else R.ctypes.data_as(C.c_void_p), float(eps), _ptr(F if want_F else None), _ptr(I),
        _stream(stream))
    check(rc, "spsim_hip_sf_nfft_molecules")
    return (F if want_F else None), I


def fft_pruned(x: torch.Tensor, M: int, m_lo: int, nf: int, L: int, l_lo: int, stream=None) -> torch.Tensor:
    """Pruned forward FFT along axis 1 of a complex64 device tensor [outer, M, inner]
    (spsim_hip_fft_pruned); returns [outer, L, inner]."""
    if x.dtype != torch.complex64 or x.dim() != 3 or x.shape[1] != M:
        raise ValueError("x must be complex64 [outer, M, inner]")
    x = x.contiguous()
    outer, _, inner = x.shape
    out = torch.empty((outer, L, inner), dtype=torch.complex64, device=x.device)
    check(_lib.load().spsim_hip_fft_pruned(_ptr(x), outer, inner, M, m_lo, nf, L, l_lo, _ptr(out), _stream(stream)),
          "spsim_hip_fft_pruned")
    return out


def hkl_detector(nx, ny, width, height, distance, wavelength, center_x=0.0, center_y=0.0, spherical=0, device=None,
                 stream=None) -> torch.Tensor:
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    out = torch.empty((nx * ny, 3), dtype=torch.float32, device=dev)
    check(_lib.load().spsim_hip_hkl_detector(nx, ny, width, height, distance, center_x, center_y, int(spherical),
                                             wavelength, _ptr(out), _stream(stream)), "spsim_hip_hkl_detector")
    return out


def hkl_detector_3d(nx, ny, nz, width, height, depth, distance, wavelength, center=(0.0, 0.0, 0.0), device=None,
                    stream=None) -> torch.Tensor:
    """Uniform q grid of a 3-D detector volume, [nz*ny*nx, 3]."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    out = torch.empty((nx * ny * nz, 3), dtype=torch.float32, device=dev)
    check(_lib.load().spsim_hip_hkl_detector_3d(nx, ny, nz, width, height, depth, center[0], center[1], center[2],
                                                distance, wavelength, _ptr(out), _stream(stream)),
          "spsim_hip_hkl_detector_3d")
    return out


def rotate_hkl(hkl: torch.Tensor, R, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    R = np.ascontiguousarray(R, np.float64).reshape(9)
    if out is None:
        out = torch.empty_like(hkl)
    check(_lib.load().spsim_hip_rotate_hkl(_ptr(hkl), int(hkl.shape[0]), R.ctypes.data_as(C.c_void_p), _ptr(out),
                                           _stream(stream)), "spsim_hip_rotate_hkl")
    return out


def detector_params(**kw) -> DetectorParams:
    p = DetectorParams()
    p.bin_x = p.bin_y = p.bin_z = 1
    p.nz = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def detector_chain(det: DetectorParams, I: torch.Tensor, npatterns: int = 1, seed: int = 0, pattern_base: int = 0,
                   outputs=("real_output", "noiseless_output"), stream=None) -> dict:
    """Runs the fused detector chain; returns the requested outputs as tensors."""
    nz, bz = max(det.nz, 1), max(det.bin_z, 1)
    npix = det.nx * det.ny * nz
    nb = (det.nx // det.bin_x) * (det.ny // det.bin_y) * (nz // bz)
    names = ["thomson", "solid_angle", "photons", "count", "electrons", "real_output", "noiseless_output"]
    sizes = [npix] * 5 + [nb] * 2
    res = {}
    ptrs = []
    for n, s in zip(names, sizes):
        if n in outputs:
            t = torch.empty((npatterns, s), dtype=torch.float32, device=I.device)
            res[n] = t
            ptrs.append(_ptr(t))
        else:
            ptrs.append(None)
    check(_lib.load().spsim_hip_detector_chain(C.byref(det), _ptr(I), npatterns, seed, pattern_base, *ptrs,
                                               _stream(stream)), "spsim_hip_detector_chain")
    return res


def detector_chain_binned(det: DetectorParams, I: torch.Tensor, npatterns: int = 1, seed: int = 0,
                          pattern_base: int = 0, outputs=("real_output", "noiseless_output"), stream=None) -> dict:
    """Binned outputs of the detector chain with one photon count drawn per bin
    (spsim_hip_detector_chain_binned); returns the requested outputs."""
    nz, bz = max(det.nz, 1), max(det.bin_z, 1)
    nb = (det.nx // det.bin_x) * (det.ny // det.bin_y) * (nz // bz)
    res = {n: torch.empty((npatterns, nb), dtype=torch.float32, device=I.device)
           for n in ("real_output", "noiseless_output") if n in outputs}
    check(_lib.load().spsim_hip_detector_chain_binned(C.byref(det), _ptr(I), npatterns, seed, pattern_base,
                                                      _ptr(res.get("real_output")), _ptr(res.get("noiseless_output")),
                                                      _stream(stream)), "spsim_hip_detector_chain_binned")
    return res
