"""ctypes bindings of libspsim_amd.so (the C-ABI in include/spsim_hip.h and the
drop-in API in include/*.h).

The library is loaded from the package directory; there is no fallback: if it
is missing or no gfx950 device is present, calls fail with RuntimeError.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libspsim_amd.so"

SPSIM_HIP_OK = 0
MAX_SPECTRUM = 65  # SPSIM_HIP_MAX_SPECTRUM
NSPECIES = 101
UNIT_SPECIES = 100
FF_STRIDE = 9
PACK_PLAIN = 0
PACK_VECTOR = 1
PACK_NFFT = 2


class Atom(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Segment(C.Structure):
    _fields_ = [("start", C.c_int), ("count", C.c_int), ("species", C.c_int), ("reserved", C.c_int)]


class Tile(C.Structure):
    _fields_ = [("pixels_per_lane", C.c_int), ("waves", C.c_int), ("chunk_atoms", C.c_int),
                ("atom_splits", C.c_int)]


class DetectorParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("bin_x", C.c_int), ("bin_y", C.c_int),
        ("width", C.c_float), ("height", C.c_float), ("distance", C.c_float),
        ("pixel_width", C.c_float), ("pixel_height", C.c_float),
        ("center_x", C.c_float), ("center_y", C.c_float),
        ("spherical", C.c_int), ("polarization", C.c_int),
        ("beam_intensity", C.c_float), ("quantum_efficiency", C.c_float), ("wavelength", C.c_float),
        ("electron_hole_production_energy", C.c_float), ("dark_current", C.c_float),
        ("exposure_time", C.c_float), ("readout_noise", C.c_float), ("linear_full_well", C.c_float),
        ("maximum_value", C.c_float), ("nz", C.c_int), ("bin_z", C.c_int),
    ]


_vp = C.c_void_p
_i = C.c_int
_f = C.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    "spsim_hip_last_error": (C.c_char_p, []),
    "spsim_hip_device_count": (_i, [C.POINTER(_i)]),
    "spsim_hip_cu_count": (_i, [C.POINTER(_i)]),
    "spsim_hip_synchronize": (_i, [_vp]),
    "spsim_hip_default_form_factors": (_i, [_vp]),
    "spsim_hip_pack_atoms": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, C.POINTER(_i), C.POINTER(_i)]),
    "spsim_hip_hkl_detector": (_i, [_i, _i, _f, _f, _f, _f, _f, _i, _f, _vp, _vp]),
    "spsim_hip_rotate_hkl": (_i, [_vp, _i, _vp, _vp, _vp]),
    "spsim_hip_hkl_detector_3d": (_i, [_i, _i, _i, _f, _f, _f, _f, _f, _f, _f, _f, _vp, _vp]),
    "spsim_hip_sf_pattern": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, _vp, _vp, C.POINTER(Tile), _vp]),
    "spsim_hip_sf_grid": (_i, [_vp, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, _vp, _vp, C.POINTER(Tile), _vp]),
    "spsim_hip_sf_grid_noalloc": (_i, [_vp, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, _vp, _vp, C.POINTER(Tile),
                                       _vp]),
    "spsim_hip_sf_spectrum": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, _i, _vp, _vp, _vp, _vp,
                                   C.POINTER(Tile), _vp]),
    "spsim_hip_spectrum_samples": (_i, [C.c_double, C.c_double, _i, _vp, _vp]),
    "spsim_hip_sf_nfft": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, C.c_double, _vp, _vp, _vp, _vp]),
    "spsim_hip_sf_nfft_batch": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _i, _vp, C.c_double, _vp, _vp,
                                     _vp, _vp]),
    "spsim_hip_sf_nfft_molecules": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _f, _vp, C.c_double, _vp, _vp,
                                         _vp]),
    "spsim_hip_sf_spectrum_nfft": (_i, [_vp, _i, _vp, _vp, _i, _i, _vp, _f, _vp, _i, _vp, _vp, C.c_double, _vp,
                                        _vp, _vp, _vp, _vp]),
    "spsim_hip_fft_pruned": (_i, [_vp, C.c_longlong, C.c_longlong, _i, _i, _i, _i, _i, _vp, _vp]),
    "spsim_hip_release": (_i, []),
    "spsim_hip_detector_chain": (_i, [C.POINTER(DetectorParams), _vp, _i, C.c_uint64, C.c_uint64,
                                      _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "spsim_hip_detector_chain_binned": (_i, [C.POINTER(DetectorParams), _vp, _i, C.c_uint64, C.c_uint64, _vp, _vp,
                                             _vp]),
    "spsim_hip_sf_adjoint": (_i, [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _f, _vp, _vp, _vp, _vp, _vp]),
    "spsim_hip_poisson_nll": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "spsim_hip_batch_create": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _f,
                                    C.POINTER(Tile), C.POINTER(_vp)]),
    "spsim_hip_batch_uses_grid": (_i, [_vp]),
    "spsim_hip_batch_run": (_i, [_vp, _vp, _vp, _i, _vp]),
    "spsim_hip_batch_run_plain": (_i, [_vp, _vp, _vp, _vp]),
    "spsim_hip_batch_terms": (_i, [_vp, C.POINTER(C.c_double)]),
    "spsim_hip_batch_set_tile": (_i, [_vp, C.POINTER(Tile)]),
    "spsim_hip_batch_get_tile": (_i, [_vp, C.POINTER(Tile), C.POINTER(_i)]),
    "spsim_hip_batch_set_item_atoms": (_i, [_vp, _i]),
    "spsim_hip_batch_describe": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "spsim_hip_batch_autotune": (_i, [_vp, _vp, C.POINTER(Tile), C.POINTER(C.c_double)]),
    "spsim_hip_batch_destroy": (None, [_vp]),
}

_LIB: C.CDLL | None = None


def hip_runtimes_mapped(maps_text: str | None = None) -> list[str]:
    """Distinct HIP runtime files (libamdhip64*) mapped into this process,
    by resolved path, from /proc/self/maps (or the given text in its format)."""
    if maps_text is None:
        try:
            maps_text = Path("/proc/self/maps").read_text()
        except OSError:
            return []
    found = set()
    for line in maps_text.splitlines():
        fields = line.split(None, 5)
        if len(fields) < 6:
            continue
        path = fields[5].strip()
        if path.endswith(" (deleted)"):
            path = path[: -len(" (deleted)")]
        if Path(path).name.startswith("libamdhip64"):
            try:
                found.add(str(Path(path).resolve()))
            except OSError:
                found.add(path)
    return sorted(found)


def check_single_hip_runtime(maps_text: str | None = None) -> None:
    """Raises RuntimeError when two different HIP runtimes are mapped: device
    pointers and streams of one are invalid in the other."""
    libs = hip_runtimes_mapped(maps_text)
    if len(libs) > 1:
        raise RuntimeError(
            "two HIP runtimes are loaded in this process (" + ", ".join(libs) + "); device pointers and streams "
            "would cross between them. Import torch before anything loads libamdhip64 (in Python, import torch "
            "or spsim_amd first), or load only one runtime.")


def load() -> C.CDLL:
    """Loads libspsim_amd.so and declares the C-ABI signatures."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m spsim_amd.build`")
        # One HIP runtime per process. The library needs libamdhip64.so.7;
        # PyTorch ships its own copy (torch/lib/libamdhip64.so, the same
        # SONAME) and loads it by path, so if the system runtime were mapped
        # first (this library or another HIP program loaded before torch),
        # importing torch would map a second HIP and HSA runtime, and device
        # pointers and streams would cross between them. Importing torch first
        # makes this library bind to torch's runtime; the maps are checked
        # before and after, so the mixed state is an error, not a crash.
        check_single_hip_runtime()
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        check_single_hip_runtime()
        L = C.CDLL(str(LIB_PATH))
        check_single_hip_runtime()
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def last_error() -> str:
    return (load().spsim_hip_last_error() or b"").decode()


def check(rc: int, what: str) -> None:
    if rc != SPSIM_HIP_OK:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")
