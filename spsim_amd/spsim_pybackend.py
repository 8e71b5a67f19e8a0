"""spsim_pybackend: the reference's Python binding surface over libspsim_amd.so.

The reference builds this module with SWIG (src/spsim_pybackend.i), which
wraps every function and struct of spsim.h, config.h, diffraction.h and
molecule.h. SWIG is not available here, so this module mirrors the same names
with ctypes:

* structs are ctypes Structures with the reference's field order
  (include/config.h:44-181, include/diffraction.h:46-62);
* functions that return a struct pointer return a proxy that reads and writes
  the C struct in place (`opts.detector.nx = 256`), as SWIG proxies do, and
  None for a NULL pointer;
* typemaps as in src/spsim_pybackend.i:26-97: `Complex` <-> Python complex;
  `float[3]` fields -> NumPy views; `sp_c3matrix*` / `sp_i3matrix*` /
  `sp_3matrix*` -> zero-copy NumPy views with the 2-D dims ordered (y, x);
* other `float*` / `int*` fields are ctypes pointers (opaque in SWIG too).

The compute functions run on the MI355X; a NULL result means the device call
failed and the reason was printed to stderr (last_error() returns it).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

# ---------------------------------------------------------------- constants (include/config.h:31-42)
CHEM_FORMULA = 1
PDB = 2
BOX_SPHERICAL = 1
BOX_PARALLEL = 2
POLARIZATION_IGNORE = 0
POLARIZATION_VERTICAL = 1
POLARIZATION_HORIZONTAL = 2
POLARIZATION_UNPOLARIZED = 3
VERSION = "2.0"


# ---------------------------------------------------------------- libspimage value types
class Complex(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


class sp_matrix(C.Structure):
    _fields_ = [("rows", C.c_uint), ("cols", C.c_uint), ("data", C.POINTER(C.c_float))]


SpRotation = sp_matrix


class sp_i3matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", C.POINTER(C.c_int))]


class sp_3matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", C.POINTER(C.c_float))]


class sp_c3matrix(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint), ("data", C.POINTER(Complex))]


class Detector(C.Structure):
    _fields_ = [("image_center", C.c_float * 3), ("pixel_size", C.c_float * 3), ("detector_distance", C.c_float),
                ("wavelength", C.c_float), ("orientation", C.POINTER(sp_matrix))]


class Image(C.Structure):
    _fields_ = [("phased", C.c_int), ("shifted", C.c_int), ("scaled", C.c_int), ("image", C.POINTER(sp_c3matrix)),
                ("mask", C.POINTER(sp_i3matrix)), ("detector", C.POINTER(Detector)), ("num_dimensions", C.c_int)]


# ---------------------------------------------------------------- spsim structs
class Chem_Formula(C.Structure):
    _fields_ = [("n_elements", C.c_int), ("atomic_number", C.POINTER(C.c_int)), ("quantity", C.POINTER(C.c_int))]


class Molecule(C.Structure):
    _fields_ = [("natoms", C.c_int), ("atomic_number", C.POINTER(C.c_int)), ("pos", C.POINTER(C.c_float)),
                ("delta_atoms", C.c_int)]


class Experiment(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("wavelength", "bandwidth", "exposure_time", "beam_intensity",
                                         "beam_center_x", "beam_center_y", "beam_fwhm", "beam_energy",
                                         "photon_energy", "focal_diameter")] + [("polarization", C.c_int)]


class CCD(C.Structure):
    _fields_ = ([(n, C.POINTER(C.c_float)) for n in ("photons_per_pixel", "thomson_correction", "solid_angle",
                                                       "photon_count", "electrons_per_pixel", "real_output",
                                                       "noiseless_output")]
                + [(n, C.c_float) for n in ("distance", "width", "height", "depth", "pixel_width", "pixel_height",
                                            "pixel_depth", "quantum_efficiency", "electron_hole_production_energy",
                                            "readout_noise", "dark_current", "linear_full_well")]
                + [(n, C.c_int) for n in ("binning_x", "binning_y", "binning_z", "nx", "ny", "nz")]
                + [("maximum_value", C.c_float), ("spherical", C.c_int), ("gaussian_blurring", C.c_float),
                   ("real_space_blurring", C.c_float), ("center_x", C.c_float), ("center_y", C.c_float),
                   ("center_z", C.c_float), ("n_patterns", C.c_int)])


class Options(C.Structure):
    _fields_ = ([("input_type", C.c_int), ("chem_formula", C.POINTER(Chem_Formula)),
                 ("experiment", C.POINTER(Experiment)), ("detector", C.POINTER(CCD)),
                 ("pdb_filename", C.c_char * 1024), ("sf_filename", C.c_char * 1024),
                 ("hkl_grid_filename", C.c_char * 1024), ("box_type", C.c_int), ("box_dimension", C.c_float),
                 ("use_fft_for_sf", C.c_int), ("use_nfft_for_sf", C.c_int), ("b_factor", C.c_float),
                 ("origin_to_com", C.c_int), ("euler_orientation", C.c_float * 3),
                 ("random_orientation", C.c_int), ("n_patterns", C.c_int), ("vectorize", C.c_int),
                 ("delta_atoms", C.c_int), ("fast_exit", C.c_int), ("crystal_size", C.c_int * 3),
                 ("crystal_cell", C.c_float * 6), ("use_cuda", C.c_int), ("wavelength_samples", C.c_int),
                 ("random_seed", C.c_int)]
                + [(n, C.c_int) for n in ("output_sf_vtk", "output_scatt_int", "output_ewald_vtk",
                                          "output_intensities", "output_noiseless_photons", "output_photons",
                                          "output_electron_pixel_vtk", "output_noiseless_count", "output_count",
                                          "output_realspace_histogram", "output_scattering_factors",
                                          "output_solid_angles", "output_real_space", "verbosity_level")])


class Diffraction_Pattern(C.Structure):
    _fields_ = [("ints", C.POINTER(C.c_float)), ("F", C.POINTER(Complex)), ("N_pix", C.c_int),
                ("max_s", C.c_float * 3), ("d_s", C.c_float * 3), ("HKL_list", C.POINTER(C.c_float)),
                ("HKL_list_size", C.c_int), ("rot", C.POINTER(sp_matrix))]


_MATRIX_VIEWS = {sp_c3matrix: np.complex64, sp_3matrix: np.float32, sp_i3matrix: np.int32}


class Proxy:
    """A SWIG-style proxy of a C struct owned by the library (no ownership)."""

    __slots__ = ("_ptr",)

    def __init__(self, ptr):
        object.__setattr__(self, "_ptr", ptr)

    @property
    def _as_parameter_(self):
        return self._ptr

    @property
    def this(self) -> int:
        return C.cast(self._ptr, C.c_void_p).value or 0

    def __eq__(self, other):
        return isinstance(other, Proxy) and other.this == self.this

    def __hash__(self):
        return hash(self.this)

    def __repr__(self):
        return f"<{self._ptr._type_.__name__} proxy at 0x{self.this:x}>"

    def __dir__(self):
        return [f for f, _ in self._ptr._type_._fields_]

    def __getattr__(self, name):
        s = self._ptr.contents
        v = getattr(s, name)
        ftype = dict(s._fields_)[name]
        if isinstance(v, C.Array) and not issubclass(ftype._type_, C.c_char):
            return np.ctypeslib.as_array(v)  # float[3], int[3], float[6]: a view
        if isinstance(v, Complex):
            return complex(v.re, v.im)
        if hasattr(ftype, "contents") or (isinstance(ftype, type) and issubclass(ftype, C._Pointer)):
            tgt = ftype._type_
            if not v:
                return None
            if tgt in _MATRIX_VIEWS:
                m = v.contents
                dims = (m.y, m.x) if m.z == 1 else (m.z, m.y, m.x)
                n = int(np.prod(dims))
                if n == 0:
                    return np.zeros(dims, _MATRIX_VIEWS[tgt])
                buf = C.cast(m.data, C.POINTER(C.c_char * (n * np.dtype(_MATRIX_VIEWS[tgt]).itemsize))).contents
                return np.frombuffer(buf, dtype=_MATRIX_VIEWS[tgt]).reshape(dims)
            if issubclass(tgt, C.Structure):
                return wrap(v)
            return v
        return v

    def __setattr__(self, name, value):
        s = self._ptr.contents
        ftype = dict(s._fields_)[name]
        if isinstance(value, Proxy):
            value = value._ptr
        if isinstance(ftype, type) and issubclass(ftype, C.Array):
            if issubclass(ftype._type_, C.c_char):
                setattr(s, name, value.encode() if isinstance(value, str) else value)
                return
            arr = getattr(s, name)
            vals = list(np.ravel(value))
            for i in range(len(arr)):
                arr[i] = vals[i]
            return
        if ftype is Complex and isinstance(value, complex):
            value = Complex(value.real, value.imag)
        setattr(s, name, value)


def wrap(ptr):
    if not ptr:
        return None
    return Proxy(ptr)


def _arg(x):
    if isinstance(x, Proxy):
        return x._ptr
    if isinstance(x, str):
        return x.encode()
    if hasattr(x, "__fspath__"):
        return str(x).encode()
    return x


# ---------------------------------------------------------------- functions
_P = C.POINTER
_SIGS = {
    # config.h
    "set_defaults": (_P(Options), []),
    "read_options_file": (None, [C.c_char_p, _P(Options)]),
    "write_options_file": (None, [C.c_char_p, _P(Options)]),
    # molecule.h
    "alloc_mol": (_P(Molecule), []),
    "add_atom_to_mol": (None, [_P(Molecule), C.c_int, C.c_float, C.c_float, C.c_float]),
    "free_mol": (None, [_P(Molecule)]),
    "make_mol": (_P(Molecule), [_P(Image), _P(Image)]),
    "origin_to_center_of_mass": (None, [_P(Molecule)]),
    "get_Molecule_from_formula": (_P(Molecule), [_P(Chem_Formula), _P(Options)]),
    "get_Molecule_from_pdb": (_P(Molecule), [C.c_char_p]),
    "write_pdb_from_mol": (None, [C.c_char_p, _P(Molecule)]),
    # diffraction.h
    "get_HKL_list_for_detector": (_P(C.c_float), [_P(CCD), _P(Experiment), _P(C.c_int)]),
    "get_HKL_list_for_3d_detector": (_P(C.c_float), [_P(CCD), _P(Experiment), _P(C.c_int)]),
    "compute_pattern_on_list": (_P(Diffraction_Pattern), [_P(Molecule), _P(C.c_float), C.c_int, C.c_float,
                                                           _P(Experiment), _P(Options)]),
    "vector_compute_pattern_on_list": (_P(Diffraction_Pattern), [_P(Molecule), _P(C.c_float), C.c_int, C.c_float,
                                                                  _P(Experiment), _P(Options)]),
    "spsim_spectrum_pattern_on_list": (_P(Diffraction_Pattern), [_P(Molecule), _P(C.c_float), C.c_int, C.c_float,
                                                                  _P(Experiment), _P(Options)]),
    "compute_pattern_on_list_by_nfft": (_P(Diffraction_Pattern), [_P(Molecule), _P(C.c_float), C.c_int, _P(CCD),
                                                                C.c_float, _P(Options)]),
    "calculate_pixel_solid_angle": (None, [_P(CCD)]),
    "calculate_thomson_correction": (None, [_P(CCD), _P(Experiment)]),
    "calculate_photons_per_pixel": (None, [_P(Diffraction_Pattern), _P(Options)]),
    "apply_orientation_to_HKL_list": (_P(sp_matrix), [_P(_P(C.c_float)), _P(C.c_int), _P(Options)]),
    # noise.h / amplification.h
    "generate_poisson_noise": (None, [_P(CCD)]),
    "calculate_electrons_per_pixel": (None, [_P(CCD), _P(Experiment)]),
    "calculate_real_detector_output": (None, [_P(CCD), _P(Experiment)]),
    "calculate_noiseless_detector_output": (None, [_P(CCD), _P(Experiment)]),
    # spsim.h
    "get_molecule": (_P(Molecule), [_P(Options)]),
    "get_HKL_list": (_P(C.c_float), [_P(Options), _P(C.c_int)]),
    "compute_sf": (_P(Diffraction_Pattern), [_P(Molecule), _P(C.c_float), C.c_int, _P(Options)]),
    "make_image": (_P(Image), [_P(C.c_float), _P(sp_matrix), _P(Options)]),
    "make_cimage": (_P(Image), [_P(Complex), _P(sp_matrix), _P(Options)]),
    "array_to_image": (None, [_P(C.c_float), _P(Image)]),
    "iarray_to_image": (None, [_P(C.c_int), _P(Image)]),
    "gaussian_blur_pattern": (None, [_P(Options), _P(Diffraction_Pattern)]),
    "simulate_shot": (_P(Diffraction_Pattern), [_P(Molecule), _P(Options)]),
    "free_stuff": (None, [_P(Diffraction_Pattern), _P(Options)]),
    "free_diffraction_pattern": (None, [_P(Diffraction_Pattern)]),
    "free_output_in_options": (None, [_P(Options)]),
    "output_array": (None, [C.c_char_p, C.c_int, _P(C.c_float), _P(sp_matrix), _P(Options)]),
    "output_carray": (None, [C.c_char_p, C.c_int, _P(Complex), _P(sp_matrix), _P(Options)]),
    "output_files": (None, [_P(Diffraction_Pattern), _P(Options), C.c_int]),
    # libspimage subset
    "sp_image_alloc": (_P(Image), [C.c_int, C.c_int, C.c_int]),
    "sp_image_free": (None, [_P(Image)]),
    "sp_rot_alloc": (_P(sp_matrix), []),
    "sp_rot_euler": (_P(sp_matrix), [C.c_float, C.c_float, C.c_float]),
    "sp_rot_uniform": (_P(sp_matrix), []),
    "sp_srand": (None, [C.c_int]),
    "p_drand48": (C.c_double, []),
}


def _make(name, res, args):
    fn = getattr(_lib.load(), name)
    fn.restype = res
    fn.argtypes = args

    def call(*a):
        out = fn(*[_arg(x) for x in a])
        if isinstance(res, type) and issubclass(res, C._Pointer) and issubclass(res._type_, C.Structure):
            return wrap(out)
        if isinstance(res, type) and issubclass(res, C._Pointer):
            return out if out else None
        return out

    call.__name__ = name
    call.__doc__ = f"{name}: libspsim_amd.so (reference prototype of the same name)"
    return call


for _name, (_res, _args) in _SIGS.items():
    globals()[_name] = _make(_name, _res, _args)


def last_error() -> str:
    return _lib.last_error()
