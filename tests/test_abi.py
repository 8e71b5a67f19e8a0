"""The C-ABI library loads and exports every function that include/*.h
declares, and the drop-in structs keep the reference's field layout
(include/config.h:44-181, include/diffraction.h:46-62 of the reference)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "spsim_amd" / "libspsim_amd.so"

PROTO = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*(?:\s*\*)*\s+\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        text = re.sub(r"//.*", "", text)
        for m in PROTO.finditer(text):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.add(m.group(1))
    return sorted(names)


def test_library_loads_and_exports_everything_declared():
    assert LIB.exists(), "run `python -m spsim_amd.build` first"
    lib = C.CDLL(str(LIB))
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


# Field lists written from the reference headers (include/config.h,
# include/diffraction.h), as ctypes sees them on this platform's C ABI.
P = C.c_void_p
REF_LAYOUTS = {
    "Molecule": [("natoms", C.c_int), ("atomic_number", P), ("pos", P), ("delta_atoms", C.c_int)],
    "Experiment": [(n, C.c_float) for n in ("wavelength", "bandwidth", "exposure_time", "beam_intensity",
                                             "beam_center_x", "beam_center_y", "beam_fwhm", "beam_energy",
                                             "photon_energy", "focal_diameter")] + [("polarization", C.c_int)],
    "CCD": [(n, P) for n in ("photons_per_pixel", "thomson_correction", "solid_angle", "photon_count",
                             "electrons_per_pixel", "real_output", "noiseless_output")]
    + [(n, C.c_float) for n in ("distance", "width", "height", "depth", "pixel_width", "pixel_height", "pixel_depth",
                                "quantum_efficiency", "electron_hole_production_energy", "readout_noise",
                                "dark_current", "linear_full_well")]
    + [(n, C.c_int) for n in ("binning_x", "binning_y", "binning_z", "nx", "ny", "nz")]
    + [("maximum_value", C.c_float), ("spherical", C.c_int), ("gaussian_blurring", C.c_float),
       ("real_space_blurring", C.c_float), ("center_x", C.c_float), ("center_y", C.c_float),
       ("center_z", C.c_float), ("n_patterns", C.c_int)],
    "Options": [("input_type", C.c_int), ("chem_formula", P), ("experiment", P), ("detector", P),
                ("pdb_filename", C.c_char * 1024), ("sf_filename", C.c_char * 1024),
                ("hkl_grid_filename", C.c_char * 1024), ("box_type", C.c_int), ("box_dimension", C.c_float),
                ("use_fft_for_sf", C.c_int), ("use_nfft_for_sf", C.c_int), ("b_factor", C.c_float),
                ("origin_to_com", C.c_int), ("euler_orientation", C.c_float * 3), ("random_orientation", C.c_int),
                ("n_patterns", C.c_int), ("vectorize", C.c_int), ("delta_atoms", C.c_int), ("fast_exit", C.c_int),
                ("crystal_size", C.c_int * 3), ("crystal_cell", C.c_float * 6), ("use_cuda", C.c_int),
                ("wavelength_samples", C.c_int), ("random_seed", C.c_int)]
    + [(n, C.c_int) for n in ("output_sf_vtk", "output_scatt_int", "output_ewald_vtk", "output_intensities",
                              "output_noiseless_photons", "output_photons", "output_electron_pixel_vtk",
                              "output_noiseless_count", "output_count", "output_realspace_histogram",
                              "output_scattering_factors", "output_solid_angles", "output_real_space",
                              "verbosity_level")],
    "Diffraction_Pattern": [("ints", P), ("F", P), ("N_pix", C.c_int), ("max_s", C.c_float * 3),
                            ("d_s", C.c_float * 3), ("HKL_list", P), ("HKL_list_size", C.c_int), ("rot", P)],
}


def c_layout(struct: str, fields):
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "spsim.h"', "int main(void){",
           f'printf("%zu\\n", sizeof({struct}));']
    src += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f, _ in fields]
    src += ["return 0;}"]
    return "\n".join(src)


@pytest.mark.parametrize("name", sorted(REF_LAYOUTS))
def test_struct_layout_matches_reference(tmp_path, name):
    fields = REF_LAYOUTS[name]
    ref = type(name, (C.Structure,), {"_fields_": fields})
    exe = tmp_path / "layout"
    (tmp_path / "l.c").write_text(c_layout(name, fields))
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(tmp_path / "l.c"), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got[0] == C.sizeof(ref)
    assert got[1:] == [getattr(ref, f).offset for f, _ in fields]


def test_complex_and_rotation_layout():
    # Complex {re, im} of float (src/spsim_pybackend.i:26-33); SpRotation {rows, cols, data}
    lib = C.CDLL(str(LIB))
    lib.sp_rot_euler.restype = C.c_void_p
    lib.sp_rot_euler.argtypes = [C.c_float] * 3

    class Rot(C.Structure):
        _fields_ = [("rows", C.c_uint), ("cols", C.c_uint), ("data", C.POINTER(C.c_float))]

    p = lib.sp_rot_euler(0.0, 0.0, 0.0)
    r = Rot.from_address(p)
    assert (r.rows, r.cols) == (3, 3)
    assert [r.data[i] for i in range(9)] == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    C.CDLL(None).free(C.c_void_p(p))  # one block, released by a single free()
