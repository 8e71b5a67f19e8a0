"""Pattern output files (output_array / output_carray / output_files,
reference src/spsim.c:146-279) on the CPU. The CXI files are read back with HDF5's own h5dump (raw binary dumps
of each dataset), so every value written is checked bit for bit."""
import ctypes as C
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from spsim_amd import spsim_pybackend as sp

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if Path("/opt/conda/bin/h5dump").exists() else None)
HDF5_LIB = "/opt/conda/lib/libhdf5.so"
needs_hdf5 = pytest.mark.skipif(H5DUMP is None or not (Path(HDF5_LIB).exists() or shutil.which("h5cc")),
                                reason="HDF5 tools not installed")


@pytest.fixture(autouse=True)
def _hdf5_env(monkeypatch, tmp_path):
    if Path(HDF5_LIB).exists():
        monkeypatch.setenv("SPSIM_HDF5_LIBRARY", HDF5_LIB)
    monkeypatch.chdir(tmp_path)


def h5read(path, dset, dtype, shape=None):
    out = Path(str(path) + dset.replace("/", "_") + ".bin")
    r = subprocess.run([H5DUMP, "-b", "LE", "-d", dset, "-o", str(out), str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    a = np.fromfile(out, dtype=dtype)
    return a.reshape(shape) if shape else a


def h5text(path, dset):
    """All numbers of a dataset from h5dump's text output (9 significant digits
    restore every float32)."""
    import re

    r = subprocess.run([H5DUMP, "-m", "%.9g", "-y", "-w", "0", "-d", dset, str(path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    body = r.stdout[r.stdout.index("DATA {") + 6:]
    return np.array(re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", body), np.float64)


def h5has(path, name):
    r = subprocess.run([H5DUMP, "-n", str(path)], capture_output=True, text=True)
    return any(l.split()[-1] == name for l in r.stdout.splitlines() if l.strip())


def make_opts(nx=6, ny=4, nz=1, bx=1, by=1, bz=1):
    o = sp.set_defaults()
    d = o.detector
    d.nx, d.ny, d.nz = nx, ny, nz
    d.binning_x, d.binning_y, d.binning_z = bx, by, bz
    d.pixel_width, d.pixel_height, d.pixel_depth = 75e-6, 50e-6, 20e-6
    d.distance = 0.1
    o.experiment.wavelength = 0.2e-9
    return o


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@needs_hdf5
def test_output_array_layout():
    o = make_opts()
    a = np.arange(24, dtype=np.float32) * np.float32(0.5) - 3
    sp.output_array("img", 7, fptr(a), None, o)
    f = Path("img-00007.cxi")
    assert f.exists()
    np.testing.assert_array_equal(h5read(f, "/entry_1/image_1/data", "<f4", (4, 6)), a.reshape(4, 6))
    assert np.all(h5read(f, "/entry_1/image_1/mask", "<i4") == 1)
    assert h5read(f, "/cxi_version", "<i4")[0] == 150
    det = "/entry_1/instrument_1/detector_1/"
    assert h5read(f, det + "distance", "<f8")[0] == np.float32(0.1)
    assert h5read(f, det + "x_pixel_size", "<f8")[0] == np.float32(75e-6)
    assert h5read(f, det + "y_pixel_size", "<f8")[0] == np.float32(50e-6)
    lam = h5read(f, "/entry_1/instrument_1/source_1/wavelength", "<f8")[0]
    assert lam == np.float32(0.2e-9)
    np.testing.assert_allclose(h5read(f, "/entry_1/instrument_1/source_1/energy", "<f8")[0],
                               1.98644521e-25 / lam, rtol=1e-15)
    # the image group links to the detector and source
    np.testing.assert_array_equal(h5read(f, "/entry_1/image_1/detector_1/distance", "<f8"),
                                  h5read(f, det + "distance", "<f8"))
    assert not h5has(f, "/entry_1/image_1/orientation")


@needs_hdf5
def test_output_carray_and_orientation():
    o = make_opts(4, 3)
    rng = np.random.default_rng(1)
    F = (rng.normal(size=12) + 1j * rng.normal(size=12)).astype(np.complex64)
    rot = sp.sp_rot_euler(0.3, 1.1, -0.7)
    R = np.array([rot.data[i] for i in range(9)], np.float64)
    sp.output_carray("amp", 0, F.ctypes.data_as(C.POINTER(sp.Complex)), rot, o)
    f = Path("amp-00000.cxi")
    got = h5text(f, "/entry_1/image_1/data").astype(np.float32).view(np.complex64)
    np.testing.assert_array_equal(got, F)
    np.testing.assert_array_equal(h5read(f, "/entry_1/image_1/orientation", "<f8"), R)


@needs_hdf5
def test_output_files_flags_and_binned_dims():
    o = make_opts(8, 6, 1, 2, 3, 1)
    d = o.detector
    n, nb = 8 * 6, (8 // 2) * (6 // 3)
    arrays = {k: (np.arange(n if k not in ("real_output", "noiseless_output") else nb, dtype=np.float32) + i)
              for i, k in enumerate(["photons_per_pixel", "solid_angle", "photon_count", "real_output",
                                     "noiseless_output"])}
    for k, a in arrays.items():
        setattr(d, k, fptr(a))
    ints = np.linspace(0, 1, n, dtype=np.float32)
    Fv = (ints + 1j * ints[::-1]).astype(np.complex64)
    pat = sp.Diffraction_Pattern()
    pat.ints = fptr(ints)
    pat.F = Fv.ctypes.data_as(C.POINTER(sp.Complex))
    pat.rot = None
    o.output_noiseless_photons = o.output_photons = o.output_count = o.output_noiseless_count = 1
    o.output_solid_angles = o.output_scattering_factors = o.output_intensities = o.output_sf_vtk = 1
    sp.output_files(C.pointer(pat), o, 3)
    names = {"noiseless_photon_out": ("photons_per_pixel", (6, 8)), "photon_out": ("photon_count", (6, 8)),
             "solid_angles": ("solid_angle", (6, 8)), "counts_out": ("real_output", (2, 4)),
             "noiseless_counts_out": ("noiseless_output", (2, 4))}
    for base, (k, shape) in names.items():
        f = Path(f"{base}-00003.cxi")
        np.testing.assert_array_equal(h5read(f, "/entry_1/image_1/data", "<f4", shape), arrays[k].reshape(shape))
        # pixel size times binning, as make_image records it
        assert h5read(f, "/entry_1/instrument_1/detector_1/x_pixel_size", "<f8")[0] == np.float32(75e-6) * 2
        assert h5read(f, "/entry_1/instrument_1/detector_1/y_pixel_size", "<f8")[0] == np.float32(50e-6) * 3
    np.testing.assert_array_equal(h5read(Path("scattering_int-00003.cxi"), "/entry_1/image_1/data", "<f4"), ints)
    np.testing.assert_array_equal(
        h5text(Path("scattering_factors-00003.cxi"), "/entry_1/image_1/data").astype(np.float32).view(np.complex64),
        Fv)
    assert not Path("scattering_factor-00003.vtk").exists()  # VTK output is not built
    for k in arrays:  # the arrays belong to this test, not to the options
        setattr(d, k, None)


def test_missing_hdf5_is_reported(monkeypatch, capfd):
    """Without an HDF5 library the writers say so and write nothing (checked in
    a child process, since the loader state is per process)."""
    code = ("import numpy as np, ctypes as C\n"
            "from spsim_amd import spsim_pybackend as sp\n"
            "o = sp.set_defaults(); o.detector.nx = 2; o.detector.ny = 2; o.detector.nz = 1\n"
            "a = np.zeros(4, np.float32)\n"
            "sp.output_array('x', 1, a.ctypes.data_as(C.POINTER(C.c_float)), None, o)\n")
    env = dict(os.environ, SPSIM_HDF5_LIBRARY="/nonexistent/libhdf5.so", LD_LIBRARY_PATH="/nonexistent")
    env["PYTHONPATH"] = str(Path(__file__).resolve().parents[1])
    r = subprocess.run(["python", "-c", code], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr
    assert "HDF5" in r.stderr and "not found" in r.stderr
    assert not Path("x-00001.cxi").exists()
