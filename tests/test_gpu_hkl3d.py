"""3-D detector q grid (spsim_hip_hkl_detector_3d, get_HKL_list_for_3d_detector,
reference src/diffraction.c:325-369) on the GPU: bit-exact against the oracle,
and the structure factor over the volume through the drop-in API."""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import device as dv
from spsim_amd import spsim
from spsim_amd import workloads as wl
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,center", [((16, 16, 16), (0.0, 0.0, 0.0)), ((7, 5, 3), (1e-5, -3e-5, 2e-5)),
                                          ((1, 1, 9), (0.0, 0.0, 0.0))])
def test_hkl_3d_bit_exact(gpu, shape, center):
    nx, ny, nz = shape
    args = (nx * 4e-5, ny * 4e-5, nz * 4e-5, 0.08, 0.3e-9)
    q = dv.hkl_detector_3d(nx, ny, nz, *args, center=center, device=gpu).cpu().numpy()
    ref = oracle.hkl_3d(nx, ny, nz, *args, *center)
    assert np.array_equal(q.view(np.uint32), ref.view(np.uint32))


def test_volume_through_dropin(gpu):
    opts = spsim.set_defaults()
    d = opts.detector
    d.nx, d.ny, d.nz = 12, 10, 8
    d.pixel_width = d.pixel_height = d.pixel_depth = 5e-5
    d.width, d.height, d.depth = (float(np.float32(n * 5e-5)) for n in (12, 10, 8))
    d.distance = 0.05
    opts.experiment.wavelength = 0.25e-9
    opts.verbosity_level = 0
    n = C.c_int(0)
    ptr = spsim.get_HKL_list(opts, C.byref(n))
    assert n.value == 12 * 10 * 8
    q = np.ctypeslib.as_array(ptr, shape=(n.value * 3,)).reshape(-1, 3).copy()
    ref = oracle.hkl_3d(12, 10, 8, d.width, d.height, d.depth, d.distance, opts.experiment.wavelength)
    assert np.array_equal(q.view(np.uint32), ref.view(np.uint32))
    rng = np.random.default_rng(2)
    Z, pos, _ = wl.random_molecule(400, 3e-9, rng)
    mol = spsim.get_molecule_from_atoms(Z, pos)
    pat = spsim.compute_pattern_on_list(mol, ptr, n.value, 0.0, opts.experiment, opts)
    assert rel_l2(spsim.float_array(pat.ints, n.value), oracle.sf_f64(q, Z, pos)[1]) < 1e-5
    spsim.free_diffraction_pattern(pat)
