"""compute_sf's structure-factor dispatch decisions that need no GPU
(src/spsim.c:110-144): the real-space FFT option is reported as unsupported
before any device work."""
import ctypes as C

import numpy as np

from spsim_amd import spsim
from spsim_amd import workloads as wl


def test_use_fft_for_sf_is_reported_unsupported(capfd):
    opts = spsim.set_defaults()
    opts.use_fft_for_sf = 1
    Z, pos, _ = wl.random_molecule(10, 1e-9, np.random.default_rng(0))
    mol = spsim.get_molecule_from_atoms(Z, pos)
    hkl = np.zeros(3 * 4, np.float32)
    pat = spsim.compute_sf(mol, hkl.ctypes.data_as(C.POINTER(C.c_float)), 4, opts)
    assert not pat
    assert "not supported" in capfd.readouterr().err
