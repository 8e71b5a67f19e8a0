"""spsim: the reference's Python convenience module (src/spsim.py) over the
MI355X library. `from spsim_amd.spsim import *` gives every name of
spsim_pybackend plus the helpers below."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .spsim_pybackend import *  # noqa: F401,F403
from . import spsim_pybackend as _b

__version__ = "0.1.0"


def get_molecule_from_opts(opts):
    """src/spsim.py:19-21"""
    return _b.get_molecule(opts)


def get_molecule_from_atoms(atomic_numbers=None, atomic_positions=None):
    """src/spsim.py:23-27: a Molecule built atom by atom with add_atom_to_mol."""
    mol = _b.alloc_mol()
    for j, (x, y, z) in zip(atomic_numbers, atomic_positions):
        _b.add_atom_to_mol(mol, int(j), float(x), float(y), float(z))
    return mol


def get_atoms_from_molecule(mol):
    """src/spsim.py:29-41: (atomic numbers int32 [N], positions float64 [N, 3]).
    Reads the Molecule's arrays directly instead of round-tripping through
    an Image; the values are the same."""
    n = mol.natoms
    if n == 0:
        return np.zeros(0, np.int32), np.zeros((0, 3), np.float64)
    pos = np.ctypeslib.as_array(mol.pos, shape=(n * 3,)).astype(np.float64).reshape(n, 3)
    anum = np.ctypeslib.as_array(mol.atomic_number, shape=(n,)).astype(np.int32).copy()
    return anum, pos


def fetch_pdb(pdb_id):
    """src/spsim.py:43-53 downloads from rcsb.org; this build has no network
    access, so the call is refused."""
    raise NotImplementedError("fetch_pdb needs network access, which this build does not have")


def float_array(ptr, n: int) -> np.ndarray:
    """Copy of n floats behind a float* field (e.g. pattern.ints)."""
    if not ptr or n <= 0:
        return np.zeros(0, np.float32)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


def complex_array(ptr, n: int) -> np.ndarray:
    """Copy of n Complex values behind a Complex* field (e.g. pattern.F)."""
    if not ptr or n <= 0:
        return np.zeros(0, np.complex64)
    raw = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_float)), shape=(2 * n,))
    return (raw[0::2] + 1j * raw[1::2]).astype(np.complex64)
