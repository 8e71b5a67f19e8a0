"""CPU checks of the NFFT-path semantics: the oracle's sf_nfft_f64 (the sums
compute_pattern_on_list_by_nfft approximates, src/diffraction.c:528-675) and
the PACK_NFFT packing mode."""
import ctypes as C

import numpy as np
import pytest

import oracle
from spsim_amd import _lib
from spsim_amd import device as dv
from spsim_amd import workloads as wl
from tests.conftest import rel_l2


def qlist(n=60, seed=0, scale=2e9):
    rng = np.random.default_rng(seed)
    return rng.uniform(-scale, scale, (n, 3)).astype(np.float32)


def brute(q, pos, f):
    """sum_j f_j(q) exp(-2 pi i q.r_j) in FP64, f an [nq, natoms] array."""
    q = q.astype(np.float64)
    pos = pos.astype(np.float32).astype(np.float64)
    return (f * np.exp(-2j * np.pi * (q @ pos.T))).sum(1)


def test_equals_plain_oracle_without_z0():
    rng = np.random.default_rng(1)
    Z, pos, _ = wl.random_molecule(80, 2e-9, rng)
    q = qlist()
    F, I = oracle.sf_nfft_f64(q, Z, pos, B=3.0)
    F2, I2 = oracle.sf_f64(q, Z, pos, B=3.0)
    np.testing.assert_allclose(F, F2, rtol=0, atol=1e-12 * np.abs(F2).max())
    np.testing.assert_allclose(I, I2, rtol=1e-12)


@pytest.mark.parametrize("B", [0.0, 7.5])
def test_z0_atom_scatters_with_row0(B):
    """Row 0 of the table is (-1, 0, ...), so one Z == 0 atom at the origin gives
    F = -exp(-B s^2), s = |q| / 2 in 1/A (the reference's scatt_factor, Z = 0)."""
    q = qlist(40, seed=2)
    F, I = oracle.sf_nfft_f64(q, [0], np.zeros((1, 3)), B=B)
    s2 = (np.linalg.norm(q.astype(np.float64), axis=1) * 1e-10) ** 2 / 4
    np.testing.assert_allclose(F.real, -np.exp(-B * s2), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(F.imag, 0.0, atol=1e-15)


def test_mixed_closed_form():
    """A C atom and a Z == 0 atom off the origin: the sum of two terms with their
    own form factors and phases."""
    q = qlist(50, seed=3)
    pos = np.array([[1e-9, -2e-9, 0.5e-9], [-0.7e-9, 0.3e-9, 1.1e-9]], np.float32)
    B = 2.0
    d = np.linalg.norm(q.astype(np.float64), axis=1)
    fC = np.array([oracle.form_factor(x, 6, B) for x in d])
    f0 = np.array([oracle.form_factor(x, 0, B) for x in d])
    F, _ = oracle.sf_nfft_f64(q, [6, 0], pos, B=B)
    ref = brute(q, pos, np.stack([fC, f0], 1))
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_delta_counts_every_atom():
    rng = np.random.default_rng(4)
    pos = rng.uniform(-2e-9, 2e-9, (30, 3)).astype(np.float32)
    Z = rng.integers(0, 10, 30).astype(np.int32)
    Z[:5] = 0
    q = qlist(40, seed=5)
    F, _ = oracle.sf_nfft_f64(q, Z, pos, delta=True, B=9.0)
    ref = brute(q, pos, np.ones((len(q), 30)))
    np.testing.assert_allclose(F, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_pack_nfft_mode():
    Z = np.array([8, 0, 6, 0, 8], np.int32)
    pos = np.arange(15, dtype=np.float32).reshape(5, 3)
    amp = np.full(5, 0.5, np.float32)
    p = dv.pack_atoms(Z, pos, amp=amp, mode=_lib.PACK_NFFT)
    assert [tuple(s[:3]) for s in p.segs] == [(0, 2, 0), (2, 1, 6), (3, 2, 8)]
    assert np.all(p.atoms[:, 3] == 1.0)  # unit strengths
    np.testing.assert_array_equal(p.perm, [1, 3, 2, 0, 4])
    d = dv.pack_atoms(Z, pos, amp=amp, delta_atoms=True, mode=_lib.PACK_NFFT)
    assert [tuple(s[:3]) for s in d.segs] == [(0, 5, _lib.UNIT_SPECIES)]
    # the plain mode drops the Z == 0 atoms
    assert dv.pack_atoms(Z, pos, mode=_lib.PACK_PLAIN).atoms.shape[0] == 3
