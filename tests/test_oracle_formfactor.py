"""Pins the oracle's Cromer-Mann form factor (src/diffraction.c:183-193) with
values that do not come from the oracle's own formula."""
import math

import numpy as np
import pytest

import oracle

# f0(sin(theta)/lambda) of neutral atoms, International Tables for
# Crystallography Vol. C, Table 6.1.1.1 (the data the Cromer-Mann fit
# approximates to about 0.04 electrons).
ITC_F0 = {
    6: [(0.1, 5.126), (0.2, 3.581), (0.3, 2.502), (0.4, 1.950), (0.5, 1.685)],
    7: [(0.1, 6.203), (0.2, 4.600), (0.3, 3.241), (0.4, 2.397), (0.5, 1.944)],
    8: [(0.1, 7.250), (0.2, 5.634), (0.3, 4.094), (0.4, 3.010), (0.5, 2.338)],
}


def d_of(s_inv_angstrom: float) -> float:
    """|q| in m^-1 for sin(theta)/lambda = s (|q| = 2 sin(theta)/lambda)."""
    return 2.0 * s_inv_angstrom * 1e10


@pytest.mark.parametrize("Z,expect", [(1, 1), (6, 6), (7, 7), (8, 8), (15, 15), (16, 16), (26, 26), (79, 79)])
def test_forward_scattering_counts_electrons(Z, expect):
    # f0(0) = a1+a2+a3+a4+c = number of electrons of the neutral atom
    assert abs(oracle.form_factor(0.0, Z) - expect) < 0.05


@pytest.mark.parametrize("Z", sorted(ITC_F0))
def test_international_tables_values(Z):
    for s, f_itc in ITC_F0[Z]:
        assert abs(oracle.form_factor(d_of(s), Z) - f_itc) < 0.05, (Z, s)


def test_debye_waller_identity():
    # f(B) = f0 * exp(-B s^2) with s = d/2 in 1/A, for every Z and B
    for Z in (1, 6, 8, 16, 79):
        for s in (0.0, 0.05, 0.21, 0.6):
            for B in (0.0, 7.5, 40.0):
                f0 = oracle.form_factor(d_of(s), Z)
                fb = oracle.form_factor(d_of(s), Z, B)
                assert fb == pytest.approx(f0 * math.exp(-B * s * s), rel=1e-12, abs=1e-14)


def test_unit_species_and_placeholder_row():
    ff = oracle.default_form_factors()
    assert ff.shape == (101, 9)
    # row 100 is the point scatterer used for delta_atoms: f = 1 everywhere
    for s in (0.0, 0.3, 1.0):
        assert oracle.form_factor(d_of(s), 100, 0.0, ff) == pytest.approx(1.0)
    # rows 0 and 99 are placeholders {-1, 0, ..., 0}: f = -exp(-B s^2)
    assert np.array_equal(ff[0], np.array([-1, 0, 0, 0, 0, 0, 0, 0, 0], np.float32))
    assert np.array_equal(ff[99], ff[0])
    assert oracle.form_factor(d_of(0.2), 0, 10.0, ff) == pytest.approx(-math.exp(-10.0 * 0.04))


def test_monotone_decay():
    for Z in (1, 6, 7, 8, 16, 26, 79):
        vals = [oracle.form_factor(d_of(s), Z) for s in np.linspace(0, 1.5, 31)]
        assert all(b <= a + 1e-12 for a, b in zip(vals, vals[1:])), Z
