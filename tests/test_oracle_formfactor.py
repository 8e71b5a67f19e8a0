"""Pins the oracle's Cromer-Mann form factor (src/diffraction.c:183-193) and its
coefficient table (include/atomsf.cdata, src/diffraction.c:58-60) with values
that do not come from the oracle's own formula or from the product's copy of
the table, and tests the ./atomsf.lib override (src/diffraction.c:118-177)."""
import ctypes as C
import hashlib
import json
import math
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

import oracle

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
LIB = ROOT / "spsim_amd" / "libspsim_amd.so"
REFERENCE_CDATA = Path("/root/reference/include/atomsf.cdata")

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


def product_default_table() -> np.ndarray:
    """The product's built-in table through spsim_hip_default_form_factors
    (host memory only, no device call)."""
    lib = C.CDLL(str(LIB))
    out = np.zeros(101 * 9, np.float32)
    assert lib.spsim_hip_default_form_factors(out.ctypes.data_as(C.c_void_p)) == 0
    return out.reshape(101, 9)


# ---------------------------------------------------------------- the table itself

@pytest.mark.skipif(not REFERENCE_CDATA.exists(), reason="the reference tree is not present on this machine")
def test_fixture_is_the_reference_table():
    sys.path.insert(0, str(GOLDEN))
    try:
        from make_atomsf_fixture import parse_cdata
    finally:
        sys.path.pop(0)
    ref = parse_cdata(REFERENCE_CDATA)
    fix = oracle.reference_atomsf()
    assert np.array_equal(ref.view(np.uint32), fix.view(np.uint32))


def test_fixture_digest_and_metadata():
    meta = json.loads((GOLDEN / "atomsf_table.json").read_text())
    fix = oracle.reference_atomsf()
    assert hashlib.sha256(np.ascontiguousarray(fix).tobytes()).hexdigest() == meta["sha256_float32_bytes"]
    sums = fix[:, :4].sum(axis=1, dtype=np.float64) + fix[:, 8]
    np.testing.assert_allclose(sums, meta["f0_at_zero"], atol=1e-6)


def test_product_table_is_the_reference_table():
    """Every coefficient the product ships equals the reference's, bit for bit
    (the oracle reads the reference's values from the fixture, not from the
    product), and row 100 is the unit scatterer of delta atoms."""
    assert LIB.exists(), "run `python -m spsim_amd.build` first"
    prod = product_default_table()
    fix = oracle.reference_atomsf()
    bad = np.argwhere(prod[:100].view(np.uint32) != fix.view(np.uint32))
    assert bad.size == 0, [(int(z), int(c), float(prod[z, c]), float(fix[z, c])) for z, c in bad[:10]]
    assert np.array_equal(prod[100], np.array([0, 0, 0, 0, 0, 0, 0, 0, 1], np.float32))
    assert np.array_equal(prod.view(np.uint32), oracle.default_form_factors().view(np.uint32))


@pytest.mark.parametrize("Z", range(1, 99))
def test_forward_scattering_counts_electrons(Z):
    # f0(0) = a1+a2+a3+a4+c = number of electrons of the neutral atom; the fits
    # hold it to 0.06 electrons for every element 1..98 (Tl and Pb are the worst)
    assert abs(oracle.form_factor(0.0, Z) - Z) < 0.06


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


# ---------------------------------------------------------------- ./atomsf.lib override

# Records in the layout the reference reads (src/diffraction.c:134-176): an
# element name line; Z first with c at column 23; a1..a4 and b1..b4 at
# 16-character columns; one more line. A second record for the same Z is skipped.
OVERRIDE_ROWS = {
    16: (7.125, 5.0, 1.5, 1.25, 1.5, 22.0, 0.25, 56.0, 1.125),
    26: (11.5, 7.25, 3.5, 2.25, 4.75, 0.3125, 15.25, 76.5, 1.0),
}
DUPLICATE_S = (1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)


def atomsf_lib_text() -> str:
    def record(name, z, v):
        a = "".join(f"{x:16.6f}" for x in v[:4])
        b = "".join(f"{x:16.6f}" for x in v[4:8])
        zl = f"{z:5d}{float(z):12.6f}".ljust(23) + f"{v[8]:.6f}"
        return f"{name}\n{zl}\n{a}\n{b}\n{'0.0':>10}\n"

    head = "AD  Cromer-Mann coefficients (test file)\nAD  a1..a4, b1..b4, c\n"
    return head + record("S", 16, OVERRIDE_ROWS[16]) + record("Fe", 26, OVERRIDE_ROWS[26]) + \
        record("S", 16, DUPLICATE_S)


def override_table(default: np.ndarray) -> np.ndarray:
    """What the reference's fill_ff_tables leaves in atomsf: every row's a1
    set to -1 (src/diffraction.c:130-133), then the first record of each Z
    in the file copied over its row."""
    t = default.copy()
    t[:100, 0] = -1.0
    for z, v in OVERRIDE_ROWS.items():
        t[z] = np.array(v, np.float32)
    return t


def read_table_in(cwd: Path) -> np.ndarray:
    """spsim_form_factor_table() (the drop-in API's table, host only) in a fresh
    process whose working directory is cwd."""
    code = textwrap.dedent(f"""
        import ctypes as C, sys
        import numpy as np
        lib = C.CDLL({str(LIB)!r})
        lib.spsim_form_factor_table.restype = C.POINTER(C.c_float)
        t = np.ctypeslib.as_array(lib.spsim_form_factor_table(), shape=(101 * 9,)).copy()
        sys.stdout.buffer.write(t.tobytes())
    """)
    r = subprocess.run([sys.executable, "-c", code], cwd=cwd, capture_output=True, timeout=120)
    assert r.returncode == 0, r.stderr.decode()
    return np.frombuffer(r.stdout, np.float32).reshape(101, 9)


def test_atomsf_lib_absent_keeps_the_builtin_table(tmp_path):
    t = read_table_in(tmp_path)
    assert np.array_equal(t.view(np.uint32), oracle.default_form_factors().view(np.uint32))


def test_atomsf_lib_override(tmp_path):
    (tmp_path / "atomsf.lib").write_text(atomsf_lib_text())
    t = read_table_in(tmp_path)
    expect = override_table(oracle.default_form_factors())
    assert np.array_equal(t.view(np.uint32), expect.view(np.uint32)), np.argwhere(t != expect)[:10]
    # rows not in the file keep a2..c but lose a1 (-1), as in the reference
    assert t[6, 0] == -1.0 and np.array_equal(t[6, 1:], oracle.default_form_factors()[6, 1:])
    # the first S record wins over the later duplicate
    assert t[16, 0] == np.float32(7.125)
