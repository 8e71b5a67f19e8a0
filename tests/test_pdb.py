"""PDB input and output (get_Molecule_from_pdb / write_pdb_from_mol,
reference src/molecule.c:65-392) on the CPU through the C library.

What pins it: the PDB fixed-column layout (coordinates in columns 31-54,
element in 77-78, REMARK 350 BIOMT rows), known atomic numbers of element
symbols, closed-form assembly operators (identity, a two-fold axis, a pure
translation) and write -> read round trips. When the reference tree is
present its example molecule (examples/DNA.pdb) is read as well and compared
with an independent column parse.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np
import pytest

from spsim_amd import _lib
from spsim_amd import spsim_pybackend as sp

f32 = np.float32
REF_DNA = Path("/root/reference/examples/DNA.pdb")


@pytest.fixture(autouse=True)
def _sigs():
    L = _lib.load()
    L.spsim_read_pdb.restype = C.c_int
    L.spsim_read_pdb.argtypes = [C.c_char_p, C.POINTER(C.POINTER(sp.Molecule)), C.c_char_p, C.c_int]
    L.spsim_write_pdb.restype = C.c_int
    L.spsim_write_pdb.argtypes = [C.c_char_p, C.POINTER(sp.Molecule)]
    L.spsim_element_from_symbol.restype = C.c_int
    L.spsim_element_from_symbol.argtypes = [C.c_char_p]


def atom_line(serial, name, x, y, z, elem, record="ATOM  "):
    """A PDB ATOM/HETATM record with the standard column layout."""
    s = f"{record}{serial:5d} {name:<4s} ALA A   1    {x:8.3f}{y:8.3f}{z:8.3f}  1.00  0.00          {elem:>2s}"
    assert len(s) == 78
    return s


def biomt_lines(op, R, t):
    return [f"REMARK 350   BIOMT{r + 1}{op:4d}{R[r][0]:10.6f}{R[r][1]:10.6f}{R[r][2]:10.6f}     {t[r]:10.5f}"
            for r in range(3)]


def read_pdb(path):
    mol = C.POINTER(sp.Molecule)()
    err = C.create_string_buffer(512)
    rc = _lib.load().spsim_read_pdb(str(path).encode(), C.byref(mol), err, 512)
    if rc != 0:
        return rc, None, None, err.value.decode()
    m = mol.contents
    Z = np.ctypeslib.as_array(m.atomic_number, shape=(m.natoms,)).copy() if m.natoms else np.zeros(0, np.int32)
    P = np.ctypeslib.as_array(m.pos, shape=(m.natoms * 3,)).reshape(-1, 3).copy() if m.natoms else np.zeros((0, 3))
    sp.free_mol(mol)
    return 0, Z, P.astype(np.float32), ""


def metres(text):
    """The reader's conversion: float(text) -> float32 -> *1e-10 in double -> float32."""
    return f32(float(f32(float(text))) * 1e-10)


def test_atoms_and_elements(tmp_path):
    lines = [
        "HEADER    TEST",
        atom_line(1, "N", 1.0, 2.0, 3.0, "N"),
        atom_line(2, "CA", -11.5, 0.25, 100.125, "C"),
        atom_line(3, "FE", 0.001, -0.002, 9999.999, "FE", record="HETATM"),
        atom_line(4, "CA", 5.0, 6.0, 7.0, "Ca", record="HETATM"),     # lower case: calcium
        atom_line(5, "O", -1.0, -2.0, -3.0, "o"),
        atom_line(6, "X", 1.0, 1.0, 1.0, "Q"),                         # no such element: skipped
        atom_line(7, "S", 2.0, 2.0, 2.0, "S ")[:77],                   # left-justified, trailing blank cut
        atom_line(8, "S", 2.0, 2.0, 2.0, "S")[:70],                    # no element columns: skipped
        atom_line(9, "H", 0.5, 0.5, 0.5, "HX"),                        # not a symbol: first letter, H
        "TER",
        "END",
    ]
    p = tmp_path / "t.pdb"
    p.write_text("\n".join(lines) + "\n")
    rc, Z, P, err = read_pdb(p)
    assert rc == 0, err
    assert Z.tolist() == [7, 6, 26, 20, 8, 16, 1]
    want = [(1.0, 2.0, 3.0), (-11.5, 0.25, 100.125), (0.001, -0.002, 9999.999), (5.0, 6.0, 7.0),
            (-1.0, -2.0, -3.0), (2.0, 2.0, 2.0), (0.5, 0.5, 0.5)]
    for i, w in enumerate(want):
        assert [P[i, d] for d in range(3)] == [metres(f"{v:.3f}") for v in w], i


def test_element_symbols():
    L = _lib.load()
    known = {"H": 1, "HE": 2, "C": 6, "N": 7, "O": 8, "NA": 11, "P": 15, "S": 16, "CL": 17, "K": 19, "CA": 20,
             "FE": 26, "ZN": 29 + 1, "SE": 34, "I": 53, "XE": 54, "W": 74, "PT": 78, "AU": 79, "HG": 80, "PB": 82,
             "U": 92, "PU": 94, "LR": 103}
    for sym, z in known.items():
        for field in (sym.rjust(2), sym.ljust(2), sym.lower().rjust(2), sym.capitalize().ljust(2)):
            assert L.spsim_element_from_symbol(field.encode()) == z, (field, z)
    assert L.spsim_element_from_symbol(b"  ") == 0
    assert L.spsim_element_from_symbol(b"12") == 0
    assert L.spsim_element_from_symbol(b"1C") == 6          # a digit before the letter is dropped
    assert L.spsim_element_from_symbol(b"CX") == 6          # unknown pair: its first letter
    assert L.spsim_element_from_symbol(b"QQ") == 0
    zs = [L.spsim_element_from_symbol(bytes([a, b])) for a in range(65, 91) for b in [32] + list(range(65, 91))]
    assert sorted(set(zs) - {0}) == list(range(1, 106))      # every Z in 1..105 has a symbol


def rotz(deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def apply_ops(P, ops):
    """Assembly expansion in float32, ((r0 x + r1 y) + r2 z) + t, operator 0 in place."""
    P = P.astype(np.float32).copy()
    n = P.shape[0]
    out = np.zeros((n * len(ops), 3), np.float32)
    R0, t0 = ops[0]
    for i in range(n):
        for r in range(3):  # operator 0 reads the row it is updating
            acc = f32(P[i, 0] * R0[r, 0])
            acc = f32(acc + f32(P[i, 1] * R0[r, 1]))
            acc = f32(acc + f32(P[i, 2] * R0[r, 2]))
            P[i, r] = f32(acc + t0[r])
    out[:n] = P
    for e, (R, t) in enumerate(ops[1:], start=1):
        for i in range(n):
            for r in range(3):
                acc = f32(P[i, 0] * R[r, 0])
                acc = f32(acc + f32(P[i, 1] * R[r, 1]))
                acc = f32(acc + f32(P[i, 2] * R[r, 2]))
                out[e * n + i, r] = f32(acc + t[r])
    return out


@pytest.mark.parametrize("first_identity", [True, False])
def test_biomt_assembly(tmp_path, first_identity):
    rng = np.random.default_rng(3)
    xyz = np.round(rng.uniform(-30, 30, (6, 3)), 3)
    elems = ["C", "N", "O", "S", "P", "FE"]
    ops = [(np.eye(3) if first_identity else rotz(90.0), np.zeros(3)),
           (rotz(180.0), np.array([0.0, 0.0, 0.0])),
           (np.eye(3), np.array([10.0, -20.0, 33.5]))]
    text = ["REMARK 350 BIOMOLECULE: 1"]
    for op, (R, t) in enumerate(ops, start=1):
        text += biomt_lines(op, np.round(R, 6), t)
    text += [atom_line(i + 1, e, *xyz[i], e) for i, e in enumerate(elems)]
    p = tmp_path / "asm.pdb"
    p.write_text("\n".join(text) + "\nEND\n")
    rc, Z, P, err = read_pdb(p)
    assert rc == 0, err
    assert Z.tolist() == [6, 7, 8, 16, 15, 26] * 3
    base = np.array([[metres(f"{v:.3f}") for v in row] for row in xyz], np.float32)
    Rf = [(np.array([[f32(float(f"{v:10.6f}")) for v in row] for row in np.round(R, 6)], np.float32),
           np.array([f32(float(f"{v:10.5f}") * 1e-10) for v in t], np.float32)) for R, t in ops]
    want = apply_ops(base, Rf)
    assert np.array_equal(P, want)
    if first_identity:
        # closed forms: copy 1 unchanged, copy 2 a two-fold about z, copy 3 shifted
        n = len(elems)
        np.testing.assert_array_equal(P[:n], base)
        np.testing.assert_allclose(P[n:2 * n], base * np.array([-1, -1, 1], np.float32), atol=1e-16)
        np.testing.assert_allclose(P[2 * n:], base + np.array([10, -20, 33.5], np.float32) * 1e-10, rtol=1e-6)


def test_round_trip_through_writer(tmp_path):
    rng = np.random.default_rng(11)
    mol = sp.alloc_mol()
    Zs = rng.choice([1, 6, 7, 8, 16, 26, 79], 40)
    xyz = rng.uniform(-9e-9, 9e-9, (40, 3)).astype(np.float32)
    for z, (x, y, w) in zip(Zs, xyz):
        sp.add_atom_to_mol(mol, int(z), float(x), float(y), float(w))
    path = tmp_path / "w.pdb"
    sp.write_pdb_from_mol(str(path), mol)
    sp.free_mol(mol)
    lines = path.read_text().splitlines()
    assert lines[-1] == "END" and len(lines) == 41
    assert lines[0].startswith("ATOM      0  ") and len(lines[0]) == 78
    rc, Z, P, err = read_pdb(path)
    assert rc == 0, err
    assert Z.tolist() == Zs.tolist()
    for i in range(40):
        assert [P[i, d] for d in range(3)] == [metres(f"{float(xyz[i, d]) * 1e10:8.3f}") for d in range(3)]


def test_missing_file(tmp_path):
    rc, *_, err = read_pdb(tmp_path / "missing.pdb")
    assert rc == -1 and "missing.pdb" in err


def test_empty_file_and_assembly_without_atoms(tmp_path):
    p = tmp_path / "e.pdb"
    p.write_text("\n".join(biomt_lines(1, np.eye(3), np.zeros(3))) + "\n")
    rc, Z, P, err = read_pdb(p)
    assert rc == 0 and Z.size == 0


def test_get_molecule_from_options(tmp_path):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        Path("in.pdb").write_text("\n".join(atom_line(i + 1, "C", i, 2 * i, 3 * i, "C") for i in range(5)) + "\n")
        o = sp.set_defaults()
        o.input_type = sp.PDB
        o.pdb_filename = "in.pdb"
        m = sp.get_molecule(o)
        assert m.natoms == 5 and m.atomic_number[4] == 6 and m.pos[14] == metres("12.000")
        sp.free_mol(m)
        # a formula molecule is also written to molout.pdb, which reads back
        o = sp.set_defaults()
        o.input_type = sp.CHEM_FORMULA
        Zs = (C.c_int * 2)(6, 8)
        Qs = (C.c_int * 2)(30, 12)
        o.chem_formula.n_elements = 2
        o.chem_formula._ptr.contents.atomic_number = C.cast(Zs, C.POINTER(C.c_int))
        o.chem_formula._ptr.contents.quantity = C.cast(Qs, C.POINTER(C.c_int))
        o.box_dimension = 2e-9
        m = sp.get_molecule(o)
        o.chem_formula._ptr.contents.atomic_number = None
        o.chem_formula._ptr.contents.quantity = None
        assert m.natoms == 42
        sp.free_mol(m)
        rc, Z, P, err = read_pdb(Path("molout.pdb"))
        assert rc == 0 and Z.tolist() == [6] * 30 + [8] * 12
        assert np.all(np.linalg.norm(P, axis=1) <= 1e-9 + 1e-12)
    finally:
        os.chdir(cwd)


@pytest.mark.skipif(not REF_DNA.exists(), reason="reference tree not mounted")
def test_reference_dna_example():
    rc, Z, P, err = read_pdb(REF_DNA)
    assert rc == 0, err
    rows = [l.rstrip("\n") for l in REF_DNA.read_text().splitlines() if l.startswith(("ATOM", "HETATM"))]
    sym = {"H": 1, "C": 6, "N": 7, "O": 8, "P": 15}
    assert Z.tolist() == [sym[l[76:78].strip()] for l in rows]
    want = np.array([[metres(l[a:a + 8]) for a in (30, 38, 46)] for l in rows], np.float32)
    assert np.array_equal(P, want)
