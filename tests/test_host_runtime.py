"""Host-side checks of the Python boundary (no GPU needed): one HIP runtime per
process, and the reference's own module names `spsim_pybackend` (SWIG
%module, src/spsim_pybackend.i:1) and `spsim` (src/spsim.py)."""
import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

from spsim_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
SYSTEM_HIP = Path("/opt/rocm/lib/libamdhip64.so.7")

MAPS_ONE = """\
7f0000000000-7f0000001000 r--p 00000000 08:01 11 /usr/lib/x86_64-linux-gnu/libc.so.6
7f0000001000-7f0000002000 r-xp 00001000 08:01 12 /opt/torch/lib/libamdhip64.so
7f0000002000-7f0000003000 r--p 00002000 08:01 12 /opt/torch/lib/libamdhip64.so
7f0000003000-7f0000004000 rw-p 00000000 00:00 0
"""
MAPS_TWO = MAPS_ONE + "7f0000005000-7f0000006000 r-xp 00000000 08:01 13 /opt/rocm-7.2.0/lib/libamdhip64.so.7.2.70200\n"


def test_runtime_parser():
    assert _lib.hip_runtimes_mapped(MAPS_ONE) == ["/opt/torch/lib/libamdhip64.so"]
    _lib.check_single_hip_runtime(MAPS_ONE)
    assert len(_lib.hip_runtimes_mapped(MAPS_TWO)) == 2
    with pytest.raises(RuntimeError, match="two HIP runtimes"):
        _lib.check_single_hip_runtime(MAPS_TWO)
    _lib.check_single_hip_runtime("")


def run_child(code: str):
    return subprocess.run([sys.executable, "-c", textwrap.dedent(code)], cwd=ROOT, capture_output=True, text=True,
                          timeout=300)


def test_load_binds_to_a_single_runtime():
    r = run_child("""
        from spsim_amd import _lib
        _lib.load()
        libs = _lib.hip_runtimes_mapped()
        assert len(libs) == 1, libs
        print("ok", libs[0])
    """)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


@pytest.mark.skipif(not SYSTEM_HIP.exists(), reason="no system HIP runtime in /opt/rocm")
def test_load_refuses_a_second_runtime():
    # The system runtime is mapped first (as a C program that loaded
    # libamdhip64.so.7 would have it); importing torch then maps torch's own
    # copy, and load() must refuse instead of mixing the two.
    r = run_child(f"""
        import ctypes
        ctypes.CDLL({str(SYSTEM_HIP)!r})
        from spsim_amd import _lib
        try:
            _lib.load()
        except RuntimeError as e:
            print("refused:", e)
        else:
            print("loaded", _lib.hip_runtimes_mapped())
    """)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "refused: two HIP runtimes" in r.stdout, r.stdout


def test_reference_module_names():
    r = run_child("""
        import spsim, spsim_pybackend
        from spsim_pybackend import *   # the reference's own import, src/spsim.py:6
        from spsim_amd import spsim_pybackend as impl
        for name in impl._SIGS:
            assert getattr(spsim_pybackend, name) is getattr(impl, name), name
            assert getattr(spsim, name) is getattr(impl, name), name
        for name in ("Options", "Molecule", "Experiment", "CCD", "Diffraction_Pattern", "Image", "Complex",
                     "get_molecule_from_opts", "get_molecule_from_atoms", "get_atoms_from_molecule", "fetch_pdb"):
            assert hasattr(spsim, name), name
        assert compute_pattern_on_list is impl.compute_pattern_on_list
        mol = spsim.get_molecule_from_atoms([6, 16], [[0, 0, 0], [1e-10, 2e-10, 3e-10]])
        Z, pos = spsim.get_atoms_from_molecule(mol)
        assert list(Z) == [6, 16] and abs(pos[1, 2] - 3e-10) < 1e-16
        print("ok")
    """)
    assert r.returncode == 0, r.stdout + r.stderr
