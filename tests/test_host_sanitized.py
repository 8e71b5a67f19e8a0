"""The host-side parsers (libconfig-format spsim.conf reader/writer and the PDB
reader/writer) built with AddressSanitizer and UndefinedBehaviorSanitizer (gcc,
host C only) and run on truncated and mutated inputs: every input must be
either read or rejected with an error, with no memory error, undefined
behaviour or leak."""
import os
import random
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "spsim_amd" / "csrc" / "host"
SOURCES = ["config.c", "config_file.c", "pdb.c", "molecule.c", "random.c"]

DRIVER = r"""
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "config.h"
#include "molecule.h"
int main(int argc, char **argv) {
  char err[512];
  for (int i = 2; i < argc; ++i) {
    const char *f = argv[i];
    size_t n = strlen(f);
    if (n > 4 && strcmp(f + n - 4, ".pdb") == 0) {
      Molecule *m = NULL;
      int rc = spsim_read_pdb(f, &m, err, sizeof err);
      printf("pdb %d %d\n", rc, m ? m->natoms : -1);
      if (m) {
        spsim_write_pdb(argv[1], m);
        free_mol(m);
      }
    } else {
      Options *o = set_defaults();
      int rc = spsim_read_options_file(f, o, err, sizeof err);
      printf("conf %d\n", rc);
      if (rc == 0) spsim_write_options_file(argv[1], o);
      free(o->chem_formula);
      free(o->detector);
      free(o->experiment);
      free(o);
    }
  }
  return 0;
}
"""

CONF = """# synthetic configuration
number_of_dimensions = 2;
input_type = "chem_formula";
chem_formula = { C = 10; N = 3; O = 4; S = 1; };
detector_distance = 0.1;
detector_width = 0.0768;
detector_height = 0.0768;
detector_pixel_width = 7.5e-05;
detector_pixel_height = 7.5e-05;
detector_quantum_efficiency = 0.15;
experiment_wavelength = 1.3e-09;
experiment_beam_intensity = 1e+25;
binning_x = 4; binning_y = 4;
use_cuda = 1;
b_factor = 12.5;
euler_orientation = [ 0.1, 0.2, 0.3 ];
group : { nested = ( 1, 2.0, "x" ); };
"""


def pdb_text(n=60, seed=0):
    rng = random.Random(seed)
    lines = ["REMARK 350 BIOMT1   1  1.000000  0.000000  0.000000        0.00000",
             "REMARK 350 BIOMT2   1  0.000000  1.000000  0.000000        0.00000",
             "REMARK 350 BIOMT3   1  0.000000  0.000000  1.000000        0.00000",
             "REMARK 350 BIOMT1   2 -1.000000  0.000000  0.000000       10.00000",
             "REMARK 350 BIOMT2   2  0.000000 -1.000000  0.000000        0.00000",
             "REMARK 350 BIOMT3   2  0.000000  0.000000  1.000000        5.00000"]
    for i in range(n):
        el = rng.choice(["C", "N", "O", "S", "FE", "P"])
        x, y, z = (rng.uniform(-50, 50) for _ in range(3))
        lines.append(f"ATOM  {i + 1:5d}  {el:<3s} ALA A{i % 999 + 1:4d}    {x:8.3f}{y:8.3f}{z:8.3f}  1.00 20.00"
                     f"          {el:>2s}")
    return "\n".join(lines) + "\nEND\n"


def mutate(text, rng, nmut, alphabet):
    b = bytearray(text.encode())
    for _ in range(nmut):
        b[rng.randrange(len(b))] = rng.choice(alphabet)
    return bytes(b)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_parsers_under_asan_and_ubsan(tmp_path):
    driver = tmp_path / "driver.c"
    driver.write_text(DRIVER)
    exe = tmp_path / "driver"
    cmd = ["gcc", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-std=gnu11", f"-I{ROOT / 'include'}", f"-I{HOST}", str(driver),
           *[str(HOST / s) for s in SOURCES], "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower():
        pytest.skip("the sanitizer runtime is not installed")
    assert r.returncode == 0, r.stderr
    rng = random.Random(7)
    pdb = pdb_text()
    cases = [("ok.conf", CONF.encode()), ("ok.pdb", pdb.encode()), ("empty.conf", b""), ("empty.pdb", b""),
             ("unterminated.conf", b'a = "abc'), ("comment.conf", b"/* never ends"),
             ("deep.conf", ("a = " + "{ b = " * 150 + "1" + "; }" * 150 + ";").encode()),
             ("long.conf", ('pdb_filename = "' + "x" * 50000 + '";').encode()),
             ("short.pdb", b"ATOM\nHETATM 12\nATOM      1  C\nREMARK 350 BIOMT1\n")]
    for i in range(40):
        cases.append((f"trunc{i}.conf", CONF.encode()[:rng.randrange(len(CONF))]))
        cases.append((f"mut{i}.conf", mutate(CONF, rng, 6, b'"{}[]();=:,@\\\n\t#/*0123456789.eE+-abc ')))
        cases.append((f"trunc{i}.pdb", pdb.encode()[:rng.randrange(len(pdb))]))
        cases.append((f"mut{i}.pdb", mutate(pdb, rng, 25, b" \n0123456789.-ABCDEFGHIJKLMNOPQRSTUVWXYZ")))
    files = []
    for name, data in cases:
        f = tmp_path / name
        f.write_bytes(data)
        files.append(str(f))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out_file = tmp_path / "written"
    for i in range(0, len(files), 20):  # several processes so a report names a small batch
        batch = files[i:i + 20]
        r = subprocess.run([str(exe), str(out_file), *batch], capture_output=True, text=True, env=env,
                           cwd=tmp_path, timeout=120)
        report = r.stdout + r.stderr
        assert r.returncode == 0 and "runtime error" not in report and "Sanitizer" not in report, \
            (batch, report[-3000:])
        assert report.count("\n") >= len(batch)  # every input produced its line
