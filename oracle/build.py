"""ORACLE build recipe (test infrastructure only): compiles oracle/oracle.c into
oracle/liboracle.so with gcc. FMA contraction is off so FP32 expressions round
exactly as written; OpenMP threads the pixel loops.

The reference C sources are not built here: their hot-path translation unit
needs libspimage (spimage.h), FFTW and GSL headers that this image lacks.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-Wall", "-o", str(tmp), str(HERE / "oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
