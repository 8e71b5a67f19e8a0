"""Build recipe for libspsim_amd.so (HIP kernels for gfx950 + the C drop-in API).

The library is built in-tree (spsim_amd/libspsim_amd.so) with hipcc for the
.hip sources and gcc for the host C sources, then linked with hipcc. Object
files go to _build/ at the repository root.

    python -m spsim_amd.build          # build (always recompiles)
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build" / "spsim_amd"
LIB = PKG / "libspsim_amd.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # Packed FP32 (v_pk_*_f32) is disabled for all device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{INCLUDE}", f"-I{CSRC / 'host'}"]


def sources() -> tuple[list[Path], list[Path]]:
    hip = sorted((CSRC / "hip").glob("*.hip"))
    c = sorted((CSRC / "host").glob("*.c"))
    return hip, c


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compiles every source and links LIB. Returns the library path."""
    if not Path(HIPCC).exists() and shutil.which(HIPCC) is None:
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(parents=True, exist_ok=True)
    hip, c = sources()
    jobs_list = []
    objs = []
    for src in hip:
        obj = BUILD / (src.stem + ".hip.o")
        jobs_list.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in c:
        obj = BUILD / (src.stem + ".c.o")
        jobs_list.append([CC, *C_FLAGS, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    n = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=n) as ex:
        for cmd, fut in [(cmd, ex.submit(_run, cmd)) for cmd in jobs_list]:
            fut.result()
            if verbose:
                print("compiled", cmd[-3])
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), "-lm", "-ldl"])
    os.replace(tmp, LIB)
    if verbose:
        print("linked", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
