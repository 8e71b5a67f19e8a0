"""Writes tests/golden/atomsf_table.npy and atomsf_table.json from the
reference's compiled-in Cromer-Mann table, include/atomsf.cdata (the
initialiser of `static float atomsf[100][9]`, src/diffraction.c:58-60).

Each value is a C double literal converted to float when the reference is
compiled; numpy's float64 -> float32 rounding does the same (round to nearest
even). The .npy holds the 100 x 9 float32 rows (a1..a4, b1..b4, c; b in A^2);
the .json records the source, the SHA-256 of the array's bytes and the row
sums f0(0) = a1 + a2 + a3 + a4 + c.

    python tests/golden/make_atomsf_fixture.py [/root/reference]
"""
from __future__ import annotations

import hashlib
import json
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def parse_cdata(path: Path) -> np.ndarray:
    text = path.read_text()
    body = text[text.index("{") + 1:]
    rows = re.findall(r"\{([^{}]*)\}", body)
    vals = np.array([[float(v) for v in r.split(",")] for r in rows], dtype=np.float64)
    if vals.shape != (100, 9):
        raise ValueError(f"{path}: expected 100 x 9 values, found {vals.shape}")
    return vals.astype(np.float32)


def main(ref_root: str = "/root/reference") -> None:
    src = Path(ref_root) / "include" / "atomsf.cdata"
    table = np.ascontiguousarray(parse_cdata(src))
    np.save(HERE / "atomsf_table.npy", table, allow_pickle=False)
    meta = {
        "source": "include/atomsf.cdata (initialiser of atomsf[100][9], src/diffraction.c:58-60)",
        "shape": [100, 9],
        "columns": ["a1", "a2", "a3", "a4", "b1", "b2", "b3", "b4", "c"],
        "dtype": "float32",
        "sha256_float32_bytes": hashlib.sha256(table.tobytes()).hexdigest(),
        "f0_at_zero": [round(float(table[z, :4].sum(dtype=np.float64) + table[z, 8]), 6) for z in range(100)],
        "generator": "tests/golden/make_atomsf_fixture.py",
    }
    (HERE / "atomsf_table.json").write_text(json.dumps(meta, indent=1) + "\n")
    print(f"wrote {HERE / 'atomsf_table.npy'} (sha256 {meta['sha256_float32_bytes'][:16]}...)")


if __name__ == "__main__":
    main(*sys.argv[1:])
