"""Synthetic workloads of the benchmark configurations.

Every generator is seeded (NumPy PCG64, `numpy.random.default_rng(seed)`) so
the same molecule, detector and orientations come back on any machine.

* Molecule: positions uniform in a ball of the given radius (rejection from the
  enclosing cube, as src/molecule.c:46-54 does for BOX_SPHERICAL), elements in
  exact C:N:O:S = 62:17:20:1 counts, shuffled.
* Orientations: uniform unit quaternions (Shoemake 1992) turned into row-major
  rotation matrices.
* Detector: flat, centred on the beam, `n x n` pixels.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

CNOS = (6, 7, 8, 16)
CNOS_FRACTIONS = (0.62, 0.17, 0.20, 0.01)


@dataclass(frozen=True)
class Detector:
    n: int  # pixels per side
    pixel: float  # m
    distance: float  # m
    wavelength: float  # m

    @property
    def npix(self) -> int:
        return self.n * self.n

    @property
    def width(self) -> float:
        return np.float32(self.n * self.pixel).item()


def element_counts(natoms: int) -> list[int]:
    counts = [int(round(f * natoms)) for f in CNOS_FRACTIONS[:-1]]
    counts.append(natoms - sum(counts))
    return counts


def random_molecule(natoms: int, radius: float, rng: np.random.Generator, atom_b_range=None):
    """Returns (Z int32 [N], pos float32 [N,3] in metres, B float32 [N] or None)."""
    pos = np.empty((natoms, 3), np.float64)
    filled = 0
    while filled < natoms:
        m = max(64, int((natoms - filled) * 2.0))
        cand = rng.uniform(-radius, radius, size=(m, 3))
        cand = cand[np.einsum("ij,ij->i", cand, cand) <= radius * radius]
        take = min(len(cand), natoms - filled)
        pos[filled:filled + take] = cand[:take]
        filled += take
    Z = np.repeat(np.array(CNOS, np.int32), element_counts(natoms))
    rng.shuffle(Z)
    B = None
    if atom_b_range is not None:
        B = rng.uniform(atom_b_range[0], atom_b_range[1], size=natoms).astype(np.float32)
    return Z, pos.astype(np.float32), B


def quaternion_to_matrix(q: np.ndarray) -> np.ndarray:
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ], np.float64)


def uniform_rotations(n: int, rng: np.random.Generator) -> np.ndarray:
    """[n, 3, 3] rotation matrices from uniform unit quaternions (Shoemake)."""
    u = rng.random((n, 3))
    a, b = np.sqrt(1 - u[:, 0]), np.sqrt(u[:, 0])
    q = np.stack([b * np.cos(2 * np.pi * u[:, 2]), a * np.sin(2 * np.pi * u[:, 1]),
                  a * np.cos(2 * np.pi * u[:, 1]), b * np.sin(2 * np.pi * u[:, 2])], axis=1)
    return np.stack([quaternion_to_matrix(qq) for qq in q])


# The benchmark configurations (unstated parameters are marked "assumed").
CFG1_DET = Detector(256, 75e-6, 0.1, 0.2e-9)
CFG2_DET = Detector(1024, 75e-6, 0.1, 0.15e-9)
CFG3_DET = Detector(512, 75e-6, 0.1, 0.2e-9)  # pixel size and distance assumed
CFG4_DET = Detector(512, 75e-6, 0.1, 0.2e-9)  # pixel size, distance, wavelength assumed
CFG5_DET = Detector(1024, 20e-6, 0.05, 1.3e-9)


def config1(seed: int = 1):
    rng = np.random.default_rng(seed)
    Z, pos, _ = random_molecule(5000, 5e-9, rng)
    return CFG1_DET, Z, pos, None


def config2(seed: int = 2):
    rng = np.random.default_rng(seed)
    Z, pos, B = random_molecule(100_000, 12e-9, rng, atom_b_range=(10.0, 40.0))
    return CFG2_DET, Z, pos, B


def config3_counts(npatterns: int = 512, seed: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(2e3), np.log(2e5), size=npatterns))).astype(np.int64)


def config3(npatterns: int = 512, seed: int = 3):
    """One molecule per pattern, N log-uniform in [2e3, 2e5], radius 5 nm (N/5000)^(1/3)."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(2e3), np.log(2e5), size=npatterns))).astype(np.int64)
    mols = []
    for n in counts:
        radius = 5e-9 * (n / 5000.0) ** (1.0 / 3.0)
        Z, pos, _ = random_molecule(int(n), radius, rng)
        mols.append((Z, pos))
    rots = uniform_rotations(npatterns, rng)
    return CFG3_DET, mols, rots


def molecules(counts, rng: np.random.Generator):
    """One C/N/O/S molecule per count, radius 5 nm (N/5000)^(1/3) (the batched
    configuration's size rule), and a uniform random rotation each."""
    mols = []
    for n in counts:
        Z, pos, _ = random_molecule(int(n), 5e-9 * (n / 5000.0) ** (1.0 / 3.0), rng)
        mols.append((Z, pos))
    return mols, uniform_rotations(len(counts), rng)


def small_slice(npatterns: int = 256, seed: int = 33):
    """The batched configuration's small end: N log-uniform in [2e3, 1e4] on the
    same 512 x 512 detector (where one launch per pattern is short)."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(np.log(2e3), np.log(1e4), size=npatterns))).astype(np.int64)
    mols, rots = molecules(counts, rng)
    return CFG3_DET, mols, rots


def skewed_batch(seed: int = 34):
    """The batched configuration's extremes in one run: one 200,000-atom
    molecule among 63 of 2,000 atoms, 512 x 512 detector."""
    rng = np.random.default_rng(seed)
    mols, rots = molecules([200_000] + [2_000] * 63, rng)
    return CFG3_DET, mols, rots


def config4(seed: int = 4):
    rng = np.random.default_rng(seed)
    Z, pos, _ = random_molecule(20_000, 7e-9, rng)
    return CFG4_DET, Z, pos


def config5(npatterns: int = 256, seed: int = 5):
    rng = np.random.default_rng(seed)
    Z, pos, _ = random_molecule(50_000, 10e-9, rng)
    rots = uniform_rotations(npatterns, rng)
    return CFG5_DET, Z, pos, rots
