import os
import sys
from pathlib import Path

import numpy as np
import pytest
import torch  # noqa: F401  (first, so that every native library binds to torch's HIP and OpenMP runtimes)

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libspsim_amd.so")


@pytest.fixture(scope="session")
def gpu():
    """The current torch device, after checking that the HIP library loads."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test run without a visible GPU")
    from spsim_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def ff_table():
    import oracle

    return oracle.default_form_factors()


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def oracle_threads() -> int:
    return min(16, os.cpu_count() or 1)
