"""spsim_amd: MI355X-native (gfx950) structure-factor hot path of spsim.

* libspsim_amd.so — HIP kernels + the reference-shaped C API (include/*.h) and
  the device C-ABI (include/spsim_hip.h).
* spsim_amd.device — device-resident Python front end over the C-ABI.
* spsim_amd.workloads — seeded synthetic molecules, detectors and orientations.
"""
from ._lib import LIB_PATH, load  # noqa: F401

__all__ = ["LIB_PATH", "load"]
