"""spsim_pybackend under the reference's own module name (SWIG `%module
spsim_pybackend`, src/spsim_pybackend.i:1), so that the reference's import
`from spsim_pybackend import *` (src/spsim.py:6) works unchanged. Every name
comes from spsim_amd.spsim_pybackend: the structs and functions of spsim.h,
config.h, diffraction.h and molecule.h over libspsim_amd.so."""
from spsim_amd.spsim_pybackend import *  # noqa: F401,F403
from spsim_amd.spsim_pybackend import last_error  # noqa: F401
