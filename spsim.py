"""spsim under the reference's own module name (src/spsim.py), so that
`import spsim` gives the reference's names: every name of spsim_pybackend plus
the helpers get_molecule_from_opts, get_molecule_from_atoms,
get_atoms_from_molecule and fetch_pdb (src/spsim.py:19-53), all from
spsim_amd.spsim over libspsim_amd.so."""
from spsim_amd.spsim import *  # noqa: F401,F403
from spsim_amd.spsim import __version__, complex_array, float_array, last_error  # noqa: F401
