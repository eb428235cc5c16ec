"""nbodyssey_amd -- MI355X-native N-body force and integration engine.

The compute path is libnbody.so (hand-written HIP for gfx950, C ABI in
include/nbody.h); this package is its Python host side: ctypes binding
(``_lib``), the ``Simulation`` handle and reference-compatible functions
(``engine``), initial-condition generators (``ic``) and file formats
(``formats``).
"""
import os

from . import formats, ic
from ._lib import LIB_PATH, NbError, load

__all__ = ["formats", "ic", "LIB_PATH", "NbError", "load", "GALAXY_SIM"]

# nbodyssey_amd.engine (the Simulation handle) loads libnbody.so on first use.

GALAXY_SIM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "galaxy_sim")

