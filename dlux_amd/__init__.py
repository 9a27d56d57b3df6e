"""dlux_amd -- an MI355X-native (gfx950) implementation of dLux's diffraction hot path.

Hand-written HIP kernels (``dlux_amd/csrc``) behind a C ABI
(``include/dlux_amd.h``), driven from a Python layer that keeps dLux's
optical-system / layer / propagator API.
"""

from . import runtime
from .runtime import set_precision, get_precision, precision

__version__ = "0.1.0"
