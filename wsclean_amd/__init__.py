"""MI355X-native w-gridder / degridder behind WSClean's w-gridder boundary.

The product is libwgrid_hip.so (HIP kernels for gfx950 + a C-ABI,
include/wgrid_hip.h); this package holds its build script and a Python
mirror of the ``WGriddingGridder_Simple`` interface.
"""
from ._lib import WgError, load  # noqa: F401
from .gridder import WGriddingGridder_Simple, fft_lines  # noqa: F401

__all__ = ["WGriddingGridder_Simple", "WgError", "fft_lines", "load"]
