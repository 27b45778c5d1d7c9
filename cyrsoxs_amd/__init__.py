"""MI355X-native (gfx950, HIP) CyRSoXS Born-approximation scattering pipeline.

The compute path is the HIP library lib/libcyrsoxs.so behind the C ABI in
include/cyrsoxs.h; `capi` is its ctypes binding.  The pybind11 module
`CyRSoXS` (built next to this file) mirrors the reference's Python API.
"""
import os

__all__ = ["capi", "synthetic"]
PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
