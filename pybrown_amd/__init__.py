"""pybrown_amd: pyBrown's hydrodynamic step (RPY / Smith-Ewald mobility,
Jeffrey-Onishi lubrication, Cholesky noise, Ermak-McCammon and midpoint
propagation with flux and wrap) on one AMD Instinct MI355X (gfx950).

The compute path is libpybrown_amd.so (hand-written HIP, C ABI in
include/pybrown_amd.h); this package is its ctypes host side with the
reference's Python interfaces (Bead, Box, diffusion.*).
"""
__version__ = "0.1.0"


def library_path():
    from pybrown_amd._lib import LIB_PATH
    return LIB_PATH
