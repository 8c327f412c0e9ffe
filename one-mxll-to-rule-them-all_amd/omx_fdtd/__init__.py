"""omx_fdtd -- Python host layer of the MI355X FDTD engine for the 3D
field-propagation hot path of One-Mxll-to-Rule-Them-All.

The engine itself is the C-ABI library libomx_fdtd_hip.so (HIP kernels for
gfx950); this package loads it with ctypes and mirrors the reference's
solver interface (TMxll, src/FDTD/mxll_base_mod.f90:7-69) in `Grid3D`, and
its 2D / 1D solvers (TMxll_2D, TMxll_1D) in `Grid2D` / `Grid1D`.
"""
from .units import *  # noqa: F401,F403
from ._lib import load_library, OmxError  # noqa: F401
from .grid import Grid3D, Batch  # noqa: F401
from .lowdim import Grid2D, Grid1D  # noqa: F401
from .grid import medium as grid_medium, library_medium  # noqa: F401,E402
from . import cases, host  # noqa: F401,E402
