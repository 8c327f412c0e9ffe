"""Atomic-unit constants and index conventions of the reference.

Values follow src/FDTD/constants_mod.f90:13-47.  Several of the reference's
parameters are written as default-REAL literals, so the compiled reference
holds their single-precision roundings; they are reproduced here with
numpy.float32 so that every derived quantity (dr, dt, coefficients) agrees
bit for bit with the reference build.
"""
import numpy as np

__all__ = [
    "PI0", "C0", "EPS0", "MU0", "NM_TO_AU", "AU_TO_NM", "EV_TO_AU", "AU_TO_EV",
    "FS_TO_AU", "AU_TO_FS", "DEBYE_TO_AU",
    "BC_CLOSE", "BC_PERIODIC", "BC_CPML",
    "MEDIUM_NONE", "MEDIUM_DIELECTRIC", "MEDIUM_DRUDE", "MEDIUM_DL",
    "DET_POINT", "DET_LINE_X", "DET_LINE_Y", "DET_LINE_Z", "DET_PLANE_XY",
    "DET_PLANE_YZ", "DET_PLANE_ZX", "DET_VOLUME",
    "EX", "EY", "EZ", "HX", "HY", "HZ", "FIELD_CODES", "DETECTOR_CODES",
    "grid_dr_dt", "index_to_nm",
]

_f32 = np.float32

PI0 = 3.141592653589793
C0 = 137.0
EPS0 = 1.0 / (4.0 * PI0)
MU0 = 1.0 / (EPS0 * C0 ** 2)
NM_TO_AU = float(_f32(18.897261259077823))          # constants_mod.f90:36
AU_TO_NM = 1.0 / NM_TO_AU                            # constants_mod.f90:37
EV_TO_AU = float(_f32(1.0) / _f32(27.2114))          # constants_mod.f90:39
AU_TO_EV = float(_f32(27.2114))                      # constants_mod.f90:40
FS_TO_AU = 1.0 / 2.4188843265864e-2                  # constants_mod.f90:45
AU_TO_FS = 2.4188843265864e-2                        # constants_mod.f90:46
DEBYE_TO_AU = float(_f32(3.0) / _f32(7.63))          # constants_mod.f90:43

# codes of constants_mod.f90:58-102
BC_CLOSE, BC_PERIODIC, BC_CPML = 1, 2, 3
MEDIUM_NONE, MEDIUM_DIELECTRIC, MEDIUM_DRUDE, MEDIUM_DL = 0, 1, 2, 3
(DET_POINT, DET_LINE_X, DET_LINE_Y, DET_LINE_Z, DET_PLANE_XY, DET_PLANE_YZ,
 DET_PLANE_ZX, DET_VOLUME) = range(1, 9)
EX, EY, EZ, HX, HY, HZ = range(1, 7)
FIELD_CODES = {"Ex": EX, "Ey": EY, "Ez": EZ, "Hx": HX, "Hy": HY, "Hz": HZ}
DETECTOR_CODES = {"point": DET_POINT, "line_x": DET_LINE_X, "line_y": DET_LINE_Y,
                  "line_z": DET_LINE_Z, "plane_xy": DET_PLANE_XY, "plane_yz": DET_PLANE_YZ,
                  "plane_zx": DET_PLANE_ZX, "volume": DET_VOLUME}


def grid_dr_dt(dx_nm, mxll_dt=10.0e10):
    """dr and dt in au as read_input_file derives them (input_mod.f90:121-122)."""
    dr = dx_nm * NM_TO_AU
    dt = min(dr / (2.0 * C0), mxll_dt)
    return dr, dt


def nm_to_index_dr(x_nm, dr, n_grid):
    """Index of a coordinate: int(x/dr) + int(N/2), truncating toward zero
    (classical_medium_mod.f90:663-665, detector_mod.f90:376-378)."""
    x = x_nm * NM_TO_AU
    return int(x / dr) + int(n_grid / 2)


def index_to_nm(idx, dx_nm, n_grid):
    """A coordinate (nm) that nm_to_index_dr maps back to `idx`: the centre of
    the interval of coordinates that truncation sends to idx, taken on the
    side away from the origin (for idx - N/2 = 0 the positive side)."""
    m = idx - int(n_grid / 2)
    return (m + 0.5) * dx_nm if m >= 0 else (m - 0.5) * dx_nm


__all__.append("nm_to_index_dr")
