"""Host-side set-up of sources and detectors, as the reference builds them.

These run once at initialisation on the host and hand index lists and
kernels to the engine:
  source kernel      compute_kernel, 3D      (src/FDTD/source_mod.f90:251-327)
  source indices     determine_indx_and_ranks, non-MPI (source_mod.f90:347-356)
  detector indices   init_detector, 3D       (src/FDTD/detector_mod.f90:373-707)
"""
import math

import numpy as np

from .units import (NM_TO_AU, FS_TO_AU, EV_TO_AU, DET_POINT, DET_LINE_X, DET_LINE_Y,
                    DET_LINE_Z, DET_PLANE_XY, DET_PLANE_YZ, DET_PLANE_ZX, DET_VOLUME)

A_BH = (0.353222222, -0.488, 0.145, -0.010222222)   # constants_mod.f90:28


def source_kernel(n_ker, dr):
    """ker_mat(-n:n,-n:n,-n:n): Blackman-Harris radial window clipped at 0 and
    normalised to unit sum."""
    r_max = (n_ker * 2 + 2) * dr
    r0 = r_max / 2.0
    w = 2 * n_ker + 1
    ker = np.zeros((w, w, w), order="F")
    norm = 0.0
    for k in range(-n_ker, n_ker + 1):
        for j in range(-n_ker, n_ker + 1):
            for i in range(-n_ker, n_ker + 1):
                x, y, z = i * dr, j * dr, k * dr
                r = math.sqrt(x * x + y * y + z * z)
                v = (A_BH[0] + A_BH[1] * math.cos(2.0 * math.pi * (r - r0) / r_max)
                     + A_BH[2] * math.cos(2.0 * math.pi * 2.0 * (r - r0) / r_max)
                     + A_BH[3] * math.cos(2.0 * math.pi * 3.0 * (r - r0) / r_max))
                if v < 0.0:
                    v = 0.0
                ker[i + n_ker, j + n_ker, k + n_ker] = v
                norm = norm + v
    return ker / norm


def source_indices(n_ker, r0_au, dr, grid):
    """(w,w,w,3) grid indices of the kernel points: i + int(r0/dr) + int(N/2)."""
    w = 2 * n_ker + 1
    ind = np.zeros((w, w, w, 3), dtype=np.int32, order="F")
    for k in range(-n_ker, n_ker + 1):
        for j in range(-n_ker, n_ker + 1):
            for i in range(-n_ker, n_ker + 1):
                ind[i + n_ker, j + n_ker, k + n_ker, 0] = i + int(r0_au[0] / dr) + int(grid[0] / 2)
                ind[i + n_ker, j + n_ker, k + n_ker, 1] = j + int(r0_au[1] / dr) + int(grid[1] / 2)
                ind[i + n_ker, j + n_ker, k + n_ker, 2] = k + int(r0_au[2] / dr) + int(grid[2] / 2)
    return ind


def source_from_line(direction, j_amp, freq_ev, t0_fs, tau_fs, x_nm, y_nm, z_nm, radius_nm,
                     t_init_fs, t_final_fs, phase, dr, grid):
    """One line of sources.in converted as read_init_sources does
    (source_mod.f90:194-209); returns the keyword arguments of Grid3D.add_source."""
    r0 = np.array([x_nm, y_nm, z_nm]) * NM_TO_AU
    radius = radius_nm * NM_TO_AU
    n_ker = int(radius / dr)
    return dict(direction=direction, j_amp=j_amp, w0=freq_ev * EV_TO_AU, t0=t0_fs * FS_TO_AU,
                tau=tau_fs * FS_TO_AU, t_init=t_init_fs * FS_TO_AU, t_final=t_final_fs * FS_TO_AU,
                phase=phase, n_ker=n_ker, ind=source_indices(n_ker, r0, dr, grid),
                ker=source_kernel(n_ker, dr))


def detector_indices(det_type, lo_nm, hi_nm, dr, grid):
    """indx_list(nd,3) of init_detector for one 3D detector on a single
    domain: cells of the index box [int(lo/dr)+N/2, int(hi/dr)+N/2] (one
    plane/line for the lower-dimensional types), listed k, j, i (i fastest)."""
    lo = [int(lo_nm[a] * NM_TO_AU / dr) + int(grid[a] / 2) for a in range(3)]
    hi = [int(hi_nm[a] * NM_TO_AU / dr) + int(grid[a] / 2) for a in range(3)]
    if det_type == DET_POINT:
        return np.array([lo], dtype=np.int32)
    rng = [[lo[a]] for a in range(3)]
    span = {DET_LINE_X: (0,), DET_LINE_Y: (1,), DET_LINE_Z: (2,), DET_PLANE_XY: (0, 1),
            DET_PLANE_YZ: (1, 2), DET_PLANE_ZX: (0, 2), DET_VOLUME: (0, 1, 2)}[det_type]
    for a in span:
        rng[a] = [v for v in range(1, grid[a] + 1) if lo[a] <= v <= hi[a]]
    out = [(i, j, k) for k in rng[2] for j in rng[1] for i in rng[0]]
    return np.array(out, dtype=np.int32).reshape(-1, 3)
