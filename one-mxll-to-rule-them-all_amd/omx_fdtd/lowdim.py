"""Grid2D and Grid1D: the reference's 2D and 1D solver interfaces over the
HIP engine (FP64).

Method names and argument meanings follow TMxll_2D (src/FDTD/mxll_2D_mod.f90:9-50)
and TMxll_1D (src/FDTD/mxll_1D_mod.f90:9-46): ``init`` arguments in the
constructor, ``kill``, ``td_propagate_H_field``, ``td_propagate_E_field(t)``;
sources, detectors and field access are device calls.  Field arrays keep the
reference's bounds: 2D Ex(0:nx,1:ny+1), Ey(1:nx+1,0:ny), Ez(1:nx+1,1:ny+1),
Hx(1:nx,0:ny), Hy(0:nx,1:ny), Hz(0:nx,0:ny); 1D Ex(1:nz), Hy(1:nz-1).
Errors raise OmxError.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from .units import EPS0, MU0, NM_TO_AU, FS_TO_AU, EV_TO_AU, FIELD_CODES

TMZ_2D_MODE, TEZ_2D_MODE, FULL_2D_MODE = 1, 2, 3
A_BH = (0.353222222, -0.488, 0.145, -0.010222222)   # constants_mod.f90:28


def field_shape_2d(field, nx, ny):
    """Shape of a 2D field array (mxll_2D_mod.f90:110-141)."""
    return {1: (nx + 1, ny + 1), 2: (nx + 1, ny + 1), 3: (nx + 1, ny + 1),
            4: (nx, ny + 1), 5: (nx + 1, ny), 6: (nx + 1, ny + 1)}[field]


def field_shape_1d(field, nz):
    return (nz,) if field == 1 else (nz - 1,)


def source_kernel(dims, n_ker, dr):
    """compute_kernel for 1D and 2D (source_mod.f90:251-300): the
    Blackman-Harris window over x (1D, not clipped) or over r (2D, clipped at
    0), normalised to unit sum; Fortran order."""
    r_max = (n_ker * 2 + 2) * dr
    r0 = r_max / 2.0
    w = 2 * n_ker + 1

    def bh(x):
        return (A_BH[0] + A_BH[1] * math.cos(2.0 * math.pi * (x - r0) / r_max)
                + A_BH[2] * math.cos(2.0 * math.pi * 2.0 * (x - r0) / r_max)
                + A_BH[3] * math.cos(2.0 * math.pi * 3.0 * (x - r0) / r_max))

    norm = 0.0
    if dims == 1:
        ker = np.zeros(w)
        for i in range(-n_ker, n_ker + 1):
            ker[i + n_ker] = bh(i * dr)
            norm = norm + ker[i + n_ker]
        return ker / norm
    ker = np.zeros((w, w), order="F")
    for j in range(-n_ker, n_ker + 1):
        for i in range(-n_ker, n_ker + 1):
            x, y = i * dr, j * dr
            v = bh(math.sqrt(x * x + y * y))
            v = max(v, 0.0)
            ker[i + n_ker, j + n_ker] = v
            norm = norm + v
    return ker / norm


def source_indices(dims, n_ker, r0_au, dr, grid):
    """determine_indx_and_ranks, non-MPI (source_mod.f90:336-346): kernel
    point i sits at i + int(r0/dr) + int(N/2); 1D uses the x coordinate."""
    w = 2 * n_ker + 1
    if dims == 1:
        return np.array([i + int(r0_au[0] / dr) + int(grid[0] / 2) for i in range(-n_ker, n_ker + 1)],
                        dtype=np.int32)
    ind = np.zeros((w, w, 2), dtype=np.int32, order="F")
    for j in range(-n_ker, n_ker + 1):
        for i in range(-n_ker, n_ker + 1):
            ind[i + n_ker, j + n_ker, 0] = i + int(r0_au[0] / dr) + int(grid[0] / 2)
            ind[i + n_ker, j + n_ker, 1] = j + int(r0_au[1] / dr) + int(grid[1] / 2)
    return ind


def source_from_line(dims, line, dr, grid):
    """One sources.in line (source_mod.f90:194-209) -> add_source keyword
    arguments: dir j_amp freq(eV) t0 tau(fs) x y z radius(nm) t_init t_final(fs) phase."""
    v = line.split()
    d, rest = v[0], [float(x) for x in v[1:]]
    j_amp, freq, t0, tau, x, y, z, radius, t_init, t_final, phase = rest
    n_ker = int(radius * NM_TO_AU / dr)
    r0 = np.array([x, y, z]) * NM_TO_AU
    return dict(direction=d, j_amp=j_amp, w0=freq * EV_TO_AU, t0=t0 * FS_TO_AU, tau=tau * FS_TO_AU,
                t_init=t_init * FS_TO_AU, t_final=t_final * FS_TO_AU, phase=phase, n_ker=n_ker,
                ind=source_indices(dims, n_ker, r0, dr, grid), ker=source_kernel(dims, n_ker, dr))


class _LowGrid:
    dims = 0

    def __init__(self, grid, npml, boundaries, dt, dr, mode=0):
        self.lib = L.load_library()
        self.grid = tuple(int(v) for v in grid)
        self.npml = int(npml)
        self.boundaries = tuple(int(b) for b in boundaries)
        self.mode = int(mode)
        self.dt, self.dr = float(dt), float(dr)
        self.dt_eps0 = self.dt / EPS0
        self.dt_mu0 = self.dt / MU0
        self.time = 0.0
        self.n_samples = 0
        d = L.LowDesc()
        d.dims = self.dims
        d.mode = self.mode
        d.n1 = self.grid[0]
        d.n2 = self.grid[1] if self.dims == 2 else 1
        d.npml = self.npml
        for a in range(len(self.boundaries)):
            d.boundaries[a] = self.boundaries[a]
        d.precision = 8
        d.dt, d.dr = self.dt, self.dr
        h = C.c_void_p()
        L.call("omxl_create", C.byref(d), C.byref(h))
        self.h = h

    def kill(self):
        if getattr(self, "h", None):
            self.lib.omxl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.kill()

    def set_media(self, media, aux):
        """media: omx_medium list (medium 1 first); aux: the cell medium map."""
        arr = (L.Medium * max(1, len(media)))(*media)
        a = np.ascontiguousarray(np.asarray(aux, dtype=np.int8).ravel(order="F"))
        L.call("omxl_set_media", self.h, len(media), C.cast(arr, C.c_void_p), a.ctypes.data_as(C.POINTER(C.c_int8)))

    def td_propagate_H_field(self):
        L.call("omxl_step_h", self.h)

    def td_propagate_E_field(self, t=0):
        L.call("omxl_step_e", self.h)

    def add_source(self, direction, j_amp, w0, t0, tau, t_init, t_final, phase, n_ker, ind, ker):
        dirc = {"x": 1, "y": 2, "z": 3}.get(direction, direction)
        ind = np.ascontiguousarray(np.asarray(ind).ravel(order="F"), dtype=np.int32)
        ker = np.ascontiguousarray(np.asarray(ker).ravel(order="F"), dtype=np.float64)
        L.call("omxl_add_source", self.h, int(dirc), j_amp, w0, t0, tau, t_init, t_final, phase, int(n_ker),
               ind.ctypes.data_as(L._IP), ker.ctypes.data_as(L._DP))

    def inject_sources(self, time):
        L.call("omxl_inject_sources", self.h, float(time))

    def add_detector(self, field, cells):
        """Samples of `field` ("Ex".. or 1..6) at cells (n, dims), 1-based."""
        f = FIELD_CODES.get(field, field) if isinstance(field, str) else int(field)
        idx = np.asarray(cells, dtype=np.int32).reshape(-1, self.dims)
        flat = np.ascontiguousarray(idx.ravel(order="F"))
        L.call("omxl_add_detector", self.h, int(f), idx.shape[0], flat.ctypes.data_as(L._IP))
        self.n_samples += idx.shape[0]

    def sample(self):
        out = np.zeros(max(self.n_samples, 1))
        L.call("omxl_sample", self.h, out.ctypes.data_as(L._DP))
        return out[:self.n_samples]

    def run(self, tt_first, n_steps, record_every=0):
        """Steps tt_first..tt_first+n_steps-1 in the driver loop's order;
        returns the detector rows of the steps with mod(tt, record_every) == 0."""
        nrow = 0
        if record_every > 0 and self.n_samples > 0:
            nrow = sum(1 for tt in range(tt_first, tt_first + n_steps) if tt % record_every == 0)
        rec = np.zeros((max(nrow, 1), max(self.n_samples, 1)))
        L.call("omxl_run", self.h, int(tt_first), int(n_steps), int(record_every), rec.ctypes.data_as(L._DP))
        return rec[:nrow, :self.n_samples]

    def _shape(self, f):
        raise NotImplementedError

    def get_field(self, name):
        f = FIELD_CODES[name] if isinstance(name, str) else int(name)
        shp = self._shape(f)
        out = np.zeros(int(np.prod(shp)))
        L.call("omxl_copy_field", self.h, f, out.ctypes.data_as(L._DP), 0)
        return out.reshape(shp, order="F")

    def set_field(self, name, arr):
        f = FIELD_CODES[name] if isinstance(name, str) else int(name)
        flat = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel(order="F"))
        assert flat.size == int(np.prod(self._shape(f)))
        L.call("omxl_copy_field", self.h, f, flat.ctypes.data_as(L._DP), 1)

    def synchronize(self):
        L.call("omxl_synchronize", self.h)

    def device_bytes(self):
        b = C.c_int64()
        L.call("omxl_device_bytes", self.h, C.byref(b))
        return b.value


class Grid2D(_LowGrid):
    """TMxll_2D: grid (nx, ny), boundaries per axis, mode 1 TMz / 2 TEz / 3 both."""
    dims = 2

    def __init__(self, grid_Ndims, npml, boundaries, dt, dr, mode):
        super().__init__(grid_Ndims[:2], npml, boundaries[:2], dt, dr, mode)
        self.nx, self.ny = self.grid

    def _shape(self, f):
        return field_shape_2d(f, self.nx, self.ny)


class Grid1D(_LowGrid):
    """TMxll_1D: the z line of nz cells (grid_Ndims(1)), Ex at k and Hy at k+1/2."""
    dims = 1

    def __init__(self, grid_Ndims, npml, boundaries, dt, dr):
        super().__init__(grid_Ndims[:1], npml, boundaries[:1], dt, dr, 0)
        self.nz = self.grid[0]

    def _shape(self, f):
        return field_shape_1d(f, self.nz)
