"""ctypes wrapper of liboracle.so, the FP64 C restatement of the reference's
3D FDTD path (omx_oracle.c).  TEST INFRASTRUCTURE ONLY: imported by tests/,
by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
checker, never as the thing measured or shipped.

OracleGrid mirrors the method names of omx_fdtd.Grid3D, and OracleLowGrid
those of omx_fdtd.Grid2D / Grid1D, so that a parity test drives both with
the same calls.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_lib = None


class OrcMedium(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_poles", C.c_int32), ("eps_r", C.c_double),
                ("omega_d", C.c_double), ("gamma_d", C.c_double),
                ("omega_k", C.c_double * 8), ("gamma_k", C.c_double * 8), ("a_k", C.c_double * 8)]


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"oracle not built at {LIB_PATH}: run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        V, P, I, D = None, C.c_void_p, C.c_int, C.c_double
        sig = {
            "orc_create": ([I, I, I, I, _IP, D, D], P),
            "orc_destroy": ([P], V),
            "orc_get_cpml": ([P, _DP, _DP, _DP, _DP, _DP], V),
            "orc_library": ([C.c_char_p, D, C.POINTER(OrcMedium)], I),
            "orc_set_media": ([P, I, C.POINTER(OrcMedium), C.POINTER(C.c_int8)], I),
            "orc_get_media_map": ([P, _IP], V),
            "orc_get_medium_coeffs": ([P, I, _DP], V),
            "orc_step_h": ([P], V),
            "orc_step_e": ([P, D, D], V),
            "orc_source_kernel": ([I, D, _DP], V),
            "orc_source_indices": ([I, _DP, D, _IP, _IP], V),
            "orc_add_source": ([P, I, D, D, D, D, D, D, D, I, _IP, _DP], I),
            "orc_inject_sources": ([P, D], V),
            "orc_add_detector": ([P, I, I, I, _IP], I),
            "orc_sample": ([P, _DP], V),
            "orc_set_q_sites": ([P, I, _IP, D, D], V),
            "orc_set_q_groups": ([P, I, _IP, I, _IP, _IP, D, D], V),
            "orc_gather_e": ([P, _DP], V),
            "orc_scatter_j": ([P, _DP], V),
            "orc_set_two_level": ([P, _DP, _DP, _DP, _DP], V),
            "orc_two_level_step": ([P], V),
            "orc_get_two_level": ([P, _DP, _DP], V),
            "orc_copy_field": ([P, I, _DP, I], V),
            "orc_set_tfsf": ([P, _IP, _IP, D, D, D, D, I], V),
            "orc_get_tfsf_line": ([P, _DP, _DP, _IP, _IP], V),
            "orc_n_samples": ([P], C.c_long),
            "orc_run": ([P, I, I, I, _DP, _IP], V),
            "orc_num_threads": ([], I),
            "orl_create": ([I, I, I, I, I, _IP, D, D], P),
            "orl_destroy": ([P], V),
            "orl_get_den": ([P, _DP], V),
            "orl_set_media": ([P, I, C.POINTER(OrcMedium), C.POINTER(C.c_int8)], I),
            "orl_get_media_map": ([P, _IP], V),
            "orl_step_h": ([P], V),
            "orl_step_e": ([P], V),
            "orl_source_kernel": ([I, I, D, _DP], V),
            "orl_add_source": ([P, I, D, D, D, D, D, D, D, I, _IP, _DP], I),
            "orl_inject_sources": ([P, D], V),
            "orl_add_detector": ([P, I, I, _IP], I),
            "orl_sample": ([P, _DP], V),
            "orl_n_samples": ([P], C.c_long),
            "orl_run": ([P, I, I, I, _DP, _IP], V),
            "orl_copy_field": ([P, I, _DP, I], V),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(_DP)


def _i(a):
    return a.ctypes.data_as(_IP)


def to_orc_medium(m):
    """Copy an omx_fdtd medium (same field layout) into the oracle struct."""
    o = OrcMedium()
    C.memmove(C.byref(o), C.byref(m), C.sizeof(OrcMedium))
    return o


def library_medium(name, eps_r):
    m = OrcMedium()
    if load().orc_library(name.encode(), eps_r, C.byref(m)) != 0:
        raise ValueError(name)
    return m


def source_kernel(n_ker, dr):
    w = 2 * n_ker + 1
    ker = np.zeros(w ** 3)
    load().orc_source_kernel(n_ker, dr, _d(ker))
    return ker.reshape((w, w, w), order="F")


def source_indices(n_ker, r0_au, dr, grid):
    w = 2 * n_ker + 1
    ind = np.zeros(3 * w ** 3, dtype=np.int32)
    r0 = np.ascontiguousarray(r0_au, dtype=np.float64)
    g = np.ascontiguousarray(grid, dtype=np.int32)
    load().orc_source_indices(n_ker, _d(r0), dr, _i(g), _i(ind))
    return ind.reshape((w, w, w, 3), order="F")


def field_shape(field, nx, ny, nz):
    n = [nx, ny, nz]
    ext = [n[a] + 1 for a in range(3)]
    if field >= 4:
        ext[field - 4] = n[field - 4]
    return tuple(ext)


class OracleGrid:
    def __init__(self, grid_Ndims, npml, boundaries, dt, dr):
        self.lib = load()
        self.nx, self.ny, self.nz = (int(v) for v in grid_Ndims)
        self.npml = int(npml)
        self.dt, self.dr = float(dt), float(dr)
        bc = np.ascontiguousarray(boundaries, dtype=np.int32)
        self.h = self.lib.orc_create(self.nx, self.ny, self.nz, self.npml, _i(bc), self.dt, self.dr)
        self.time = 0.0
        self.t_skip = 0.0
        self.n_sites = 0

    def kill(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass

    def cpml_tables(self):
        p = self.npml
        be, ce = np.zeros(max(p, 1)), np.zeros(max(p, 1))
        bh, ch = np.zeros(max(p - 1, 1)), np.zeros(max(p - 1, 1))
        den = np.zeros(2 * (self.nx + self.ny + self.nz))
        self.lib.orc_get_cpml(self.h, _d(be), _d(ce), _d(bh), _d(ch), _d(den))
        n = [self.nx, self.ny, self.nz]
        out = {"be": be[:p], "ce": ce[:p], "bh": bh[:p - 1], "ch": ch[:p - 1]}
        o = 0
        for q, name in enumerate(("den_ex", "den_ey", "den_ez", "den_hx", "den_hy", "den_hz")):
            out[name] = den[o:o + n[q % 3]].copy()
            o += n[q % 3]
        return out

    def set_media(self, media, aux):
        arr = (OrcMedium * max(1, len(media)))(*[to_orc_medium(m) if not isinstance(m, OrcMedium) else m
                                                  for m in media])
        a = np.ascontiguousarray(np.asarray(aux, dtype=np.int8).ravel(order="F"))
        self.lib.orc_set_media(self.h, len(media), arr, a.ctypes.data_as(C.POINTER(C.c_int8)))

    def media_map(self):
        m = np.zeros(self.nx * self.ny * self.nz * 3, dtype=np.int32)
        self.lib.orc_get_media_map(self.h, _i(m))
        return m.reshape((self.nx, self.ny, self.nz, 3), order="F")

    def medium_coeffs(self, m):
        out = np.zeros(47)
        self.lib.orc_get_medium_coeffs(self.h, int(m), _d(out))
        return {"A1": out[0], "A2": out[1], "C1": out[2], "C2": out[3], "C3": out[4], "C4": out[5],
                "inv_eps": out[6], "alpha_k": out[7:15], "beta_k": out[15:23], "gamma_k": out[23:31],
                "B1_k": out[31:39], "B2_k": out[39:47]}

    def td_propagate_H_field(self):
        self.lib.orc_step_h(self.h)

    def td_propagate_E_field(self, t=0):
        self.lib.orc_step_e(self.h, self.time, self.t_skip)

    def add_source(self, direction, j_amp, w0, t0, tau, t_init, t_final, phase, n_ker, ind, ker):
        dirc = {"x": 1, "y": 2, "z": 3}.get(direction, direction)
        ind = np.ascontiguousarray(np.asarray(ind).ravel(order="F"), dtype=np.int32)
        ker = np.ascontiguousarray(np.asarray(ker).ravel(order="F"), dtype=np.float64)
        self.lib.orc_add_source(self.h, int(dirc), j_amp, w0, t0, tau, t_init, t_final, phase, int(n_ker),
                                _i(ind), _d(ker))

    def inject_sources(self, time):
        self.lib.orc_inject_sources(self.h, float(time))

    def add_detector(self, det_type, field, indx_list):
        idx = np.asarray(indx_list, dtype=np.int32).reshape(-1, 3)
        flat = np.ascontiguousarray(idx.ravel(order="F"))
        self.lib.orc_add_detector(self.h, int(det_type), int(field), idx.shape[0], _i(flat))

    @property
    def n_samples(self):
        return int(self.lib.orc_n_samples(self.h))

    def sample(self):
        out = np.zeros(self.n_samples)
        self.lib.orc_sample(self.h, _d(out))
        return out

    def set_q_sites(self, ijk, density, dt_q):
        idx = np.asarray(ijk, dtype=np.int32).reshape(-1, 3)
        flat = np.ascontiguousarray(idx.ravel(order="F"))
        self.lib.orc_set_q_sites(self.h, idx.shape[0], _i(flat), float(density), float(dt_q))
        self.n_sites = idx.shape[0]

    def set_q_groups(self, group_types, group_of, ijk, density, dt_q):
        """group_types: 1 material / 2 single per group; group_of: each
        system's group (0-based, in group order); ijk: (n,3) cells."""
        idx = np.asarray(ijk, dtype=np.int32).reshape(-1, 3)
        flat = np.ascontiguousarray(idx.ravel(order="F"))
        gt = np.ascontiguousarray(group_types, dtype=np.int32)
        gof = np.ascontiguousarray(group_of, dtype=np.int32)
        self.lib.orc_set_q_groups(self.h, gt.size, _i(gt), idx.shape[0], _i(gof), _i(flat), float(density),
                                  float(dt_q))
        self.n_sites = idx.shape[0]

    def gather_E(self):
        e = np.zeros(3 * self.n_sites)
        self.lib.orc_gather_e(self.h, _d(e))
        return e.reshape((self.n_sites, 3), order="F")

    def scatter_J(self, dpdt):
        d = np.ascontiguousarray(np.asarray(dpdt, dtype=np.float64).reshape(self.n_sites, 3).ravel(order="F"))
        self.lib.orc_scatter_j(self.h, _d(d))

    def set_two_level(self, omega0, mu, pol, gamma2):
        o, m, g = (np.ascontiguousarray(v, dtype=np.float64) for v in (omega0, mu, gamma2))
        p = np.ascontiguousarray(np.asarray(pol, dtype=np.float64).reshape(self.n_sites, 3).ravel(order="F"))
        self.lib.orc_set_two_level(self.h, _d(o), _d(m), _d(p), _d(g))

    def two_level_step(self):
        self.lib.orc_two_level_step(self.h)

    def two_level_state(self):
        dip = np.zeros(3 * self.n_sites)
        rho = np.zeros(4 * self.n_sites)
        self.lib.orc_get_two_level(self.h, _d(dip), _d(rho))
        return dip.reshape((self.n_sites, 3), order="F"), rho.reshape((self.n_sites, 4), order="F")

    def get_field(self, name):
        f = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz").index(name) + 1
        shp = field_shape(f, self.nx, self.ny, self.nz)
        out = np.zeros(int(np.prod(shp)))
        self.lib.orc_copy_field(self.h, f, _d(out), 0)
        return out.reshape(shp, order="F")

    def set_field(self, name, arr):
        f = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz").index(name) + 1
        flat = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel(order="F"))
        self.lib.orc_copy_field(self.h, f, _d(flat), 1)

    def tfsf_line(self):
        """(kb, ex1, hy1): the incident line's Ex at k = kb + m and Hy at
        k + 1/2."""
        kb, n = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
        self.lib.orc_get_tfsf_line(self.h, None, None, _i(kb), _i(n))
        ex1, hy1 = np.zeros(int(n[0])), np.zeros(int(n[0]))
        self.lib.orc_get_tfsf_line(self.h, _d(ex1), _d(hy1), _i(kb), _i(n))
        return int(kb[0]), ex1, hy1

    def set_tfsf(self, lo, hi, e0, w0, t0, tau, max_steps):
        lo = np.ascontiguousarray(lo, dtype=np.int32)
        hi = np.ascontiguousarray(hi, dtype=np.int32)
        self.lib.orc_set_tfsf(self.h, _i(lo), _i(hi), e0, w0, t0, tau, int(max_steps))

    def run(self, tt_first, n_steps, record_every=0):
        """Returns the recorded detector rows (rows x n_samples)."""
        ns = self.n_samples
        nrow = 0
        if record_every > 0:
            nrow = sum(1 for tt in range(tt_first, tt_first + n_steps) if tt % record_every == 0)
        rec = np.zeros((max(nrow, 1), max(ns, 1)))
        rows = C.c_int32()
        self.lib.orc_run(self.h, int(tt_first), int(n_steps), int(record_every), _d(rec), C.byref(rows))
        return rec[:rows.value, :ns]


class OracleLowGrid:
    """The 2D (dims 2, mode 1 TMz / 2 TEz / 3 both) or 1D (dims 1) grid of
    omx_oracle.c, with the method names of omx_fdtd.Grid2D / Grid1D."""

    def __init__(self, dims, grid, npml, boundaries, dt, dr, mode=0):
        self.lib = load()
        self.dims, self.mode = int(dims), int(mode)
        self.grid = tuple(int(v) for v in grid)[:self.dims]
        self.npml = int(npml)
        self.dt, self.dr = float(dt), float(dr)
        bc = np.ascontiguousarray(list(boundaries)[:self.dims] + [1] * (2 - self.dims), dtype=np.int32)
        n2 = self.grid[1] if self.dims == 2 else 1
        self.h = self.lib.orl_create(self.dims, self.mode, self.grid[0], n2, self.npml, _i(bc), self.dt, self.dr)

    def kill(self):
        if getattr(self, "h", None):
            self.lib.orl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass

    def den(self):
        """den_ex, den_hx (1..n1), den_ey, den_hy (1..n2) as one array."""
        n1 = self.grid[0]
        n2 = self.grid[1] if self.dims == 2 else 1
        out = np.zeros(2 * (n1 + n2))
        self.lib.orl_get_den(self.h, _d(out))
        return out

    def set_media(self, media, aux):
        arr = (OrcMedium * max(1, len(media)))(*[to_orc_medium(m) if not isinstance(m, OrcMedium) else m
                                                  for m in media])
        a = np.ascontiguousarray(np.asarray(aux, dtype=np.int8).ravel(order="F"))
        self.lib.orl_set_media(self.h, len(media), arr, a.ctypes.data_as(C.POINTER(C.c_int8)))

    def media_map(self):
        if self.dims == 1:
            m = np.zeros(self.grid[0], dtype=np.int32)
            self.lib.orl_get_media_map(self.h, _i(m))
            return m
        m = np.zeros(self.grid[0] * self.grid[1] * 3, dtype=np.int32)
        self.lib.orl_get_media_map(self.h, _i(m))
        return m.reshape(self.grid + (3,), order="F")

    def td_propagate_H_field(self):
        self.lib.orl_step_h(self.h)

    def td_propagate_E_field(self, t=0):
        self.lib.orl_step_e(self.h)

    def add_source(self, direction, j_amp, w0, t0, tau, t_init, t_final, phase, n_ker, ind, ker):
        dirc = {"x": 1, "y": 2, "z": 3}.get(direction, direction)
        ind = np.ascontiguousarray(np.asarray(ind).ravel(order="F"), dtype=np.int32)
        ker = np.ascontiguousarray(np.asarray(ker).ravel(order="F"), dtype=np.float64)
        self.lib.orl_add_source(self.h, int(dirc), j_amp, w0, t0, tau, t_init, t_final, phase, int(n_ker),
                                _i(ind), _d(ker))

    def inject_sources(self, time):
        self.lib.orl_inject_sources(self.h, float(time))

    def add_detector(self, field, cells):
        f = {"Ex": 1, "Ey": 2, "Ez": 3, "Hx": 4, "Hy": 5, "Hz": 6}.get(field, field)
        idx = np.asarray(cells, dtype=np.int32).reshape(-1, self.dims)
        flat = np.ascontiguousarray(idx.ravel(order="F"))
        self.lib.orl_add_detector(self.h, int(f), idx.shape[0], _i(flat))

    @property
    def n_samples(self):
        return int(self.lib.orl_n_samples(self.h))

    def sample(self):
        out = np.zeros(max(self.n_samples, 1))
        self.lib.orl_sample(self.h, _d(out))
        return out[:self.n_samples]

    def run(self, tt_first, n_steps, record_every=0):
        ns = self.n_samples
        nrow = 0
        if record_every > 0:
            nrow = sum(1 for tt in range(tt_first, tt_first + n_steps) if tt % record_every == 0)
        rec = np.zeros((max(nrow, 1), max(ns, 1)))
        rows = C.c_int32()
        self.lib.orl_run(self.h, int(tt_first), int(n_steps), int(record_every), _d(rec), C.byref(rows))
        return rec[:rows.value, :ns]

    def _shape(self, f):
        if self.dims == 1:
            return (self.grid[0],) if f == 1 else (self.grid[0] - 1,)
        nx, ny = self.grid
        return {1: (nx + 1, ny + 1), 2: (nx + 1, ny + 1), 3: (nx + 1, ny + 1),
                4: (nx, ny + 1), 5: (nx + 1, ny), 6: (nx + 1, ny + 1)}[f]

    def get_field(self, name):
        f = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz").index(name) + 1
        shp = self._shape(f)
        out = np.zeros(int(np.prod(shp)))
        self.lib.orl_copy_field(self.h, f, _d(out), 0)
        return out.reshape(shp, order="F")

    def set_field(self, name, arr):
        f = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz").index(name) + 1
        flat = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).ravel(order="F"))
        self.lib.orl_copy_field(self.h, f, _d(flat), 1)


def source_kernel_lowdim(dims, n_ker, dr):
    w = 2 * n_ker + 1
    ker = np.zeros(w ** dims)
    load().orl_source_kernel(int(dims), int(n_ker), float(dr), _d(ker))
    return ker if dims == 1 else ker.reshape((w, w), order="F")


def rasterize(shapes, nx, ny, nz):
    """Cell medium map of a shape list (index units, cell (i,j,k) at (i,j,k));
    first medium wins, medium 0 carves.  The engine rasterises on the device."""
    aux = np.zeros((nx, ny, nz), dtype=np.int8)
    I, J, K = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    for typ, med, p in shapes:
        p = list(p) + [0.0] * (15 - len(p))
        if typ == 1:
            inside = (I - p[0]) ** 2 + (J - p[1]) ** 2 + (K - p[2]) ** 2 <= p[3] ** 2
        elif typ == 2:
            dx, dy, dz = I - p[0], J - p[1], K - p[2]
            u = p[6] * dx + p[7] * dy + p[8] * dz
            w = p[9] * dx + p[10] * dy + p[11] * dz
            t = p[12] * dx + p[13] * dy + p[14] * dz
            inside = (u / p[3]) ** 2 + (w / p[4]) ** 2 + (t / p[5]) ** 2 <= 1.0
        elif typ == 3:
            inside = (I >= p[0]) & (J >= p[1]) & (K >= p[2]) & (I <= p[3]) & (J <= p[4]) & (K <= p[5])
        elif typ == 4:
            inside = ((I - p[0]) ** 2 + (J - p[1]) ** 2 <= p[2] ** 2) & (K >= p[3]) & (K <= p[4])
        else:
            raise ValueError(typ)
        if med == 0:
            aux[inside] = 0
        else:
            aux[inside & (aux == 0)] = med
    return aux


def make_medium(kind, eps_r, omega_d, gamma_d, poles):
    m = OrcMedium()
    m.kind, m.eps_r, m.omega_d, m.gamma_d, m.n_poles = kind, eps_r, omega_d, gamma_d, len(poles)
    for q, (w, g, a) in enumerate(poles):
        m.omega_k[q], m.gamma_k[q], m.a_k[q] = w, g, a
    return m
