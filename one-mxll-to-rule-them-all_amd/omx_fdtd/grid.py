"""Grid3D: the reference's 3D solver interface over the HIP engine.

Method names and argument meanings follow TMxll / TMxll_3D
(src/FDTD/mxll_base_mod.f90:7-69, src/FDTD/mxll_3D_mod.f90:9-57):
``init``, ``kill``, ``td_propagate_H_field``, ``td_propagate_E_field(t)``,
with the state scalars ``time``, ``t_skip``, ``dt_eps0``, ``dt_mu0`` kept on
the object exactly as the reference driver sets them.  The couplings that
the reference performs by touching the solver's arrays directly (sources,
E gathering, J scattering, detector sampling) are device calls here.
Errors raise OmxError (the reference stops with a message).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .units import EPS0, MU0, BC_CPML

FIELD_NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def field_shape(field, nx, ny, nz):
    """Shape of the reference array of a field (mxll_3D_mod.f90:123-131)."""
    n = [nx, ny, nz]
    if field <= 3:
        ext = [n[a] + 1 for a in range(3)]          # 0:n along own axis, 1:n+1 across
    else:
        ext = [n[a] + 1 for a in range(3)]          # 1:n along own axis, 0:n across
        ext[field - 4] = n[field - 4]
    return tuple(ext)


def medium(kind, eps_r=1.0, omega_d=0.0, gamma_d=0.0, poles=()):
    """An omx_medium.  poles: iterable of (omega_k, gamma_k, a_k) in au."""
    m = L.Medium()
    m.kind = kind
    m.eps_r = eps_r
    m.omega_d = omega_d
    m.gamma_d = gamma_d
    m.n_poles = len(poles)
    for q, (w, g, a) in enumerate(poles):
        m.omega_k[q] = w
        m.gamma_k[q] = g
        m.a_k[q] = a
    return m


def library_medium(name, eps_r):
    """Drude-Lorentz pole data of the reference library (Ag, Al, Au)."""
    m = L.Medium()
    L.call("omx_medium_library", name.encode(), eps_r, C.byref(m))
    return m


class Grid3D:
    """One 3D Yee grid on the GPU; precision 4 (FP32) or 8 (FP64)."""

    def __init__(self, grid_Ndims, npml, boundaries, dt, dr, precision=8):
        self.lib = L.load_library()
        self.nx, self.ny, self.nz = (int(v) for v in grid_Ndims)
        self.npml = int(npml)
        self.boundaries = tuple(int(b) for b in boundaries)
        self.dt = float(dt)
        self.dr = float(dr)
        self.precision = int(precision)
        self.dt_eps0 = self.dt / EPS0
        self.dt_mu0 = self.dt / MU0
        self.time = 0.0
        self.t_skip = 0.0
        self.n_samples = 0
        self.n_sites = 0
        d = L.Desc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.npml = self.npml
        for a in range(3):
            d.boundaries[a] = self.boundaries[a]
        d.precision = self.precision
        d.dt, d.dr = self.dt, self.dr
        h = C.c_void_p()
        L.call("omx3d_create", C.byref(d), C.byref(h))
        self.h = h

    # -- lifecycle --------------------------------------------------------
    def kill(self):
        if getattr(self, "h", None):
            self.lib.omx3d_destroy(self.h)
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

    def set_tiling(self, rows_y=0, xchunk=0, tile_order=-1):
        """rows_y / xchunk 0 and tile_order -1 keep the current value."""
        L.call("omx3d_set_tiling", self.h, int(rows_y), int(xchunk), int(tile_order))

    def set_addressing(self, offset_bits=32):
        """32: byte offsets where every array is below 4 GiB (default); 64: element indices."""
        L.call("omx3d_set_addressing", self.h, int(offset_bits))

    def tiling(self):
        r, x, o = C.c_int32(), C.c_int32(), C.c_int32()
        L.call("omx3d_get_tiling", self.h, C.byref(r), C.byref(x), C.byref(o))
        return r.value, x.value, o.value

    def set_source_patch(self, radius):
        """FP64 near-field boxes of `radius` cells around point sources (FP32 grids); 0: off."""
        L.call("omx3d_set_source_patch", self.h, int(radius))

    def source_patch(self):
        """(radius, boxes, cells) of the near-field patches in use."""
        r, nb, nc = C.c_int32(), C.c_int32(), C.c_int64()
        L.call("omx3d_get_source_patch", self.h, C.byref(r), C.byref(nb), C.byref(nc))
        return r.value, nb.value, nc.value

    def set_step_mode(self, mode=0, rows=0, xchunk=0, ywaves=None):
        """mode 0: fused H+E pass where it applies; 1: separate H and E
        kernels.  Fused tile: rows (1/2) per wavefront, xchunk planes per
        workgroup, ywaves wavefronts along y per workgroup (0 automatic)."""
        L.call("omx3d_set_step_mode", self.h, int(mode), int(rows), int(xchunk))
        if ywaves is not None:
            L.call("omx3d_set_fused_waves", self.h, int(ywaves))

    def fused_shape(self):
        """(rows per wavefront, planes per chunk, wavefronts along y) of the fused step."""
        r, x, y = C.c_int32(), C.c_int32(), C.c_int32()
        L.call("omx3d_get_fused_shape", self.h, C.byref(r), C.byref(x), C.byref(y))
        return r.value, x.value, y.value

    def last_run_fused(self):
        f = C.c_int32()
        L.call("omx3d_get_step_mode", self.h, C.byref(f))
        return bool(f.value)

    def autotune(self, steps=4, rows=0):
        """Pick the fastest tile shape (rows 1/2/4: keep that row count);
        clears the fields (call before a run)."""
        r, x, o, ms = C.c_int32(int(rows)), C.c_int32(), C.c_int32(), C.c_double()
        L.call("omx3d_autotune", self.h, int(steps), C.byref(r), C.byref(x), C.byref(o), C.byref(ms))
        return r.value, x.value, o.value, ms.value

    def autotune_step(self, steps=8):
        """Keep the faster of the separate H/E kernels and the fused step
        (several fused tile shapes tried); clears the fields (call before a
        run).  Returns (mode, ms_separate, ms_fused or None)."""
        m, ts, tf = C.c_int32(), C.c_double(), C.c_double()
        L.call("omx3d_autotune_step", self.h, int(steps), C.byref(m), C.byref(ts), C.byref(tf))
        return m.value, ts.value, (tf.value if tf.value >= 0.0 else None)

    # -- coefficients ------------------------------------------------------
    def cpml_tables(self):
        p = self.npml
        be, ce = np.zeros(max(p, 1)), np.zeros(max(p, 1))
        bh, ch = np.zeros(max(p - 1, 1)), np.zeros(max(p - 1, 1))
        den = np.zeros(2 * (self.nx + self.ny + self.nz))
        L.call("omx3d_get_cpml", self.h, L.dptr(be), L.dptr(ce), L.dptr(bh), L.dptr(ch), L.dptr(den))
        n = [self.nx, self.ny, self.nz]
        out = {"be": be[:p], "ce": ce[:p], "bh": bh[:p - 1], "ch": ch[:p - 1]}
        o = 0
        for q, name in enumerate(("den_ex", "den_ey", "den_ez", "den_hx", "den_hy", "den_hz")):
            out[name] = den[o:o + n[q % 3]].copy()
            o += n[q % 3]
        return out

    # -- media -------------------------------------------------------------
    def set_media(self, media, aux):
        """media: list of omx_medium (ids 1..n); aux: (nx,ny,nz) cell medium ids."""
        arr = (L.Medium * max(1, len(media)))(*media)
        a = np.asfortranarray(np.asarray(aux, dtype=np.int8))
        if a.shape != (self.nx, self.ny, self.nz):
            raise ValueError("aux map must have shape (nx, ny, nz)")
        flat = np.ascontiguousarray(a.ravel(order="F"))
        L.call("omx3d_set_media", self.h, len(media), arr, flat.ctypes.data_as(C.POINTER(C.c_int8)))

    def set_media_shapes(self, media, shapes):
        arr = (L.Medium * max(1, len(media)))(*media)
        sh = (L.Shape * max(1, len(shapes)))()
        for s, (typ, med, params) in enumerate(shapes):
            sh[s].type = typ
            sh[s].medium = med
            for q, v in enumerate(params):
                sh[s].p[q] = v
        L.call("omx3d_set_media_shapes", self.h, len(media), arr, len(shapes), sh)

    def media_map(self):
        m = np.zeros(self.nx * self.ny * self.nz * 3, dtype=np.int32)
        L.call("omx3d_get_media_map", self.h, L.iptr(m))
        return m.reshape((self.nx, self.ny, self.nz, 3), order="F")

    def medium_coeffs(self, m):
        c = L.MediumCoeffs()
        L.call("omx3d_get_medium_coeffs", self.h, int(m), C.byref(c))
        return c

    # -- time stepping (TMxll interface) -------------------------------------
    def td_propagate_H_field(self):
        L.call("omx3d_step_h", self.h)

    def td_propagate_E_field(self, t=0):
        L.call("omx3d_step_e", self.h, self.time, self.t_skip)

    def run(self, tt_first, n_steps, record_every=0):
        L.call("omx3d_run", self.h, int(tt_first), int(n_steps), int(record_every))

    def synchronize(self):
        L.call("omx3d_synchronize", self.h)

    # -- sources -------------------------------------------------------------
    def add_source(self, direction, j_amp, w0, t0, tau, t_init, t_final, phase, n_ker, ind, ker):
        dirc = {"x": 1, "y": 2, "z": 3}.get(direction, direction)
        ind = L.i32(np.asarray(ind).ravel(order="F"))
        ker = L.f64(np.asarray(ker).ravel(order="F"))
        L.call("omx3d_add_source", self.h, int(dirc), j_amp, w0, t0, tau, t_init, t_final, phase,
               int(n_ker), L.iptr(ind), L.dptr(ker))

    def inject_sources(self, time):
        L.call("omx3d_inject_sources", self.h, float(time))

    # -- detectors -----------------------------------------------------------
    def add_detector(self, det_type, field, indx_list):
        idx = np.asarray(indx_list, dtype=np.int32).reshape(-1, 3)
        flat = L.i32(idx.ravel(order="F"))
        L.call("omx3d_add_detector", self.h, int(det_type), int(field), idx.shape[0], L.iptr(flat))
        self.n_samples += idx.shape[0]

    def set_record_capacity(self, rows):
        L.call("omx3d_set_record_capacity", self.h, int(rows))

    def record(self):
        L.call("omx3d_record", self.h)

    def read_records(self):
        rows = C.c_int32()
        L.call("omx3d_read_records", self.h, None, C.byref(rows))
        out = np.zeros((rows.value, self.n_samples))
        if rows.value:
            L.call("omx3d_read_records", self.h, L.dptr(out), C.byref(rows))
        return out

    def sample(self):
        out = np.zeros(self.n_samples)
        L.call("omx3d_sample", self.h, L.dptr(out))
        return out

    # -- quantum systems -------------------------------------------------------
    def set_q_sites(self, ijk, density, dt_q):
        idx = np.asarray(ijk, dtype=np.int32).reshape(-1, 3)
        flat = L.i32(idx.ravel(order="F"))
        L.call("omx3d_set_q_sites", self.h, idx.shape[0], L.iptr(flat), float(density), float(dt_q))
        self.n_sites = idx.shape[0]

    def set_q_groups(self, group_types, group_of, ijk, density, dt_q):
        """Quantum systems in q-groups (omx3d_set_q_groups): group_types 1
        (material) or 2 (single) per group, group_of each system's group
        (0-based, systems in group order), ijk (n,3) cells."""
        idx = np.asarray(ijk, dtype=np.int32).reshape(-1, 3)
        flat = L.i32(idx.ravel(order="F"))
        gt = L.i32(group_types)
        gof = L.i32(group_of)
        L.call("omx3d_set_q_groups", self.h, gt.size, L.iptr(gt), idx.shape[0], L.iptr(gof), L.iptr(flat),
               float(density), float(dt_q))
        self.n_sites = idx.shape[0]

    def gather_E(self):
        e = np.zeros(3 * self.n_sites)
        L.call("omx3d_gather_e", self.h, L.dptr(e))
        return e.reshape((self.n_sites, 3), order="F")

    def scatter_J(self, dpdt):
        d = L.f64(np.asarray(dpdt, dtype=np.float64).reshape(self.n_sites, 3).ravel(order="F"))
        L.call("omx3d_scatter_j", self.h, L.dptr(d))

    def set_two_level(self, omega0, mu, pol, gamma2):
        o, m, g = (L.f64(np.asarray(v, dtype=np.float64)) for v in (omega0, mu, gamma2))
        p = L.f64(np.asarray(pol, dtype=np.float64).reshape(self.n_sites, 3).ravel(order="F"))
        L.call("omx3d_set_two_level", self.h, L.dptr(o), L.dptr(m), L.dptr(p), L.dptr(g))

    def two_level_step(self):
        L.call("omx3d_two_level_step", self.h)

    def two_level_state(self):
        dip = np.zeros(3 * self.n_sites)
        rho = np.zeros(4 * self.n_sites)
        L.call("omx3d_get_two_level", self.h, L.dptr(dip), L.dptr(rho))
        return dip.reshape((self.n_sites, 3), order="F"), rho.reshape((self.n_sites, 4), order="F")

    # -- fields, reductions ------------------------------------------------------
    def get_field(self, name):
        f = FIELD_NAMES.index(name) + 1
        shp = field_shape(f, self.nx, self.ny, self.nz)
        out = np.zeros(int(np.prod(shp)))
        L.call("omx3d_copy_field", self.h, f, L.dptr(out), 0)
        return out.reshape(shp, order="F")

    def set_field(self, name, arr):
        f = FIELD_NAMES.index(name) + 1
        shp = field_shape(f, self.nx, self.ny, self.nz)
        a = np.asarray(arr, dtype=np.float64)
        if a.shape != shp:
            raise ValueError(f"{name} must have shape {shp}")
        flat = L.f64(a.ravel(order="F"))
        L.call("omx3d_copy_field", self.h, f, L.dptr(flat), 1)

    def energy(self):
        w64, w32 = C.c_double(), C.c_double()
        L.call("omx3d_energy", self.h, C.byref(w64), C.byref(w32))
        return w64.value, w32.value

    def tfsf_line(self):
        """(kb, ex1, hy1): the incident line's Ex at k = kb + m and Hy at
        k + 1/2 (omx3d_get_tfsf_line)."""
        kb, n = L.i32([0]), L.i32([0])
        L.call("omx3d_get_tfsf_line", self.h, None, None, L.iptr(kb), L.iptr(n))
        ex1, hy1 = np.zeros(int(n[0])), np.zeros(int(n[0]))
        L.call("omx3d_get_tfsf_line", self.h, L.dptr(ex1), L.dptr(hy1), L.iptr(kb), L.iptr(n))
        return int(kb[0]), ex1, hy1

    def set_tfsf(self, lo, hi, e0, w0, t0, tau, max_steps):
        lo = L.i32(lo)
        hi = L.i32(hi)
        L.call("omx3d_set_tfsf", self.h, L.iptr(lo), L.iptr(hi), e0, w0, t0, tau, int(max_steps))

    # -- profiling -----------------------------------------------------------------
    def set_profiling(self, on, every=1):
        """Events around the H and E launches of every `every`-th step."""
        L.call("omx3d_set_profiling", self.h, int(every) if on else 0)

    def profile(self):
        mh, me, n = C.c_double(), C.c_double(), C.c_int32()
        L.call("omx3d_get_profile", self.h, C.byref(mh), C.byref(me), C.byref(n))
        return mh.value, me.value, n.value

    def device_bytes(self):
        b = C.c_int64()
        L.call("omx3d_device_bytes", self.h, C.byref(b))
        return b.value

    @property
    def has_cpml(self):
        return any(b == BC_CPML for b in self.boundaries)


class Batch:
    """Independent grids advanced together: one launch per phase for all."""

    def __init__(self, grids):
        self.lib = L.load_library()
        self.grids = list(grids)
        arr = (C.c_void_p * len(self.grids))(*[g.h for g in self.grids])
        h = C.c_void_p()
        L.call("omxb_create", len(self.grids), arr, C.byref(h))
        self.h = h

    def run(self, tt_first, n_steps, record_every=0, graph=True):
        L.call("omxb_run", self.h, int(tt_first), int(n_steps), int(record_every), 1 if graph else 0)

    def kill(self):
        if getattr(self, "h", None):
            self.lib.omxb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.kill()
        except Exception:
            pass
