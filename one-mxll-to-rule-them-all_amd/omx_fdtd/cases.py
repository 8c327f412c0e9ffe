"""Case descriptions shared by the engine, the parity tests and the bench.

A Case holds what the reference reads from its input files -- the namelist
(input_mod.f90:46-195), sources.in (source_mod.f90:151-212), medium_XXXX.in
(classical_medium_mod.f90:546-692), detectors.in (output_mod.f90:18-80) --
plus the extensions the synthetic workloads need (custom Drude-Lorentz
poles, shapes rasterised on the device, two-level emitters, TF/SF plane
wave).  `apply(case, grid)` sets up any object with the Grid3D method names.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from .units import (NM_TO_AU, EV_TO_AU, FS_TO_AU, DEBYE_TO_AU, BC_CPML, BC_PERIODIC, BC_CLOSE,
                    MEDIUM_DIELECTRIC, MEDIUM_DRUDE, MEDIUM_DL, DET_POINT, DET_PLANE_XY,
                    FIELD_CODES, grid_dr_dt, index_to_nm)
from . import host

SEED_BASE = 20261203


@dataclass
class Case:
    name: str
    grid: tuple
    dx_nm: float
    npml: int
    boundaries: tuple
    steps: int
    record_every: int = 1
    media: list = field(default_factory=list)     # medium specs (see medium_struct)
    shapes: list = field(default_factory=list)    # (type, medium, params) in index units
    aux: object = None                            # explicit (nx,ny,nz) int8 cell map
    sources: list = field(default_factory=list)   # sources.in lines as tuples
    detectors: list = field(default_factory=list) # (type, field, indx_list)
    q_sites: object = None                        # (n,3) int indices
    q_density: float = 0.0
    q_skip: int = 1                               # quantum step dt_q = q_skip * dt (n_skip_steps)
    two_level: dict = None                        # omega0, mu, pol, gamma2 (au)
    tfsf: dict = None                             # lo, hi, e0, w0, t0, tau
    info: dict = field(default_factory=dict)

    @property
    def dr_dt(self):
        return grid_dr_dt(self.dx_nm)

    @property
    def n_cells(self):
        return int(np.prod(self.grid))


def medium_struct(spec, make):
    """spec -> medium via `make(kind, eps_r, omega_d, gamma_d, poles)` or a
    library call.  Specs (energies in eV, as in medium_XXXX.in):
      ("dielectric", eps_r)
      ("drude", omega_d_eV, gamma_d_eV, eps_r)
      ("drude-lorentz", material, eps_r)            library Ag / Al / Au
      ("dl-custom", eps_inf, omega_p_eV, gamma_d_eV, [(f_k, omega_k_eV, gamma_k_eV), ...])
    """
    kind = spec[0]
    if kind == "dielectric":
        return make(MEDIUM_DIELECTRIC, spec[1], 0.0, 0.0, ())
    if kind == "drude":
        return make(MEDIUM_DRUDE, spec[3], spec[1] * EV_TO_AU, spec[2] * EV_TO_AU, ())
    if kind == "drude-lorentz":
        return ("library", spec[1], spec[2])
    if kind == "dl-custom":
        eps_inf, wp, gd, poles = spec[1], spec[2] * EV_TO_AU, spec[3] * EV_TO_AU, spec[4]
        pl = [(w * EV_TO_AU, g * EV_TO_AU, f * wp * wp) for (f, w, g) in poles]
        return make(MEDIUM_DL, eps_inf, wp, gd, pl)
    raise ValueError(f"unknown medium spec {spec!r}")


def apply(case, g, library=None, make_medium=None, rasterize=None, record=True):
    """Configure grid object `g` (Grid3D or the oracle) for `case`."""
    dr, dt = case.dr_dt
    if case.media:
        meds = []
        for spec in case.media:
            m = medium_struct(spec, make_medium)
            if isinstance(m, tuple) and m[0] == "library":
                m = library(m[1], m[2])
            meds.append(m)
        if case.aux is not None:
            g.set_media(meds, case.aux)
        elif rasterize is None:
            g.set_media_shapes(meds, case.shapes)
        else:
            g.set_media(meds, rasterize(case.shapes, *case.grid))
    for line in case.sources:
        kw = host.source_from_line(*line, dr=dr, grid=case.grid)
        g.add_source(**kw)
    for (dtype, fld, idx) in case.detectors:
        g.add_detector(dtype, fld, idx)
    if case.q_sites is not None:
        g.set_q_sites(case.q_sites, case.q_density, case.q_skip * dt)
        if case.two_level is not None:
            t = case.two_level
            g.set_two_level(t["omega0"], t["mu"], t["pol"], t["gamma2"])
    if case.tfsf is not None:
        t = case.tfsf
        g.set_tfsf(t["lo"], t["hi"], t["e0"], t["w0"], t["t0"], t["tau"], case.steps + 8)
    if record and hasattr(g, "set_record_capacity") and case.record_every > 0:
        g.set_record_capacity(case.steps // case.record_every + 1)
    return g


def n_records(case, tt_first=1):
    if case.record_every <= 0:
        return 0
    return sum(1 for tt in range(tt_first, tt_first + case.steps) if tt % case.record_every == 0)


# ---------------------------------------------------------------------------
# Synthetic workloads
# ---------------------------------------------------------------------------
def point_detectors(positions, fields=("Ex", "Ey", "Ez")):
    out = []
    for p in positions:
        for f in fields:
            out.append((DET_POINT, FIELD_CODES[f], np.array([p], dtype=np.int32)))
    return out


def plane_xy(grid, k, i_range=None, j_range=None):
    nx, ny, _ = grid
    i0, i1 = i_range or (1, nx)
    j0, j1 = j_range or (1, ny)
    idx = [(i, j, k) for j in range(j0, j1 + 1) for i in range(i0, i1 + 1)]
    return np.array(idx, dtype=np.int32)


def config_vacuum_point(n=64, dx_nm=5.0, npml=10, steps=1000, seed_no=1, n_det=8):
    """Plain 3D vacuum grid with CPML on all faces and one Ez point source at
    the centre (Gaussian-modulated 800 nm carrier, 5 fs width) and seeded
    point detectors recording Ex, Ey, Ez every step."""
    rng = np.random.default_rng(SEED_BASE + seed_no)
    grid = (n, n, n)
    pos = rng.integers(npml + 2, n - npml - 1, size=(n_det, 3), endpoint=True)
    src = ("z", 1.0, 1.549802, 5.0, 5.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0e6, 0.0)
    return Case(name=f"vacuum{n}", grid=grid, dx_nm=dx_nm, npml=npml, boundaries=(BC_CPML,) * 3,
                steps=steps, record_every=1, sources=[src], detectors=point_detectors(pos),
                info={"detector_cells": pos.tolist()})


def dl_sphere_poles(rng):
    poles = [(float(rng.uniform(0.5, 2.0)), float(rng.uniform(2.5, 4.0)), float(rng.uniform(0.1, 0.5)))
             for _ in range(2)]
    return ("dl-custom", 5.0, 9.0, 0.07, poles)


def config_dl_sphere_tfsf(n=256, dx_nm=2.0, npml=12, steps=4000, seed_no=2, record_every=20,
                          radius_nm=None):
    """3D grid with one seeded Drude-Lorentz sphere (eps_inf 5, Drude 9 eV /
    0.07 eV, two seeded Lorentz poles) in vacuum, lit by an x-polarised TF/SF
    plane wave along +z (Gaussian pulse, 600 nm carrier); one plane_xy
    detector each for Ex, Ey, Ez on the transmitted side."""
    rng = np.random.default_rng(SEED_BASE + seed_no)
    grid = (n, n, n)
    c = rng.uniform(0.25 * n, 0.75 * n, size=3)
    r_nm = float(rng.uniform(20.0, 40.0)) if radius_nm is None else radius_nm
    r = r_nm / dx_nm
    med = dl_sphere_poles(rng)
    margin = npml + 4
    lo = (margin, margin, margin)
    hi = (n - margin, n - margin, n - margin)
    kdet = int(min(c[2] + r + 8, hi[2] - 3))
    pl = plane_xy(grid, kdet)
    dets = [(DET_PLANE_XY, FIELD_CODES[f], pl) for f in ("Ex", "Ey", "Ez")]
    _, dt = grid_dr_dt(dx_nm)
    tfsf = dict(lo=lo, hi=hi, e0=1.0, w0=2.066 * EV_TO_AU, t0=5.0 * FS_TO_AU, tau=1.5 * FS_TO_AU)
    return Case(name=f"dlsphere{n}", grid=grid, dx_nm=dx_nm, npml=npml, boundaries=(BC_CPML,) * 3,
                steps=steps, record_every=record_every, media=[med],
                shapes=[(1, 1, [float(c[0]), float(c[1]), float(c[2]), r])], detectors=dets, tfsf=tfsf,
                info={"sphere_centre_cells": c.tolist(), "sphere_radius_nm": r_nm, "detector_k": kdet,
                      "medium": med})


def config_batched(n_grids=64, edges=(96, 128, 160, 192), npml=8, steps=2000, dx_nm=2.0, seed_no=4,
                   record_every=10):
    """Independent grids with edges drawn from `edges` per axis, CPML on every
    face, 1-8 seeded Drude-Lorentz spheres (radius 5-30 cells, clipped to
    the non-CPML interior) and one dipole source of seeded polarisation (three
    point sources x, y, z scaled by the polarisation components); one point
    detector (Ex, Ey, Ez) per grid."""
    rng = np.random.default_rng(SEED_BASE + seed_no)
    out = []
    for g in range(n_grids):
        n = tuple(int(v) for v in rng.choice(edges, size=3))
        med = dl_sphere_poles(rng)
        shapes = []
        for _ in range(int(rng.integers(1, 9))):
            rmax = min(30.0, (min(n) - 2 * npml - 4) / 2.0)
            r = float(rng.uniform(5.0, max(5.0, rmax)))
            c = [float(rng.uniform(npml + r + 2, n[a] - npml - r - 2)) for a in range(3)]
            shapes.append((1, 1, c + [r]))
        p = rng.normal(size=3)
        p /= np.linalg.norm(p)
        src_cell = [int(rng.integers(npml + 4, n[a] - npml - 4)) for a in range(3)]
        pos_nm = [index_to_nm(src_cell[a], dx_nm, n[a]) for a in range(3)]
        srcs = [(d, float(p[q]), 2.066, 4.0, 1.5, pos_nm[0], pos_nm[1], pos_nm[2], 0.0, 0.0, 1.0e6, 0.0)
                for q, d in enumerate(("x", "y", "z"))]
        det_cell = [int(rng.integers(npml + 2, n[a] - npml - 2)) for a in range(3)]
        out.append(Case(name=f"batch{g}", grid=n, dx_nm=dx_nm, npml=npml, boundaries=(BC_CPML,) * 3,
                        steps=steps, record_every=record_every, media=[med], shapes=shapes, sources=srcs,
                        detectors=point_detectors([det_cell]),
                        info={"polarisation": p.tolist(), "n_spheres": len(shapes)}))
    return out


def random_rotation(rng):
    """Uniform random rotation (unit quaternion) as a row-major 3x3 matrix."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]


def config_ellipsoids(grid=(768, 768, 512), dx_nm=1.0, npml=16, steps=10000, n_ell=64, seed_no=3,
                      record_every=100):
    """Large grid: 64 seeded Drude-Lorentz ellipsoids (semi-axes 5-25 nm,
    uniform random orientations, centres in the upper three quarters) above a
    dielectric substrate (eps 2.25, lower quarter of the grid); a soft
    z-dipole source above the substrate; 32 point detectors recording Ez; the
    volume energy integral is taken every `record_every` steps.

    The substrate stops where the CPML begins: the reference adds the CPML
    term dt/eps0 * psi to E without the 1/eps_r of the medium
    (mxll_3D_mod.f90:826, :897 and the other faces), and with eps_r != 1
    inside the CPML that update grows without bound (max |Ez| 1e24 after 800
    steps of a 64x64x48 version on the oracle)."""
    rng = np.random.default_rng(SEED_BASE + seed_no)
    nx, ny, nz = grid
    med = dl_sphere_poles(rng)
    k_sub = nz // 4
    shapes = []
    for _ in range(n_ell):
        ax = rng.uniform(5.0, 25.0, size=3) / dx_nm
        ax = np.minimum(ax, (nz - k_sub - npml - 6) / 2.0)
        amax = float(ax.max())
        c = [float(rng.uniform(npml + amax + 2, nx - npml - amax - 2)),
             float(rng.uniform(npml + amax + 2, ny - npml - amax - 2)),
             float(rng.uniform(k_sub + amax + 2, nz - npml - amax - 2))]
        shapes.append((2, 1, c + ax.tolist() + random_rotation(rng)))
    shapes.append((3, 2, [npml + 1, npml + 1, npml + 1, nx - npml, ny - npml, k_sub]))
    src_cell = [nx // 2, ny // 2, int(0.6 * nz)]
    pos_nm = [index_to_nm(src_cell[a], dx_nm, grid[a]) for a in range(3)]
    # one-cell soft z dipole (its near field is kept in FP64 by the engine's
    # near-field patch on FP32 grids)
    srcs = [("z", 1.0, 2.066, 3.0, 1.0, pos_nm[0], pos_nm[1], pos_nm[2], 0.0, 0.0, 1.0e6, 0.0)]
    det = rng.integers([npml + 2, npml + 2, npml + 2], [nx - npml - 2, ny - npml - 2, nz - npml - 2], size=(32, 3))
    return Case(name="ellipsoids", grid=grid, dx_nm=dx_nm, npml=npml, boundaries=(BC_CPML,) * 3, steps=steps,
                record_every=record_every, media=[med, ("dielectric", 2.25)], shapes=shapes, sources=srcs,
                detectors=point_detectors(det.tolist(), ("Ez",)), info={"substrate_k": k_sub})


def config_emitters(grid=(512, 512, 384), dx_nm=2.0, npml=12, steps=20000, n_emit=4096, seed_no=5,
                    record_every=100, slab_nm=40.0):
    """Silver (library Ag) slab with a seeded nanohole array, 4096 two-level
    emitters (transition 2.0-2.2 eV, dipole 5-15 D, dephasing 10-50 meV) at
    seeded cells in a 20 nm layer above the slab, local-dipole coupling; a
    soft z-dipole source above the layer excites them.  The slab stops where
    the x and y CPML begin (a medium inside the CPML is unstable under the
    reference's CPML update, see config_ellipsoids)."""
    rng = np.random.default_rng(SEED_BASE + seed_no)
    nx, ny, nz = grid
    t = int(round(slab_nm / dx_nm))
    k0 = nz // 2 - t
    k1 = k0 + t - 1
    shapes = [(3, 1, [npml + 1, npml + 1, k0, nx - npml, ny - npml, k1])]
    pitch = float(rng.uniform(80.0, 120.0)) / dx_nm
    nh = int((nx - 2 * npml) // pitch)
    for a in range(nh):
        for b in range(nh):
            r = float(rng.uniform(10.0, 20.0)) / dx_nm
            cx = npml + (a + 0.5) * pitch
            cy = npml + (b + 0.5) * pitch
            shapes.append((4, 0, [cx, cy, r, k0, k1]))
    layer = int(round(20.0 / dx_nm))
    cells = np.stack([rng.integers(npml + 2, nx - npml - 2, size=n_emit),
                      rng.integers(npml + 2, ny - npml - 2, size=n_emit),
                      rng.integers(k1 + 1, k1 + 1 + layer, size=n_emit)], axis=1).astype(np.int32)
    pol = rng.normal(size=(n_emit, 3))
    pol /= np.linalg.norm(pol, axis=1)[:, None]
    tl = dict(omega0=rng.uniform(2.0, 2.2, n_emit) * EV_TO_AU,
              mu=rng.uniform(5.0, 15.0, n_emit) * DEBYE_TO_AU,
              pol=pol, gamma2=rng.uniform(10e-3, 50e-3, n_emit) * EV_TO_AU)
    src_cell = [nx // 2, ny // 2, min(nz - npml - 4, k1 + layer + 20)]
    pos_nm = [index_to_nm(src_cell[a], dx_nm, grid[a]) for a in range(3)]
    srcs = [("z", 1.0, 2.1, 3.0, 1.0, pos_nm[0], pos_nm[1], pos_nm[2], 0.0, 0.0, 1.0e6, 0.0)]   # one cell
    det = cells[:8].tolist()
    return Case(name="emitters", grid=grid, dx_nm=dx_nm, npml=npml, boundaries=(BC_CPML,) * 3, steps=steps,
                record_every=record_every, media=[("drude-lorentz", "Ag", 1.0)], shapes=shapes, sources=srcs,
                detectors=point_detectors(det), q_sites=cells, q_density=1.0e-3 / NM_TO_AU ** 3,
                two_level=tl, info={"slab_k": [k0, k1], "holes": len(shapes) - 1, "pitch_cells": pitch})
