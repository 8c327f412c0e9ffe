"""Near-field patches of FP32 grids: FP64 copies of E and H in boxes around
point sources.  Checked here: where the boxes go (clipped outside the CPML,
merged when they touch, none for FP64 grids, TF/SF grids or radius 0), that
they bring a one-cell source's far-field FP32 series within relL2 1e-5 of the
FP64 oracle where plain FP32 is several times further off, in both step
modes, and that fields written from the host reach the FP64 copies."""
import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G
from omx_fdtd.units import BC_CPML, BC_CLOSE, EV_TO_AU, FS_TO_AU, index_to_nm
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def src_line(cell, grid, dx, direction="z", freq=2.066, t0=1.0, tau=0.35, r_nm=0.0):
    pos = [index_to_nm(cell[a], dx, grid[a]) for a in range(3)]
    return (direction, 1.0, freq, t0, tau, pos[0], pos[1], pos[2], r_nm, 0.0, 1.0e6, 0.0)


def point_case(n=80, dx=1.0, npml=8, steps=1300, cells=None, bc=(BC_CPML,) * 3):
    """Vacuum cube, dx 1 nm (600 nm carrier: the near field is ~1e4 times the
    radiated field one cell from the source), one-cell z sources, point
    detectors 18-26 cells from the centre."""
    grid = (n, n, n)
    cells = cells or [(n // 2, n // 2, n // 2)]
    c0 = np.array((n // 2, n // 2, n // 2))
    offs = ((20, 3, 0), (-5, 22, 4), (6, -18, 15), (0, 0, 25), (-24, -2, -9))
    dets = [tuple(int(v) for v in c0 + d) for d in offs]
    dets = [d for d in dets if all(npml + 1 <= v <= n - npml for v in d)]
    return cases.Case(name="nearfield", grid=grid, dx_nm=dx, npml=npml, boundaries=bc, steps=steps,
                      record_every=3, sources=[src_line(c, grid, dx) for c in cells],
                      detectors=cases.point_detectors(dets, ("Ex", "Ez")))


def engine(case, precision, radius=None, mode=1):
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    if radius is not None:
        g.set_source_patch(radius)
    g.set_step_mode(mode)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    return g


def oracle_series(case):
    dr, dt = case.dr_dt
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    want = o.run(1, case.steps, case.record_every)
    o.kill()
    return want


def test_patch_boxes():
    print("[patch] box construction", flush=True)
    case = point_case(n=60, npml=8, cells=[(30, 30, 30), (36, 31, 29)])
    g = engine(case, 4)
    assert g.source_patch()[0] == 16             # the default radius
    g.kill()
    g = engine(case, 4, radius=8)
    r, nb, nc = g.source_patch()
    assert (r, nb) == (8, 1)                     # two touching boxes merge
    assert nc == (36 + 8 - 22 + 1) * (31 + 8 - 22 + 1) * (30 + 8 - 21 + 1)
    g.kill()
    case = point_case(n=80, npml=8, cells=[(20, 20, 20), (60, 60, 60)])
    g = engine(case, 4, radius=5)
    assert g.source_patch() == (5, 2, 2 * 11 ** 3)
    g.kill()
    # clipped to the cells outside the CPML faces (9..n-8 with npml 8)
    case = point_case(n=60, npml=8, cells=[(12, 30, 50)])
    g = engine(case, 4, radius=8)
    assert g.source_patch() == (8, 1, (20 - 9 + 1) * 17 * (52 - 42 + 1))
    g.kill()
    # a source inside the CPML gets no box
    case = point_case(n=60, npml=8, cells=[(4, 30, 30)])
    g = engine(case, 4, radius=3)
    assert g.source_patch()[1:] == (0, 0)
    g.kill()
    for prec, radius in ((8, None), (4, 0)):
        g = engine(point_case(n=40, npml=6), prec, radius)
        assert g.source_patch()[1:] == (0, 0)
        g.kill()
    c = point_case(n=48, npml=6)
    c.tfsf = dict(lo=(12, 12, 12), hi=(36, 36, 36), e0=1.0, w0=2.066 * EV_TO_AU, t0=0.6 * FS_TO_AU, tau=0.2 * FS_TO_AU)
    g = engine(c, 4)
    assert g.source_patch()[1:] == (0, 0)
    g.kill()


@pytest.mark.parametrize("mode", [1, 0])
def test_patch_brings_point_source_within_tolerance(mode):
    case = point_case()
    print(f"[patch] {case.grid} dx=1 nm one-cell source, {case.steps} steps, mode {mode}", flush=True)
    want = oracle_series(case)
    assert np.abs(want).max() > 0
    res = {}
    for radius in (0, 8, 16):
        g = engine(case, 4, radius, mode)
        g.run(1, case.steps, case.record_every)
        if mode == 0:
            assert g.last_run_fused()
        res[radius] = rel_l2(g.read_records(), want)
        g.kill()
    print(f"[patch] FP32 vs oracle relL2: plain {res[0]:.3e}, r=8 {res[8]:.3e}, r=16 {res[16]:.3e}", flush=True)
    assert res[8] <= TOL32 and res[16] <= TOL32
    assert res[0] >= 5 * res[8]


def test_patch_follows_fields_written_from_host():
    case = point_case(n=56, npml=6, steps=240)
    dr, dt = case.dr_dt
    rng = np.random.default_rng(7)
    init = {}
    for q, name in enumerate(G.FIELD_NAMES):
        shp = G.field_shape(q + 1, *case.grid)
        a = np.zeros(shp)
        a[2:-2, 2:-2, 2:-2] = rng.uniform(-1e-3, 1e-3, size=tuple(v - 4 for v in shp))
        init[name] = a if name[0] == "E" else a / 137.0
    g = engine(case, 4)
    g.run(1, 10, 0)                    # shadows in use, then overwritten from the host
    for name in G.FIELD_NAMES:
        g.set_field(name, init[name])
    g.run(11, case.steps, case.record_every)
    got = g.read_records()
    g.kill()
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium)
    for name in G.FIELD_NAMES:
        o.set_field(name, init[name])
    want = o.run(11, case.steps, case.record_every)
    o.kill()
    r = rel_l2(got, want)
    print(f"[patch] after host writes: FP32 vs oracle relL2 {r:.3e}", flush=True)
    assert r <= TOL32


@pytest.mark.parametrize("mode", [1, 0])
def test_patch_over_dispersive_cells(mode):
    """The source inside a Drude-Lorentz sphere (library Ag) next to a
    dielectric block, both reaching into the box: their cells in the box
    update in FP64 too (the ADE state of the box's dispersive cells is kept
    in FP64)."""
    case = point_case(n=72, npml=8, steps=1100)
    c = case.grid[0] // 2
    case.media = [("drude-lorentz", "Ag", 1.0), ("dielectric", 2.25)]
    case.shapes = [(1, 1, [c + 2.0, c - 1.0, c + 1.0, 5.5]), (3, 2, [c - 9, c - 9, c - 12, c + 9, c + 9, c - 6])]
    print(f"[patch] source in Ag, dielectric block, {case.grid}, mode {mode}", flush=True)
    want = oracle_series(case)
    res = {}
    for radius in (0, 8):
        g = engine(case, 4, radius, mode)
        g.run(1, case.steps, case.record_every)
        res[radius] = rel_l2(g.read_records(), want)
        g.kill()
    print(f"[patch] FP32 vs oracle relL2: plain {res[0]:.3e}, patch r=8 {res[8]:.3e}", flush=True)
    assert res[8] <= TOL32
    assert res[0] >= 2 * res[8]
