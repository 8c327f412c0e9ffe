"""The 2D and 1D solvers of the HIP engine (omxl_* entry points, FP64)
against the golden vectors of the reference's own mxll_2D_mod / mxll_1D_mod
builds and against the oracle: detector series and whole-field snapshots
within relL2 1e-12 (the engine contracts a*b+c into fused multiply-adds,
the reference does not), the FULL mode as both families at once, field
copies, and the error paths."""
import numpy as np
import pytest

import _golden_lowdim as G
from omx_fdtd import Grid1D, Grid2D, OmxError, grid_medium, library_medium, lowdim
from omx_fdtd.units import grid_dr_dt
from oracle import oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def engine(desc):
    dr, dt = G.dr_dt(desc)
    if desc["dims"] == 1:
        return Grid1D(desc["grid"], desc["npml"], desc["bc"], dt, dr)
    return Grid2D(desc["grid"], desc["npml"], desc["bc"], dt, dr, desc["mode"])


@pytest.mark.parametrize("name", G.NAMES)
def test_engine_matches_reference(name):
    desc, fx = G.load(name)
    print(f"[lowdim] {name} grid={desc['grid']} dims={desc['dims']}", flush=True)
    g = G.setup(engine(desc), desc, fx, library_medium, grid_medium)
    series, snaps = G.drive(g, desc)
    g.kill()
    r, rf = rel_l2(series, fx["series"]), rel_l2(snaps, fx["fields"])
    print(f"{name}: engine vs reference relL2 series {r:.3e} fields {rf:.3e}")
    assert series.shape == fx["series"].shape and snaps.shape == fx["fields"].shape
    assert r <= TOL and rf <= TOL


def test_full_mode_runs_both_families():
    """mode 3 carries TMz and TEz fields together; each family matches the
    oracle (and so the reference's own TMz / TEz runs)."""
    desc, fx = G.load("tmz2d")
    dr, dt = G.dr_dt(desc)
    src = desc["sources"] + ["x 1.0 1.8 0.5 0.2 -30.0 10.0 0.0 7.5 0.0 1.0e6 0.0",
                             "y 0.5 2.2 0.6 0.3 12.0 -8.0 0.0 2.5 0.0 1.0e6 0.4"]
    full = dict(desc, mode=3, sources=src, steps=120, field_every=60,
                detectors=desc["detectors"] + [[f, [30, 26]] for f in ("Ex", "Ey", "Hz")])
    g = G.setup(Grid2D(desc["grid"], desc["npml"], desc["bc"], dt, dr, 3), full, fx, library_medium, grid_medium)
    o = G.setup(oracle.OracleLowGrid(2, desc["grid"], desc["npml"], desc["bc"], dt, dr, 3), full, fx,
                oracle.library_medium, oracle.make_medium)
    s_g = g.run(1, full["steps"], 1)
    s_o = o.run(1, full["steps"], 1)
    assert rel_l2(s_g, s_o) <= TOL
    for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        a, b = g.get_field(f), o.get_field(f)
        assert np.abs(b).max() > 0, f
        assert rel_l2(a, b) <= TOL, f
    g.kill()


def test_field_copy_round_trip_and_step_from_state():
    """set_field / get_field keep the reference bounds; stepping from a
    planted state matches the oracle stepped from the same state."""
    rng = np.random.default_rng(7)
    dr, dt = grid_dr_dt(5.0)
    for dims, mode in ((2, 1), (2, 2), (1, 0)):
        grid = (33, 29) if dims == 2 else (91,)
        bc = (3, 2) if dims == 2 else (3,)
        g = Grid2D(grid, 5, bc, dt, dr, mode) if dims == 2 else Grid1D(grid, 5, bc, dt, dr)
        o = oracle.OracleLowGrid(dims, grid, 5, bc, dt, dr, mode)
        names = {1: ("Ez", "Hx", "Hy"), 2: ("Ex", "Ey", "Hz"), 0: ("Ex", "Hy")}[mode]
        for f in names:
            a = rng.standard_normal(g.get_field(f).shape) * 1e-3
            g.set_field(f, a)
            o.set_field(f, a)
            np.testing.assert_array_equal(g.get_field(f), a)
        g.run(1, 25)
        o.run(1, 25)
        for f in names:
            assert rel_l2(g.get_field(f), o.get_field(f)) <= TOL, (dims, mode, f)
        g.kill()


def test_error_paths():
    dr, dt = grid_dr_dt(5.0)
    with pytest.raises(OmxError):
        lowdim._LowGrid.__init__(object.__new__(Grid2D), (10, 10), 2, (3, 3), dt, dr, 7)
    with pytest.raises(OmxError):
        Grid2D((10, 10), 6, (3, 3), dt, dr, 1)          # CPML thicker than half the grid
    g = Grid2D((20, 20), 4, (3, 3), dt, dr, 1)
    with pytest.raises(OmxError):
        g.add_detector("Ex", [(5, 5)])                 # no Ex in TMz
    with pytest.raises(OmxError):
        g.add_source("z", 1.0, 0.05, 10.0, 5.0, 0.0, 1e6, 0.0, 0, np.array([[25], [3]], dtype=np.int32).T,
                     np.ones(1))                        # outside the grid
    g.kill()
    g = Grid1D((50,), 4, (3,), dt, dr)
    with pytest.raises(OmxError):
        g.add_source("y", 1.0, 0.05, 10.0, 5.0, 0.0, 1e6, 0.0, 0, np.array([25], dtype=np.int32), np.ones(1))
    g.kill()
