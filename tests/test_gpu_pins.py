"""The engine against what the mathematics fixes, as the oracle is pinned in
test_cpu_pins.py: the eps_r-weighted volume energy against math.fsum, the
exact conservation of the discrete Yee energy in lossless boxes, and the
TF/SF total field inside the box against the incident line."""
import math

import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G
from omx_fdtd.units import BC_CLOSE, BC_CPML, BC_PERIODIC, grid_dr_dt
from oracle import oracle
from test_cpu_pins import EPS0, MU0, discrete_energy, interior, lossless_box, tfsf_case, tfsf_check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [8, 4])
def test_energy_with_media_matches_fsum(precision):
    """k_energy<T, true>: eps_r of each E component's medium (the staggered
    media map: dielectric 2.25, a Drude-Lorentz sphere with eps_inf 3, a Drude
    block with eps_inf 1.5)."""
    n = (20, 18, 16)
    dr, dt = grid_dr_dt(5.0)
    g = omx_fdtd.Grid3D(n, 4, (BC_CPML,) * 3, dt, dr, precision=precision)
    c = cases.Case(name="w", grid=n, dx_nm=5.0, npml=4, boundaries=(BC_CPML,) * 3, steps=0,
                   media=[("dielectric", 2.25), ("drude-lorentz", "Ag", 3.0), ("drude", 9.0, 0.07, 1.5)],
                   shapes=[(3, 1, [2, 2, 2, 10, 9, 6]), (1, 2, [12.0, 10.0, 9.0, 4.0]), (3, 3, [14, 3, 10, 18, 7, 14])])
    cases.apply(c, g, library=G.library_medium, make_medium=G.medium)
    eps_of = np.array([1.0, 2.25, 3.0, 1.5])
    rng = np.random.default_rng(11)
    terms = []
    mm = g.media_map()
    assert set(np.unique(mm)) == {0, 1, 2, 3}
    for q, name in enumerate(G.FIELD_NAMES):
        v = rng.uniform(-1, 1, size=g.get_field(name).shape) * (1.0 if q < 3 else 1.0 / 137.0)
        if precision == 4:
            v = v.astype(np.float32).astype(np.float64)
        g.set_field(name, v)
        core = interior(v, q, n)
        if q < 3:
            terms.extend((EPS0 * eps_of[mm[..., q]] * core ** 2).ravel().tolist())
        else:
            terms.extend((MU0 * core ** 2).ravel().tolist())
    want = 0.5 * dr ** 3 * math.fsum(terms)
    w64, w32 = g.energy()
    g.kill()
    print(f"FP{8 * precision}: energy w64 rel {abs(w64 - want) / want:.2e}, w32c rel {abs(w32 - want) / want:.2e}")
    assert abs(w64 - want) <= 1e-12 * want
    assert abs(w32 - want) <= 1e-6 * want


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_CLOSE])
def test_lossless_box_conserves_discrete_energy(bc, precision):
    n = (16, 14, 12)
    dr, dt = grid_dr_dt(5.0)
    g = omx_fdtd.Grid3D(n, 4, (bc,) * 3, dt, dr, precision=precision)
    epsr = lossless_box(g, n, bc)
    w0 = discrete_energy(g, n, dr, epsr)
    g.run(1, 300, 0)
    w1 = discrete_energy(g, n, dr, epsr)
    g.kill()
    print(f"bc {bc} FP{8 * precision}: W' drift over 300 steps {(w1 - w0) / w0:.2e}")
    assert abs(w1 - w0) <= (1e-12 if precision == 8 else 1e-6) * w0


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("mode", [1, 0])
def test_tfsf_total_field_is_the_incident_line(mode, precision):
    n, npml, lo, hi, tf = tfsf_case()
    dr, dt = grid_dr_dt(2.0)
    g = omx_fdtd.Grid3D(n, npml, (BC_CPML,) * 3, dt, dr, precision=precision)
    g.set_step_mode(mode)
    g.set_tfsf(lo, hi, tf["e0"], tf["w0"], tf["t0"], tf["tau"], 200)
    g.run(1, 150, 0)
    assert g.last_run_fused() == (mode == 0)
    tfsf_check(g, n, lo, hi, 1e-12 if precision == 8 else 1e-5)
    # the engine's incident line is the oracle's
    o = oracle.OracleGrid(n, npml, (BC_CPML,) * 3, dt, dr)
    o.set_tfsf(lo, hi, tf["e0"], tf["w0"], tf["t0"], tf["tau"], 200)
    o.run(1, 150, 0)
    kb, ex1, hy1 = g.tfsf_line()
    kbo, ex1o, hy1o = o.tfsf_line()
    o.kill()
    g.kill()
    assert kb == kbo
    np.testing.assert_allclose(ex1, ex1o, rtol=0, atol=1e-13 * np.abs(ex1o).max())
    np.testing.assert_allclose(hy1, hy1o, rtol=0, atol=1e-13 * np.abs(hy1o).max())
