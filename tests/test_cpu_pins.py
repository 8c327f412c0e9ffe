"""Pins of two oracle paths against what the mathematics fixes (CPU only).

Discrete energy: the Yee update conserves W' = 1/2 dr^3 sum(eps0 eps_r
E^n.E^n + mu0 H^{n-1/2}.H^{n+1/2}) exactly in a lossless box (periodic or
closed, dielectric included), because the E->H curl is minus the adjoint of
the H->E curl; a wrong sign, index or coefficient in either breaks it.

TF/SF: the incident wave comes from a 1D Yee line with the 3D grid's dt and
dr, so in vacuum it is an exact solution of the 3D scheme; the total field
inside the box must equal the line (Ex at k, Hy at k+1/2), the transverse
components must vanish, and nothing may leak outside.
"""
import math

import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G
from omx_fdtd.units import BC_CLOSE, BC_CPML, BC_PERIODIC, EV_TO_AU, FS_TO_AU, grid_dr_dt
from oracle import oracle

EPS0, MU0 = omx_fdtd.EPS0, omx_fdtd.MU0


def interior(v, q, n):
    """Cells 1..n of reference array q (Ex, Ey, Ez, Hx, Hy, Hz bounds)."""
    lo = [0 if (q < 3 and a == q) else 1 for a in range(3)] if q < 3 else [1 if a == q - 3 else 0 for a in range(3)]
    return v[tuple(slice(1 - lo[a], 1 - lo[a] + n[a]) for a in range(3))]


def lossless_box(g, n, bc, seed=1):
    """A dielectric block (eps_r 2.25) and random fields on the cells 1..n."""
    c = cases.Case(name="box", grid=n, dx_nm=5.0, npml=4, boundaries=(bc,) * 3, steps=0,
                   media=[("dielectric", 2.25)], shapes=[(3, 1, [3, 3, 3, 9, 8, 7])])
    mod = oracle if isinstance(g, oracle.OracleGrid) else G
    cases.apply(c, g, library=mod.library_medium if mod is oracle else G.library_medium,
                make_medium=oracle.make_medium if mod is oracle else G.medium,
                rasterize=oracle.rasterize if mod is oracle else None)
    rng = np.random.default_rng(seed)
    for q, name in enumerate(G.FIELD_NAMES):
        v = np.zeros_like(g.get_field(name))
        cv = interior(v, q, n)
        cv[...] = rng.uniform(-1, 1, size=cv.shape) * (1.0 if q < 3 else 1.0 / 137.0)
        g.set_field(name, v)
    epsr = np.ones(n + (3,))
    epsr[g.media_map() == 1] = 2.25
    return epsr


def discrete_energy(g, n, dr, epsr):
    """W' at the current step: one extra H half step gives H^{n+1/2}, then H
    is put back."""
    E = [interior(g.get_field(nm), q, n) for q, nm in enumerate(("Ex", "Ey", "Ez"))]
    H = [g.get_field(nm) for nm in ("Hx", "Hy", "Hz")]
    g.td_propagate_H_field()
    H1 = [interior(g.get_field(nm), q + 3, n) for q, nm in enumerate(("Hx", "Hy", "Hz"))]
    for nm, v in zip(("Hx", "Hy", "Hz"), H):
        g.set_field(nm, v)
    t = []
    for q in range(3):
        t.extend((EPS0 * epsr[..., q] * E[q] ** 2).ravel())
        t.extend((MU0 * interior(H[q], q + 3, n) * H1[q]).ravel())
    return 0.5 * dr ** 3 * math.fsum(t)


@pytest.mark.parametrize("bc", [BC_PERIODIC, BC_CLOSE])
def test_oracle_conserves_discrete_energy(bc):
    n = (16, 14, 12)
    dr, dt = grid_dr_dt(5.0)
    o = oracle.OracleGrid(n, 4, (bc,) * 3, dt, dr)
    epsr = lossless_box(o, n, bc)
    w0 = discrete_energy(o, n, dr, epsr)
    for tt in range(1, 301):
        o.td_propagate_H_field()
        o.td_propagate_E_field(tt)
    w1 = discrete_energy(o, n, dr, epsr)
    o.kill()
    print(f"bc {bc}: W' drift over 300 steps {(w1 - w0) / w0:.2e}")
    assert abs(w1 - w0) <= 1e-13 * w0


def tfsf_case():
    n, npml = (40, 36, 48), 6
    lo, hi = (12, 10, 12), (28, 26, 36)
    tf = dict(lo=lo, hi=hi, e0=1.0, w0=2.066 * EV_TO_AU, t0=0.6 * FS_TO_AU, tau=0.2 * FS_TO_AU)
    return n, npml, lo, hi, tf


def tfsf_check(g, n, lo, hi, tol):
    """Total field inside the box against the incident line; everything else
    (transverse components, the scattered region) against zero."""
    kb, ex1, hy1 = g.tfsf_line()
    (i0, j0, k0), (i1, j1, k1) = lo, hi
    F = {nm: g.get_field(nm) for nm in G.FIELD_NAMES}
    # reference bounds: Ex(0:nx,1:ny+1,1:nz+1), Hy(0:nx,1:ny,0:nz)
    ex = F["Ex"][i0:i1, j0 - 1:j1, k0 - 1:k1]
    hy = F["Hy"][i0:i1, j0 - 1:j1, k0:k1]
    scale = np.abs(ex1).max()
    assert scale > 0
    d_ex = np.abs(ex - ex1[k0 - kb:k1 + 1 - kb][None, None, :]).max() / scale
    d_hy = np.abs(hy - hy1[k0 - kb:k1 - kb][None, None, :]).max() / np.abs(hy1).max()
    mask = np.ones_like(F["Ex"], dtype=bool)
    mask[i0:i1, j0 - 1:j1, k0 - 1:k1] = False
    leak = np.abs(F["Ex"][mask]).max() / scale
    trans = max(np.abs(F[nm]).max() for nm in ("Ey", "Hx")) / scale
    print(f"TF/SF: Ex vs line {d_ex:.2e}, Hy vs line {d_hy:.2e}, leakage {leak:.2e}, Ey/Hx {trans:.2e}")
    assert d_ex <= tol and d_hy <= tol and leak <= tol and trans <= tol


def test_oracle_tfsf_total_field_is_the_incident_line():
    n, npml, lo, hi, tf = tfsf_case()
    dr, dt = grid_dr_dt(2.0)
    o = oracle.OracleGrid(n, npml, (BC_CPML,) * 3, dt, dr)
    o.set_tfsf(lo, hi, tf["e0"], tf["w0"], tf["t0"], tf["tau"], 200)
    o.run(1, 150, 0)
    tfsf_check(o, n, lo, hi, 1e-12)
    o.kill()
