"""CPU tests: C-ABI exports, host-side set-up against the oracle, and the
oracle's two-level emitter against the mathematics (no reference code
exists for it)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

import omx_fdtd
from omx_fdtd import cases, host, units
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "omx_fdtd.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(omx\w+)\s*\(", hdr, re.M)))
    assert len(declared) >= 30
    lib = omx_fdtd.load_library()
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(omx_fdtd._lib.EXPORTED) == declared
    assert lib.omx_version() == 1


def test_constants_are_the_reference_single_roundings():
    assert units.NM_TO_AU == 18.897260665893555
    assert units.EV_TO_AU == 0.036749303340911865
    assert units.AU_TO_EV == 27.21139907836914
    assert units.DEBYE_TO_AU == 0.3931847810745239
    dr, dt = units.grid_dr_dt(5.0)
    assert dr == 5.0 * 18.897260665893555
    assert dt == dr / 274.0


@pytest.mark.parametrize("n_ker", [0, 1, 2, 4])
def test_source_kernel_matches_oracle(n_ker):
    dr = 2.0 * units.NM_TO_AU
    a = host.source_kernel(n_ker, dr)
    b = oracle.source_kernel(n_ker, dr)
    np.testing.assert_allclose(a, b, rtol=1e-15, atol=1e-17)
    assert abs(a.sum() - 1.0) < 1e-14
    if n_ker == 0:
        assert a.ravel()[0] == 1.0


def test_source_indices_match_oracle():
    dr = 5.0 * units.NM_TO_AU
    for r0 in ((0.0, 0.0, 0.0), (12.4, -7.6, 3.1), (-0.2, 0.2, -11.0)):
        r0au = np.array(r0) * units.NM_TO_AU
        a = host.source_indices(2, r0au, dr, (40, 36, 30))
        b = oracle.source_indices(2, r0au, dr, (40, 36, 30))
        np.testing.assert_array_equal(a, b)


def test_index_convention_roundtrip():
    for dx in (1.0, 2.0, 5.0):
        dr = dx * units.NM_TO_AU
        for n in (63, 64, 256):
            for idx in range(1, n + 1):
                assert units.nm_to_index_dr(units.index_to_nm(idx, dx, n), dr, n) == idx


def _tls_oracle(E0, omega0, mu, gamma2, pol, nsteps, dt):
    n = (6, 6, 6)
    o = oracle.OracleGrid(n, 2, (1, 1, 1), dt, 1.0)
    o.set_q_sites(np.array([[3, 3, 3]]), 1.0, dt)
    o.set_two_level([omega0], [mu], np.array([pol]), [gamma2])
    for name, comp in (("Ex", 0), ("Ey", 1), ("Ez", 2)):
        f = o.get_field(name)
        o.set_field(name, np.full(f.shape, E0[comp]))
    dips = []
    for _ in range(nsteps):
        o.two_level_step()
        dips.append(o.two_level_state()[0][0].copy())
    return o.two_level_state()[1][0], np.array(dips)


def test_two_level_rabi_closed_form():
    """Constant field, no dephasing: rho(t) = U rho0 U^+ with U = exp(-iHt)."""
    omega0, mu, dt = 2.1 * units.EV_TO_AU, 10.0 * units.DEBYE_TO_AU, 0.3448
    pol = np.array([0.0, 0.6, 0.8])
    E0 = np.array([0.0, 0.0, 2e-3])
    n = 800
    rho, dips = _tls_oracle(E0, omega0, mu, 0.0, pol, n, dt)
    V = -mu * float(E0 @ pol)
    H = np.array([[-omega0 / 2, V], [V, omega0 / 2]], dtype=complex)  # basis (g, e)
    U = expm(-1j * H * dt * n)
    r = U @ np.diag([1.0, 0.0]).astype(complex) @ U.conj().T
    assert abs(rho[0] - r[0, 0].real) < 1e-7
    assert abs(rho[1] - r[1, 1].real) < 1e-7
    assert abs(rho[2] - r[1, 0].real) < 1e-7
    assert abs(rho[3] - r[1, 0].imag) < 1e-7
    assert abs(rho[0] + rho[1] - 1.0) < 1e-12
    # dipole along the polarisation with magnitude 2 mu Re(rho_eg)
    np.testing.assert_allclose(dips[-1], 2 * mu * r[1, 0].real * pol, atol=1e-7 * mu)


def test_two_level_dephasing_against_master_equation():
    """Constant field with dephasing: compare with the Lindblad equation in
    matrix form, L[rho] = -i[H, rho] + gamma2 (P_e rho P_g + P_g rho P_e)*(-1)
    (pure dephasing of the coherences), integrated by scipy."""
    omega0, mu, g2, dt = 2.0 * units.EV_TO_AU, 8.0 * units.DEBYE_TO_AU, 0.03 * units.EV_TO_AU, 0.3448
    pol = np.array([1.0, 0.0, 0.0])
    E0 = np.array([5e-3, 0.0, 0.0])
    n = 600
    rho, _ = _tls_oracle(E0, omega0, mu, g2, pol, n, dt)
    V = -mu * E0[0]
    H = np.array([[-omega0 / 2, V], [V, omega0 / 2]], dtype=complex)

    def rhs(t, y):
        r = y.reshape(2, 2)
        d = -1j * (H @ r - r @ H)
        d[0, 1] -= g2 * r[0, 1]
        d[1, 0] -= g2 * r[1, 0]
        return d.ravel()

    sol = solve_ivp(rhs, (0, n * dt), np.array([1, 0, 0, 0], dtype=complex), rtol=1e-11, atol=1e-13)
    r = sol.y[:, -1].reshape(2, 2)
    assert abs(rho[1] - r[1, 1].real) < 1e-7
    assert abs(rho[2] - r[1, 0].real) < 1e-7
    assert abs(rho[3] - r[1, 0].imag) < 1e-7


def test_engine_errors_are_reported():
    d = omx_fdtd._lib.Desc()
    d.nx, d.ny, d.nz, d.npml = 10, 10, 10, 8
    d.boundaries[0] = d.boundaries[1] = d.boundaries[2] = 3
    d.precision, d.dt, d.dr = 8, 0.1, 1.0
    h = C.c_void_p()
    rc = omx_fdtd.load_library().omx3d_create(C.byref(d), C.byref(h))
    assert rc == 1
    assert b"CPML" in omx_fdtd.load_library().omx_last_error()
    d.npml, d.precision = 2, 3
    assert omx_fdtd.load_library().omx3d_create(C.byref(d), C.byref(h)) == 1
    with pytest.raises(omx_fdtd.OmxError):
        omx_fdtd.Grid3D((10, 10, 10), 8, (3, 3, 3), 0.1, 1.0)


def test_workload_media_stay_out_of_the_cpml():
    """The reference adds dt/eps0 * psi to E inside the CPML whatever the
    medium (mxll_3D_mod.f90:826), which grows without bound where a medium
    with eps_r != 1 reaches into the CPML.  The workload builders keep their
    substrate and slab out of it; on the oracle the scaled ellipsoid case
    stays bounded, and the same case with the substrate reaching into the
    CPML does not."""
    import dataclasses
    from oracle import oracle as O
    base = cases.config_ellipsoids(grid=(48, 48, 40), n_ell=4, npml=6, steps=600, record_every=0)
    nx, ny, nz = base.grid
    sub = [s for s in base.shapes if s[0] == 3][0]
    assert sub[2][:3] == [base.npml + 1] * 3 and sub[2][3:5] == [nx - base.npml, ny - base.npml]
    slab = cases.config_emitters(grid=(64, 64, 48), n_emit=4, npml=6, steps=10).shapes[0]
    assert slab[2][:2] == [7, 7] and slab[2][3:5] == [58, 58]
    into = dataclasses.replace(base, shapes=[(3, 2, [1, 1, 1, nx, ny, base.info["substrate_k"]])])
    peaks = []
    for c in (base, into):
        dr, dt = c.dr_dt
        g = O.OracleGrid(c.grid, c.npml, c.boundaries, dt, dr)
        cases.apply(c, g, library=O.library_medium, make_medium=O.make_medium, rasterize=O.rasterize)
        g.run(1, 600, 0)
        peaks.append(float(np.abs(g.get_field("Ez")).max()))
        g.kill()
    print(f"max |Ez| after 600 steps: substrate outside the CPML {peaks[0]:.3e}, inside {peaks[1]:.3e}")
    assert peaks[0] < 1e-3
    assert peaks[1] > 1e3
