"""The oracle (C restatement) against golden vectors produced by the
reference's own Fortran modules (tests/golden/make_golden.py)."""
import numpy as np
import pytest

import _golden
from omx_fdtd import cases, units
from oracle import oracle


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def build(name):
    desc, fx = _golden.load(name)
    case = _golden.to_case(desc, fx)
    dr, dt = case.dr_dt
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    return desc, fx, case, o


@pytest.mark.parametrize("name", _golden.NAMES)
def test_tables_match_reference(name):
    desc, fx, case, o = build(name)
    t = o.cpml_tables()
    den = np.concatenate([t[k] for k in ("den_ex", "den_ey", "den_ez", "den_hx", "den_hy", "den_hz")])
    np.testing.assert_array_equal(den, fx["den"])
    if "be" in fx:
        for k in ("be", "ce", "bh", "ch"):
            np.testing.assert_array_equal(t[k], fx[k], err_msg=k)
    if desc["media"]:
        np.testing.assert_array_equal(o.media_map(), fx["media_map"])
        raw = fx["media_coeffs"]
        p = 0
        for m, (head, params) in enumerate(desc["media"], start=1):
            c = o.medium_coeffs(m)
            six = raw[p:p + 6]
            p += 6
            if head == "dielectric":
                continue          # the reference leaves these unset for dielectrics
            np.testing.assert_array_equal([c[k] for k in ("A1", "A2", "C1")], six[:3], err_msg=head)
            if head == "drude":
                np.testing.assert_array_equal([c["C3"], c["C4"]], six[4:6], err_msg=head)
            if head == "drude-lorentz":
                np.testing.assert_array_equal(six[3:6], [c["C2"], c["C3"], c["C4"]], err_msg=head)
                lib = oracle.library_medium(params.split()[0], float(params.split()[1]))
                P = lib.n_poles
                blk = raw[p:p + 2 + 8 * P]
                p += 2 + 8 * P
                assert blk[0] == lib.omega_d and blk[1] == lib.gamma_d
                np.testing.assert_array_equal(blk[2:2 + P], list(lib.omega_k)[:P])
                np.testing.assert_array_equal(blk[2 + P:2 + 2 * P], list(lib.gamma_k)[:P])
                np.testing.assert_array_equal(blk[2 + 2 * P:2 + 3 * P], list(lib.a_k)[:P])
                for q, k in enumerate(("alpha_k", "beta_k", "gamma_k", "B1_k", "B2_k")):
                    np.testing.assert_array_equal(blk[2 + (3 + q) * P:2 + (4 + q) * P], c[k][:P], err_msg=k)


@pytest.mark.parametrize("name", _golden.NAMES)
def test_time_series_match_reference(name):
    desc, fx, case, o = build(name)
    series, es = _golden.drive(o, desc, case)
    assert series.shape == fx["series"].shape
    r = rel_l2(series, fx["series"])
    print(f"{name}: oracle vs reference relL2 = {r:.3e}")
    assert r <= 1e-13
    if es is not None:
        assert rel_l2(es, fx["qsites_e"]) <= 1e-13


def test_quantum_move_schedule_matches_reference():
    """The steps on which the reference driver loop moved the quantum sites
    (recorded by the reference build) against the schedule the tests drive."""
    for name in ("qsites", "qsites_skip3"):
        desc, fx = _golden.load(name)
        want = fx["qmoves"].tolist()
        assert _golden.q_moves(desc.get("n_skip", 1), desc["steps"]) == want, name
    assert _golden.q_moves(3, 10) == [1, 3, 6, 9]


def test_oracle_run_loop_moves_emitters_every_n_skip_steps():
    """orc_run's own sub-step bookkeeping against the schedule driven call by
    call: two-level emitters with dt_q = 3 dt."""
    desc, fx = _golden.load("qsites")
    case = _golden.to_case(desc, fx)
    dr, dt = case.dr_dt
    ijk = np.array([q[:3] for q in desc["qsites"][1:]], dtype=np.int32)
    n = len(ijk)
    tl = dict(omega0=np.full(n, 0.08), mu=np.linspace(2.0, 4.0, n), pol=np.eye(3)[np.arange(n) % 3] + 0.1,
              gamma2=np.full(n, 1e-3))
    res = []
    for manual in (False, True):
        o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
        cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                    rasterize=oracle.rasterize)
        o.set_q_sites(ijk, 1.0e-3 / units.NM_TO_AU ** 3, 3 * dt)
        o.set_two_level(tl["omega0"], tl["mu"], tl["pol"], tl["gamma2"])
        if manual:
            moves = set(_golden.q_moves(3, case.steps))
            rows = []
            for tt in range(1, case.steps + 1):
                o.time = dt * (tt - 0.5)
                if tt in moves:
                    o.t_skip = o.time
                o.td_propagate_H_field()
                o.inject_sources(o.time)
                o.td_propagate_E_field(tt)
                if tt in moves:
                    o.two_level_step()
                rows.append(o.sample())
            res.append(np.array(rows))
        else:
            res.append(o.run(1, case.steps, 1))
        o.kill()
    assert np.abs(res[0]).max() > 0
    np.testing.assert_array_equal(res[0], res[1])


def test_planted_mistake_is_caught():
    """A one-sign change in the oracle's E update must fail the golden check:
    run the oracle with the wrong sign of dt_eps0 on a copy of the case."""
    desc, fx, case, o = build("media_mix")
    case2 = _golden.to_case(desc, fx)
    dr, dt = case2.dr_dt
    bad = oracle.OracleGrid(case2.grid, case2.npml, case2.boundaries, dt * (1 + 1e-9), dr)
    cases.apply(case2, bad, library=oracle.library_medium, make_medium=oracle.make_medium)
    series, _ = _golden.drive(bad, desc, case2)
    assert rel_l2(series, fx["series"]) > 1e-9
