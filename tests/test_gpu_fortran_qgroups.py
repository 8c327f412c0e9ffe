"""The Fortran driver on a quantum-group deck: mol_group_XXXXXXX.in files of
a "single" group (7^3 Blackman-Harris kernels) and a "material" group with
selected printing, two-level systems moving every second step (dt_q =
2 dt, q_group_mod.f90 / mxim_mxll.f90:66-132), against the FP64 oracle
driven with the same groups and emitters: detector series at 1e-12 and the
energies and dipoles of output_q_group_XXXX/energy_and_dipole.dat
(output_mod.f90:236-388) at every quantum print step."""
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import omx_fdtd  # noqa: F401  (puts the package's units on the path)
from omx_fdtd import cases
from omx_fdtd.units import BC_CPML, EV_TO_AU, FIELD_CODES, index_to_nm, grid_dr_dt
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

EXE = os.path.join(os.path.dirname(HERE), "one-mxll-to-rule-them-all_amd", "omx_fdtd_3d")
AU_TO_FS = 2.4188843265864e-2
NM_TO_AU = 18.897261259077823
AU_TO_EV = float(np.float32(27.2114))

pytestmark = pytest.mark.gpu

GRID, DX, NPML, STEPS = (24, 22, 20), 5.0, 4, 90
# (group, mol_id, id_file, cell, print_on); group 1 single, group 2 material
SYSTEMS = [(1, 7, 1, (9, 9, 9), True), (1, 8, 2, (11, 10, 10), True),
           (2, 21, 3, (14, 12, 10), True), (2, 22, 4, (10, 9, 10), False), (2, 23, 1, (16, 8, 12), True)]
# two_level_XXXXXXX.in per id_file: omega0 (eV), mu (au), pol, gamma2 (eV)
TLS = {1: (1.55, 2.0, 1.0, 0.2, 0.1, 0.01), 2: (1.60, 1.5, 0.0, 0.3, 1.0, 0.02),
       3: (1.50, 2.5, 0.5, 1.0, 0.0, 0.005), 4: (1.45, 1.0, 1.0, 1.0, 1.0, 0.0)}
DETS = [("point", "Ex", (12, 11, 10)), ("point", "Ez", (12, 11, 10)), ("point", "Ez", (8, 14, 12)),
        ("plane_xy", "Ez", (1, 1, 11), (24, 22, 11))]


def nm(cell):
    return [index_to_nm(cell[a], DX, GRID[a]) for a in range(3)]


def write_deck(d, dt):
    with open(os.path.join(d, "inp"), "w") as f:
        f.write("&OMxRTA\n mxll_boundaries = 'cpml', 'cpml', 'cpml'\n mxll_dimensions = 3\n")
        f.write(f" mxll_npml = {NPML}\n mxll_n_src = 1\n mxll_n_media = 0\n mxll_n_detectors = {len(DETS)}\n")
        f.write(" mxll_box_size = %r, %r, %r\n mxll_dr = %r\n" % (GRID[0] * DX, GRID[1] * DX, GRID[2] * DX, DX))
        f.write(" mxll_total_time = %r\n mxll_dt_det_print = %r\n" % ((STEPS + 0.5) * dt * AU_TO_FS, 1.5 * dt * AU_TO_FS))
        # dt_q is taken in a.u. as given (input_mod.f90:123): two Maxwell steps
        f.write(" mxll_n_q_groups = 2\n mxll_dt_q = %r\n mxll_dt_q_print = %r\n/\n" % (2.0 * dt, 3.5 * dt * AU_TO_FS))
    with open(os.path.join(d, "sources.in"), "w") as f:
        x, y, z = nm((12, 12, 12))
        f.write(f"z 1.0 1.549802 0.6 0.3 {x!r} {y!r} {z!r} 0.0 0.0 1.0e6 0.0\n")
    with open(os.path.join(d, "detectors.in"), "w") as f:
        for det in DETS:
            lo = det[2]
            hi = det[3] if len(det) > 3 else lo
            a, b = nm(lo), nm(hi)
            f.write(f"{det[0]} {det[1]} {a[0]!r} {b[0]!r} {a[1]!r} {b[1]!r} {a[2]!r} {b[2]!r}\n")
    for g, (kind, popt) in ((1, ("single", "print_all")), (2, ("material", "print_selected"))):
        sysg = [s for s in SYSTEMS if s[0] == g]
        with open(os.path.join(d, f"mol_group_{g:07d}.in"), "w") as f:
            f.write(f"{kind} two_level {len(sysg)} {popt}\n")
            for (_, mol, idf, cell, pon) in sysg:
                x, y, z = nm(cell)
                extra = (" print_on" if pon else " print_off") if popt == "print_selected" else ""
                f.write(f"{mol} {idf} {x!r} {y!r} {z!r}{extra}\n")
    for idf, v in TLS.items():
        with open(os.path.join(d, f"two_level_{idf:07d}.in"), "w") as f:
            f.write(" ".join(repr(x) for x in v) + "\n")


def parse_energy_dipole(path):
    """(time, [(mol_id, energy eV, dipole xyz)]) per print step (a list-
    directed record may wrap onto a second line)."""
    steps, tok = [], []
    for line in open(path):
        if line.startswith("# Time"):
            steps.append((float(line.split()[3]), []))
        elif not line.startswith("#") and line.strip():
            tok += line.split()
            while len(tok) >= 5:
                v, tok = tok[:5], tok[5:]
                steps[-1][1].append((int(v[0]), float(v[1]), [float(x) for x in v[2:5]]))
    assert not tok
    return steps


def test_fortran_driver_q_groups_match_oracle():
    assert os.path.exists(EXE), "omx_fdtd_3d not built"
    dr, dt = grid_dr_dt(DX)
    with tempfile.TemporaryDirectory() as d:
        write_deck(d, dt)
        print(f"[fortran-q] deck in {d}", flush=True)
        p = subprocess.run([EXE], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           env=dict(os.environ, OMX_FDTD_PRECISION="8"), timeout=300)
        print(p.stdout[-1500:])
        assert p.returncode == 0
        series = make_golden.parse_series(d, {"detectors": DETS}, STEPS)
        groups = [parse_energy_dipole(os.path.join(d, f"output_q_group_{g:04d}", "energy_and_dipole.dat"))
                  for g in (1, 2)]

    # the oracle with the same groups, emitters, source and detectors
    case = cases.Case(name="qgroups_deck", grid=GRID, dx_nm=DX, npml=NPML, boundaries=(BC_CPML,) * 3, steps=STEPS,
                      record_every=1, sources=[("z", 1.0, 1.549802, 0.6, 0.3, *nm((12, 12, 12)), 0.0, 0.0, 1.0e6, 0.0)],
                      detectors=cases.point_detectors([(12, 11, 10)], ("Ex", "Ez")) +
                      cases.point_detectors([(8, 14, 12)], ("Ez",)) +
                      [(5, FIELD_CODES["Ez"], cases.plane_xy(GRID, 11))])
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    ijk = np.array([s[3] for s in SYSTEMS], dtype=np.int32)
    o.set_q_groups([2, 1], [s[0] - 1 for s in SYSTEMS], ijk, 1.0e-3 / NM_TO_AU ** 3, 2.0 * dt)
    par = np.array([TLS[s[2]] for s in SYSTEMS])
    o.set_two_level(par[:, 0] * EV_TO_AU, par[:, 1], par[:, 2:5], par[:, 5] * EV_TO_AU)
    rows, states = [], []
    for tt in range(1, STEPS + 1):
        rows.append(o.run(tt, 1, 1))
        if tt % 3 == 0:
            states.append(o.two_level_state())
    o.kill()
    want = np.concatenate(rows, axis=0)
    r = np.linalg.norm(series - want) / np.linalg.norm(want)
    print(f"q-group deck: Fortran driver (FP64 GPU) vs oracle relL2 = {r:.3e}")
    assert np.abs(want).max() > 0
    assert r <= 1e-12

    # energy_and_dipole.dat: every third step, the printed systems of each group
    omega0 = par[:, 0] * EV_TO_AU
    for g, steps in zip((1, 2), groups):
        assert len(steps) == len(states)
        idx = [q for q, s in enumerate(SYSTEMS) if s[0] == g and s[4]]
        for (time, recs), (dip, rho) in zip(steps, states):
            assert [rec[0] for rec in recs] == [SYSTEMS[q][1] for q in idx]
            for rec, q in zip(recs, idx):
                e = 0.5 * omega0[q] * (rho[q, 1] - rho[q, 0]) * AU_TO_EV
                np.testing.assert_allclose(rec[1], e, rtol=1e-10, atol=1e-14)
                np.testing.assert_allclose(rec[2], dip[q], rtol=1e-8, atol=1e-12 * max(np.abs(dip).max(), 1e-30))
    print(f"energy_and_dipole.dat: {len(states)} print steps match the oracle")
