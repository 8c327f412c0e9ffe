"""The Fortran driver (omx_fdtd_3d, maxwell_factory(2) and (1)) on 2D and 1D
decks written in the reference's own input formats: its point.dat series
against the golden vectors of the reference's own mxll_2D_mod /
mxll_1D_mod builds (relL2 1e-12), and its line / plane files
(write_fields_subs_mod.f90:16-628 names, headers and coordinates) against
the oracle sampled at the same cells."""
import glob
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import _golden_lowdim as G
from omx_fdtd.units import AU_TO_NM, grid_dr_dt, index_to_nm
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden_lowdim  # noqa: E402

EXE = os.path.join(os.path.dirname(HERE), "one-mxll-to-rule-them-all_amd", "omx_fdtd_3d")
AU_TO_FS = 2.4188843265864e-2
BC = {1: "close", 2: "periodic", 3: "cpml"}
pytestmark = pytest.mark.gpu


def write_inp(d, case, dt, n_det):
    dims, grid, dx = case["dims"], case["grid"], case["dx"]
    bc = [BC[b] for b in case["bc"]] + ["close"] * (3 - len(case["bc"]))
    box = [g * dx for g in grid] + [dx] * (3 - len(grid))
    mode = {1: "TMz", 2: "TEz", 3: "full"}.get(case.get("mode", 0), "none")
    with open(os.path.join(d, "inp"), "w") as f:
        f.write("&OMxRTA\n")
        f.write(" mxll_boundaries = " + ", ".join(f"'{b}'" for b in bc) + "\n")
        f.write(f" mxll_dimensions = {dims}\n mxll_2D_mode = '{mode}'\n mxll_npml = {case['npml']}\n")
        f.write(f" mxll_n_src = {len(case['sources'])}\n mxll_n_media = {len(case['media'])}\n")
        f.write(f" mxll_n_detectors = {n_det}\n")
        f.write(" mxll_box_size = %r, %r, %r\n mxll_dr = %r\n" % (box[0], box[1], box[2], dx))
        f.write(" mxll_total_time = %r\n" % ((case["steps"] + 0.5) * dt * AU_TO_FS))
        f.write(" mxll_dt_det_print = %r\n/\n" % ((case["record_every"] + 0.5) * dt * AU_TO_FS))


def run_deck(d):
    p = subprocess.run([EXE], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(p.stdout[-800:])
    assert p.returncode == 0


def point_series(d, n_det):
    cols = []
    for n in range(1, n_det + 1):
        rows = [l.split() for l in open(os.path.join(d, f"output_detector_{n:07d}", "point.dat"), encoding="latin-1")
                if not l.startswith("#")]
        cols.append([float(r[-1]) for r in rows])
    return np.array(cols).T


@pytest.mark.parametrize("name", ["tmz2d", "tez2d", "tmz2d_periodic", "tez2d_ycpml", "line1d", "line1d_close"])
def test_driver_deck_matches_reference(name):
    assert os.path.exists(EXE), "omx_fdtd_3d not built"
    desc, fx = G.load(name)
    case = make_golden_lowdim.golden_cases()[name]
    _, dt = grid_dr_dt(case["dx"])
    with tempfile.TemporaryDirectory() as d:
        make_golden_lowdim.write_inputs(case, d)
        write_inp(d, case, dt, len(case["detectors"]))
        print(f"[fortran-lowdim] {name} deck in {d}", flush=True)
        run_deck(d)
        series = point_series(d, len(case["detectors"]))
        with open(os.path.join(d, "output_detector_0000001", "point.dat"), encoding="latin-1") as f:
            head = f.readline()
    r = np.linalg.norm(series - fx["series"]) / np.linalg.norm(fx["series"])
    print(f"{name}: Fortran driver vs reference relL2 = {r:.3e}")
    assert series.shape == fx["series"].shape and r <= 1e-12
    assert head.startswith("# Time (a.u.)              " + case["detectors"][0][0] + " (a.u.),      x = ")


def _rows(path, ncol):
    """The two header lines and the rows (list-directed records may wrap)."""
    with open(path, encoding="latin-1") as f:
        lines = f.read().splitlines()
    tok = [float(v) for l in lines[2:] for v in l.split()]
    assert len(tok) % ncol == 0, path
    return lines[:2], np.array(tok).reshape(-1, ncol)


def test_driver_line_and_plane_files():
    """2D line_x / line_y / plane_xy and 1D line_z files: names, headers,
    coordinates and values (the oracle sampled at the same cells)."""
    for name, dets in (("tmz2d", [("line_x", "Ez", (1, 26), (60, 26)), ("line_y", "Hx", (30, 5), (30, 40)),
                                  ("plane_xy", "Hy", (20, 20), (27, 24))]),
                       ("tez2d", [("line_x", "Hz", (3, 30), (50, 30)), ("plane_xy", "Ex", (10, 10), (14, 13))]),
                       ("line1d", [("line_z", "Ex", (90,), (130,)), ("line_z", "Hy", (1,), (12,))])):
        case = dict(make_golden_lowdim.golden_cases()[name], steps=60, record_every=20)
        desc, fx = G.load(name)
        dr, dt = grid_dr_dt(case["dx"])
        grid = case["grid"]
        o = oracle.OracleLowGrid(case["dims"], grid, case["npml"], case["bc"], dt, dr, case.get("mode", 0))
        G.setup(o, dict(desc, detectors=[]), fx, oracle.library_medium, oracle.make_medium)
        cells = []
        for (t, fld, lo, hi) in dets:
            axes = [range(lo[q], hi[q] + 1) for q in range(len(lo))]
            c = [(i,) for i in axes[0]] if len(lo) == 1 else [(i, j) for j in axes[1] for i in axes[0]]
            cells.append(c)
            o.add_detector(fld, c)
        with tempfile.TemporaryDirectory() as d:
            make_golden_lowdim.write_inputs(dict(case, detectors=[]), d)
            with open(os.path.join(d, "detectors.in"), "w") as f:
                for (t, fld, lo, hi) in dets:
                    if case["dims"] == 1:
                        a, b = index_to_nm(lo[0], case["dx"], grid[0]), index_to_nm(hi[0], case["dx"], grid[0])
                        f.write(f"{t} {fld} 0.0 0.0 0.0 0.0 {a!r} {b!r}\n")
                    else:
                        a = [index_to_nm(lo[q], case["dx"], grid[q]) for q in range(2)]
                        b = [index_to_nm(hi[q], case["dx"], grid[q]) for q in range(2)]
                        f.write(f"{t} {fld} {a[0]!r} {b[0]!r} {a[1]!r} {b[1]!r} 0.0 0.0\n")
            write_inp(d, case, dt, len(dets))
            print(f"[fortran-lowdim] {name} line/plane deck", flush=True)
            run_deck(d)
            for p in range(3):
                want_all = o.run(20 * p + 1, 20, 20)[0]
                off = 0
                for n, (t, fld, lo, hi) in enumerate(dets, start=1):
                    want = want_all[off:off + len(cells[n - 1])]
                    off += len(cells[n - 1])
                    kind = "plane_xy" if t == "plane_xy" else ("line_z" if case["dims"] == 1 and fld == "Hy" else "line")
                    path = os.path.join(d, f"output_detector_{n:07d}", f"{fld}_{kind}_{p:09d}.dat")
                    ncol = 3 if t == "plane_xy" else 2
                    head, rows = _rows(path, ncol)
                    assert head[0].startswith("# Time = "), head
                    assert head[1].startswith("# " + {"line_x": "x", "line_y": "y", "plane_xy": "x", "line_z": "z"}[t]
                                              + " (nm)"), head
                    assert rows.shape[0] == want.size, (path, rows.shape, want.shape)
                    np.testing.assert_allclose(rows[:, -1], want, rtol=1e-12, atol=1e-12 * np.abs(want).max() + 1e-300)
                    coord = np.array([[(c[q] - int(grid[q] / 2)) * dr * AU_TO_NM for q in range(len(lo))]
                                      for c in cells[n - 1]])
                    if t == "line_y":
                        coord = coord[:, 1:]
                    elif t == "line_x":
                        coord = coord[:, :1]
                    np.testing.assert_allclose(rows[:, :-1], coord, rtol=1e-6, atol=1e-6)
        o.kill()
