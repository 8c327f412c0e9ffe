"""Writes the golden fixtures of tests/golden/ by running the reference's own
3D solver modules, compiled from the reference tree into oracle/_ref/ by
`make -C oracle ref`, through the test driver oracle/ref_driver.f90.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py NAME ...   # some

Each case is described in the reference's own input formats (medium_XXXX.in,
sources.in, detectors.in) plus the driver's case.nml.  A fixture
(NAME.npz, no pickles) holds the case description as plain arrays and what
the reference produced: CPML tables, den arrays, medium coefficients, the
staggered media map, the detector time series parsed from the reference's
output_detector_XXXXXXX/ text files, and E at quantum sites.
"""
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "one-mxll-to-rule-them-all_amd"), ROOT]

from omx_fdtd import host  # noqa: E402
from omx_fdtd.units import index_to_nm, DETECTOR_CODES, FIELD_CODES, grid_dr_dt  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "omx_ref_driver")


def cell_nm(idx, dx, grid):
    return [index_to_nm(int(idx[a]), dx, grid[a]) for a in range(3)]


def golden_cases():
    """Case descriptions.  media: list of (file header line, parameter line,
    cell predicate over (I,J,K) index arrays).  detectors: (type, field,
    lo_cell, hi_cell).  sources: sources.in lines."""
    out = {}
    rng = np.random.default_rng(20261203 + 1)
    n, npml = 64, 10
    pos = rng.integers(npml + 2, n - npml - 1, size=(8, 3), endpoint=True)
    dets = [("point", f, tuple(int(v) for v in p), tuple(int(v) for v in p)) for p in pos for f in ("Ex", "Ey", "Ez")]
    out["vacuum64"] = dict(
        grid=(n, n, n), npml=npml, bc=(3, 3, 3), dx=5.0, steps=1000, record_every=1, media=[],
        sources=["z 1.0 1.549802 5.0 5.0 0.0 0.0 0.0 0.0 0.0 1.0e6 0.0"], detectors=dets, qsites=None)

    g = (24, 22, 20)
    dets = []
    for p in ((12, 11, 14), (7, 15, 9), (17, 6, 12)):
        dets += [("point", f, p, p) for f in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")]
    dets += [("line_x", "Ez", (1, 12, 11), (24, 12, 11)), ("line_y", "Hx", (9, 1, 13), (9, 22, 13)),
             ("line_z", "Ey", (14, 8, 1), (14, 8, 20)), ("plane_xy", "Hy", (1, 1, 12), (24, 22, 12)),
             ("plane_yz", "Ex", (13, 1, 1), (13, 22, 20)), ("plane_zx", "Hz", (1, 10, 1), (24, 10, 20)),
             ("volume", "Ez", (9, 8, 7), (12, 11, 10))]
    out["media_mix"] = dict(
        grid=g, npml=4, bc=(3, 3, 3), dx=5.0, steps=120, record_every=2,
        media=[("drude-lorentz", "Ag 1.0", lambda I, J, K: (I - 12) ** 2 + (J - 11) ** 2 + (K - 10) ** 2 <= 20.0),
               ("dielectric", "2.25", lambda I, J, K: (K <= 4) & (I >= 3) & (I <= 21) & (J >= 3) & (J <= 19)),
               ("drude", "9.0 0.07 1.5", lambda I, J, K: (I >= 17) & (I <= 20) & (J >= 4) & (J <= 8) & (K >= 14) & (K <= 17))],
        sources=["x 2.0 1.2 0.5 0.25 -12.0 8.0 -3.0 5.0 0.05 1.0e6 0.3",
                 "y 1.0 2.0 0.7 0.3 15.0 -10.0 8.0 0.0 0.0 1.0e6 0.0",
                 "z 1.5 1.8 0.4 0.2 0.0 0.0 0.0 10.0 0.0 1.0e6 1.0"],
        detectors=dets, qsites=None)

    g = (20, 18, 22)
    dets = [("point", f, (9, 8, 11), (9, 8, 11)) for f in ("Ex", "Ey", "Ez")]
    dets += [("plane_xy", "Ex", (1, 1, 6), (20, 18, 6))]
    out["au_al_periodic"] = dict(
        grid=g, npml=4, bc=(2, 3, 1), dx=2.0, steps=100, record_every=1,
        media=[("drude-lorentz", "Au 2.0", lambda I, J, K: (I - 10) ** 2 + (J - 9) ** 2 + (K - 14) ** 2 <= 12.0),
               ("drude-lorentz", "Al 1.0", lambda I, J, K: (I >= 3) & (I <= 7) & (J >= 10) & (J <= 14) & (K >= 5) & (K <= 9))],
        sources=["z 1.0 2.5 0.2 0.1 0.0 0.0 -6.0 2.0 0.0 1.0e6 0.0"],
        detectors=dets, qsites=None)

    g = (20, 18, 16)
    dets = [("point", f, (10, 9, 9), (10, 9, 9)) for f in ("Ex", "Ey", "Ez", "Hz")]
    out["qsites"] = dict(
        grid=g, npml=4, bc=(3, 3, 3), dx=5.0, steps=60, record_every=1,
        media=[("dielectric", "2.0", lambda I, J, K: K <= 4),
               ("drude-lorentz", "Ag 1.0", lambda I, J, K: (I >= 13) & (I <= 15) & (J >= 6) & (J <= 8) & (K >= 11) & (K <= 13))],
        sources=["z 1.0 1.549802 0.6 0.3 0.0 0.0 0.0 0.0 0.0 1.0e6 0.0"],
        detectors=dets,
        qsites=[(8, 9, 8, 1e-2, -2e-2, 5e-3), (9, 9, 8, 3e-3, 1e-2, -1e-2), (12, 6, 10, -4e-3, 2e-3, 7e-3),
                (8, 9, 8, 2e-2, 1e-3, -3e-3), (14, 7, 12, 1e-2, 1e-2, 1e-2)])

    # the same sites with the quantum systems moved every 3rd step
    # (n_skip_steps = 3: J interpolated over tau = time - t_skip in between)
    out["qsites_skip3"] = dict(out["qsites"], steps=61, n_skip=3)

    # three q-groups (q_group_mod.f90): "single" (7^3 kernels, overlapping),
    # "material" (one site inside the first group's kernels, one inside the
    # third's), "single" again (overlapping both); qsites rows carry their
    # 1-based group.  Cells written by the first groups only keep J_old = J,
    # the last group's cells take J_old from the groups before it.
    g = (24, 22, 20)
    out["qgroups"] = dict(
        grid=g, npml=4, bc=(3, 3, 3), dx=5.0, steps=60, record_every=1,
        media=[("dielectric", "2.0", lambda I, J, K: K <= 3),
               ("drude-lorentz", "Ag 1.0", lambda I, J, K: (I >= 17) & (I <= 19) & (J >= 13) & (J <= 15) & (K >= 13) & (K <= 15))],
        sources=["z 1.0 1.549802 0.6 0.3 0.0 0.0 0.0 0.0 0.0 1.0e6 0.0"],
        detectors=[("point", f, (12, 11, 10), (12, 11, 10)) for f in ("Ex", "Ey", "Ez", "Hx")] +
                  [("point", f, (8, 14, 12), (8, 14, 12)) for f in ("Ex", "Ez")] +
                  [("plane_xy", "Ez", (1, 1, 10), (24, 22, 10))],
        q_groups=[2, 1, 2],
        qsites=[(9, 9, 9, 1e-2, -2e-2, 5e-3, 1), (11, 10, 9, 3e-3, 1e-2, -1e-2, 1),
                (14, 12, 10, -4e-3, 2e-3, 7e-3, 2), (10, 9, 10, 2e-2, 1e-3, -3e-3, 2), (16, 8, 12, 1e-2, 1e-2, 1e-2, 2),
                (15, 12, 11, -1e-2, 5e-3, 2e-3, 3), (12, 11, 10, 6e-3, -4e-3, 1e-2, 3)])
    return out


def aux_map(case):
    nx, ny, nz = case["grid"]
    I, J, K = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    aux = np.zeros((nx, ny, nz), dtype=np.int8)
    for m, (_, _, pred) in enumerate(case["media"], start=1):
        sel = pred(I, J, K) & (aux == 0)
        aux[sel] = m
    return aux


def write_inputs(case, d):
    nx, ny, nz = case["grid"]
    dx = case["dx"]
    aux = aux_map(case)
    for m, (head, params, _) in enumerate(case["media"], start=1):
        with open(os.path.join(d, f"medium_{m:04d}.in"), "w") as f:
            f.write(head + "\n" + params + "\n")
            for (i, j, k) in zip(*np.nonzero(aux == m)):
                x, y, z = cell_nm((i + 1, j + 1, k + 1), dx, case["grid"])
                f.write(f"{x!r} {y!r} {z!r}\n")
    with open(os.path.join(d, "sources.in"), "w") as f:
        for s in case["sources"]:
            f.write(s + "\n")
    with open(os.path.join(d, "detectors.in"), "w") as f:
        for (t, fld, lo, hi) in case["detectors"]:
            a, b = cell_nm(lo, dx, case["grid"]), cell_nm(hi, dx, case["grid"])
            f.write(f"{t} {fld} {a[0]!r} {b[0]!r} {a[1]!r} {b[1]!r} {a[2]!r} {b[2]!r}\n")
    nq = 0
    if case["qsites"]:
        nq = len(case["qsites"])
        with open(os.path.join(d, "qsites.in"), "w") as f:
            for q in case["qsites"]:
                f.write(" ".join(repr(v) for v in q) + "\n")
    diel = [1.0] * 15
    for m, (head, params, _) in enumerate(case["media"], start=1):
        if head == "dielectric":
            diel[m - 1] = float(params.split()[0])
    with open(os.path.join(d, "case.nml"), "w") as f:
        f.write("&case\n")
        f.write(f" nx={nx}, ny={ny}, nz={nz}, npml={case['npml']}, bc={case['bc'][0]},{case['bc'][1]},{case['bc'][2]},\n")
        f.write(f" n_steps={case['steps']}, record_every={case['record_every']}, n_media={len(case['media'])},\n")
        f.write(f" n_src={len(case['sources'])}, n_det={len(case['detectors'])}, n_q={nq}, timing=0,\n")
        f.write(f" n_skip={case.get('n_skip', 1)},\n")
        if case.get("q_groups"):
            f.write(f" n_qg={len(case['q_groups'])}, qg_type=" + ",".join(str(t) for t in case["q_groups"]) + ",\n")
        f.write(f" dx_nm={dx!r}, diel_eps=" + ",".join(repr(v) for v in diel) + "\n/\n")
    return aux


def parse_series(d, case, nrec):
    cols = []
    for n in range(1, len(case["detectors"]) + 1):
        dd = os.path.join(d, f"output_detector_{n:07d}")
        if os.path.exists(os.path.join(dd, "point.dat")):
            rows = [l.split() for l in open(os.path.join(dd, "point.dat"), encoding="latin-1") if not l.startswith("#")]
            cols.append(np.array([[float(r[-1])] for r in rows]))
        else:
            files = sorted(glob.glob(os.path.join(dd, "*.dat")), key=lambda p: int(re.findall(r"_(\d+)\.dat$", p)[0]))
            block = []
            for p in files:
                block.append([float(l.split()[-1]) for l in open(p, encoding="latin-1") if l.strip() and not l.startswith("#")])
            cols.append(np.array(block))
    series = np.concatenate(cols, axis=1)
    assert series.shape[0] == nrec, (series.shape, nrec)
    return series


def detector_headers(d, case):
    """The header lines the reference wrote for each detector: point.dat's
    first line, or the two header lines of the first and of the last print
    step's file (text as written, NUL bytes included)."""
    out = []
    for n in range(1, len(case["detectors"]) + 1):
        dd = os.path.join(d, f"output_detector_{n:07d}")
        if os.path.exists(os.path.join(dd, "point.dat")):
            with open(os.path.join(dd, "point.dat"), encoding="latin-1") as f:
                out.append([f.readline()])
            continue
        files = sorted(glob.glob(os.path.join(dd, "*.dat")), key=lambda p: int(re.findall(r"_(\d+)\.dat$", p)[0]))
        hdr = []
        for p in (files[0], files[-1]):
            with open(p, encoding="latin-1") as f:
                hdr.append(f.readline() + f.readline())
        out.append(hdr)
    return out


def run_case(name, case):
    d = tempfile.mkdtemp(prefix=f"omxgold_{name}_")
    try:
        aux = write_inputs(case, d)
        env = dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 1))
        p = subprocess.run([DRIVER], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env)
        if p.returncode != 0:
            raise RuntimeError(p.stdout)
        print(name, p.stdout.strip().splitlines()[-1])
        nx, ny, nz = case["grid"]
        npml = case["npml"]
        co = np.fromfile(os.path.join(d, "coeffs.bin"), dtype=np.float64)
        den_len = 2 * (nx + ny + nz)
        cp = co[:-den_len]
        out = dict(den=co[-den_len:], aux=aux)
        if cp.size:
            out.update(be=cp[:npml], ce=cp[npml:2 * npml], bh=cp[2 * npml:3 * npml - 1], ch=cp[3 * npml - 1:])
        raw = np.fromfile(os.path.join(d, "media.bin"), dtype=np.uint8)
        nmap = nx * ny * nz * 3 * 4
        out["media_map"] = raw[-nmap:].view(np.int32).reshape((nx, ny, nz, 3), order="F") if nmap <= raw.size else np.zeros(0)
        out["media_coeffs"] = raw[:-nmap].view(np.float64) if raw.size > nmap else np.zeros(0)
        nrec = sum(1 for tt in range(1, case["steps"] + 1) if tt % case["record_every"] == 0)
        out["series"] = parse_series(d, case, nrec)
        if case["qsites"]:
            e = np.fromfile(os.path.join(d, "qsites_e.bin"), dtype=np.float64)
            moves = np.fromfile(os.path.join(d, "qmoves.bin"), dtype=np.int32)
            out["qsites_e"] = e.reshape((moves.size, 3, len(case["qsites"]))).transpose(0, 2, 1)
            out["qmoves"] = moves
        out["headers_json"] = np.array(json.dumps(detector_headers(d, case)))
        desc = {k: v for k, v in case.items() if k != "media"}
        desc["media"] = [[h, p] for (h, p, _) in case["media"]]
        out["case_json"] = np.array(json.dumps(desc))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    finally:
        shutil.rmtree(d, ignore_errors=True)


if __name__ == "__main__":
    if not os.path.exists(DRIVER):
        sys.exit("reference build missing: make -C oracle ref")
    cs = golden_cases()
    for name in (sys.argv[1:] or list(cs)):
        run_case(name, cs[name])
