"""Loads the golden fixtures written by tests/golden/make_golden.py and drives
an engine (Grid3D or the oracle) through the same case."""
import json
import os

import numpy as np

from omx_fdtd import cases, host
from omx_fdtd.units import DETECTOR_CODES, FIELD_CODES, NM_TO_AU, index_to_nm, grid_dr_dt

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("vacuum64", "media_mix", "au_al_periodic", "qsites", "qsites_skip3", "qgroups")


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    fx = {k: z[k] for k in z.files}
    desc = json.loads(str(fx["case_json"]))
    return desc, fx


def media_specs(desc):
    out = []
    for head, params in desc["media"]:
        p = params.split()
        if head == "dielectric":
            out.append(("dielectric", float(p[0])))
        elif head == "drude":
            out.append(("drude", float(p[0]), float(p[1]), float(p[2])))
        else:
            out.append(("drude-lorentz", p[0], float(p[1])))
    return out


def to_case(desc, fx):
    grid = tuple(desc["grid"])
    dx = desc["dx"]
    dr, _ = grid_dr_dt(dx)
    dets = []
    for (t, f, lo, hi) in desc["detectors"]:
        lo_nm = [index_to_nm(lo[a], dx, grid[a]) for a in range(3)]
        hi_nm = [index_to_nm(hi[a], dx, grid[a]) for a in range(3)]
        code = DETECTOR_CODES[t]
        dets.append((code, FIELD_CODES[f], host.detector_indices(code, lo_nm, hi_nm, dr, grid)))
    srcs = []
    for line in desc["sources"]:
        p = line.split()
        srcs.append((p[0],) + tuple(float(v) for v in p[1:]))
    return cases.Case(name=desc.get("name", "golden"), grid=grid, dx_nm=dx, npml=desc["npml"],
                      boundaries=tuple(desc["bc"]), steps=desc["steps"], record_every=desc["record_every"],
                      media=media_specs(desc), aux=fx["aux"] if desc["media"] else None, sources=srcs,
                      detectors=dets)


def q_moves(n_skip, steps):
    """Steps on which the quantum systems move in the driver loop of
    mxim_mxll.f90:76-77,97-104: tq_step = 1 + int((tt - tt_0)/n_skip_steps),
    tt_0 = 1 for n_skip_steps = 1 and 0 otherwise, a move whenever tq_step
    grows."""
    tt_0 = 1 if n_skip == 1 else 0
    old, out = 0, []
    for tt in range(1, steps + 1):
        tq = 1 + (tt - tt_0) // n_skip
        if tq > old:
            old = tq
            out.append(tt)
    return out


def drive(g, desc, case):
    """Runs the case; with quantum sites the reference loop is driven call by
    call with the driver's prescribed dPt/dt, moving the quantum sites every
    n_skip steps.  Returns (series, E at sites on the moving steps)."""
    dr, dt = case.dr_dt
    if not desc["qsites"]:
        if hasattr(g, "read_records"):
            g.run(1, case.steps, case.record_every)
            return g.read_records(), None
        return g.run(1, case.steps, case.record_every), None
    q = np.array(desc["qsites"], dtype=float)
    ijk = q[:, :3].astype(np.int32)
    amp = q[:, 3:6]
    n_skip = desc.get("n_skip", 1)
    if desc.get("q_groups"):
        g.set_q_groups(desc["q_groups"], q[:, 6].astype(np.int32) - 1, ijk, 1.0e-3 / NM_TO_AU ** 3, n_skip * dt)
    else:
        g.set_q_sites(ijk, 1.0e-3 / NM_TO_AU ** 3, n_skip * dt)
    moves = set(q_moves(n_skip, case.steps))
    rows, es = [], []
    for tt in range(1, case.steps + 1):
        time = dt * (tt - 0.5)
        g.time = time
        if tt in moves:
            g.t_skip = time
        g.td_propagate_H_field()
        g.inject_sources(time)
        g.td_propagate_E_field(tt)
        if tt in moves:
            es.append(g.gather_E())
            dpdt = amp * np.sin(0.3 * tt + np.arange(3))[None, :]
            g.scatter_J(dpdt)
        if tt % case.record_every == 0:
            rows.append(g.sample())
    return np.array(rows), np.array(es)
