"""Loads the 2D / 1D golden fixtures written by
tests/golden/make_golden_lowdim.py and drives a grid (omx_fdtd.Grid2D /
Grid1D or the oracle's OracleLowGrid) through the same case."""
import json
import os

import numpy as np

import _golden
from omx_fdtd import cases, lowdim
from omx_fdtd.units import grid_dr_dt

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("tmz2d", "tez2d", "tmz2d_periodic", "tez2d_ycpml", "line1d", "line1d_periodic", "line1d_close")
# fields of one snapshot, in the order the reference driver wrote them
SNAP_FIELDS = {1: ("Ez", "Hx", "Hy"), 2: ("Ex", "Ey", "Hz"), 0: ("Ex", "Hy")}


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False)
    fx = {k: z[k] for k in z.files}
    return json.loads(str(fx["case_json"])), fx


def dr_dt(desc):
    return grid_dr_dt(desc["dx"])


def setup(g, desc, fx, library, make_medium):
    """Media, sources and detectors of the case on grid `g`."""
    dims = desc["dims"]
    dr, _ = dr_dt(desc)
    if desc["media"]:
        meds = []
        for spec in _golden.media_specs(desc):
            m = cases.medium_struct(spec, make_medium)
            if isinstance(m, tuple) and m[0] == "library":
                m = library(m[1], m[2])
            meds.append(m)
        g.set_media(meds, fx["aux"])
    for line in desc["sources"]:
        g.add_source(**lowdim.source_from_line(dims, line, dr, tuple(desc["grid"])))
    for (f, cell) in desc["detectors"]:
        g.add_detector(f, [cell])
    return g


def drive(g, desc):
    """Runs the case: detector rows of every recorded step and the field
    snapshots (flattened as the reference driver wrote them)."""
    fe, steps = desc["field_every"], desc["steps"]
    fields = SNAP_FIELDS[desc.get("mode", 0) if desc["dims"] == 2 else 0]
    rows, snaps = [], []
    tt = 1
    while tt <= steps:
        n = min(fe - (tt - 1) % fe, steps - tt + 1)
        rows.append(g.run(tt, n, desc["record_every"]))
        tt += n
        if (tt - 1) % fe == 0:
            snaps.append(np.concatenate([g.get_field(f).ravel(order="F") for f in fields]))
    return np.concatenate(rows, axis=0), np.array(snaps)
