"""The FP64 engine against the oracle at the real shape of the headline
workload: 256^3 cells, dx = 2 nm, CPML 12 on every face, the seeded
Drude-Lorentz sphere and the TF/SF plane wave, 100 steps.  The six fields
start from the same seeded random values (zero on the arrays' outer planes)
so that every path -- bulk, CPML faces, the sphere's ADE cells, the TF/SF
corrections -- is active from the first step.  Both step kernels run: the
separate H and E launches and the fused step (which the step tuner keeps on
this workload in FP64 too).  Bar: relL2 <= 1e-12 for each final field and
for the detector series.  This is what makes the FP64 engine a valid
stand-in for the oracle on the full-length runs of the benchmark, whose FP32
error is stated against it."""
import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / nb


def seeded_fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    out = {}
    for q, name in enumerate(FIELDS):
        shp = G.field_shape(q + 1, nx, ny, nz)
        a = rng.uniform(-1.0, 1.0, size=shp)
        a[0], a[-1] = 0.0, 0.0
        a[:, 0], a[:, -1] = 0.0, 0.0
        a[:, :, 0], a[:, :, -1] = 0.0, 0.0
        # H is a factor c0 smaller than E in a plane wave (atomic units)
        out[name] = a if name[0] == "E" else a / 137.0
    return out


@pytest.mark.timeout(1200)
def test_fp64_engine_equals_oracle_at_headline_shape():
    steps = 100
    case = cases.config_dl_sphere_tfsf(n=256, steps=steps, record_every=20)
    dr, dt = case.dr_dt
    print(f"[fullshape] {case.grid} npml={case.npml} steps={steps} FP64 vs oracle", flush=True)
    init = seeded_fields(*case.grid, seed=20261203)
    runs = {}
    for mode in (1, 0):
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
        g.set_step_mode(mode)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        for name in FIELDS:
            g.set_field(name, init[name])
        g.run(1, steps, case.record_every)
        assert g.last_run_fused() == (mode == 0)
        runs[mode] = ({name: g.get_field(name) for name in FIELDS}, g.read_records())
        g.kill()
        print(f"[fullshape] engine done (step mode {mode})", flush=True)
    print("[fullshape] running the oracle", flush=True)
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    for name in FIELDS:
        o.set_field(name, init[name])
    want_series = o.run(1, steps, case.record_every)
    want = {name: o.get_field(name) for name in FIELDS}
    o.kill()
    for mode, (got, got_series) in runs.items():
        worst = 0.0
        for name in FIELDS:
            assert np.abs(want[name]).max() > 0
            r = rel_l2(got[name], want[name])
            worst = max(worst, r)
            print(f"[fullshape] step mode {mode} {name}: relL2 = {r:.3e}", flush=True)
            assert r <= 1e-12, (mode, name)
        rs = rel_l2(got_series, want_series)
        print(f"[fullshape] step mode {mode} detector series relL2 = {rs:.3e}, worst field {worst:.3e}",
              flush=True)
        assert rs <= 1e-12
