"""The four measured workloads, scaled down so that the oracle finishes in
seconds, against the oracle: the same case builders, media, sources,
detectors and (emitters) two-level systems as the full-size runs of the
bench.  FP64 detector series within relL2 <= 1e-12, FP32 within 1e-5; the
separate H/E kernels and the fused step both."""
import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def engine(case, precision, mode):
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    g.set_step_mode(mode)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    return g


def reference(case):
    dr, dt = case.dr_dt
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    return o


def check(case, precision, mode, extra=None):
    print(f"[workload] {case.name} grid={case.grid} steps={case.steps} precision={precision} mode={mode}",
          flush=True)
    o = reference(case)
    want = o.run(1, case.steps, case.record_every)
    g = engine(case, precision, mode)
    g.run(1, case.steps, case.record_every)
    got = g.read_records()
    assert got.shape == want.shape and np.abs(want).max() > 0
    r = rel_l2(got, want)
    print(f"{case.name} FP{8 * precision} mode {mode}: relL2 = {r:.3e} (fused: {g.last_run_fused()})", flush=True)
    assert r <= (TOL64 if precision == 8 else TOL32)
    if extra:
        extra(g, o, precision)
    g.kill()
    o.kill()


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("precision", [8, 4])
def test_dl_sphere_tfsf_scaled(precision, mode):
    case = cases.config_dl_sphere_tfsf(n=48, npml=4, steps=1200, record_every=10, radius_nm=12.0)
    check(case, precision, mode)


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("precision", [8, 4])
def test_ellipsoids_scaled(precision, mode):
    case = cases.config_ellipsoids(grid=(64, 64, 48), n_ell=6, npml=8, steps=500, record_every=10)
    check(case, precision, mode)


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("q_skip", [1, 3])
@pytest.mark.parametrize("mode", [1, 0])
def test_emitters_scaled(mode, q_skip, precision):
    """q_skip 3: the emitters move every third step over dt_q = 3 dt, the
    current interpolated over time - t_skip in between (mxim_mxll.f90:73-104)."""
    case = cases.config_emitters(grid=(96, 96, 64), n_emit=16, npml=8, steps=300, record_every=10)
    case.q_skip = q_skip

    def tls(g, o, precision):
        (dg, rg), (do, ro) = g.two_level_state(), o.two_level_state()
        assert np.abs(do).max() > 0
        if precision == 8:
            np.testing.assert_allclose(rg, ro, rtol=1e-10, atol=1e-14)
            np.testing.assert_allclose(dg, do, rtol=1e-8, atol=1e-11 * np.abs(do).max())
        else:
            # FP32 fields, FP64 density matrices: the emitters' dipoles within
            # relL2 1e-5 of the FP64 oracle's
            r = rel_l2(dg, do)
            print(f"emitter dipoles FP32 vs oracle relL2 = {r:.3e}", flush=True)
            assert r <= TOL32

    check(case, precision, mode, extra=tls)


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("graph", [True, False])
def test_batched_scaled(graph, precision):
    cs = cases.config_batched(n_grids=4, edges=(32, 40, 48), npml=4, steps=200, record_every=4)
    print(f"[workload] batched grids={[c.grid for c in cs]} graph={graph} precision={precision}", flush=True)
    grids = [engine(c, precision, 1) for c in cs]
    # each grid with its own autotuned planes and order at a common row
    # count, as the batched bench runs them
    for g in grids:
        rows, xc, order, ms = g.autotune(steps=2, rows=2)
        assert rows == 2
        print(f"[workload] autotuned rows {rows} xchunk {xc} order {order} ({ms:.3f} ms/step)", flush=True)
    b = omx_fdtd.Batch(grids)
    b.run(1, cs[0].steps, cs[0].record_every, graph=graph)
    for c, g in zip(cs, grids):
        o = reference(c)
        want = o.run(1, c.steps, c.record_every)
        got = g.read_records()
        assert np.abs(want).max() > 0
        r = rel_l2(got, want)
        print(f"{c.name} {c.grid}: relL2 = {r:.3e}", flush=True)
        assert r <= (TOL64 if precision == 8 else TOL32)
        o.kill()
    b.kill()
    for g in grids:
        g.kill()
