"""Batched launches: every grid of a batch computes exactly (bitwise) what
its own single-grid run computes, with and without the captured graph, each
grid under its own tiling (rows, planes, tile order)."""
import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G

pytestmark = pytest.mark.gpu


def make(case, precision, tiling=None):
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    if tiling:
        g.set_tiling(*tiling)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    return g


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_equals_single_grid_runs(precision, graph):
    cs = cases.config_batched(n_grids=6, edges=(24, 32, 40), npml=4, steps=60, record_every=3)
    # a second source sharing a cell with the first of grid 0 exercises source passes
    cs[0].sources.append(cs[0].sources[2])
    # every grid keeps its own rows, planes and tile order inside the batch
    tilings = [(4, 16, 1 if graph else 2), (2, 5, 0), (1, 64, 1), (4, 3, 0), (4, 8, 1), (1, 1, 2)]
    print(f"[batch] precision={precision} graph={graph} grids={[c.grid for c in cs]}", flush=True)
    singles = []
    for c, t in zip(cs, tilings):
        g = make(c, precision)
        g.set_step_mode(1)   # the batch runs the separate H and E kernels
        g.run(1, c.steps, c.record_every)
        singles.append(g.read_records())
        g.kill()
    grids = [make(c, precision, t) for c, t in zip(cs, tilings)]
    if not graph:
        grids[3].set_addressing(64)   # one grid on 64-bit indices: the whole batch takes them
    b = omx_fdtd.Batch(grids)
    b.run(1, 30, 3, graph=graph)
    b.run(31, 30, 3, graph=graph)
    for c, g, ref in zip(cs, grids, singles):
        got = g.read_records()
        assert got.shape == ref.shape
        assert np.abs(ref).max() > 0
        np.testing.assert_array_equal(got, ref)
    b.kill()
    for g in grids:
        g.kill()


def test_autotune_keeps_results_identical():
    c = cases.config_batched(n_grids=1, edges=(40,), npml=4, steps=40, record_every=2)[0]
    g = make(c, 4)
    g.run(1, c.steps, c.record_every)
    ref = g.read_records()
    g.kill()
    g = make(c, 4)
    rows, xc, order, ms = g.autotune(steps=2)
    print(f"autotune: rows_y={rows} xchunk={xc} tile order={order} {ms:.4f} ms/step")
    assert g.tiling() == (rows, xc, order)
    g.run(1, c.steps, c.record_every)
    np.testing.assert_array_equal(g.read_records(), ref)


def test_batch_follows_grid_changes_and_full_buffers():
    """A grid advanced on its own between batch runs (the fused step swaps its
    field buffers) is picked up by the next batch run; recording into a grid
    whose record buffer is full (or absent) fails as omx3d_run does."""
    cs = cases.config_batched(n_grids=2, edges=(32, 40), npml=4, steps=40, record_every=4)
    refs = []
    for c in cs:
        g = make(c, 8)
        g.run(1, c.steps, c.record_every)
        refs.append(g.read_records())
        g.kill()
    grids = [make(c, 8) for c in cs]
    print(f"[batch] grids={[c.grid for c in cs]} with a fused single-grid run between batch runs", flush=True)
    b = omx_fdtd.Batch(grids)
    b.run(1, 12, 4)
    grids[0].set_step_mode(0)       # grid 0's source is in vacuum: the fused step applies
    grids[0].run(13, 8, 4)
    assert grids[0].last_run_fused()
    grids[1].run(13, 8, 4)
    b.run(21, 20, 4)
    for g, ref in zip(grids, refs):
        got = g.read_records()
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    with pytest.raises(omx_fdtd.OmxError, match="record buffer full"):
        b.run(41, 8, 4)
    b.kill()
    for g in grids:
        g.kill()
