"""The oracle's 2D and 1D grids (C restatement) against golden vectors
produced by the reference's own mxll_2D_mod / mxll_1D_mod modules
(tests/golden/make_golden_lowdim.py): CPML tables through the den arrays,
the staggered media maps, point-detector series and whole-field snapshots."""
import numpy as np
import pytest

import _golden_lowdim as G
from oracle import oracle


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def build(name):
    desc, fx = G.load(name)
    dr, dt = G.dr_dt(desc)
    o = oracle.OracleLowGrid(desc["dims"], desc["grid"], desc["npml"], desc["bc"], dt, dr, desc.get("mode", 0))
    G.setup(o, desc, fx, oracle.library_medium, oracle.make_medium)
    return desc, fx, o


@pytest.mark.parametrize("name", G.NAMES)
def test_tables_match_reference(name):
    desc, fx, o = build(name)
    den = o.den()
    n1 = desc["grid"][0]
    if desc["dims"] == 1:
        # reference order: den_ez(1:nz) den_hz(1:nz); entries nz of the loops' k = 1..nz-1 stay 1/dr
        np.testing.assert_array_equal(np.concatenate([den[:n1], den[n1:2 * n1]]), fx["den"])
        np.testing.assert_array_equal(o.media_map(), fx["media_map"])
    else:
        n2 = desc["grid"][1]
        ex, hx, ey, hy = den[:n1], den[n1:2 * n1], den[2 * n1:2 * n1 + n2], den[2 * n1 + n2:]
        np.testing.assert_array_equal(np.concatenate([ex, ey, hx, hy]), fx["den"])
        np.testing.assert_array_equal(o.media_map(), fx["media_map"])


@pytest.mark.parametrize("name", G.NAMES)
def test_series_and_fields_match_reference(name):
    desc, fx, o = build(name)
    series, snaps = G.drive(o, desc)
    assert series.shape == fx["series"].shape
    assert snaps.shape == fx["fields"].shape
    r = rel_l2(series, fx["series"])
    rf = rel_l2(snaps, fx["fields"])
    print(f"{name}: oracle vs reference relL2 series {r:.3e} fields {rf:.3e}")
    assert np.abs(fx["series"]).max() > 0
    assert r <= 1e-13 and rf <= 1e-13


def test_kernels_match_host_restatement():
    """The oracle's 1D / 2D source kernels and the package's host-side ones
    (both restating compute_kernel) agree, and sum to 1."""
    from omx_fdtd import lowdim
    for dims in (1, 2):
        for n_ker in (0, 1, 3):
            a = oracle.source_kernel_lowdim(dims, n_ker, 94.5)
            b = lowdim.source_kernel(dims, n_ker, 94.5)
            np.testing.assert_array_equal(a, b)
            assert abs(a.sum() - 1.0) < 1e-14
