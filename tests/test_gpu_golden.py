"""The HIP engine against golden vectors of the reference's own Fortran
modules: FP64 within relL2 1e-12, FP32 within relL2 1e-5.  The golden cases
drive one-cell sources; the FP32 engine keeps the near field of each source
in FP64 (near-field patches), without which FP32 rounding of that near field
alone puts the series ~1e-5 to 1e-4 off."""
import numpy as np
import pytest

import _golden
import omx_fdtd
from omx_fdtd import cases, grid as G

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", [8, 4])
@pytest.mark.parametrize("name", _golden.NAMES)
def test_engine_matches_reference(name, precision):
    desc, fx = _golden.load(name)
    case = _golden.to_case(desc, fx)
    dr, dt = case.dr_dt
    print(f"[golden] {name} precision={precision}", flush=True)
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    if desc["media"]:
        np.testing.assert_array_equal(g.media_map(), fx["media_map"])
    t = g.cpml_tables()
    den = np.concatenate([t[k] for k in ("den_ex", "den_ey", "den_ez", "den_hx", "den_hy", "den_hz")])
    np.testing.assert_allclose(den, fx["den"], rtol=1e-15, atol=0)
    if "be" in fx:
        for k in ("be", "ce", "bh", "ch"):
            np.testing.assert_allclose(t[k], fx[k], rtol=2e-15, atol=0, err_msg=k)
    series, es = _golden.drive(g, desc, case)
    r = rel_l2(series, fx["series"])
    print(f"{name} FP{8 * precision}: engine vs reference relL2 = {r:.3e}", flush=True)
    assert r <= (1e-12 if precision == 8 else 1e-5)
    if es is not None:
        assert rel_l2(es, fx["qsites_e"]) <= (1e-12 if precision == 8 else 1e-5)
    g.kill()
