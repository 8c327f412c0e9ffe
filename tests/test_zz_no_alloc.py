"""omx3d_energy_noalloc (the energy reduction with the library's fixed
scratch array instead of an allocation) against omx3d_energy, through the
C-ABI: two calls in a row, the second on a smaller grid, each bit for bit
equal to omx3d_energy on the same fields, and no device memory taken by the
context for it.  The output buffer is device memory allocated through the
same HIP runtime the engine links (libamdhip64, by ctypes)."""
import ctypes as C

import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G

pytestmark = pytest.mark.gpu


def grid_with_fields(n, precision, seed):
    case = cases.config_dl_sphere_tfsf(n=n, npml=4, steps=30, record_every=0, radius_nm=0.3 * n)
    case.media.append(("dielectric", 2.25))
    case.shapes.append((3, 2, [6, 6, 6, n - 6, n - 6, n // 3]))
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    rng = np.random.default_rng(seed)
    for q, name in enumerate(G.FIELD_NAMES):
        shp = G.field_shape(q + 1, *case.grid)
        g.set_field(name, rng.uniform(-1.0, 1.0, size=shp) * (1.0 if name[0] == "E" else 1.0 / 137.0))
    g.run(1, case.steps, 0)
    return g


def hip_runtime():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


@pytest.mark.parametrize("precision", [4, 8])
def test_energy_noalloc_matches_energy(precision):
    lib = omx_fdtd.load_library()
    hip = hip_runtime()
    print(f"[noalloc] precision={precision}: grids 48^3 then 32^3", flush=True)
    grids = [grid_with_fields(n, precision, seed) for n, seed in ((48, 1), (32, 2))]
    out = C.c_void_p()
    assert hip.hipMalloc(C.byref(out), 16) == 0
    nan2 = np.full(2, np.nan)
    assert hip.hipMemcpy(out, nan2.ctypes.data, 16, 1) == 0     # host to device
    for g in grids:
        n = g.nx
        g.synchronize()
        before = g.device_bytes()
        rc = lib.omx3d_energy_noalloc(g.h, out)
        assert rc == 0, lib.omx_last_error()
        g.synchronize()
        got = np.zeros(2)
        assert hip.hipMemcpy(got.ctypes.data, out, 16, 2) == 0   # device to host
        assert g.device_bytes() == before          # nothing allocated by the context
        w64, w32 = g.energy()                       # the allocating entry point
        print(f"[noalloc] n={n}: noalloc {got[0]:.17g} {got[1]:.9g}, energy {w64:.17g} {w32:.9g}", flush=True)
        assert w64 > 0
        assert got[0] == w64 and got[1] == w32
        g.kill()
    assert hip.hipFree(out) == 0


def test_energy_noalloc_rejects_host_output():
    g = grid_with_fields(24, 4, 3)
    lib = omx_fdtd.load_library()
    host = (C.c_double * 2)()
    assert lib.omx3d_energy_noalloc(g.h, C.cast(host, C.c_void_p)) == 1      # OMX_ERR_ARG
    assert b"device memory" in lib.omx_last_error()
    g.kill()
