"""GPU parity tests: the HIP engine (through its C-ABI) against the oracle.

Bar: FP64 detector time series within relL2 <= 1e-12 of the oracle (the
engine contracts multiply-adds, the oracle does not, so bitwise equality is
not expected); integer/index work (media maps, classes) bit-exact; FP32 runs
within relL2 <= 1e-5 of the FP64 oracle (point sources with the engine's
FP64 near-field patches, the default of FP32 grids).
"""
import numpy as np
import pytest

import omx_fdtd
from omx_fdtd import cases, grid as G
from omx_fdtd.units import (BC_CPML, BC_PERIODIC, BC_CLOSE, DET_POINT, DET_LINE_X, DET_PLANE_XY,
                            DET_PLANE_ZX, DET_VOLUME, EX, EY, EZ, HX, HY, HZ, EV_TO_AU, FS_TO_AU)
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def both(case, precision=8):
    dr, dt = case.dr_dt
    print(f"[case] {case.name} grid={case.grid} steps={case.steps} precision={precision}", flush=True)
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    return g, o


def run_both(case, precision=8):
    g, o = both(case, precision)
    g.run(1, case.steps, case.record_every)
    got = g.read_records()
    want = o.run(1, case.steps, case.record_every)
    return g, o, got, want


def small_case(name, n=(22, 20, 18), npml=4, bc=(BC_CPML,) * 3, steps=60, media=(), shapes=(),
               sources=None, dets=None, dx=5.0):
    nx, ny, nz = n
    if sources is None:
        sources = [("z", 1.0, 1.549802, 0.6, 0.3, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0e6, 0.0)]
    if dets is None:
        pts = [(nx // 2 + 2, ny // 2 - 1, nz // 2 + 1), (nx // 2 - 3, ny // 2 + 2, nz // 2 - 2),
               (npml + 1, ny // 2, nz // 2)]
        dets = cases.point_detectors(pts, ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
        dets.append((DET_LINE_X, EZ, np.array([(i, ny // 2, nz // 2 + 1) for i in range(1, nx + 1)])))
        dets.append((DET_PLANE_XY, HX, cases.plane_xy(n, nz // 2 - 1)))
    return cases.Case(name=name, grid=n, dx_nm=dx, npml=npml, boundaries=bc, steps=steps, record_every=1,
                      media=list(media), shapes=list(shapes), sources=sources, detectors=dets)


# ---------------------------------------------------------------------------
def test_cpml_tables_match_oracle():
    for npml, n in ((4, (20, 18, 16)), (10, (64, 64, 64)), (12, (40, 30, 26))):
        case = small_case("cpml", n=n, npml=npml, bc=(BC_CPML, BC_CPML, BC_CPML))
        g, o = both(case)
        a, b = g.cpml_tables(), o.cpml_tables()
        for k in b:
            np.testing.assert_allclose(a[k], b[k], rtol=2e-15, atol=0, err_msg=k)


def test_media_map_and_coeffs_match_oracle():
    n = (26, 24, 22)
    shapes = [(1, 1, [9.0, 10.0, 11.0, 5.2]), (3, 2, [15, 3, 3, 24, 20, 8]),
              (2, 3, [14.0, 14.0, 14.0, 6.0, 3.0, 4.0, 0.8, 0.6, 0.0, -0.6, 0.8, 0.0, 0.0, 0.0, 1.0]),
              (4, 0, [20.0, 12.0, 2.5, 1, 22])]
    media = [("drude-lorentz", "Ag", 1.0), ("dielectric", 2.25), ("drude", 9.0, 0.07, 1.0)]
    case = small_case("media", n=n, media=media, shapes=shapes)
    g, o = both(case)
    np.testing.assert_array_equal(g.media_map(), o.media_map())
    for m in (1, 2, 3):
        c, w = g.medium_coeffs(m), o.medium_coeffs(m)
        for k in ("A1", "A2", "C1", "C2", "C3", "C4", "inv_eps"):
            np.testing.assert_allclose(getattr(c, k), w[k], rtol=1e-15, atol=0, err_msg=f"{m}:{k}")
        for k in ("alpha_k", "beta_k", "gamma_k", "B1_k", "B2_k"):
            np.testing.assert_allclose(np.array(getattr(c, k)), w[k], rtol=1e-15, atol=0, err_msg=f"{m}:{k}")


def test_library_media_match_oracle():
    for name in ("Ag", "Al", "Au"):
        a = G.library_medium(name, 1.3)
        b = oracle.library_medium(name, 1.3)
        assert a.n_poles == b.n_poles
        for k in ("eps_r", "omega_d", "gamma_d"):
            assert getattr(a, k) == getattr(b, k), (name, k)
        for k in ("omega_k", "gamma_k", "a_k"):
            assert list(getattr(a, k)) == list(getattr(b, k)), (name, k)


@pytest.mark.parametrize("variant", ["cpml", "close", "periodic", "mixed"])
def test_vacuum_fp64(variant):
    bc = {"cpml": (BC_CPML,) * 3, "close": (BC_CLOSE,) * 3, "periodic": (BC_PERIODIC,) * 3,
          "mixed": (BC_PERIODIC, BC_CPML, BC_CLOSE)}[variant]
    case = small_case(f"vacuum-{variant}", bc=bc, steps=80)
    g, o, got, want = run_both(case)
    assert np.abs(want).max() > 0
    r = rel_l2(got, want)
    print(f"relL2 = {r:.3e}")
    assert r <= TOL64


@pytest.mark.parametrize("direction", ["x", "y", "z"])
def test_source_kernel_fp64(direction):
    src = [(direction, 2.0, 1.2, 0.5, 0.25, 1.0, -2.0, 0.5, 10.0, 0.05, 1.0e6, 0.3)]
    case = small_case(f"kernel-{direction}", n=(24, 22, 20), sources=src, steps=50)
    g, o, got, want = run_both(case)
    r = rel_l2(got, want)
    print(f"relL2 = {r:.3e}")
    assert r <= TOL64


@pytest.mark.parametrize("medium", ["Ag", "Au", "Al", "drude", "dielectric", "custom"])
def test_media_fp64(medium):
    spec = {"Ag": ("drude-lorentz", "Ag", 1.0), "Au": ("drude-lorentz", "Au", 2.0),
            "Al": ("drude-lorentz", "Al", 1.0), "drude": ("drude", 9.0, 0.07, 1.5),
            "dielectric": ("dielectric", 2.25),
            "custom": ("dl-custom", 5.0, 9.0, 0.07, [(1.2, 3.1, 0.3), (0.7, 2.7, 0.15)])}[medium]
    shapes = [(1, 1, [11.0, 10.0, 9.0, 4.5]), (3, 1, [3, 3, 2, 20, 18, 4])]
    case = small_case(f"medium-{medium}", media=[spec], shapes=shapes, steps=100)
    g, o, got, want = run_both(case)
    r = rel_l2(got, want)
    print(f"relL2 = {r:.3e}")
    assert r <= TOL64


@pytest.mark.parametrize("precision", [8, 4])
def test_cpml_faces_adjacent(precision):
    """Grids exactly 2 npml cells along x and z: the low and high CPML faces
    touch, every cell of those axes lies in one of them and a z tile holds
    both; vacuum, dielectric and Drude-Lorentz cells."""
    spec = ("drude-lorentz", "Ag", 1.0)
    shapes = [(1, 1, [8.0, 9.0, 8.0, 3.5]), (3, 2, [2, 2, 2, 15, 17, 3])]
    case = small_case("faces-adjacent", n=(16, 18, 16), npml=8, media=[spec, ("dielectric", 2.25)],
                      shapes=shapes, steps=80)
    g, o, got, want = run_both(case, precision=precision)
    assert np.abs(want).max() > 0
    r = rel_l2(got, want)
    print(f"FP{8 * precision} relL2 = {r:.3e}")
    assert r <= (TOL64 if precision == 8 else TOL32)


@pytest.mark.parametrize("precision", [8, 4])
def test_zface_psi_batches(precision):
    """z-face CPML psi held in registers a batch of planes at a time: nz =
    133 with npml 6 puts the high H face (k = 128..132) across the z tiles
    ending at 128 and 129, chunks of 13 and 40 planes take several batches
    (12 planes per batch in the first z tile, 2 in the last), xchunk 1 one
    plane per batch; a Drude-Lorentz sphere inside the low z face and a
    dielectric box inside the high one (the dispersive workgroups own the
    sphere's psi).  Every tiling against the FP64 oracle."""
    n = (14, 13, 133)
    spec = ("drude-lorentz", "Ag", 1.0)
    shapes = [(1, 1, [7.0, 6.0, 4.0, 3.5]), (3, 2, [2, 2, 126, 13, 12, 132])]
    pts = [(7, 6, 3), (7, 6, 130), (5, 7, 128), (8, 5, 66), (3, 3, 131), (11, 10, 2)]
    dets = cases.point_detectors(pts, ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
    # point sources 8 cells inside each z face (positions in nm from the centre cell 66)
    sources = [("x", 1.0, 1.549802, 0.6, 0.3, 0.0, 0.0, -232.0, 0.0, 0.0, 1.0e6, 0.0),
               ("y", 1.0, 1.549802, 0.6, 0.3, 4.0, 0.0, 232.0, 0.0, 0.0, 1.0e6, 0.0)]
    case = small_case("zface-batches", n=n, npml=6, media=[spec, ("dielectric", 2.25)], shapes=shapes,
                      steps=70, sources=sources, dets=dets, dx=4.0)
    dr, dt = case.dr_dt
    o = oracle.OracleGrid(case.grid, case.npml, case.boundaries, dt, dr)
    cases.apply(case, o, library=oracle.library_medium, make_medium=oracle.make_medium,
                rasterize=oracle.rasterize)
    want = o.run(1, case.steps, case.record_every)
    assert np.abs(want).max() > 0
    for rows, xc in ((4, 4), (1, 1), (2, 13), (4, 40)):
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
        g.set_step_mode(1)
        g.set_tiling(rows, xc)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, case.record_every)
        got = g.read_records()
        r = rel_l2(got, want)
        # each detector point on its own (six components): the weak points
        # near the faces count as much as the strong ones
        rp = max(rel_l2(got[:, 6 * q:6 * q + 6], want[:, 6 * q:6 * q + 6]) for q in range(len(pts)))
        print(f"FP{8 * precision} rows {rows} xchunk {xc}: relL2 = {r:.3e}, worst point {rp:.3e}", flush=True)
        assert r <= (TOL64 if precision == 8 else TOL32)
        if precision == 8:
            assert rp <= 1e-10
        g.kill()


@pytest.mark.parametrize("seed", [7, 11])
def test_seeded_random_grids(seed):
    """Seeded random grids (sizes, CPML thickness, tile shape and order,
    Drude-Lorentz sphere, dielectric layer, two point sources, detectors
    near both z faces): FP64 against the oracle, per detector point."""
    rng = np.random.default_rng(seed)
    for it in range(4):
        npml = int(rng.integers(3, 9))
        n = (int(rng.integers(2 * npml + 1, 2 * npml + 50)), int(rng.integers(2 * npml + 1, 2 * npml + 50)),
             int(rng.integers(max(2 * npml + 1, 20), 150)))
        dx = 4.0
        pts = [tuple(int(rng.integers(1, n[a] + 1)) for a in range(3)) for _ in range(4)]
        pts += [(n[0] // 2, n[1] // 2, 2), (n[0] // 2, n[1] // 2, n[2] - 1)]
        dets = cases.point_detectors(pts, ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"))
        srcs = []
        for _ in range(2):
            pos = [float((rng.integers(1, n[a] + 1) - n[a] // 2) * dx) for a in range(3)]
            srcs.append((str(rng.choice(["x", "y", "z"])), 1.0, 1.549802, 0.6, 0.3, *pos, 0.0, 0.0, 1.0e6, 0.0))
        centre = [float(rng.uniform(1, n[a])) for a in range(3)]
        shapes = [(1, 1, centre + [float(rng.uniform(2, 6))]), (3, 2, [1, 1, 1, n[0], n[1], int(rng.integers(1, 5))])]
        media = [("drude-lorentz", str(rng.choice(["Ag", "Au", "Al"])), 1.0), ("dielectric", 2.25)]
        case = small_case(f"random-{seed}-{it}", n=n, npml=npml, media=media, shapes=shapes, steps=50,
                          sources=srcs, dets=dets, dx=dx)
        rows, xc, order = int(rng.choice([1, 2, 4])), int(rng.integers(1, 20)), int(rng.integers(0, 3))
        g, o = both(case)
        g.set_step_mode(1)
        g.set_tiling(rows, xc, order)
        g.run(1, case.steps, case.record_every)
        got = g.read_records()
        want = o.run(1, case.steps, case.record_every)
        r = rel_l2(got, want)
        rp = max(rel_l2(got[:, 6 * q:6 * q + 6], want[:, 6 * q:6 * q + 6]) for q in range(len(pts))
                 if np.abs(want[:, 6 * q:6 * q + 6]).max() > 1e-30)
        print(f"grid {n} npml {npml} tiling {rows}x{xc} order {order}: relL2 {r:.2e}, worst point {rp:.2e}",
              flush=True)
        assert np.abs(want).max() > 0
        assert r <= TOL64 and rp <= 1e-10
        g.kill()


def test_dl_sphere_fp32_against_fp64_oracle():
    spec = ("dl-custom", 5.0, 9.0, 0.07, [(1.2, 3.1, 0.3), (0.7, 2.7, 0.15)])
    case = small_case("fp32-dl", n=(28, 26, 24), media=[spec], shapes=[(1, 1, [14.0, 13.0, 12.0, 5.0])],
                      steps=150)
    g, o, got, want = run_both(case, precision=4)
    r = rel_l2(got, want)
    print(f"FP32 vs FP64 oracle relL2 = {r:.3e}")
    assert r <= TOL32


def test_tiling_variants_identical():
    spec = ("drude-lorentz", "Ag", 1.0)
    case = small_case("tiling", n=(30, 27, 70), media=[spec], shapes=[(1, 1, [10.0, 9.0, 22.0, 5.0])],
                      steps=40)
    ref = None
    # separate H/E kernels under several tile shapes
    for rows, xc, order in ((4, 16, 0), (1, 1, 1), (2, 7, 2), (4, 64, 1), (1, 4, 0), (4, 2, 2)):
        dr, dt = case.dr_dt
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
        g.set_step_mode(1)
        g.set_tiling(rows, xc, order)
        assert g.tiling() == (rows, xc, order)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, 1)
        assert not g.last_run_fused()
        rec = g.read_records()
        if ref is None:
            ref = rec
        else:
            np.testing.assert_array_equal(rec, ref)
        g.kill()
    # fused H+E step under several tile shapes (rows per wavefront, planes
    # per workgroup chunk, wavefronts along y): each equal to the separate
    # kernels up to rounding (the point source enters E at its own cell after
    # the update, and each row count is its own kernel, contracted to
    # multiply-adds on its own)
    for rows, xc, ny in ((2, 8, 0), (1, 1, 1), (2, 3, 3), (1, 64, 8), (2, 5, 1), (1, 7, 2)):
        dr, dt = case.dr_dt
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
        g.set_step_mode(0, rows, xc, ny)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, 1)
        assert g.last_run_fused()
        rec = g.read_records()
        r = rel_l2(rec, ref)
        m = np.abs(rec - ref).max() / np.abs(ref).max()
        print(f"fused rows {rows} xchunk {xc} y-waves {ny}: relL2 vs separate {r:.3e}, max |diff|/max|ref| {m:.3e}")
        assert r <= 1e-13 and m <= 1e-13
        g.kill()


@pytest.mark.parametrize("precision", [8, 4])
def test_autotune_step_leaves_the_grid_as_found(precision):
    """omx3d_autotune_step times both step modes on the grid itself (TF/SF
    line, DL sphere, CPML), keeps one and clears the state: the run after it
    equals a fresh grid's run with the separate kernels to rounding."""
    case = cases.config_dl_sphere_tfsf(n=40, npml=6, steps=60, record_every=2, radius_nm=10.0)
    dr, dt = case.dr_dt
    ref = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    ref.set_step_mode(1)
    cases.apply(case, ref, library=G.library_medium, make_medium=G.medium)
    ref.run(1, case.steps, case.record_every)
    want = ref.read_records()
    ref.kill()
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    mode, t_sep, t_fus = g.autotune_step(steps=3)
    print(f"FP{8 * precision}: mode {mode}, separate {t_sep:.4f} ms, fused {t_fus} ms", flush=True)
    assert mode in (0, 1) and t_sep > 0.0 and t_fus is not None and t_fus > 0.0
    g.run(1, case.steps, case.record_every)
    assert g.last_run_fused() == (mode == 0)
    got = g.read_records()
    g.kill()
    assert np.abs(want).max() > 0
    r = rel_l2(got, want)
    print(f"after the tuner vs fresh separate run: relL2 {r:.3e}", flush=True)
    assert r <= (1e-13 if precision == 8 else 5e-6)


def test_autotune_step_periodic_grid_keeps_separate():
    """A periodic axis: the fused step does not apply, the tuner keeps the
    separate kernels and reports no fused time."""
    case = small_case("periodic", n=(20, 18, 16), bc=(BC_PERIODIC, BC_CPML, BC_CPML), steps=10)
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    mode, t_sep, t_fus = g.autotune_step(steps=2)
    assert mode == 1 and t_fus is None and t_sep > 0.0
    g.run(1, case.steps, 1)
    assert not g.last_run_fused()
    g.kill()


@pytest.mark.parametrize("precision", [8, 4])
def test_fused_dispersive_cells_on_cpml_faces(precision):
    """A Drude-Lorentz sphere across the x-low CPML face and nothing but
    vacuum besides: the fused step runs without its media variant, steps the
    sphere's E psi itself and the dispersive pass adds it as it is.  Must
    equal the separate kernels (which step the psi in the dispersive pass)."""
    spec = ("drude-lorentz", "Ag", 1.0)
    case = small_case("dispcpml", n=(30, 28, 40), npml=6, media=[spec], shapes=[(1, 1, [3.0, 14.0, 20.0, 8.0])],
                      steps=80)
    out = []
    for mode in (1, 0):
        dr, dt = case.dr_dt
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
        g.set_step_mode(mode)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        mm = g.media_map()
        assert (mm[:case.npml] != 0).any()          # dispersive cells on the CPML face
        g.run(1, case.steps, 1)
        assert g.last_run_fused() == (mode == 0)
        out.append(g.read_records())
        g.kill()
    assert np.abs(out[0]).max() > 0
    r = rel_l2(out[1], out[0])
    print(f"FP{8 * precision}: fused vs separate relL2 = {r:.3e}")
    assert r <= (1e-13 if precision == 8 else 5e-6)


@pytest.mark.parametrize("precision", [8, 4])
def test_fused_z_face_across_two_wavefronts(precision):
    """nz = 70 with CPML 8: the high z face (k = 63..69) spans the first and
    the second wavefront along z of a fused workgroup.  The fused step must
    equal the separate kernels there (same arithmetic per cell)."""
    spec = ("drude-lorentz", "Ag", 1.0)
    case = small_case("zface2w", n=(20, 18, 70), npml=8, media=[spec], shapes=[(1, 1, [10.0, 9.0, 40.0, 3.0])],
                      steps=60)
    case.detectors.append((DET_PLANE_XY, EZ, cases.plane_xy(case.grid, 65)))
    out = []
    for mode in (1, 0):
        dr, dt = case.dr_dt
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
        g.set_step_mode(mode, 2 if mode == 0 else 0, 4 if mode == 0 else 0)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, 1)
        assert g.last_run_fused() == (mode == 0)
        out.append(g.read_records())
        g.kill()
    assert np.abs(out[0]).max() > 0
    r = rel_l2(out[1], out[0])
    print(f"FP{8 * precision}: fused vs separate relL2 = {r:.3e}")
    assert r <= (1e-13 if precision == 8 else 5e-6)


@pytest.mark.parametrize("precision", [8, 4])
def test_tiling_addressing_variants_identical_tfsf(precision):
    """TF/SF corrections, CPML and a Drude-Lorentz sphere under several tile
    shapes and orders, with 32-bit byte offsets and with 64-bit indices: the
    same arithmetic on every cell, so bitwise equal records."""
    case = cases.config_dl_sphere_tfsf(n=40, npml=6, steps=80, record_every=2, radius_nm=10.0)
    ref = None
    for rows, xc, order, bits in ((4, 4, 2, 32), (4, 4, 2, 64), (1, 3, 0, 32), (2, 16, 1, 64), (4, 1, 2, 32),
                                  (2, 5, 0, 64)):
        dr, dt = case.dr_dt
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
        g.set_step_mode(1)
        g.set_tiling(rows, xc, order)
        g.set_addressing(bits)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, case.record_every)
        rec = g.read_records()
        print(f"rows {rows} xchunk {xc} order {order} offsets {bits}: max |rec| {np.abs(rec).max():.3e}")
        if ref is None:
            ref = rec
            assert np.abs(ref).max() > 0
        else:
            np.testing.assert_array_equal(rec, ref)
        g.kill()


@pytest.mark.parametrize("precision", [8, 4])
def test_fused_equals_separate_kernels_tfsf(precision):
    """No point sources: the fused pass and the separate kernels evaluate the
    same expressions in the same order; they differ only where the compiler
    contracts multiply-adds differently in the two kernels (last-bit level)."""
    case = cases.config_dl_sphere_tfsf(n=48, npml=6, steps=120, record_every=4, radius_nm=12.0)
    out = []
    for mode in (1, 0):
        dr, dt = case.dr_dt
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
        g.set_step_mode(mode)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, case.record_every)
        assert g.last_run_fused() == (mode == 0)
        out.append(g.read_records())
        g.kill()
    assert np.abs(out[0]).max() > 0
    r = rel_l2(out[1], out[0])
    print(f"fused vs separate kernels relL2 = {r:.3e}")
    assert r <= (1e-13 if precision == 8 else 5e-6)


def test_fused_tile_shapes_tfsf():
    """The TF/SF plane wave in the fused step across tile shapes (rows, planes
    per chunk, wavefronts along y): the H corrections of the face lanes, the
    extra workgroup that advances the incident line's H after the last box,
    and the E corrections of the post kernel, against the separate kernels
    (FP64)."""
    case = cases.config_dl_sphere_tfsf(n=40, npml=6, steps=80, record_every=4, radius_nm=10.0)
    dr, dt = case.dr_dt
    ref = None
    for shape in (None, (2, 8, 0), (1, 1, 1), (2, 3, 3), (1, 7, 2), (2, 16, 4)):
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
        if shape is None:
            g.set_step_mode(1)
        else:
            g.set_step_mode(0, *shape)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, case.record_every)
        assert g.last_run_fused() == (shape is not None)
        got = g.read_records()
        g.kill()
        if ref is None:
            ref = got
            assert np.abs(ref).max() > 0
            continue
        r = rel_l2(got, ref)
        print(f"fused shape {shape} vs separate kernels relL2 = {r:.3e}", flush=True)
        assert r <= 1e-13


def test_fused_long_dispersive_list():
    """A dispersive list longer than the post kernel's loads-first limit
    (2^19 cells): a Drude-Lorentz block of 85 x 85 x 81 cells takes the
    per-component loads of the separate kernels' dispersive cells; the
    fused step against the separate kernels (FP64)."""
    from omx_fdtd.units import index_to_nm
    n = (100, 100, 100)
    assert 85 * 85 * 81 > 1 << 19
    dx = 5.0
    pos = [index_to_nm(c, dx, n[a]) for a, c in enumerate((50, 50, 93))]
    src = [("z", 1.0, 1.549802, 0.6, 0.3, pos[0], pos[1], pos[2], 0.0, 0.0, 1.0e6, 0.0)]
    dets = cases.point_detectors([(50, 50, 60), (30, 70, 20), (50, 50, 92)], ("Ex", "Ez", "Hy"))
    case = small_case("long-list", n=n, npml=6, steps=40, media=[("drude-lorentz", "Ag", 1.0)],
                      shapes=[(3, 1, [8.0, 8.0, 10.0, 92.0, 92.0, 90.0])], sources=src, dets=dets, dx=dx)
    dr, dt = case.dr_dt
    out = []
    for mode in (1, 0):
        g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
        g.set_step_mode(mode)
        cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
        g.run(1, case.steps, case.record_every)
        assert g.last_run_fused() == (mode == 0)
        out.append(g.read_records())
        g.kill()
    assert np.abs(out[0]).max() > 0
    r = rel_l2(out[1], out[0])
    print(f"fused vs separate kernels relL2 = {r:.3e}", flush=True)
    assert r <= 1e-13


def test_stepwise_api_with_quantum_sites():
    """Drive the reference step order call by call (mxim_mxll.f90:92-132) with
    a prescribed dPt/dt per site, checking E at the sites and the detectors."""
    n = (20, 18, 16)
    sites = np.array([(8, 9, 8), (9, 9, 8), (12, 6, 10), (8, 9, 8)], dtype=np.int32)
    case = small_case("qsites", n=n, media=[("dielectric", 2.0)], shapes=[(3, 1, [1, 1, 1, 20, 18, 4])],
                      steps=0)
    dr, dt = case.dr_dt
    g, o = both(case)
    dens = 1e-3 / omx_fdtd.NM_TO_AU ** 3
    g.set_q_sites(sites, dens, dt)
    o.set_q_sites(sites, dens, dt)
    rng = np.random.default_rng(7)
    amp = rng.uniform(-1, 1, size=(len(sites), 3)) * 1e-2
    eg_all, eo_all = [], []
    for tt in range(1, 41):
        time = dt * (tt - 0.5)
        for x in (g, o):
            x.time = time
            x.t_skip = time
            x.td_propagate_H_field()
            x.inject_sources(time)
            x.td_propagate_E_field(tt)
        eg_all.append(g.gather_E())
        eo_all.append(o.gather_E())
        dpdt = amp * np.sin(0.3 * tt + np.arange(3))[None, :]
        g.scatter_J(dpdt)
        o.scatter_J(dpdt)
    assert rel_l2(eg_all, eo_all) <= TOL64
    r = rel_l2(g.sample(), o.sample())
    print(f"relL2 = {r:.3e}")
    assert r <= TOL64


def test_two_level_emitters_fp64():
    n = (20, 18, 16)
    sites = np.array([(8, 9, 8), (11, 8, 9), (6, 12, 7)], dtype=np.int32)
    case = small_case("tls", n=n, steps=0)
    dr, dt = case.dr_dt
    g, o = both(case)
    for x in (g, o):
        x.set_q_sites(sites, 1e-3 / omx_fdtd.NM_TO_AU ** 3, dt)
        x.set_two_level(np.full(3, 2.1 * EV_TO_AU), np.full(3, 10.0 * omx_fdtd.DEBYE_TO_AU),
                        np.array([[0, 0, 1.0], [1, 1, 0], [0.2, 0.3, 0.9]]), np.full(3, 0.02 * EV_TO_AU))
    dets = cases.point_detectors([(9, 9, 9)])
    for d in dets:
        g.add_detector(*d)
        o.add_detector(*d)
    g.set_record_capacity(200)
    g.run(1, 150, 1)
    want = o.run(1, 150, 1)
    got = g.read_records()
    r = rel_l2(got, want)
    print(f"relL2 = {r:.3e}")
    assert r <= TOL64
    (dg, rg), (do, ro) = g.two_level_state(), o.two_level_state()
    np.testing.assert_allclose(rg, ro, rtol=1e-11, atol=1e-15)
    assert np.abs(do).max() > 0
    np.testing.assert_allclose(dg, do, rtol=1e-9, atol=1e-12 * np.abs(do).max())


def test_profiling_stride():
    """Events around the H and E launches of every n-th step: the launch
    count and per-launch times of the profiled steps."""
    case = small_case("profiling", steps=10)
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    for every, launches in ((1, 10), (3, 4), (8, 2), (20, 1)):
        g.set_profiling(True, every=every)
        g.run(1, 10, 0)
        mh, me, n = g.profile()
        print(f"every {every}: {n} profiled steps, H {mh / n:.4f} ms, E {me / n:.4f} ms")
        assert n == launches and mh > 0 and me > 0
    g.set_profiling(False)
    g.run(11, 5, 0)
    assert g.profile()[2] == 0
    with pytest.raises(omx_fdtd.OmxError):
        g.set_profiling(True, every=-1)
    g.kill()


def test_field_copy_roundtrip_and_energy():
    case = small_case("fields", n=(14, 12, 10), steps=0)
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=8)
    rng = np.random.default_rng(3)
    tot = []
    for name in G.FIELD_NAMES:
        f = g.get_field(name)
        v = rng.uniform(-1, 1, size=f.shape)
        g.set_field(name, v)
        np.testing.assert_array_equal(g.get_field(name), v)
        tot.append(v)
    w64, w32 = g.energy()
    s = 0.0
    import math
    terms = []
    for q, v in enumerate(tot):
        # interior cells 1..n of each reference array
        lo = [0 if (q < 3 and a == q) else 1 for a in range(3)] if q < 3 else \
             [1 if a == q - 3 else 0 for a in range(3)]
        core = v[tuple(slice(1 - lo[a], 1 - lo[a] + (g.nx, g.ny, g.nz)[a]) for a in range(3))]
        c = omx_fdtd.EPS0 if q < 3 else omx_fdtd.MU0
        terms.extend((c * core.ravel() ** 2).tolist())
    want = 0.5 * dr ** 3 * math.fsum(terms)
    assert abs(w64 - want) <= 1e-12 * want
    assert abs(w32 - want) <= 1e-6 * want


@pytest.mark.parametrize("mode", [1, 0])
@pytest.mark.parametrize("precision", [8, 4])
def test_tfsf_dispersive_cells_on_the_faces(precision, mode):
    """A Drude-Lorentz sphere across the k1 and i1 faces of the TF/SF box:
    the E corrections of its cells cannot ride in the E tiles or in the
    fused step's post kernel (the dispersive workgroups update those cells),
    so the E side of the TF/SF runs after the E launch (mode 1) or after the
    fused step (mode 0) while the H side stays inside."""
    n = (40, 40, 60)
    lo, hi = (10, 10, 12), (30, 30, 48)
    case = cases.Case(name="tfsf-disp", grid=n, dx_nm=2.0, npml=6, boundaries=(BC_CPML,) * 3, steps=300,
                      record_every=1, media=[("drude-lorentz", "Ag", 1.0)],
                      shapes=[(1, 1, [29.0, 20.0, 47.0, 4.5])],
                      tfsf=dict(lo=lo, hi=hi, e0=1.0, w0=2.066 * EV_TO_AU, t0=0.6 * FS_TO_AU, tau=0.2 * FS_TO_AU))
    case.detectors = cases.point_detectors([(20, 20, 40), (29, 20, 52), (34, 20, 44), (20, 20, 30)],
                                           ("Ex", "Ez", "Hy"))
    g, o = both(case, precision)
    g.set_step_mode(mode)
    g.run(1, case.steps, case.record_every)
    assert g.last_run_fused() == (mode == 0)
    got = g.read_records()
    want = o.run(1, case.steps, case.record_every)
    assert np.abs(want).max() > 0
    r = rel_l2(got, want)
    print(f"FP{8 * precision} relL2 vs oracle = {r:.3e}")
    assert r <= (TOL64 if precision == 8 else TOL32)


def test_tfsf_vacuum_leakage():
    n = (40, 40, 60)
    lo, hi = (10, 10, 12), (30, 30, 48)
    case = cases.Case(name="tfsf", grid=n, dx_nm=2.0, npml=6, boundaries=(BC_CPML,) * 3, steps=500,
                      record_every=1,
                      tfsf=dict(lo=lo, hi=hi, e0=1.0, w0=2.066 * EV_TO_AU, t0=0.6 * FS_TO_AU, tau=0.2 * FS_TO_AU))
    inside = (20, 20, 40)
    outside = [(20, 20, 52), (20, 4 + 2, 30), (34, 20, 30), (20, 20, 9)]
    case.detectors = cases.point_detectors([inside] + outside, ("Ex",))
    g, o, got, want = run_both(case)
    r = rel_l2(got, want)
    print(f"relL2 vs oracle = {r:.3e}")
    assert r <= TOL64
    peak = np.abs(got[:, 0]).max()
    leak = np.abs(got[:, 1:]).max()
    print(f"peak inside {peak:.3e}, max outside {leak:.3e}")
    assert peak > 1e-3
    assert leak <= 1e-3 * peak
