"""Benchmark of the MI355X FDTD engine on the 3D dispersive-media workload:
256^3 Yee grid, dx = 2 nm, CPML 12 cells on every face, one seeded
Drude-Lorentz sphere (eps_inf 5, Drude 9 eV / 0.07 eV, two seeded Lorentz
poles) in vacuum, x-polarised TF/SF plane wave along +z (600 nm carrier),
plane detectors Ex, Ey, Ez on the transmitted side.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full Yee step (H and E updates with CPML, the Drude-Lorentz
ADE, TF/SF corrections and detector recording at its cadence) over the whole
grid, inputs resident in HBM.  The FP32 engine is timed (the headline), then
the FP64 engine on the same steps; the FP32 detector series is compared
with the FP64 one (relative L2).  The dominant kernel's roofline fraction
comes from HIP events recorded around each of its launches in the timed
steps, on the stream the kernels run on.  The
CPU baseline is the reference's own Fortran/OpenMP solver (built from its
sources into oracle/_ref) timed on a bounded sample on this host.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "one-mxll-to-rule-them-all_amd"), ROOT]

import omx_fdtd  # noqa: E402
from omx_fdtd import cases, grid as G  # noqa: E402
from omx_fdtd.units import index_to_nm, MEDIUM_DL  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOAD = "dl_sphere_tfsf_256"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_grid(case, precision):
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    return g


def algorithmic_bytes(case, g, w):
    """Bytes per H and per E launch: every live array read once and written
    once (per cell: H 9 words, E 9 words), plus 4 words per CPML slab cell
    (two psi arrays, read and written), plus per Drude-Lorentz component
    cell 4 + 4P words (E_old, PD, PL_k, PL_old_k read and written), plus the
    one-byte cell tag when the grid holds media."""
    nx, ny, nz = case.grid
    N = nx * ny * nz
    p = case.npml
    slab_h = 2 * (p - 1) * (ny * nz + nx * nz + nx * ny)   # cells in H CPML slabs, per axis pair
    slab_e = 2 * (p - 1) * (ny * nz + nx * nz + nx * ny)   # E slabs: 2..npml and n+1-npml..n-1
    mm = g.media_map()
    P = case.media[0][4] and len(case.media[0][4]) or 0
    dl_comp = int((mm > 0).sum())
    b_h = w * (9 * N + 4 * slab_h)
    b_e = w * (9 * N + 4 * slab_e + (4 + 4 * P) * dl_comp) + N
    # fused pass: E and H each read once and written once (12 words), psi of
    # both, dispersive state, tag
    b_f = w * (12 * N + 4 * slab_h + 4 * slab_e + (4 + 4 * P) * dl_comp) + N
    return b_h, b_e, b_f, dl_comp


def timed_run(g, first, steps, record_every):
    g.synchronize()
    t0 = time.perf_counter()
    g.run(first, steps, record_every)
    g.synchronize()
    return time.perf_counter() - t0


def cpu_reference_baseline(case, steps):
    """Times the reference's own solver (oracle/_ref/omx_ref_driver) on
    `steps` steps of the same grid on this host's cores."""
    drv = os.path.join(ROOT, "oracle", "_ref", "omx_ref_driver")
    if not os.path.exists(drv):
        return None
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    threads = max(1, min(ncpu, 16))
    nx, ny, nz = case.grid
    c = case.info["sphere_centre_cells"]
    r = case.info["sphere_radius_nm"] / case.dx_nm
    d = tempfile.mkdtemp(prefix="omx_cpu_")
    with open(os.path.join(d, "medium_0001.in"), "w") as f:
        f.write("drude-lorentz\nAg 5.0\n")
        I, J, K = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij",
                              sparse=True)
        sel = np.nonzero((I - c[0]) ** 2 + (J - c[1]) ** 2 + (K - c[2]) ** 2 <= r * r)
        for (i, j, k) in zip(*sel):
            f.write(f"{index_to_nm(i + 1, case.dx_nm, nx)!r} {index_to_nm(j + 1, case.dx_nm, ny)!r} "
                    f"{index_to_nm(k + 1, case.dx_nm, nz)!r}\n")
    with open(os.path.join(d, "sources.in"), "w") as f:
        f.write("x 1.0 2.066 5.0 1.5 0.0 0.0 0.0 0.0 0.0 1.0e6 0.0\n")
    with open(os.path.join(d, "case.nml"), "w") as f:
        f.write(f"&case\n nx={nx}, ny={ny}, nz={nz}, npml={case.npml}, bc=3,3,3, n_steps={steps},\n"
                f" record_every=0, n_media=1, n_src=1, n_det=0, n_q=0, timing=1, dx_nm={case.dx_nm!r}\n/\n")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores", OMP_PROC_BIND="spread")
    p = subprocess.run([drv], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=600)
    m = re.search(r"REF_TIMING steps=(\d+) seconds=\s*([0-9.Ee+-]+) threads=(\d+)", p.stdout)
    if p.returncode != 0 or not m:
        log("cpu baseline failed:", p.stdout[-2000:])
        return None
    secs = float(m.group(2))
    cpu_model = ""
    try:
        cpu_model = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        cpu_model = re.search(r"Model name:\s*(.*)", cpu_model).group(1).strip()
    except Exception:
        pass
    return {"value": nx * ny * nz * steps / secs / 1e9, "unit": "Gcell-updates/s", "cores": int(m.group(3)),
            "kind": "reference",
            "sample": f"{steps} steps of the {nx}x{ny}x{nz} grid (CPML {case.npml}, same sphere cells as a "
                      f"library Ag Drude-Lorentz medium, one point source, no TF/SF, no detector I/O): "
                      f"the reference's Fortran/OpenMP td_propagate_H/E_3Dfield built with amdflang -O2, "
                      f"{secs:.2f} s on {cpu_model}"}


def copy_calibration(mib):
    """A plain device-to-device copy of a known size: its FETCH_SIZE and
    WRITE_SIZE counters give the factor that turns the counters of the FDTD
    kernels into bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert bool(torch.equal(dst[-16:], src[-16:]))
    log(f"[bench] copy calibration: 5 copies of {mib} MiB ({n} float32)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=("separate", "fused"), default="separate",
                    help="separate H and E kernels, or the fused H+E pass")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU slab "
              f"decomposition is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)
    W, K = args.warmup, args.steps
    case = cases.config_dl_sphere_tfsf(n=args.n, steps=W + K)
    nx, ny, nz = case.grid
    N = nx * ny * nz

    results = {}
    series = {}
    for prec in ((4,) if args.no_fp64 else (4, 8)):
        log(f"[bench] building FP{8 * prec} grid {case.grid}")
        g = build_grid(case, prec)
        g.set_step_mode(1 if args.mode == "separate" else 0)
        if W:
            g.run(1, W, case.record_every)
        # HIP events around every H and E launch of the timed steps, recorded
        # on the engine's stream
        g.set_profiling(True)
        secs = timed_run(g, W + 1, K, case.record_every)
        ms_h, ms_e, n = g.profile()
        g.set_profiling(False)
        series[prec] = g.read_records()
        b_h, b_e, b_f, dl_comp = algorithmic_bytes(case, g, prec)
        fused = g.last_run_fused()
        results[prec] = dict(secs=secs, gcells=N * K / secs / 1e9, ms_step=1e3 * secs / K,
                             ms_h=ms_h / n, ms_e=ms_e / n, b_h=b_h, b_e=b_e, b_f=b_f, dl_comp=dl_comp,
                             fused=fused, dev_bytes=g.device_bytes())
        log(f"[bench] FP{8 * prec}: {results[prec]['gcells']:.2f} Gcell/s, {results[prec]['ms_step']:.3f} ms/step, "
            f"H {results[prec]['ms_h']:.3f} ms, E {results[prec]['ms_e']:.3f} ms")
        g.kill()

    r32 = results[4]
    rel = None
    if 8 in series:
        a, b = series[4], series[8]
        rel = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    if r32["fused"]:
        dom, kname, b_dom, ms_dom = "e", "k_step_fused", r32["b_f"], r32["ms_e"]
        step_bytes = r32["b_f"]
    else:
        dom = "e" if r32["ms_e"] >= r32["ms_h"] else "h"
        kname = "k_e_update" if dom == "e" else "k_h_update"
        b_dom, ms_dom = r32["b_" + dom], r32["ms_" + dom]
        step_bytes = r32["b_h"] + r32["b_e"]
    achieved = b_dom / (ms_dom * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU reference baseline: {args.cpu_steps} steps")
        cpu = cpu_reference_baseline(case, args.cpu_steps)
    out = {
        "metric": "Gcell-updates/s, FP32 3D Drude-Lorentz + CPML FDTD (TF/SF plane wave)",
        "value": round(r32["gcells"], 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(r32["ms_step"], 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": WORKLOAD, "grid": list(case.grid), "dx_nm": case.dx_nm, "npml": case.npml,
                   "sphere_radius_nm": round(case.info["sphere_radius_nm"], 4),
                   "dl_component_cells": r32["dl_comp"], "record_every": case.record_every,
                   "detectors": "plane_xy Ex, Ey, Ez at k=%d" % case.info["detector_k"],
                   "step_kernels": "fused" if r32["fused"] else "separate"},
        "roofline": {"bound": "hbm", "kernel": kname,
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": b_dom, "avg_launch_ms": round(ms_dom, 5)},
        "step_roofline_frac": round(step_bytes / (r32["ms_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "cpu_baseline": cpu,
    }
    if 8 in results:
        r64 = results[8]
        b64 = r64["b_f"] if r64["fused"] else r64["b_h"] + r64["b_e"]
        out["fp64"] = {"value": round(r64["gcells"], 4), "ms_per_step": round(r64["ms_step"], 5),
                       "step_roofline_frac": round(b64 / (r64["ms_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        out["fp32_over_fp64"] = round(r32["gcells"] / r64["gcells"], 3)
        out["fp32_vs_fp64_detector_relL2"] = rel
    if cpu:
        out["gpu_over_cpu"] = round(r32["gcells"] / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
