"""Benchmark of the MI355X FDTD engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload dl_sphere_tfsf_256]

Default workload (the headline): 256^3 Yee grid, dx = 2 nm, CPML 12 cells on
every face, one seeded Drude-Lorentz sphere (eps_inf 5, Drude 9 eV / 0.07
eV, two seeded Lorentz poles) in vacuum, x-polarised TF/SF plane wave along
+z (600 nm carrier), plane detectors Ex, Ey, Ez on the transmitted side.

A step is one full Yee step (H and E updates with CPML, the Drude-Lorentz
ADE, TF/SF corrections and detector recording at its cadence) over the whole
grid, inputs resident in HBM.  Before the timed run of a single-grid workload
the engine's tuners choose, on the grid itself, the tile shape and then the
step kernels (separate H and E launches, or the fused H+E pass and its
workgroup shape); the choice is recorded in the JSON line (--mode separate /
fused forces a step mode at the default tiles).  The FP32 engine is timed (the headline), then
the FP64 engine on the same steps; the FP32 detector series is compared
with the FP64 one (relative L2).  The dominant kernel's roofline fraction
comes from HIP events recorded around its launches in every PROF_EVERY-th
timed step, on the stream the kernels run on (an event pair costs the
stream a few microseconds, so events on every step would slow the timed
run by several percent); its HBM traffic from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary of the same kernel and workload.
The CPU baseline is the reference's own Fortran/OpenMP solver (built from
its sources into oracle/_ref) timed on a bounded sample on this host.

Other workloads (each prints its own JSON line):
  ellipsoids_768x768x512   64 DL ellipsoids over a dielectric substrate,
                           energy integral every 100 steps
  batched_64               64 independent grids (edges 96-192) in one batch,
                           against the same grids one by one and against one
                           large grid of the same cell count
  emitters_512x512x384     Ag slab with a nanohole array, 4096 two-level
                           emitters coupled to the field
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
from omx_fdtd.units import index_to_nm, MEDIUM_DRUDE, MEDIUM_DL, BC_CPML  # noqa: E402

HBM_PEAK_GBS = 8000.0
PROF_EVERY = 8          # timed steps per profiled step (HIP events around its H and E launches)
HEADLINE = "dl_sphere_tfsf_256"
METRIC = "Gcell-updates/s, FP32 3D Drude-Lorentz + CPML FDTD (TF/SF plane wave)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PATCH_RADIUS = None   # --patch-radius: FP64 near-field box radius of FP32 grids (None: the engine default)
FORCED_TILES = None        # --tiles rows,xchunk,order: tile shape of the separate kernels (profiler runs)
FORCED_FUSED_SHAPE = None  # --fused-shape rows,xchunk,ywaves: the fused step's tile (profiler runs)


TUNE_STEPS = 16        # timed steps per candidate of the tile and step-kernel tuners


def build_grid(case, precision, mode="separate"):
    """A grid set up for `case`.  mode "auto": the project's tuners pick the
    tile shape (omx3d_autotune) and then the step kernels, separate H/E or
    fused, with the fused tile shape (omx3d_autotune_step), on the grid
    itself; "separate" / "fused": that step mode at the default tiles."""
    dr, dt = case.dr_dt
    g = omx_fdtd.Grid3D(case.grid, case.npml, case.boundaries, dt, dr, precision=precision)
    g.set_step_mode(0 if mode == "fused" else 1)
    if FORCED_TILES is not None:
        g.set_tiling(*FORCED_TILES)
    if FORCED_FUSED_SHAPE is not None:
        g.set_step_mode(0 if mode == "fused" else 1, *FORCED_FUSED_SHAPE)
    if PATCH_RADIUS is not None:
        g.set_source_patch(PATCH_RADIUS)
    cases.apply(case, g, library=G.library_medium, make_medium=G.medium)
    g.tuned = None
    if mode == "auto":
        t0 = time.perf_counter()
        rows, xc, order, _ = g.autotune(steps=TUNE_STEPS)
        m, t_sep, t_fus = g.autotune_step(steps=TUNE_STEPS)
        g.tuned = {"tiling": [rows, xc, order], "step_kernels": "fused" if m == 0 else "separate",
                   "fused_shape": list(g.fused_shape()) if m == 0 else None,
                   "tuner_ms_per_step": {"separate": round(t_sep, 5),
                                         "fused": None if t_fus is None else round(t_fus, 5)},
                   "tuning_s": round(time.perf_counter() - t0, 2)}
        log(f"[bench] FP{8 * precision} tuned: tiles {rows}x{xc} order {order}, "
            f"{g.tuned['step_kernels']} {g.tuned['fused_shape'] or ''} (separate {t_sep:.4f} ms, fused {t_fus} ms)")
    return g


def resolved_media(case):
    out = []
    for spec in case.media:
        m = cases.medium_struct(spec, G.medium)
        if isinstance(m, tuple) and m[0] == "library":
            m = G.library_medium(m[1], m[2])
        out.append(m)
    return out


def algorithmic_bytes(case, g, w):
    """Bytes per H and per E half step: every live array read once and
    written once (per cell: H 9 words, E 9 words), plus 4 words per CPML slab
    cell and half step (two psi arrays, read and written), plus per
    dispersive component cell 2 words (Drude: PD) or 4 + 4P words
    (Drude-Lorentz with P poles: E_old, PD, PL_k, PL_old_k), all read and
    written, plus the one-byte cell tag when the grid holds media.  Returns
    (H bytes, E bytes, dispersive-state bytes, dispersive component cells).
    The separate E launch (tiles and dispersive cells) moves the E bytes;
    the fused step moves H and E bytes but its dispersive cells run in a
    launch of their own, so the fused kernel moves H + E minus the
    dispersive state."""
    nx, ny, nz = case.grid
    N = nx * ny * nz
    p = case.npml
    slab = 0
    for n_a, bc in zip(case.grid, case.boundaries):
        if bc == BC_CPML:
            slab += 2 * (p - 1) * (N // n_a)
    disp = 0
    comp_cells = 0
    if case.media:
        mm = g.media_map()
        for m_idx, m in enumerate(resolved_media(case), start=1):
            cnt = int((mm == m_idx).sum())
            if m.kind == MEDIUM_DRUDE:
                disp += 2 * cnt
                comp_cells += cnt
            elif m.kind == MEDIUM_DL:
                disp += (4 + 4 * m.n_poles) * cnt
                comp_cells += cnt
    tag = N if case.media else 0
    b_h = w * (9 * N + 4 * slab)
    b_e = w * (9 * N + 4 * slab + disp) + tag
    return b_h, b_e, w * disp, comp_cells


def committed_traffic(workload, dtype, step_kernels, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by the
    factor a plain copy of known size gives), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}_{dtype}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"```json\n(.*?)\n```", open(path).read(), re.S)
    if not m:
        return None
    d = json.loads(m.group(1))
    if d.get("step_kernels") != step_kernels or kernel not in d.get("kernels", {}):
        return None
    return d["kernels"][kernel]["bytes"]


def timed_run(g, first, steps, record_every):
    g.synchronize()
    t0 = time.perf_counter()
    g.run(first, steps, record_every)
    g.synchronize()
    return time.perf_counter() - t0


def rel_l2(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else None


def host_cpu():
    """Host cores this process may use, the machine's logical CPUs and the CPU model."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    model = ""
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True).stdout
        model = re.search(r"Model name:\s*(.*)", out).group(1).strip()
    except Exception:
        pass
    return affinity, os.cpu_count() or affinity, model


def cpu_reference_baseline(case, steps):
    """Times the reference's own solver (oracle/_ref/o3/omx_ref_driver, its
    Fortran sources built with amdflang -O3 -fopenmp) on `steps` steps of the
    same grid, with one OpenMP thread per host core this process may use."""
    drv = os.path.join(ROOT, "oracle", "_ref", "o3", "omx_ref_driver")
    if not os.path.exists(drv):
        return None
    affinity, logical, cpu_model = host_cpu()
    threads = affinity
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
    return {"value": nx * ny * nz * steps / secs / 1e9, "unit": "Gcell-updates/s", "cores": int(m.group(3)),
            "kind": "reference", "threads": int(m.group(3)), "host_cores": affinity, "host_logical_cpus": logical,
            "cpu_model": cpu_model, "build": "amdflang -O3 -fopenmp",
            "sample": f"{steps} steps of the {nx}x{ny}x{nz} grid (CPML {case.npml}, same sphere cells as a "
                      f"library Ag Drude-Lorentz medium, one point source, no TF/SF, no detector I/O): "
                      f"the reference's Fortran/OpenMP td_propagate_H/E_3Dfield built with amdflang -O3, "
                      f"{int(m.group(3))} OpenMP threads (OMP_PLACES=cores, OMP_PROC_BIND=spread), "
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


def roofline(case, g, prec, res, workload):
    """Roofline object of the dominant kernel of a run (res: ms_h, ms_e,
    fused), and the algorithmic bytes of a whole step."""
    b_h, b_e, b_disp, _ = algorithmic_bytes(case, g, prec)
    dtype = "fp32" if prec == 4 else "fp64"
    if res["fused"]:
        kname, b_dom, ms_dom = "k_step_wg", b_h + b_e - b_disp, res["ms_e"]
    elif res["ms_e"] >= res["ms_h"]:
        kname, b_dom, ms_dom = "k_e_update", b_e, res["ms_e"]
    else:
        kname, b_dom, ms_dom = "k_h_update", b_h, res["ms_h"]
    achieved = b_dom / (ms_dom * 1e-3) / 1e9
    traffic = committed_traffic(workload, dtype, "fused" if res["fused"] else "separate", kname)
    return {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": b_dom, "avg_launch_ms": round(ms_dom, 5)}, b_h + b_e


def run_single(case, prec, W, K, mode, energy_every=0):
    """Warm-up W steps, then K timed steps with HIP events around the H and E
    launches of every PROF_EVERY-th step.  With energy_every > 0 the volume energy integral is taken
    every energy_every steps inside the timed region."""
    log(f"[bench] building FP{8 * prec} grid {case.grid}")
    g = build_grid(case, prec, mode)
    if W:
        g.run(1, W, case.record_every)
    g.set_profiling(True, every=PROF_EVERY)
    energies = []
    if energy_every > 0:
        g.synchronize()
        t0 = time.perf_counter()
        tt = W + 1
        while tt < W + 1 + K:
            n = min(energy_every, W + 1 + K - tt)
            g.run(tt, n, case.record_every)
            tt += n
            energies.append(g.energy()[0])
            if len(energies) % 10 == 0:
                log(f"[bench] FP{8 * prec}: step {tt - 1} of {W + K}")
        g.synchronize()
        secs = time.perf_counter() - t0
    else:
        secs = timed_run(g, W + 1, K, case.record_every)
    ms_h, ms_e, n = g.profile()
    g.set_profiling(False)
    res = dict(secs=secs, gcells=int(np.prod(case.grid)) * K / secs / 1e9, ms_step=1e3 * secs / K,
               ms_h=ms_h / max(n, 1), ms_e=ms_e / max(n, 1), fused=g.last_run_fused(), tuned=g.tuned,
               series=g.read_records(), energies=np.array(energies), dev_bytes=g.device_bytes())
    log(f"[bench] FP{8 * prec}: {res['gcells']:.2f} Gcell/s, {res['ms_step']:.3f} ms/step, "
        f"H {res['ms_h']:.3f} ms, E {res['ms_e']:.3f} ms")
    return g, res


def base_record(workload, K, W, value, ms_step, dtype, config):
    return {"metric": METRIC if workload == HEADLINE else f"Gcell-updates/s, FP32 3D FDTD ({workload})",
            "value": round(value, 4), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": K, "warmup": W,
            "ms_per_step": round(ms_step, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic",
            "config": dict({"workload": workload}, **config,
                           **({} if PATCH_RADIUS is None else {"patch_radius": PATCH_RADIUS}))}


def precisions(args):
    return (4,) if args.no_fp64 else (4, 8)


def add_fp64(out, results, r32):
    if 8 in results:
        r64 = results[8]
        out["fp64"] = {"value": round(r64["gcells"], 4), "ms_per_step": round(r64["ms_step"], 5),
                       "step_roofline_frac": round(r64["step_bytes"] / (r64["ms_step"] * 1e-3) / 1e9
                                                   / HBM_PEAK_GBS, 4),
                       "step_kernels": "fused" if r64["fused"] else "separate"}
        if r64.get("tuned"):
            out["fp64"]["tuning"] = r64["tuned"]
        out["fp32_over_fp64"] = round(r32["gcells"] / r64["gcells"], 3)


FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def validation_run(case, prec, mode, fields=True):
    """The whole workload (case.steps steps from step 1) at one precision:
    detector series, the final fields (all six, reference bounds) and the
    engine for further reads."""
    g = build_grid(case, prec, mode)
    g.run(1, case.steps, case.record_every)
    out = {"series": g.read_records()}
    if fields:
        out["fields"] = np.concatenate([g.get_field(f).ravel() for f in FIELDS])
    return g, out


def fp32_error(case, mode, fields=True, series=None):
    """FP32 against FP64 over the workload's full length (case.steps steps,
    whatever the timed window is): relative L2 of the detector series and of
    the six final fields.  `series` (prec -> recorded series) reuses timed
    runs that already covered the full length.  Raises when the detectors
    saw no signal, so that no result goes out without a stated error."""
    res = {}
    for prec in (4, 8):
        if series is not None and not fields:
            res[prec] = {"series": series[prec]}
            continue
        log(f"[bench] validation: FP{8 * prec}, {case.steps} steps")
        g, res[prec] = validation_run(case, prec, mode, fields)
        g.kill()
    det = rel_l2(res[4]["series"], res[8]["series"])
    if det is None:
        raise RuntimeError(f"validation window of {case.steps} steps recorded no signal at the detectors")
    out = {"steps": case.steps, "detector_relL2": det, "detector_samples": int(res[8]["series"].size)}
    if fields:
        out["field_relL2"] = rel_l2(res[4]["fields"], res[8]["fields"])
    log(f"[bench] validation: FP32 vs FP64 detector relL2 {det:.3e}"
        + (f", final-field relL2 {out['field_relL2']:.3e}" if fields else ""))
    return out


def single_grid_runs(case, args, W, K, workload, energy_every=0, extra=None):
    results = {}
    for prec in precisions(args):
        g, res = run_single(case, prec, W, K, args.mode, energy_every)
        res["roof"], res["step_bytes"] = roofline(case, g, prec, res, workload)
        res["comp_cells"] = algorithmic_bytes(case, g, prec)[3]
        if extra:
            res.update(extra(g))
        g.kill()
        results[prec] = res
    return results


def finish(out, results):
    r32 = results[4]
    if r32.get("tuned"):
        out["config"]["tuning"] = r32["tuned"]
    out["roofline"] = r32["roof"]
    out["step_roofline_frac"] = round(r32["step_bytes"] / (r32["ms_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    add_fp64(out, results, r32)
    return out


HEADLINE_STEPS = 4000   # the workload's length: the pulse crosses the detector plane within it


def bench_headline(args):
    W, K = args.warmup, args.steps
    case = cases.config_dl_sphere_tfsf(n=args.n, steps=W + K)
    results = single_grid_runs(case, args, W, K, HEADLINE)
    r32 = results[4]
    full = cases.config_dl_sphere_tfsf(n=args.n, steps=max(HEADLINE_STEPS, W + K))
    val = SKIPPED if args.no_validation else fp32_error(full, args.mode)
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU reference baseline: {args.cpu_steps} steps")
        cpu = cpu_reference_baseline(case, args.cpu_steps)
    out = base_record(HEADLINE, K, W, r32["gcells"], r32["ms_step"], "fp32",
                      {"grid": list(case.grid), "dx_nm": case.dx_nm, "npml": case.npml,
                       "sphere_radius_nm": round(case.info["sphere_radius_nm"], 4),
                       "dl_component_cells": r32["comp_cells"], "record_every": case.record_every,
                       "detectors": "plane_xy Ex, Ey, Ez at k=%d" % case.info["detector_k"],
                       "step_kernels": "fused" if r32["fused"] else "separate"})
    finish(out, results)
    out["fp32_vs_fp64_detector_relL2"] = val["detector_relL2"]
    out["fp32_vs_fp64_validation"] = val
    out["cpu_baseline"] = cpu
    if cpu:
        out["gpu_over_cpu"] = round(r32["gcells"] / cpu["value"], 1)
    return out


# --no-validation (profiler runs, whose counters slow every launch): no
# FP32-vs-FP64 validation window, so the record states no error
SKIPPED = {"detector_relL2": None, "skipped": "--no-validation"}


def workload_error(args, W, K, results, full_case):
    """FP32 vs FP64 detector error over the workload's full length: from the
    timed runs when they covered it (warm-up and timed steps record from step
    1), otherwise from validation runs of the full length."""
    if 8 in results and W + K >= full_case.steps:
        det = rel_l2(results[4]["series"], results[8]["series"])
        if det is None:
            raise RuntimeError("the detectors recorded no signal")
        return {"steps": W + K, "detector_relL2": det, "detector_samples": int(results[8]["series"].size)}
    return fp32_error(full_case, args.mode, fields=False)


def bench_ellipsoids(args):
    W = args.warmup
    K = args.steps if args.steps_given else 10000
    case = cases.config_ellipsoids(steps=W + K)
    results = single_grid_runs(case, args, W, K, "ellipsoids_768x768x512", energy_every=100)
    r32 = results[4]
    out = base_record("ellipsoids_768x768x512", K, W, r32["gcells"], r32["ms_step"], "fp32",
                      {"grid": list(case.grid), "dx_nm": case.dx_nm, "npml": case.npml, "ellipsoids": 64,
                       "dl_component_cells": r32["comp_cells"], "substrate_k": case.info["substrate_k"],
                       "detectors": "32 point Ez, every 100 steps", "energy_every": 100,
                       "step_kernels": "fused" if r32["fused"] else "separate"})
    finish(out, results)
    val = SKIPPED if args.no_validation else workload_error(args, W, K, results,
                                                            cases.config_ellipsoids(steps=max(10000, W + K)))
    out["fp32_vs_fp64_detector_relL2"] = val["detector_relL2"]
    out["fp32_vs_fp64_validation"] = val
    out["device_bytes_fp32"] = r32["dev_bytes"]
    if 8 in results:
        out["fp32_vs_fp64_energy_relL2"] = rel_l2(r32["energies"], results[8]["energies"])
    return out


def bench_emitters(args):
    W = args.warmup
    K = args.steps if args.steps_given else 20000
    case = cases.config_emitters(steps=W + K)

    def tls(g):
        dip, rho = g.two_level_state()
        return {"tls": np.concatenate([dip.ravel(), rho.ravel()])}

    results = single_grid_runs(case, args, W, K, "emitters_512x512x384", extra=tls)
    r32 = results[4]
    out = base_record("emitters_512x512x384", K, W, r32["gcells"], r32["ms_step"], "fp32",
                      {"grid": list(case.grid), "dx_nm": case.dx_nm, "npml": case.npml, "emitters": 4096,
                       "slab_k": case.info["slab_k"], "holes": case.info["holes"],
                       "dl_component_cells": r32["comp_cells"], "detectors": "8 point Ex, Ey, Ez, every 100 steps",
                       "step_kernels": "fused" if r32["fused"] else "separate"})
    finish(out, results)
    val = SKIPPED if args.no_validation else workload_error(args, W, K, results,
                                                            cases.config_emitters(steps=max(20000, W + K)))
    out["fp32_vs_fp64_detector_relL2"] = val["detector_relL2"]
    out["fp32_vs_fp64_validation"] = val
    if 8 in results:
        out["fp32_vs_fp64_emitter_state_relL2"] = rel_l2(r32["tls"], results[8]["tls"])
    return out


def batched_as_one_grid(cs):
    """One cubic grid with the cell count of the batch and the same media:
    block q of a 4x4x4 arrangement carries the spheres of grid q (positions
    scaled into the block), one dipole source per block, one point detector
    per block."""
    cells = sum(int(np.prod(c.grid)) for c in cs)
    edge = int(round(cells ** (1.0 / 3.0)))
    nb = 4
    blk = edge / nb
    shapes, srcs, dets = [], [], []
    c0 = cs[0]
    for q, c in enumerate(cs[: nb ** 3]):
        o = np.array([q % nb, (q // nb) % nb, q // (nb * nb)]) * blk
        sc = blk / np.array(c.grid, dtype=float)
        for (typ, med, p) in c.shapes:
            ctr = o + np.array(p[:3]) * sc
            shapes.append((typ, med, ctr.tolist() + [float(p[3]) * float(sc.min())]))
        cell = (o + blk / 2).astype(int)
        pos = [index_to_nm(int(cell[a]), c.dx_nm, edge) for a in range(3)]
        srcs.append((c.sources[2][0], c.sources[2][1]) + c.sources[2][2:5] + tuple(pos) + c.sources[2][8:])
        dets.append(tuple(int(v) + 3 for v in cell))
    return cases.Case(name="batch_as_one", grid=(edge, edge, edge), dx_nm=c0.dx_nm, npml=c0.npml,
                      boundaries=c0.boundaries, steps=c0.steps, record_every=c0.record_every, media=c0.media[:1],
                      shapes=shapes, sources=srcs, detectors=cases.point_detectors(dets))


def bench_batched(args):
    W = args.warmup
    K = args.steps if args.steps_given else 2000
    cs = cases.config_batched(steps=W + K)
    cells = sum(int(np.prod(c.grid)) for c in cs)
    out_prec = {}
    for prec in precisions(args):
        log(f"[bench] batch of {len(cs)} grids, {cells} cells, FP{8 * prec}")
        grids = [build_grid(c, prec) for c in cs]
        # each grid's planes and tile order tuned on its own at the batch's
        # common row count (one workgroup height for all grids; the batch
        # keeps every grid's tiling)
        if not args.no_autotune:
            rows = grids[0].tiling()[0]
            tiles = [g.autotune(steps=2, rows=rows)[:3] for g in grids]
            log(f"[bench] per-grid tilings (rows, xchunk, order): {sorted(set(tiles))}")
        b = omx_fdtd.Batch(grids)
        if W:
            b.run(1, W, cs[0].record_every)
        for g in grids:
            g.synchronize()
        t0 = time.perf_counter()
        b.run(W + 1, K, cs[0].record_every)
        for g in grids:
            g.synchronize()
        secs = time.perf_counter() - t0
        series = [g.read_records() for g in grids]
        b.kill()
        for g in grids:
            g.kill()
        out_prec[prec] = dict(gcells=cells * K / secs / 1e9, ms_step=1e3 * secs / K, series=series)
        log(f"[bench] batched FP{8 * prec}: {out_prec[prec]['gcells']:.2f} Gcell/s")
    # the same grids one after another (FP32, a bounded number of steps)
    Ks = min(K, 200)
    t_seq = 0.0
    for c in cs:
        g = build_grid(c, 4)
        g.run(1, 2, c.record_every)
        t_seq += timed_run(g, 3, Ks, c.record_every)
        g.kill()
    seq = cells * Ks / t_seq / 1e9
    log(f"[bench] one by one FP32: {seq:.2f} Gcell/s")
    # one large grid of the same cell count holding the same content: a 4x4x4
    # arrangement of blocks, block q carrying grid q's spheres (CPML 8 outside)
    big = batched_as_one_grid(cs)
    edge = big.grid[0]
    g = build_grid(big, 4)
    g.run(1, 2, big.record_every)
    t_big = timed_run(g, 3, Ks, big.record_every)
    g.kill()
    single = edge ** 3 * Ks / t_big / 1e9
    log(f"[bench] single {edge}^3 grid FP32: {single:.2f} Gcell/s")
    # the same large grid with the tuners' choice (which may be the fused
    # step, a kernel the batch does not run)
    g = build_grid(big, 4, "auto")
    big_tuning = g.tuned
    g.run(1, 2, big.record_every)
    t_big_t = timed_run(g, 3, Ks, big.record_every)
    g.kill()
    single_t = edge ** 3 * Ks / t_big_t / 1e9
    log(f"[bench] single {edge}^3 grid FP32, tuned: {single_t:.2f} Gcell/s")
    r32 = out_prec[4]
    out = base_record("batched_64", K, W, r32["gcells"], r32["ms_step"], "fp32",
                      {"grids": len(cs), "edges": [96, 128, 160, 192], "total_cells": cells, "npml": 8,
                       "record_every": cs[0].record_every, "graph": True,
                       "tiling": "defaults" if args.no_autotune else "autotuned per grid"})
    out["one_by_one_gcells"] = round(seq, 4)
    out["single_grid_gcells"] = round(single, 4)
    out["single_grid_edge"] = edge
    out["batched_over_single_grid"] = round(r32["gcells"] / single, 4)
    out["single_grid_tuned_gcells"] = round(single_t, 4)
    out["single_grid_tuning"] = big_tuning
    out["batched_over_single_grid_tuned"] = round(r32["gcells"] / single_t, 4)
    out["batched_over_one_by_one"] = round(r32["gcells"] / seq, 4)
    if 8 in out_prec:
        r64 = out_prec[8]
        out["fp64"] = {"value": round(r64["gcells"], 4), "ms_per_step": round(r64["ms_step"], 5)}
        out["fp32_over_fp64"] = round(r32["gcells"] / r64["gcells"], 3)
        rels = [rel_l2(a, b) for a, b in zip(r32["series"], r64["series"])]
        rels = [r for r in rels if r is not None]
        out["fp32_vs_fp64_detector_relL2_max"] = max(rels) if rels else None
    return out


WORKLOADS = {HEADLINE: bench_headline, "ellipsoids_768x768x512": bench_ellipsoids,
             "batched_64": bench_batched, "emitters_512x512x384": bench_emitters}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batched_64: keep the default tiling of every grid")
    ap.add_argument("--mode", choices=("auto", "separate", "fused"), default="auto",
                    help="single-grid workloads: auto (the tuners pick tiles and step kernels on the grid), "
                         "separate H and E kernels, or the fused H+E pass, at default tiles")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default=HEADLINE)
    ap.add_argument("--no-validation", action="store_true",
                    help="skip the FP32-vs-FP64 validation window (profiler runs); the record then states no error")
    ap.add_argument("--patch-radius", type=int, default=None,
                    help="FP64 near-field box radius around point sources of FP32 grids (0: off)")
    ap.add_argument("--tiles", default=None, metavar="R,X,O",
                    help="with --mode separate/fused: rows, x chunk and tile order of the separate kernels")
    ap.add_argument("--fused-shape", default=None, metavar="R,X,Y",
                    help="with --mode fused: rows per wavefront, planes per chunk, wavefronts along y "
                         "(profiler runs repeat a tuned configuration this way: a profiler slows the "
                         "launches the tuners time)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    global PATCH_RADIUS, FORCED_TILES, FORCED_FUSED_SHAPE
    PATCH_RADIUS = args.patch_radius
    if args.tiles:
        FORCED_TILES = tuple(int(v) for v in args.tiles.split(","))
    if args.fused_shape:
        FORCED_FUSED_SHAPE = tuple(int(v) for v in args.fused_shape.split(","))
    if (FORCED_TILES or FORCED_FUSED_SHAPE) and args.mode == "auto":
        print("--tiles / --fused-shape need --mode separate or fused", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU slab "
              f"decomposition is a design, not part of this build.", file=sys.stderr)
        sys.exit(2)
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = 4000
    out = WORKLOADS[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
