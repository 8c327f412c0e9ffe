"""The Fortran host layer driven by the reference's own input files: the
solver made by maxwell_factory(3) (a TMxll_3D_hip, reached through
`class is (TMxll_3D_hip)`), the reference's step order, and its detector
files, against the golden vectors of the reference's Fortran modules: the
series at 1e-12 and the file headers byte for byte.  Rows come from the
engine's detector samples, or (OMX_FDTD_DET_MIRROR=1) from the TMxll_3D host
mirrors synced at every print step and read with the reference's stencils."""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import _golden

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402

EXE = os.path.join(os.path.dirname(HERE), "one-mxll-to-rule-them-all_amd", "omx_fdtd_3d")
AU_TO_FS = 2.4188843265864e-2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows", ["samples", "mirror"])
@pytest.mark.parametrize("name", ["media_mix", "au_al_periodic", "vacuum64"])
def test_fortran_driver_matches_reference(name, rows):
    assert os.path.exists(EXE), "omx_fdtd_3d not built"
    case = make_golden.golden_cases()[name]
    desc, fx = _golden.load(name)
    dr, dt = _golden.grid_dr_dt(case["dx"])
    names = {1: "close", 2: "periodic", 3: "cpml"}
    with tempfile.TemporaryDirectory() as d:
        make_golden.write_inputs(case, d)
        g = case["grid"]
        with open(os.path.join(d, "inp"), "w") as f:
            f.write("&OMxRTA\n")
            f.write(" mxll_boundaries = " + ", ".join(f"'{names[b]}'" for b in case["bc"]) + "\n")
            f.write(" mxll_dimensions = 3\n mxll_npml = %d\n" % case["npml"])
            f.write(" mxll_n_src = %d\n mxll_n_media = %d\n mxll_n_detectors = %d\n"
                    % (len(case["sources"]), len(case["media"]), len(case["detectors"])))
            f.write(" mxll_box_size = %r, %r, %r\n" % tuple(float(v * case["dx"]) for v in g))
            f.write(" mxll_dr = %r\n" % case["dx"])
            f.write(" mxll_total_time = %r\n" % ((case["steps"] + 0.5) * dt * AU_TO_FS))
            f.write(" mxll_dt_det_print = %r\n/\n" % ((case["record_every"] + 0.5) * dt * AU_TO_FS))
        print(f"[fortran] {name} ({rows}) in {d}", flush=True)
        env = dict(os.environ, OMX_FDTD_PRECISION="8", OMX_FDTD_DET_MIRROR="1" if rows == "mirror" else "0")
        p = subprocess.run([EXE], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           env=env, timeout=300)
        print(p.stdout[-2000:])
        assert p.returncode == 0
        nrec = fx["series"].shape[0]
        series = make_golden.parse_series(d, case, nrec)
        headers = make_golden.detector_headers(d, case)
    r = np.linalg.norm(series - fx["series"]) / np.linalg.norm(fx["series"])
    print(f"{name} ({rows}): Fortran driver (FP64 GPU) vs reference relL2 = {r:.3e}")
    assert r <= 1e-12
    assert headers == json.loads(str(fx["headers_json"]))
