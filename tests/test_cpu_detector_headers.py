"""The detector files' headers written by the Fortran host layer
(omx_detector_io_mod), byte for byte against the headers the reference's own
output code wrote for the golden decks (point.dat: output_mod.f90:131-161;
line / plane / volume files: write_fields_subs_mod.f90:631-748, with the
coordinates it carries over from earlier detectors and the NUL byte of its
unset column name).  CPU only: the header writer needs no GPU."""
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import _golden

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FDIR = os.path.join(ROOT, "one-mxll-to-rule-them-all_amd", "fortran")
FC = os.environ.get("OMX_FC", "/opt/rocm/bin/amdflang")
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402


@pytest.fixture(scope="module")
def harness():
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    d = tempfile.mkdtemp(prefix="omxhdr_")
    srcs = [os.path.join(FDIR, "omx_mxll_base.f90"), os.path.join(FDIR, "omx_detector_io.f90"),
            os.path.join(HERE, "fortran", "header_harness.f90")]
    for s in srcs:
        subprocess.run([FC, "-O2", "-c", s], cwd=d, check=True, capture_output=True)
    objs = [os.path.basename(s).replace(".f90", ".o") for s in srcs]
    subprocess.run([FC, "-O2", "-o", "harness"] + objs, cwd=d, check=True, capture_output=True)
    yield os.path.join(d, "harness")
    shutil.rmtree(d, ignore_errors=True)


def written_headers(d, n_det):
    out = []
    for n in range(1, n_det + 1):
        dd = os.path.join(d, f"output_detector_{n:07d}")
        if os.path.exists(os.path.join(dd, "point.dat")):
            with open(os.path.join(dd, "point.dat"), encoding="latin-1") as f:
                out.append([f.read()])
            continue
        files = sorted(glob.glob(os.path.join(dd, "*.dat")), key=lambda p: int(re.findall(r"_(\d+)\.dat$", p)[0]))
        hdr = []
        for p in (files[0], files[-1]):
            with open(p, encoding="latin-1") as f:
                hdr.append(f.read())
        out.append(hdr)
    return out


@pytest.mark.parametrize("name", _golden.NAMES)
def test_detector_headers_match_reference(harness, name):
    case = make_golden.golden_cases()[name]
    desc, fx = _golden.load(name)
    want = json.loads(str(fx["headers_json"]))
    with tempfile.TemporaryDirectory() as d:
        make_golden.write_inputs(case, d)
        with open(os.path.join(d, "harness.nml"), "w") as f:
            g = case["grid"]
            f.write(f"&harness\n nx = {g[0]}, ny = {g[1]}, nz = {g[2]}, n_det = {len(case['detectors'])},\n"
                    f" record_every = {case['record_every']}, steps = {case['steps']}, dx_nm = {case['dx']!r}\n/\n")
        p = subprocess.run([harness], cwd=d, capture_output=True, text=True)
        assert p.returncode == 0, p.stdout + p.stderr
        got = written_headers(d, len(case["detectors"]))
    assert len(got) == len(want)
    kinds = set()
    for n, (g, w) in enumerate(zip(got, want)):
        assert g == w, f"detector {n + 1}: {g!r} != {w!r}"
        kinds.add("point" if len(w) == 1 else w[0].split("\n")[1][:4])
    print(f"{name}: {len(want)} detectors' headers identical ({sorted(kinds)})")
