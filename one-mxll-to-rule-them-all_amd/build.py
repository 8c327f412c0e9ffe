"""Build recipe of the engine: libomx_fdtd_hip.so for gfx950, in-tree.

    python one-mxll-to-rule-them-all_amd/build.py

Every HIP source is compiled with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OMX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}")


def build_engine():
    src = os.path.join(PKG, "csrc", "omx_fdtd.hip")
    out = os.path.join(PKG, "libomx_fdtd_hip.so")
    tmp = out + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-o", tmp, src])
    os.replace(tmp, out)
    return out


FC = os.environ.get("OMX_FC", "/opt/rocm/bin/amdflang")
F_SRC = ["omx_fdtd_cabi.f90", "omx_mxll_base.f90", "omx_mxll_3d_hip.f90", "omx_mxll_lowdim_hip.f90", "omx_factory.f90",
         "omx_detector_io.f90", "omx_qgroup_io.f90", "omx_fdtd_main.f90"]


def build_fortran():
    """The Fortran host layer (TMxll_3D_hip) and the omx_fdtd_3d driver,
    linked against the engine with an rpath to this directory."""
    bdir = os.path.join(PKG, "fortran", "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for f in F_SRC:
        run([FC, "-O2", "-c", os.path.join(PKG, "fortran", f)], cwd=bdir)
        objs.append(f.replace(".f90", ".o"))
    exe = os.path.join(PKG, "omx_fdtd_3d")
    run([FC, "-O2", "-o", exe] + objs + ["-L" + PKG, "-lomx_fdtd_hip", "-Wl,-rpath," + PKG,
                                          "-Wl,-rpath,$ORIGIN"], cwd=bdir)
    return exe


if __name__ == "__main__":
    build_engine()
    build_fortran()
    sys.exit(0)
