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


if __name__ == "__main__":
    build_engine()
    sys.exit(0)
