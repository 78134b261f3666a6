"""In-tree build of libofx_hip.so for gfx950 (hipcc), no cmake/ninja.

    python -m optifax_amd.build          # library only
Outputs optifax_amd/lib/libofx_hip.so.  Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed-FP32 instruction selection disabled for all HIP code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-value",
    "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_SOURCES = ["ofx_backend_hip.hip", "ofx_runtime.cpp"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the device feature flag; drop that notice
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose=False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libofx_hip.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", *[os.path.join(CSRC, s) for s in LIB_SOURCES], "-o", tmp]
    log = _run(cmd)
    if verbose and log.strip():
        print(log)
    os.replace(tmp, out)
    return out


def build_mex(verbose=False) -> str:
    """OPOmexBatch gateway over the in-repo mx-API shim (no MATLAB here)."""
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libOPOmexBatch.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           os.path.join(CSRC, "mex_gateway.cpp"), os.path.join(CSRC, "mx_shim.cpp"),
           "-I", os.path.join(ROOT, "include"), "-L", LIBDIR, "-lofx_hip",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    log = _run(cmd)
    if verbose and log.strip():
        print(log)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    paths = [build_hip(verbose)]
    if os.path.exists(os.path.join(CSRC, "mex_gateway.cpp")):
        paths.append(build_mex(verbose))
    return paths


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print(p)
    sys.exit(0)
