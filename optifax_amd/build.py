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

LIB_SOURCES = ["ofx_backend_hip.hip", "ofx_backend_rkl.hip", "ofx_runtime.cpp", "ofx_mlp.hip",
               "ofx_qpm.hip", "ofx_gnlse.hip"]


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


def build_mex(verbose=False):
    """MEX gateways (OPOmexBatch, OPOmexCavity) over the in-repo mx-API shim.

    Without MATLAB the gateways link libofx_mx.so (optifax_amd/csrc/mx_shim.cpp),
    which implements the documented mx*/mex* functions they use; with MATLAB
    the same sources build with `mex -R2018a`."""
    os.makedirs(LIBDIR, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]
    flags = ["-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-DOFX_MX_SHIM"]
    out = []
    shim = os.path.join(LIBDIR, "libofx_mx.so")
    _run(["g++", *flags, os.path.join(CSRC, "mx_shim.cpp"), *inc, "-o", shim + ".tmp"])
    os.replace(shim + ".tmp", shim)
    out.append(shim)
    for name, src in (("libOPOmexBatch.so", "mex_gateway.cpp"), ("libOPOmexCavity.so", "mex_cavity.cpp")):
        dst = os.path.join(LIBDIR, name)
        log = _run(["g++", *flags, os.path.join(CSRC, src), *inc, "-L", LIBDIR, "-lofx_mx",
                    "-lofx_hip", "-Wl,-rpath,$ORIGIN", "-o", dst + ".tmp"])
        if verbose and log.strip():
            print(log)
        os.replace(dst + ".tmp", dst)
        out.append(dst)
    return out


def build_all(verbose=False):
    return [build_hip(verbose), *build_mex(verbose)]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print(p)
    sys.exit(0)
