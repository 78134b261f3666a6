"""Inputs and calls shared by the tests of ofx_crystal_pass_noalloc (the
no-allocation single pass) against ofx_crystal_pass_host, both through the
C ABI with raw arrays in the library's layout."""
import ctypes

import numpy as np

from optifax_amd import _lib
from optifax_amd.solver import _params

from cases import crystal_case


def raw_case(n, batch, per_instance_G, seed=0, L=0.2e-3, method=0, cols=4):
    """Raw arrays of one call: Et [batch][2n] f32, alpha [n], G [nsteps *
    g_batch], beta [n], ApFT [batch][cols][2n], stepMods [batch][nsteps]."""
    sw, si, E, args = crystal_case(n=n, window=2e-12 * n / 256, L=L, seed=seed + 1)
    ns = int(args[5])
    Es = [crystal_case(n=n, window=2e-12 * n / 256, L=L, seed=seed + 1 + b)[2] for b in range(batch)]
    Et = np.ascontiguousarray(np.stack(Es).astype(np.complex64).view(np.float32))
    G = np.asarray(args[1], np.float32)
    if per_instance_G:
        G = np.concatenate([G * (1.0 + 0.05 * b) for b in range(batch)]).astype(np.float32)
    rng = np.random.default_rng(seed)
    ap = (rng.standard_normal((batch, cols, n)) + 1j * rng.standard_normal((batch, cols, n)))
    sm = rng.integers(0, 3, (batch, ns)).astype(np.float32)      # entries the pass leaves keep these
    p = _params(n, ns, args[2], args[3], args[4], args[6], args[8], args[9], args[10], cols, method)
    return dict(p=p, batch=batch, Et=Et, alpha=np.asarray(args[0], np.float32),
                G=np.ascontiguousarray(G), g_batch=batch if per_instance_G else 1,
                beta=np.asarray(args[7], np.float32),
                ApFT=np.ascontiguousarray(ap.astype(np.complex64).view(np.float32)),
                stepMods=np.ascontiguousarray(sm), status=np.zeros(batch, np.int32))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def call_host(ctx, c):
    """ofx_crystal_pass_host on copies; returns (rc, Et, ApFT, stepMods, status)."""
    out = {k: c[k].copy() for k in ("Et", "ApFT", "stepMods", "status")}
    rc = ctx.lib.ofx_crystal_pass_host(
        ctx.h, ctypes.byref(c["p"]), c["batch"], _fp(out["Et"]), _fp(c["alpha"]), _fp(c["G"]),
        c["g_batch"], _fp(c["beta"]), _fp(out["ApFT"]), _fp(out["stepMods"]),
        out["status"].ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None)
    return rc, out["Et"], out["ApFT"], out["stepMods"], out["status"]


def call_noalloc(ctx, c, ptr_of):
    """ofx_crystal_pass_noalloc with the arrays at the addresses ptr_of gives
    (host addresses for the CPU emulation, device addresses on a GPU);
    ptr_of(name, array) -> (address, fetch) where fetch() returns the array
    after the call."""
    res = {}
    addr = {}
    for k in ("Et", "alpha", "G", "beta", "ApFT", "stepMods", "status"):
        addr[k], res[k] = ptr_of(k, c[k])
    rc = ctx.lib.ofx_crystal_pass_noalloc(
        ctx.h, ctypes.byref(c["p"]), c["batch"], addr["Et"], addr["alpha"], addr["G"],
        c["g_batch"], addr["beta"], addr["ApFT"], addr["stepMods"], addr["status"])
    return rc, res["Et"](), res["ApFT"](), res["stepMods"](), res["status"]()


def host_ptrs(keep):
    def ptr_of(name, a):
        b = a.copy()
        keep.append(b)
        return b.ctypes.data, (lambda: b)
    return ptr_of


OFX_E_ARG = _lib.OFX_E_ARG
