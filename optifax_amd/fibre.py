"""Fibre GNLSE: gnlsefib.m (J. C. Travers, M. H. Frosz, J. M. Dudley 2009, as
shipped in the reference) with its integration on the device.

`gnlsefib` has the reference's signature and returns what it returns:
    Z, AT, AW, W, B = gnlsefib(T, A, w0, gamma, betas, loss, fr, RT, flength, nsaves)
The host forms the operators exactly as gnlsefib.m:22-45 does (Taylor series
of the betas, loss in dB/m, self-steepening, the Raman response in the
frequency domain, FFT order), `ofx_gnlse_run` integrates the interaction-
picture equation (gnlsefib.m:49-61) with the Dormand-Prince 5(4) pair under
ode45's step control (RelTol 1e-5, AbsTol 1e-12, NormControl on,
gnlsefib.m:71-74) and returns the field at Z = linspace(0, flength, nsaves)
(gnlsefib.m:70), and the host changes variables back (gnlsefib.m:79-86).

`gnlse_batch` runs several fibres of the same grid and output points in one
launch, one workgroup each (pump-power or fibre-length scans).

There is no CPU fallback: without libofx_hip.so or a HIP device every call
raises.  The field must fit the LDS of one workgroup: n <= 2^13 points with
FP32 fields (precision="single"), 2^12 with FP64 fields ("double").
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib

_DP = ctypes.POINTER(ctypes.c_double)


def _dp(a):
    return a.ctypes.data_as(_DP)


def operators(T, w0, gamma, betas, loss, fr, RT):
    """gnlsefib.m:22-45 -> dict(V, L, W, RW, gamma, B, dT); L, W and RW in FFT
    order, gamma divided by w0 when self-steepening is on (|w0| > eps), RW
    None without a Raman response (a scalar RT or fr = 0)."""
    T = np.asarray(T, dtype=np.float64)
    n = T.size
    dT = float(T[1] - T[0])
    V = 2 * np.pi * np.arange(-(n // 2), n - n // 2) / (n * dT)
    alpha = math.log(10.0 ** (loss / 10.0))
    B = np.zeros(n)
    for i, b in enumerate(np.atleast_1d(np.asarray(betas, dtype=np.float64))):
        B = B + b / math.factorial(i + 2) * V ** (i + 2)
    L = 1j * B - alpha / 2
    if abs(w0) > np.finfo(float).eps:
        g = gamma / w0
        W = V + w0
    else:
        g = gamma
        W = np.ones(n)
    RT = np.atleast_1d(np.asarray(RT))
    raman = RT.size > 1 and abs(fr) >= np.finfo(float).eps
    RW = n * np.fft.ifft(np.fft.fftshift(RT.astype(np.float64))) if raman else None
    return dict(V=V, L=np.fft.fftshift(L), W=np.fft.fftshift(W), RW=RW, gamma=float(g),
                B=B, dT=dT, fr=float(fr))


def raman_response(T, tau1=12.2e-15, tau2=32.0e-15):
    """Waveguide.RamanResponse (Waveguide.m:34-42) on the grid T:
    hR = (tau1^2 + tau2^2) / (tau1 tau2^2) exp(-t / tau2) sin(t / tau1), 0 for
    t < 0; the default response times are fused silica's (OpticalFibre.m:24-27,
    with RamanFraction 0.18)."""
    T = np.asarray(T, dtype=np.float64)
    h = (tau1 ** 2 + tau2 ** 2) / (tau1 * tau2 ** 2) * np.exp(-T / tau2) * np.sin(T / tau1)
    return np.where(T < 0, 0.0, h)


def gnlse_batch(T, fields, ops, flength, nsaves, precision="double", rtol=1e-5, atol=1e-12,
                device=0, max_steps=10_000_000, lib=None):
    """Integrate len(fields) fibres on the grid T (each with its own operators
    dict from `operators`) to the points linspace(0, flength, nsaves).
    Returns (Z, AWz [B, nsaves, n] interaction-picture fields, stats [B, 4],
    device ms)."""
    L_ = lib or _lib.load()
    T = np.asarray(T, dtype=np.float64)
    n = T.size
    B = len(fields)
    if B < 1 or len(ops) != B:
        raise ValueError("one operators dict per field")
    if nsaves < 3:
        raise ValueError("nsaves must be at least 3 (ode45 then returns the requested points)")
    raman = [o["RW"] is not None for o in ops]
    if any(raman) and not all(raman):
        raise ValueError("a batch is either all with or all without the Raman term")
    Z = np.linspace(0.0, float(flength), int(nsaves))
    y0 = np.ascontiguousarray(np.stack([np.fft.ifft(np.asarray(f, dtype=np.complex128))
                                        for f in fields]))
    Ls = np.ascontiguousarray(np.stack([o["L"] for o in ops]).astype(np.complex128))
    Ws = np.ascontiguousarray(np.stack([np.real(o["W"]) for o in ops]).astype(np.float64))
    RWs = (np.ascontiguousarray(np.stack([o["RW"] for o in ops]).astype(np.complex128))
           if raman[0] else None)
    g = np.array([o["gamma"] for o in ops], np.float64)
    fr = np.array([o["fr"] for o in ops], np.float64)
    dT = np.array([o["dT"] for o in ops], np.float64)
    out = np.zeros((B, int(nsaves), n), np.complex128)
    stats = np.zeros((B, 4), np.int64)
    ms = ctypes.c_double()
    prec = _lib.OFX_F64 if precision == "double" else _lib.OFX_F32
    st = L_.ofx_gnlse_run(int(device), prec, B, n, int(nsaves), _dp(Z), _dp(y0.view(np.float64)),
                          _dp(Ls.view(np.float64)), _dp(Ws),
                          None if RWs is None else _dp(RWs.view(np.float64)), _dp(g), _dp(fr),
                          _dp(dT), float(rtol), float(atol), int(max_steps),
                          _dp(out.view(np.float64)),
                          stats.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                          ctypes.byref(ms))
    if st != _lib.OFX_OK:
        raise _lib.OfxError(st, "ofx_gnlse_run failed")
    return Z, out, stats, ms.value


def lab_frame(Z, AWz, op, n, dT):
    """gnlsefib.m:79-86: AW exp(L Z) -> (AT = fft(.), AW = fftshift(.) dT n)."""
    AT = np.zeros_like(AWz)
    AW = np.zeros_like(AWz)
    for i in range(len(Z)):
        a = AWz[i] * np.exp(op["L"] * Z[i])
        AT[i] = np.fft.fft(a)
        AW[i] = np.fft.fftshift(a) * dT * n
    return AT, AW


def gnlsefib(T, A, w0, gamma, betas, loss, fr, RT, flength, nsaves, precision="double",
             rtol=1e-5, atol=1e-12, device=0, lib=None, max_steps=10_000_000):
    """(Z, AT, AW, W, B) as gnlsefib.m returns them; raises if the integration
    failed (step below 16 eps(z), more than max_steps attempted steps, or a
    non-finite field)."""
    op = operators(T, w0, gamma, betas, loss, fr, RT)
    n = np.asarray(T).size
    Z, AWz, stats, _ = gnlse_batch(T, [A], [op], flength, nsaves, precision, rtol, atol, device,
                                   max_steps=max_steps, lib=lib)
    if stats[0, 3] != 0:
        raise RuntimeError(f"gnlse: integration failed (status {int(stats[0, 3])})")
    AT, AW = lab_frame(Z, AWz[0], op, n, op["dT"])
    return Z, AT, AW, op["V"] + w0, op["B"]
