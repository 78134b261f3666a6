"""FP64 CPU oracle for the chi(2) split-step crystal pass and OPO round trips.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
`cpu_baseline` leg of bench.py may import this package, and only as the checker
(or as the timed CPU baseline).  The product library (optifax_amd) never
imports, links or calls it.

The arithmetic lives in oracle.c (plain C99, double precision), which restates
the reference's MATLAB CPU solver toolbox/lib/internal/main/NEE_CPU.m, the
RK23 control flow of toolbox/lib/MEXlib/src/ode23/OPO_CUDA_ODE23.cu and the
round-trip order of toolbox/lib/classes/OpticalSim.m:207-301.  This module is
a ctypes wrapper plus numpy conveniences.

Parity status per function (what pins it; tests/test_oracle.py,
tests/test_oracle_pins.py):
  fft            : brute-force long-double DFT, numpy.fft, committed golden
                   vectors at 2^13-2^16
  nee_heun       : linear closed form (Chi2 = 0, incl. the k bookkeeping),
                   hand-worked step tables for growth, clipping at NSteps and
                   forced rejection, Heun 2nd-order convergence, low-gain SHG
                   sinc^2 law (magnitude and L^2), low-gain difference-frequency
                   law of the 2|A|^2 term
  rk23           : 3rd-order convergence, hand-worked step tables for doubling,
                   chunk-end truncation and forced rejection (k -= s, s -= 1),
                   linear closed form
  gayer_n        : published index value of 5% MgO:CLN
  phi_calc       : numpy diff + cumulative trapezoid on the same arrays,
                   GVD sign/magnitude check
  qpm_poling     : exact square-wave structure
  opo_run        : linear cavity (Chi2 = 0) geometric series for the field,
                   the output-coupled spectrum, its energy and the
                   convergence metric (Heun and RK23); stop at conv_tol
  Trajectories of the full reference path (MATLAB/CUDA) cannot be run here:
  parity against the reference itself is unpinned beyond these checks.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_LL = ctypes.POINTER(ctypes.c_longlong)
_I = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile oracle.c into oracle/liboracle.so with gcc (no fast-math)."""
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fft.argtypes = [_D, ctypes.c_int, ctypes.c_int]
        L.orc_nee_heun.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, ctypes.c_double, _D,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, _D,
                                   ctypes.c_int, _D, _LL]
        L.orc_rk23.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int, ctypes.c_double, _D,
                               ctypes.c_double, ctypes.c_double, ctypes.c_int, _D,
                               ctypes.c_int, _D, _LL]
        L.orc_gayer_n.restype = ctypes.c_double
        L.orc_gayer_n.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_sellmeier_n.restype = ctypes.c_double
        L.orc_sellmeier_n.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_phi_calc_mat.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, _D, _D, ctypes.c_int, _D, _D, _D]
        L.orc_chirped_grating.argtypes = [_D, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, _D, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]
        L.orc_qpm_poling_widths.argtypes = [ctypes.c_int, ctypes.c_double, _D, ctypes.c_int, _D]
        L.orc_phi_calc.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D,
                                   ctypes.c_int, _D, _D, _D]
        L.orc_qpm_poling.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, _D]
        L.orc_opo_run.argtypes = ([ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D,
                                   _D, _D, ctypes.c_double, _D, _D, _D, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_double, _D, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_int, _D, _D, _D, _LL,
                                   ctypes.c_double, _I])
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _c128(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _check(st, what):
    if st == 3:
        raise FloatingPointError(f"{what}: field became NaN")
    if st:
        raise RuntimeError(f"{what} failed with status {st}")


def fft(x, inverse=False):
    """MATLAB fft (unnormalised) / ifft (1/N) of a complex vector."""
    y = _c128(x).copy()
    _check(lib().orc_fft(y.view(np.float64).ctypes.data_as(_D), y.size, 1 if inverse else 0), "fft")
    return y


def nee_heun(Et, xT_h, G33, w0, bdiffw0, h, nSteps, dt, hBarg, maxErr, minErr, sel, ApFT=None,
             n_cols=None, want_stepmods=True, allow_nonfinite=False):
    """NEE_CPU.m in FP64.  Returns (Et, ApFT, stepMods, counters)."""
    Et = _c128(Et).copy()
    n = Et.size
    if n_cols is None:
        n_cols = 0 if ApFT is None else ApFT.shape[1]
    ap = np.zeros((n_cols, n), np.complex128) if ApFT is None else _c128(ApFT.T).copy()
    sm = np.zeros(int(nSteps), np.float64) if want_stepmods else None
    cnt = np.zeros(3, np.int64)
    st = lib().orc_nee_heun(n, Et.view(np.float64).ctypes.data_as(_D), _dp(_f64(xT_h)),
                            _dp(_f64(G33)), w0, bdiffw0, h, int(nSteps), dt, _dp(_f64(hBarg)),
                            maxErr, minErr, float(sel),
                            ap.view(np.float64).ctypes.data_as(_D) if n_cols else None, n_cols,
                            _dp(sm), cnt.ctypes.data_as(_LL))
    if not (allow_nonfinite and st == 3):
        _check(st, "nee_heun")
    return Et, ap.T.copy(), sm, cnt


def rk23(Et, xT_h, G, w0, bdiffw0, h, nSteps, dt, hBarg, maxErr, minErr, sel, n_cols=0,
         want_stepmods=True, allow_nonfinite=False):
    """Bogacki-Shampine RK23 pass (OPO_CUDA_ODE23.cu control) in FP64.
    Returns (Et, ApFT, stepMods, counters); with allow_nonfinite a pass whose
    field turns NaN returns what it reached (counters count the attempt that
    found it) instead of raising."""
    Et = _c128(Et).copy()
    n = Et.size
    ap = np.zeros((max(n_cols, 1), n), np.complex128)
    sm = np.zeros(int(nSteps), np.float64) if want_stepmods else None
    cnt = np.zeros(3, np.int64)
    st = lib().orc_rk23(n, Et.view(np.float64).ctypes.data_as(_D), _dp(_f64(xT_h)), _dp(_f64(G)),
                        w0, bdiffw0, h, int(nSteps), dt, _dp(_f64(hBarg)), maxErr, minErr,
                        int(sel), ap.view(np.float64).ctypes.data_as(_D) if n_cols else None,
                        n_cols, _dp(sm), cnt.ctypes.data_as(_LL))
    if not (allow_nonfinite and st == 3):
        _check(st, "rk23")
    return Et, ap[:n_cols].T.copy(), sm, cnt


def gayer_n(lam_um, T=20.0, col=1):
    lam = np.atleast_1d(np.asarray(lam_um, dtype=np.float64))
    L = lib()
    return np.array([L.orc_gayer_n(float(x), float(T), int(col)) for x in lam])


def phi_calc(L, T, w_abs, w_rel, ref):
    """phi_calc.m + phi2GD.m; returns (phi_rel, phi_abs, gd)."""
    w_abs = _f64(w_abs)
    w_rel = _f64(w_rel)
    n = w_abs.size
    pr = np.zeros(n)
    pa = np.zeros(n)
    gd = np.zeros(n)
    _check(lib().orc_phi_calc(n, L, T, _dp(w_abs), _dp(w_rel), int(ref), _dp(pr), _dp(pa), _dp(gd)),
           "phi_calc")
    return pr, pa, gd


MATERIALS = {"PPLN": 1, "LN_e": 1, "LN_o": 2, "PPLN_undoped": 4, "OP-GaP": 10, "OPGaP": 10,
             "KTP": 11, "PPKTP": 11}


def sellmeier_n(lam_um, material, T=20.0):
    """sellmeier.m:3-21 for the crystal materials (real part)."""
    lam = np.atleast_1d(np.asarray(lam_um, dtype=np.float64))
    code = MATERIALS[material]
    L = lib()
    return np.array([L.orc_sellmeier_n(float(x), code, float(T)) for x in lam])


def phi_calc_mat(L, T, material, w_abs, w_rel, ref):
    """phi_calc.m for any crystal material ("PPLN" uses T, the others 20 C)."""
    w_abs = _f64(w_abs)
    w_rel = _f64(w_rel)
    n = w_abs.size
    pr, pa, gd = np.zeros(n), np.zeros(n), np.zeros(n)
    code = 0 if material == "PPLN" else MATERIALS[material]
    _check(lib().orc_phi_calc_mat(n, L, T, code, _dp(w_abs), _dp(w_rel), int(ref), _dp(pr),
                                  _dp(pa), _dp(gd)), "phi_calc_mat")
    return pr, pa, gd


def chirped_grating(z, P1, P2, a, tol):
    """chirpedgrating.m: domain widths."""
    z = _f64(z)
    smallest = max(min(P1, P2) / 2.0 - tol, tol)        # narrowest domain the rounding allows
    cap = int(np.ceil(z.size * (z[1] - z[0]) / smallest)) + 16
    w = np.zeros(cap)
    cnt = ctypes.c_int()
    _check(lib().orc_chirped_grating(_dp(z), z.size, P1, P2, a, tol, _dp(w), cap,
                                     ctypes.byref(cnt)), "chirped_grating")
    return w[:cnt.value].copy()


def qpm_poling_widths(nsteps, L, widths):
    P = np.zeros(int(nsteps))
    w = _f64(widths)
    _check(lib().orc_qpm_poling_widths(int(nsteps), L, _dp(w), w.size, _dp(P)),
           "qpm_poling_widths")
    return P


def qpm_poling(nsteps, L, period, duty_off=0.0):
    P = np.zeros(int(nsteps))
    _check(lib().orc_qpm_poling(int(nsteps), L, period, duty_off, _dp(P)), "qpm_poling")
    return P


def opo_run(trips, E_cav, E_pump, n_air, n_xtal, R_oc, T_oc, phi_oc, delay, w_rel, solver,
            method=0, conv_tol=0.0):
    """OpticalSim.run restated (natural-order time, centred spectra).

    solver: dict with xT_h, G33, w0, bdiffw0, h, nSteps, dt, hBarg, maxErr,
    minErr, sel, n_cols.  Returns dict(E_cav, E_oc, oc_energy, conv, steps,
    trips_done)."""
    E = _c128(E_cav).copy()
    n = E.size
    Eoc = np.zeros(n, np.complex128)
    en = np.zeros(trips)
    conv = np.zeros(trips)
    steps = np.zeros(2 * trips, np.int64)
    done = ctypes.c_int(0)
    s = dict(solver)
    s.setdefault("xT_h", s.get("alpha_w"))
    s.setdefault("hBarg", s.get("beta_op"))
    st = lib().orc_opo_run(n, method, trips, E.view(np.float64).ctypes.data_as(_D),
                           _c128(E_pump).view(np.float64).ctypes.data_as(_D), _dp(_f64(n_air)),
                           _dp(_f64(n_xtal)), _dp(_f64(R_oc)), _dp(_f64(T_oc)), _dp(_f64(phi_oc)),
                           float(delay), _dp(_f64(w_rel)), _dp(_f64(s["xT_h"])),
                           _dp(_f64(s["G33"])), s["w0"], s["bdiffw0"], s["h"], int(s["nSteps"]),
                           s["dt"], _dp(_f64(s["hBarg"])), s["maxErr"], s["minErr"],
                           float(s["sel"]), int(s.get("n_cols", 0)),
                           Eoc.view(np.float64).ctypes.data_as(_D), _dp(en), _dp(conv),
                           steps.ctypes.data_as(_LL), float(conv_tol), ctypes.byref(done))
    _check(st, "opo_run")
    t = done.value
    return dict(E_cav=E, E_oc=Eoc, oc_energy=en[:t], conv=conv[:t],
                steps=steps[: 2 * t].reshape(t, 2), trips_done=t)
