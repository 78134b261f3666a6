"""QPM gain maps: the reference's GPU phase-matching analysis (qpmgain.m,
sincgain.m / sincgain_sparse.m) with the heavy sums on the device
(ofx_qpm_zsum, ofx_sinc_gain in libofx_hip).

qpm_gain(): the small-signal parametric gain of every (signal, pump) pair
of a pump pulse through a poled crystal, qpmgain.m:1-150 (the branch with
the rectangular low-intensity idler, the one the crystal scans use,
NonlinearCrystal.m:161,179,264).  The host forms the wavelength grids,
indices, mismatches and prefactors exactly as the reference does; the sum
over the crystal, As = g * sum_z P(z) exp(i dk z) dz, runs on the device
with the sampled grating compressed into runs of equal poling.

sinc_gain(): the narrowband mismatch factor over a fan-out grating's periods
(sincgain.m:1-48), in the consistent units of sincgain_sparse.m (rad/m
throughout; sincgain.m itself mixes kcalc's rad/m with a period and length in
um).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from . import physics as ph

EPS0 = ph.EPS0
C0 = ph.C0


def pulse_descriptors(simwin: ph.SimWindow, E_t, area, rep_rate=1e9, n_medium=1.0):
    """What qpmgain reads from an OpticalPulse (OpticalPulse.m): |SpectralField|
    (centred order), PeakWavelength (the ESD maximum), WavelengthFWHM
    (findfwnm of the ESD over the wavelengths at 1/2) and AverageIntensity
    (Energy / Area / DurationCheck / 2, DurationCheck = FWHM of |E(t)|^2).
    E_t is the natural-order temporal field [V/m]."""
    Ek = ph.SimWindow.t2k(np.asarray(E_t, dtype=np.complex128))
    absk = np.abs(Ek)
    # ESD up to constant factors (its maximum and half-width are what is used)
    esd = absk ** 2
    lam = simwin.wavelengths
    peak = float(lam[int(np.argmax(esd))])
    fwhm_lam = findfwnm(lam, esd, 0.5)
    energy = ph.energy_of(simwin, E_t, area, n_medium)
    dur = findfwnm(simwin.times, np.abs(E_t) ** 2, 0.5)
    return dict(abs_spectrum=absk, peak_wavelength=peak, wavelength_fwhm=fwhm_lam,
                average_intensity=energy / area / dur / 2.0)


def findfwnm(x, y, frac):
    """findfwnm.m: full width where y exceeds frac of its maximum, with linear
    interpolation of the crossing on each side (in index space, then mapped
    onto x)."""
    q = np.asarray(y, dtype=np.float64)
    q = q / q.max()
    p = np.nonzero(q > frac)[0]
    p0, p1, p2, p3 = p.min() - 1, p.min(), p.max(), p.max() + 1
    s01 = (q[p1] - q[p0]) / (p1 - p0)
    s23 = (q[p3] - q[p2]) / (p3 - p2)
    a = p0 + (frac - q[p0]) / s01
    b = p2 + (frac - q[p2]) / s23
    w = np.interp([a, b], np.arange(len(x)), np.asarray(x, dtype=np.float64))
    return float(abs(w[1] - w[0]))


def runs_of(z, w, rtol=1e-9):
    """Compress samples (z_k, w_k) into runs of equal weight and uniform
    spacing dz (the median spacing): returns z0, len, wgt, dz.  Samples that
    break the spacing or change the weight start a new run; a run of one
    sample is the literal term."""
    z = np.asarray(z, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    d = np.diff(z)
    dz = float(np.median(d)) if d.size else 1.0
    if not dz > 0:
        dz = 1.0
    uniform = np.abs(d - dz) <= rtol * dz
    same = w[1:] == w[:-1]
    cont = np.concatenate([[False], uniform & same])     # sample k continues the run of k-1
    starts = np.nonzero(~cont)[0]
    lens = np.diff(np.concatenate([starts, [z.size]]))
    keep = w[starts] != 0.0                                # zero-weight runs add nothing
    return (np.ascontiguousarray(z[starts][keep]), np.ascontiguousarray(lens[keep].astype(np.int32)),
            np.ascontiguousarray(w[starts][keep]), dz)


def zsum(dk, z, w, device=0, lib=None, rtol=1e-9):
    """sum_k w_k exp(i dk z_k) for every entry of dk, on the device."""
    L = lib or _lib.load()
    dk = np.ascontiguousarray(np.asarray(dk, dtype=np.float64).reshape(-1))
    z0, ln, wg, dz = runs_of(z, w, rtol)
    out = np.zeros(dk.size, np.complex128)
    ms = ctypes.c_double()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    st = L.ofx_qpm_zsum(int(device), dk.size, dp(dk), z0.size, dp(z0),
                        ln.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dp(wg), dz,
                        out.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        ctypes.byref(ms))
    if st:
        raise _lib.OfxError(st, "ofx_qpm_zsum failed")
    zsum.last_ms = ms.value
    zsum.last_runs = z0.size
    return out


def _mrange_indices(n_total, n_take):
    """crystal.Polarisation(1 : NSteps/n_steps : end) as 0-based indices; the
    step must be a whole number (MATLAB indexes with the range's values)."""
    step = n_total / n_take
    if abs(step - round(step)) > 1e-9:
        raise ValueError(f"NSteps/n_steps = {step} is not a whole number")
    return np.arange(0, n_total, int(round(step)))


def qpm_gain(crystal: ph.NonlinearCrystal, simwin: ph.SimWindow, pulse: dict,
             sigrange_nm=(1100.0, 1700.0), step=1e-7, device=0):
    """qpmgain.m with the rectangular low-intensity idler.  `pulse`: the
    dict of pulse_descriptors.  Returns dict(gain [signals x pumps], pump,
    signal, idler (wavelength grids [m]), weights, p_mask, i_mask)."""
    a = qpm_arrays(crystal, simwin, pulse, sigrange_nm, step)
    As = a["g"] * zsum(a["dk"], a["z"], a["w"], device).reshape(a["dk"].shape)
    Is = a["nS"][:, None] * (np.abs(As) / a["I2A"]) ** 2
    return dict(gain=Is / a["Ip"], pump=a["pump"], signal=a["signal"], idler=a["idler"],
                weights=a["weights"], p_mask=a["p_mask"], i_mask=a["i_mask"])


def qpm_arrays(crystal: ph.NonlinearCrystal, simwin: ph.SimWindow, pulse: dict,
               sigrange_nm=(1100.0, 1700.0), step=1e-7):
    """The host side of qpm_gain (qpmgain.m:12-143): wavelength grids,
    mismatches dk, prefactors g and the crystal samples z, w = P dz."""
    sig = np.asarray(sigrange_nm, dtype=np.float64) * 1e-9
    L = crystal.length
    N = simwin.n
    n_steps = int(round(5e5 * L)) if ((L > 1e-2 and N > 2 ** 13) or N > 2 ** 16) \
        else int(round(1e6 * L))
    zl = np.linspace(0.0, L, n_steps)
    del_z = zl[1]
    dz = np.concatenate([[0.0], np.full(n_steps - 2, del_z), [0.0]])
    dz[-1] = L - np.sum(dz[:-1])
    z = np.cumsum(dz)
    NS = crystal.nsteps(step)
    P = crystal.poling(NS) * crystal.chi2                  # Polarisation (NonlinearCrystal.m:106)
    P = P[_mrange_indices(NS, n_steps)]
    if P.size != n_steps:
        raise ValueError("poling samples do not match the z samples")
    d_eff = P.max()
    P = P / P.max()
    lam = simwin.wavelengths
    nr = crystal.refractive_index(simwin)
    # pump: +-3 FWHM around the peak, at most ~100 wavelengths
    prange = pulse["wavelength_fwhm"] * np.array([-3.0, 3.0]) + pulse["peak_wavelength"]
    pid = (lam > prange[0]) & (lam < prange[1])
    pump = lam[pid]
    p_mask = pulse["abs_spectrum"] / pulse["abs_spectrum"].max()
    p_mask = p_mask[pid]
    nP = nr[pid]
    if pump.size > 100:
        f = int(np.floor(pump.size / 100 + 0.5))
        pump, p_mask, nP = pump[::f], p_mask[::f], nP[::f]
    olap = (sig > prange[0]) & (sig < prange[1])
    if olap[0]:
        sig[0] = prange[1] + 1e-9
    elif olap[1]:
        sig[1] = prange[0] - 1e-9
    # rectangular low-intensity idler
    sid = (lam > sig[0]) & (lam < sig[1])
    signal = lam[sid]
    nS = nr[sid]
    if signal.size > 2000:
        f = int(np.floor(signal.size / 2000 + 0.5))
        signal, nS = signal[::f], nS[::f]
    idler = C0 / (C0 / pump[None, :] - C0 / signal[:, None])   # dfg_lambda(pump, signal.')
    nI = np.abs(ph.sellmeier_n(idler * 1e6, crystal.material, crystal.temperature))
    i_mask = np.ones(signal.size) * p_mask.max()
    kI = 2 * np.pi * nI / idler
    sig2 = np.repeat(signal[:, None], pump.size, axis=1)
    kS = 2 * np.pi * nS[:, None] / sig2
    I2A = np.sqrt(1.0 / (2.0 * EPS0 * C0))
    Ip = pulse["average_intensity"]
    Ii = Ip / 100.0
    Ap = np.sqrt(Ip / nP) * I2A
    Ai = np.sqrt(Ii / nI) * I2A
    weights = np.outer(i_mask, p_mask)
    pump2 = np.repeat(pump[None, :], signal.size, axis=0)
    kP = 2 * np.pi * nP[None, :] / pump2
    if sig[1] < prange[0]:                                  # SFG regime: second-harmonic terms
        shg = pump2 == idler
        if np.any(shg):
            i2 = np.zeros_like(weights)
            i2[np.arange(min(weights.shape)), np.arange(min(weights.shape))] = i_mask[
                :min(weights.shape)]
            weights[shg] += np.broadcast_to(p_mask[None, :], weights.shape)[shg] ** 2
            weights[shg] += i2[shg] ** 2
    dk = np.abs(kP - kS - kI)                               # the device branch takes |dk|
    weights = Ai * Ap[None, :] * weights
    g = 1j * 2 * d_eff * (kS / nS[:, None] ** 2) * weights
    return dict(dk=dk, g=g, z=z, w=P * dz, nS=nS, I2A=I2A, Ip=Ip, pump=pump2, signal=sig2,
                idler=idler, weights=weights, p_mask=p_mask, i_mask=i_mask)


def sinc_gain(crystal: ph.NonlinearCrystal, lam_p, lam_s, lam_i, nx_pos, m=1, device=0):
    """sincgain.m for a fan-out crystal (GratingPeriod = [P1 P2]): the
    mismatch factor gain[p, s, g] = sum_m (sinc(dk L / 2) / m)^2 with
    dk = | |k_p - k_i| - k_s | - m 2 pi / period_g over nx_pos periods from P1
    to P2, and conv_eff = 2 (Chi2 L / pi)^2 / (eps0 c^3) w_s^2 / |n_p n_s n_i|.
    lam_* in um, broadcastable to [pumps, signals]; returns (gain
    [pumps, signals, gratings], grating_um, conv_eff [pumps, signals])."""
    p = np.atleast_1d(np.asarray(crystal.period, dtype=np.float64))
    if p.size != 2:
        raise ValueError("sinc_gain needs a fan-out crystal (two grating periods)")
    grating_um = np.linspace(p[0], p[1], nx_pos) * 1e6
    ms = np.atleast_1d(np.asarray(m, dtype=np.float64))

    def kcalc(lam_um):
        n = np.abs(ph.sellmeier_n(np.asarray(lam_um, dtype=np.float64), crystal.material,
                                  crystal.temperature))
        return 2 * np.pi * n / (np.asarray(lam_um) * 1e-6), n, 2 * np.pi * C0 / (np.asarray(lam_um) * 1e-6)

    kp, n_p, _ = kcalc(lam_p)
    ks, n_s, w_s = kcalc(lam_s)
    ki, n_i, _ = kcalc(lam_i)
    dk0 = np.abs(np.abs(kp - ki) - ks)
    shape = dk0.shape
    dk0 = np.ascontiguousarray(dk0.reshape(-1).astype(np.float64))
    kc = np.ascontiguousarray(2 * np.pi / (grating_um * 1e-6))
    out = np.zeros(dk0.size * kc.size)
    L = _lib.load()
    msd = np.ascontiguousarray(ms)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t = ctypes.c_double()
    st = L.ofx_sinc_gain(int(device), dk0.size, dp(dk0), kc.size, dp(kc), msd.size, dp(msd),
                         float(crystal.length), dp(out), ctypes.byref(t))
    if st:
        raise _lib.OfxError(st, "ofx_sinc_gain failed")
    sinc_gain.last_ms = t.value
    gain = np.moveaxis(out.reshape(kc.size, *shape), 0, -1)
    conv = (w_s ** 2) / np.abs(n_p * n_s * n_i)
    conv = 2 * (crystal.chi2 / np.pi * crystal.length) ** 2 / (EPS0 * C0 ** 3) * conv
    return gain, grating_um, conv
