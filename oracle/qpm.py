"""FP64 oracle of the QPM gain maps (TEST INFRASTRUCTURE ONLY: only tests/
may import it, as the checker).

qpmgain(): toolbox/lib/internal/analysis/qpm/qpmgain.m:1-150 restated in
numpy, the rectangular low-intensity idler branch (:87-104), with the sum over
the crystal taken literally, sample by sample (:145-147: As = sum(g .*
exp(1i*dk.*z) .* P .* dz, 3)).  Indices come from the C oracle's Sellmeier
(oracle.sellmeier_n), the grating from oracle.qpm_poling.

sinc_gain(): sincgain.m:1-48 for a fan-out crystal in the units of
sincgain_sparse.m:21-31 (rad/m), MATLAB's sinc(x) = sin(pi x)/(pi x).

Pins (tests/test_qpm.py): a single-domain crystal's sum is a geometric series
in closed form; a square-wave grating at exact first-order QPM sums to
(2/pi) L; sinc_gain's factor is 1 at zero mismatch and follows the
sinc^2 law at a few hand-evaluated mismatches.
"""
from __future__ import annotations

import numpy as np

import oracle

C0 = 299792458.0          # toolbox/static/CONST/c.m:2
EPS0 = 8.854e-12          # toolbox/static/CONST/eps0.m:2


def z_samples(L, N):
    """qpmgain.m:17-27: n_steps samples, z = cumsum([0, del_z ... , L - sum])."""
    n_steps = round(5e5 * L) if ((L > 1e-2 and N > 2 ** 13) or N > 2 ** 16) else round(1e6 * L)
    n_steps = int(n_steps)
    del_z = L / (n_steps - 1)
    dz = [0.0] + [del_z] * (n_steps - 2)
    dz.append(L - sum(dz))
    dz = np.array(dz)
    return n_steps, np.cumsum(dz), dz


def literal_sum(dk, z, w, chunk=4096):
    """sum_k w_k exp(i dk z_k), one exponential per sample (chunked over dk)."""
    dk = np.asarray(dk, dtype=np.float64).reshape(-1)
    out = np.zeros(dk.size, np.complex128)
    for a in range(0, dk.size, chunk):
        out[a:a + chunk] = np.exp(1j * np.outer(dk[a:a + chunk], z)) @ w
    return out


def qpmgain(L, period, chi2, material, T, step, lam, n_crystal, abs_spectrum, peak_wavelength,
            wavelength_fwhm, average_intensity, sigrange_nm=(1100.0, 1700.0)):
    """Returns (gain [signals x pumps], pump, signal, idler, dk)."""
    sig = np.array(sigrange_nm, dtype=np.float64) * 1e-9
    n_steps, z, dz = z_samples(L, lam.size)
    NS = int(np.floor(L / step))
    P = oracle.qpm_poling(NS, L, period) * chi2
    stride = NS // n_steps
    assert stride * n_steps == NS
    P = P[::stride]
    d_eff = P.max()
    P = P / P.max()
    prange = wavelength_fwhm * np.array([-3.0, 3.0]) + peak_wavelength
    pid = np.logical_and(lam > prange[0], lam < prange[1])
    pump = lam[pid]
    p_mask = (abs_spectrum / abs_spectrum.max())[pid]
    nP = n_crystal[pid]
    if pump.size > 100:
        f = int(np.floor(pump.size / 100 + 0.5))
        pump, p_mask, nP = pump[::f], p_mask[::f], nP[::f]
    if prange[0] < sig[0] < prange[1]:
        sig[0] = prange[1] + 1e-9
    elif prange[0] < sig[1] < prange[1]:
        sig[1] = prange[0] - 1e-9
    sid = np.logical_and(lam > sig[0], lam < sig[1])
    signal = lam[sid]
    nS = n_crystal[sid]
    if signal.size > 2000:
        f = int(np.floor(signal.size / 2000 + 0.5))
        signal, nS = signal[::f], nS[::f]
    ns_, np_ = signal.size, pump.size
    idler = np.zeros((ns_, np_))
    dkm = np.zeros((ns_, np_))
    I2A = np.sqrt(1.0 / (2.0 * EPS0 * C0))
    Ip = average_intensity
    Ii = Ip / 100.0
    wgt = np.zeros((ns_, np_))
    i_mask = p_mask.max()                 # rectangular idler: ones * max(p_mask)
    for i in range(ns_):
        for j in range(np_):
            li = C0 / (C0 / pump[j] - C0 / signal[i])      # dfg_lambda
            nI = abs(float(oracle.sellmeier_n(li * 1e6, material, T)[0]))
            kI = 2 * np.pi * nI / li
            kS = 2 * np.pi * nS[i] / signal[i]
            kP = 2 * np.pi * nP[j] / pump[j]
            idler[i, j] = li
            dkm[i, j] = abs(kP - kS - kI)
            wgt[i, j] = (np.sqrt(Ii / nI) * I2A) * (np.sqrt(Ip / nP[j]) * I2A) * i_mask * p_mask[j]
    As = literal_sum(dkm, z, P * dz).reshape(ns_, np_)
    kS2 = 2 * np.pi * nS[:, None] / signal[:, None]
    g = 1j * 2 * d_eff * (kS2 / nS[:, None] ** 2) * wgt
    As = g * As
    Is = nS[:, None] * (np.abs(As) / I2A) ** 2
    return Is / Ip, pump, signal, idler, dkm


def sinc_gain(lam_p, lam_s, lam_i, P1, P2, nx_pos, L, material, T, m=(1.0,)):
    """gain[p, s, g] = sum_m (sinc(dk L/2)/m)^2, dk = ||k_p - k_i| - k_s| - m 2pi/period_g."""
    def k(lam_um):
        n = np.abs(np.array([oracle.sellmeier_n(float(x), material, T)[0]
                             for x in np.ravel(lam_um)])).reshape(np.shape(lam_um))
        return 2 * np.pi * n / (np.asarray(lam_um) * 1e-6)

    dk0 = np.abs(np.abs(k(lam_p) - k(lam_i)) - k(lam_s))
    grating = np.linspace(P1, P2, nx_pos)
    out = np.zeros(dk0.shape + (nx_pos,))
    for g in range(nx_pos):
        for mj in m:
            x = (dk0 - mj * 2 * np.pi / grating[g]) * L / 2
            out[..., g] += (np.sinc(x) / mj) ** 2
    return out
