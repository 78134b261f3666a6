"""FP64 oracle of the fibre GNLSE solver (TEST INFRASTRUCTURE ONLY: only
tests/ and bench.py's CPU baseline may import it, as the checker).

gnlsefib.m:1-89 restated in numpy: the generalised NLSE in the interaction
picture (gnlsefib.m:22-61: linear operator L = i B - alpha/2 with the Taylor
series of the betas, self-steepening through W = V + w0 and gamma / w0, the
Raman response as a frequency-domain product), integrated over z with
MATLAB's ode45 at RelTol 1e-5, AbsTol 1e-12, NormControl on
(gnlsefib.m:71-74) and the solution returned at the requested points
Z = linspace(0, flength, nsaves) (gnlsefib.m:70), then changed back to the
lab frame (gnlsefib.m:79-86).

ode45 is MATLAB's, absent here.  Its published algorithm (Dormand & Prince
1980 for the 5(4) pair; Shampine & Reichelt, "The MATLAB ODE Suite", SIAM J.
Sci. Comput. 18 (1997) for the step control; Shampine 1986 for the
continuous extension used at requested points) is restated in `ode45`:
  * initial step: min(hmax, |tspan(2) - tspan(1)|), cut to 1 / rh with
    rh = (||f0|| / max(||y0||, threshold)) / (0.8 rtol^(1/5)) when larger;
    hmax = 0.1 |tfinal - t0|, threshold = atol / rtol;
  * a step is stretched to tfinal when 1.1 |h| >= |tfinal - t|;
  * error (NormControl): err = |h| ||sum_j E_j f_j|| / max(||y||, ||ynew||,
    threshold); accepted when err <= rtol;
  * first rejection: |h| *= max(0.1, 0.8 (rtol / err)^(1/5)); later ones in
    the same step: |h| /= 2;
  * after a step with no rejection: |h| /= 1.25 (err / rtol)^(1/5), or
    |h| *= 5 when that factor is <= 0.2;
  * the first-same-as-last stage carries f(tnew, ynew) to the next step;
  * at each requested point inside (t, tnew]: ynew if it is tnew, else the
    quartic continuous extension y + h f BI [s, s^2, s^3, s^4], s = (z - t)/h.

MATLAB's fft / ifft conventions are kept: gnlsefib.m works in the frequency
domain through ifft (AW = ifft(A)) and goes back with fft.

Pins (tests/test_gnlse.py): the Dormand-Prince order conditions and the
continuous extension's end values (s = 1 gives the 5th-order solution, its
derivative at s = 0 the first stage); a linear ODE y' = lambda y against
exp(lambda z) to the tolerance; dispersion only (gamma = 0) reproduces the
closed form exp(L z) exactly; self-phase modulation only (no dispersion,
Raman or shock) keeps |A(t)| and rotates the phase by gamma |A|^2 z; the
fundamental soliton keeps its shape; energy is conserved without loss and
Raman, and decays as exp(-alpha z) with loss alone.
"""
from __future__ import annotations

import math

import numpy as np

# Dormand & Prince (1980) 5(4) pair: nodes, stage matrix (column k gives the
# coefficients of stage k + 2, the last column the 5th-order solution) and
# the error weights b5 - b4 (seven stages, first same as last)
DP_A = np.array([1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])
DP_B = np.array([
    [1 / 5, 3 / 40, 44 / 45, 19372 / 6561, 9017 / 3168, 35 / 384],
    [0, 9 / 40, -56 / 15, -25360 / 2187, -355 / 33, 0],
    [0, 0, 32 / 9, 64448 / 6561, 46732 / 5247, 500 / 1113],
    [0, 0, 0, -212 / 729, 49 / 176, 125 / 192],
    [0, 0, 0, 0, -5103 / 18656, -2187 / 6784],
    [0, 0, 0, 0, 0, 11 / 84],
    [0, 0, 0, 0, 0, 0],
])
DP_E = np.array([71 / 57600, 0, -71 / 16695, 71 / 1920, -17253 / 339200, 22 / 525, -1 / 40])
# continuous extension (Shampine 1986): y(t + s h) = y + h f BI [s s^2 s^3 s^4]^T
DP_BI = np.array([
    [1, -183 / 64, 37 / 12, -145 / 128],
    [0, 0, 0, 0],
    [0, 1500 / 371, -1000 / 159, 1000 / 371],
    [0, -125 / 32, 125 / 12, -375 / 64],
    [0, 9477 / 3392, -729 / 106, 25515 / 6784],
    [0, -11 / 7, 11 / 3, -55 / 28],
    [0, 3 / 2, -4, 5 / 2],
])


def _comb(f, c):
    """sum_j c_j f_j over the stages, in stage order (the order the device uses)."""
    acc = np.zeros_like(f[0])
    for j in range(len(c)):
        if c[j] != 0.0:
            acc = acc + c[j] * f[j]
    return acc


def ode45(rhs, tspan, y0, rtol=1e-5, atol=1e-12, max_steps=10_000_000):
    """ode45 with NormControl on, outputs at every point of tspan (> 2 points).
    Returns (tout, yout [len(tspan), n], stats dict(accepted, rejected,
    evaluations))."""
    tspan = np.asarray(tspan, dtype=np.float64)
    t0, tfinal = float(tspan[0]), float(tspan[-1])
    tdir = 1.0 if tfinal >= t0 else -1.0
    y = np.asarray(y0, dtype=np.complex128).copy()
    n_out = tspan.size
    yout = np.zeros((n_out, y.size), np.complex128)
    yout[0] = y
    nxt = 1
    pow_ = 1.0 / 5.0
    threshold = atol / rtol
    hmax = 0.1 * abs(tfinal - t0)
    t = t0
    f = [None] * 7
    f[0] = rhs(t, y)
    nfev = 1
    normy = np.linalg.norm(y)
    hmin = 16 * np.spacing(t)
    absh = min(hmax, abs(tspan[1] - tspan[0]))
    rh = (np.linalg.norm(f[0]) / max(normy, threshold)) / (0.8 * rtol ** pow_)
    if absh * rh > 1:
        absh = 1.0 / rh
    absh = max(absh, hmin)
    acc = rej = 0
    done = False
    while not done:
        hmin = 16 * np.spacing(t)
        absh = min(hmax, max(hmin, absh))
        h = tdir * absh
        if 1.1 * absh >= abs(tfinal - t):
            h = tfinal - t
            absh = abs(h)
            done = True
        nofailed = True
        while True:
            hA = h * DP_A
            hB = h * DP_B
            for k in range(5):
                f[k + 1] = rhs(t + hA[k], y + _comb(f[:k + 1], hB[:k + 1, k]))
            tnew = t + hA[5]
            if done:
                tnew = tfinal
            h = tnew - t
            ynew = y + _comb(f[:6], hB[:6, 5])
            f[6] = rhs(tnew, ynew)
            nfev += 6
            normynew = np.linalg.norm(ynew)
            errwt = max(max(normy, normynew), threshold)
            err = absh * (np.linalg.norm(_comb(f, DP_E)) / errwt)
            if err > rtol:
                rej += 1
                if absh <= hmin:
                    raise RuntimeError(f"ode45: step size below hmin at t={t}")
                if nofailed:
                    nofailed = False
                    absh = max(hmin, absh * max(0.1, 0.8 * (rtol / err) ** pow_))
                else:
                    absh = max(hmin, 0.5 * absh)
                h = tdir * absh
                done = False
            else:
                break
        acc += 1
        if acc + rej > max_steps:
            raise RuntimeError("ode45: step cap reached")
        while nxt < n_out and tdir * (tnew - tspan[nxt]) >= 0:
            if tspan[nxt] == tnew:
                yout[nxt] = ynew
            else:
                s = (tspan[nxt] - t) / h
                cs = h * (DP_BI @ np.array([s, s * s, s ** 3, s ** 4]))
                yout[nxt] = y + _comb(f, cs)
            nxt += 1
        if done:
            break
        if nofailed:
            temp = 1.25 * (err / rtol) ** pow_
            absh = absh / temp if temp > 0.2 else 5.0 * absh
        t = tnew
        y = ynew
        normy = normynew
        f[0] = f[6]
    return tspan.copy(), yout, dict(accepted=acc, rejected=rej, evaluations=nfev)


def operators(T, w0, gamma, betas, loss, fr, RT):
    """gnlsefib.m:22-45: (V, L, W, RW, gamma_eff), L, W and RW in FFT order."""
    T = np.asarray(T, dtype=np.float64)
    n = T.size
    dT = T[1] - T[0]
    V = 2 * np.pi * np.arange(-n // 2, n // 2) / (n * dT)
    alpha = np.log(10.0 ** (loss / 10.0))
    B = np.zeros(n)
    for i, b in enumerate(np.atleast_1d(betas)):
        B = B + b / math.factorial(i + 2) * V ** (i + 2)
    L = 1j * B - alpha / 2
    if abs(w0) > np.finfo(float).eps:
        gamma_eff = gamma / w0
        W = V + w0
    else:
        gamma_eff = gamma
        W = np.ones(n)
    RT = np.atleast_1d(np.asarray(RT))
    RW = n * np.fft.ifft(np.fft.fftshift(RT)) if RT.size > 1 else None
    return V, np.fft.fftshift(L), np.fft.fftshift(W), RW, gamma_eff, B, dT


def rhs_factory(L, W, RW, gamma_eff, fr, dT):
    """gnlsefib.m:49-61: dAW/dz."""
    raman = RW is not None and abs(fr) >= np.finfo(float).eps

    def rhs(z, AW):
        AT = np.fft.fft(AW * np.exp(L * z))
        IT = np.abs(AT) ** 2
        if not raman:
            M = np.fft.ifft(AT * IT)
        else:
            RS = dT * fr * np.fft.fft(np.fft.ifft(IT) * RW)
            M = np.fft.ifft(AT * ((1 - fr) * IT + RS))
        return 1j * gamma_eff * W * M * np.exp(-L * z)
    return rhs


def gnlsefib(T, A, w0, gamma, betas, loss, fr, RT, flength, nsaves, rtol=1e-5, atol=1e-12):
    """(Z, AT, AW, W, B, stats) as gnlsefib.m returns them (+ the integrator's
    step counts)."""
    V, L, W, RW, g, B, dT = operators(T, w0, gamma, betas, loss, fr, RT)
    n = np.asarray(T).size
    rhs = rhs_factory(L, W, RW, g, fr, dT)
    Z = np.linspace(0.0, flength, nsaves)
    Z, AWz, stats = ode45(rhs, Z, np.fft.ifft(np.asarray(A, dtype=np.complex128)), rtol, atol)
    AT = np.zeros_like(AWz)
    AWo = np.zeros_like(AWz)
    for i in range(len(Z)):
        a = AWz[i] * np.exp(L * Z[i])
        AT[i] = np.fft.fft(a)
        AWo[i] = np.fft.fftshift(a) * dT * n
    return Z, AT, AWo, V + w0, B, stats
