"""Host-side producers of the crystal-solver and cavity inputs (FP64, numpy).

These mirror the MATLAB objects that build the arguments of the reference's
solver call (toolbox/lib/classes/OpticalSim.m:213-252): the simulation grid
(SimWindow.m), the MgO:PPLN dispersion (n_mgoppln_gayer.m, phi_calc.m,
phi2GD.m), the QPM poling (QPMcrystal.m, NonlinearCrystal.m:76-110), the pump
(Laser.m:66-77, sechPulse.m), and the cavity's spectral operators
(OpticalPulse.m:137-197,256-261).  New here: seeded one-photon-per-mode noise.

All arrays are numpy float64/complex128; the device library converts them.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C0 = 299792458.0          # toolbox/static/CONST/c.m:2
EPS0 = 8.854e-12          # toolbox/static/CONST/eps0.m:2
HBAR = 1.054571817e-34    # CODATA 2018 reduced Planck constant


class SimWindow:
    """Time-constrained simulation window (SimWindow.m:36-163).

    Index order of every spectral array here is the SimWindow order (centred:
    index N/2 is the reference frequency); `fft_order(x)` converts to the
    solver's FFT order (OpticalSim.m:223 fftshift)."""

    def __init__(self, lambda_ref: float, n_points: int, temporal_range: float,
                 spectral_limits_nm=(300.0, 6000.0)):
        if n_points < 4 or n_points & (n_points - 1):
            raise ValueError("n_points must be a power of two")
        self.reference_wave = float(lambda_ref)
        self.n = int(n_points)
        self.temporal_range = float(temporal_range)
        self.spectral_limits = tuple(float(x) for x in spectral_limits_nm)
        m = np.arange(-self.n // 2, self.n // 2, dtype=np.float64)
        self.times = m / self.n * self.temporal_range          # SimWindow.m:69
        self.dt = self.times[1] - self.times[0]                # SimWindow.m:126
        self.rel_frequencies = m / (self.n * self.dt)          # fftax.m
        f0 = C0 / self.reference_wave
        self.omegas = 2 * np.pi * (self.rel_frequencies + 1 * f0)  # SimWindow.m:102-106
        self.reference_omega = 2 * np.pi * C0 / self.reference_wave  # SimWindow.m:61-63
        self.rel_omegas = self.omegas - self.reference_omega   # SimWindow.m:108-110
        self.wavelengths = 2 * np.pi * C0 / self.omegas        # SimWindow.m:112-114
        self.reference_index = self.n // 2                     # SimWindow.m:161-163 (0-based)
        self.d_omega = self.omegas[1] - self.omegas[0]         # SimWindow.m:133-135
        self.d_nu = self.rel_frequencies[1] - self.rel_frequencies[0]

    @staticmethod
    def fft_order(x):
        return np.fft.ifftshift(x)

    @staticmethod
    def window_order(x):
        return np.fft.fftshift(x)

    def in_band(self):
        lam_nm = self.wavelengths * 1e9
        lo, hi = self.spectral_limits
        return (lam_nm >= lo) & (lam_nm <= hi) & (self.omegas > 0)

    # SpectralField / k2t (OpticalPulse.m:283-316): the 2N padded transform
    # decimated by 2 equals the N-point transform.
    @staticmethod
    def t2k(E):
        return np.fft.fftshift(np.fft.fft(np.fft.fftshift(E)))

    @staticmethod
    def k2t(Ek):
        return np.fft.ifftshift(np.fft.ifft(np.fft.ifftshift(Ek)))


def gayer_n(lam_um, T=20.0, col=1):
    """MgO:LN index, n_mgoppln_gayer.m:11-37; real part of a complex sqrt."""
    a1 = [5.756, 5.653, 5.078, 5.35583][col - 1]
    a2 = [0.0983, 0.1185, 0.0964, 0.100473][col - 1]
    a3 = [0.2020, 0.2091, 0.2065, 0.20692][col - 1]
    a4 = [189.32, 89.61, 61.16, 100][col - 1]
    a5 = [12.52, 10.85, 10.55, 11.34927][col - 1]
    a6 = [1.32e-2, 1.97e-2, 1.59e-2, 1.5334e-2][col - 1]
    b1 = [2.860e-6, 7.941e-7, 4.677e-6, 4.629e-7][col - 1]
    b2 = [4.700e-8, 3.134e-8, 7.822e-8, 3.862e-8][col - 1]
    b3 = [6.113e-8, -4.641e-9, -2.653e-8, -0.89e-8][col - 1]
    b4 = [1.516e-4, -2.188e-6, 1.096e-4, 2.657e-5][col - 1]
    rl = np.asarray(lam_um, dtype=np.float64)
    fn = (T - 24.5) * (T + 570.82)
    x = ((a1 + b1 * fn) + (a2 + b2 * fn) / (rl * rl - (a3 + b3 * fn) * (a3 + b3 * fn))
         + (a4 + b4 * fn) / (rl * rl - a5 * a5) - a6 * rl * rl)
    return np.sqrt(x.astype(np.complex128)).real


def _re_sqrt(x):
    """Real part of MATLAB's sqrt (complex for a negative argument)."""
    return np.sqrt(np.asarray(x, dtype=np.float64).astype(np.complex128)).real


def opgap_wei_n(lam_um, T=20.0):
    """OP-GaP index, Wei et al., Opt. Mater. Express 8, 485 (2018), with the
    temperature in kelvin inside (n_opgap_wei.m:1-15)."""
    Tk = T + 273.15
    l2 = np.asarray(lam_um, dtype=np.float64) ** 2
    A = 10.926 + 7.0787e-4 * Tk + 1.8594e-7 * Tk ** 2
    B = 0.53718 + 5.8035e-5 * Tk + 1.9819e-7 * Tk ** 2
    D = 1504.0 + 0.25935 * Tk - 0.00023326 * Tk ** 2
    return _re_sqrt(A + B / (l2 - 0.0911014) + D / (l2 - 758.048))


def ktp_z_n(lam_um, T=20.0):
    """KTP n_z: the Kato & Takaoka (2002) Sellmeier form at 20 C plus the
    Emanueli & Arie (2003) thermal terms in (T - 20) (n_ktp_zhao.m:1-38)."""
    lam = np.asarray(lam_um, dtype=np.float64)
    l2 = lam * lam
    n0 = _re_sqrt(2.12725 + 1.18431 / (1 - 5.14852e-2 / l2) + 0.6603 / (1 - 100.00507 / l2)
                  - 9.68956e-3 * l2)
    dT = T - 20.0
    n1 = -6.1537e-6 + 64.505e-6 / lam - 56.447e-6 / l2 + 17.169e-6 / (l2 * lam)
    n2 = -0.96751e-8 + 13.192e-8 / lam - 11.78e-8 / l2 + 3.6292e-8 / (l2 * lam)
    return n0 + n1 * dT + n2 * dT ** 2


def load_material_table(path):
    """A material table in the layout of the reference's Materials.csv
    (sellmeier.m:22-51): columns Material, six coefficients, Equivalent,
    Equation ("Schott" or "CDGM"), ...  Returns {name: (equation, coeffs)}.
    Rows whose six coefficient cells are not all numbers are skipped: the
    reference's own table tags PPKTP "Schott" with blank coefficients, and
    sellmeier() never reads that row (it dispatches PPKTP by name)."""
    import csv
    out = {}
    keys = ["B1", "B2", "B3", "C1", "C2", "C3"]
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            name = (row.get("Material") or "").strip()
            eq = (row.get("Equation") or "").strip()
            if not name or eq not in ("Schott", "CDGM"):
                continue
            try:
                coeffs = [float(row[k]) for k in keys]
            except (TypeError, ValueError, KeyError):
                continue
            if not all(np.isfinite(coeffs)):
                continue
            out[name] = (eq, coeffs)
    return out


def glass_n(lam_um, equation, coeffs):
    """sellmeier.m:27-51: the Schott form n^2 = 1 + sum B_i l^2 / (l^2 - C_i)
    (coefficients B1..B3, C1..C3, C in um^2) or the CDGM form
    n^2 = A0 + A1 l^2 + A2 l^-2 + A3 l^-4 + A4 l^-6 + A5 l^-8 (A0..A5 in the
    same six columns); real part of the square root."""
    x = np.asarray(lam_um, dtype=np.float64)
    x2 = x * x
    if equation == "Schott":
        B1, B2, B3, C1, C2, C3 = coeffs
        nf = 1 + B1 * x2 / (x2 - C1) + B2 * x2 / (x2 - C2) + B3 * x2 / (x2 - C3)
    elif equation == "CDGM":
        A0, A1, A2, A3, A4, A5 = coeffs
        nf = A0 + A1 * x2 + A2 / x2 + A3 / x2 ** 2 + A4 / x2 ** 3 + A5 / x2 ** 4
    else:
        raise ValueError(f"unknown equation {equation!r}")
    return _re_sqrt(nf)


def sellmeier_n(lam_um, material="PPLN", T=20.0, table=None):
    """sellmeier.m:3-56 (real part): the crystal materials by name, any other
    material from `table` (load_material_table of a Materials.csv-layout
    file; glasses for cavity optics, no temperature dependence)."""
    if material in ("PPLN", "LN_e"):
        return gayer_n(lam_um, T, 1)
    if material == "LN_o":
        return gayer_n(lam_um, T, 2)
    if material == "PPLN_undoped":
        return gayer_n(lam_um, T, 4)
    if material in ("OP-GaP", "OPGaP"):
        return opgap_wei_n(lam_um, T)
    if material in ("KTP", "PPKTP"):
        return ktp_z_n(lam_um, T)
    if table is not None and material in table:
        return glass_n(lam_um, *table[material])
    raise ValueError(f"unknown material {material!r} (glasses need a material table)")


def phi_calc(L, T, simwin: SimWindow, material="PPLN", table=None):
    """phi_calc.m:5-45: (phi_rel, phi_abs).  The string "PPLN" takes the
    Gayer index at T; every other material goes through sellmeier() at its
    default 20 C (phi_calc.m:17-20 passes no temperature there).

    phi_rel is defined on the 0.2-20 um knot band (NaN elsewhere): at the
    knots the reference's spline (phi_calc.m:40) returns the cumtrapz values
    themselves, and outside the 300-6000 nm spectral limits the crystal
    transmission is zero, so the extrapolated values never reach the field."""
    w_abs, w_rel = simwin.omegas, simwin.rel_omegas
    lam = 2 * np.pi * C0 / w_abs
    n_mat = gayer_n(lam * 1e6, T, 1) if material == "PPLN" else \
        sellmeier_n(lam * 1e6, material, 20.0, table)
    beta_abs = w_abs * n_mat / C0
    phi_abs = L * beta_abs
    gd_abs = np.concatenate([[0.0], np.diff(phi_abs) / np.diff(w_abs)])
    ref = simwin.reference_index
    gd_rel = gd_abs - gd_abs[ref]
    del_phi = gd_rel * np.concatenate([[0.0], np.diff(w_rel)])
    band = (lam > 2e-7) & (lam < 2e-5)
    idx = np.nonzero(band)[0]
    if idx.size == 0 or not band[ref] or np.any(np.diff(idx) != 1):
        raise ValueError("phi_calc: knot band must be contiguous and contain the reference")
    y = del_phi[idx]
    acc = np.concatenate([[0.0], np.cumsum(0.5 * (y[1:] + y[:-1]))])
    phi = np.full(simwin.n, np.nan)
    phi[idx] = acc
    return phi - phi[ref], phi_abs


def path_length(simwin: SimWindow, material, length, table=None, T=20.0):
    """Dielectric.PathLength (Dielectric.m:94-101): a number, or for a length
    given as a function of the refractive index (a prism) that function of
    the material's index as Dielectric.simulate forms it (Dielectric.m:39-46:
    sellmeier at the optic's temperature, real part, clamped at 1), one path
    per wavelength."""
    if not callable(length):
        return length
    n = np.real(sellmeier_n(simwin.wavelengths * 1e6, material, T, table))
    return np.asarray(length(np.where(n < 1, 1.0, n)), dtype=np.float64)


def prism_pb_path_length(n, L1=19.0):
    """prismfn_PB.m:1-29, the reference's Pellin-Broca prism: the path
    through the prism [m] for index n, apex angle a = 78 deg 26', Brewster
    incidence at 54 deg: 2 L1 sin(a) / cos(asin(sin(54 deg) / n)) / 1000."""
    a = math.radians(78 + 26 / 60)
    th_i = np.real(np.arcsin(np.sin(math.radians(54.0)) / np.asarray(n, dtype=np.complex128)))
    return 2 * L1 * math.sin(a) / np.cos(th_i) / 1000.0


def bulk_dispersion(simwin: SimWindow, material, path_length_, table=None, T=20.0):
    """Dielectric.Dispersion (Dielectric.m:106-114): the relative spectral
    phase of a bulk optic, phi_calc of its path length (a number, or a
    function of the index for a prism: see path_length; NaN outside the
    0.2-20 um knot band).  Its factor exp(-i phi) multiplies a cavity's
    spectral transfer where the optic sits in the chain."""
    L = path_length(simwin, material, path_length_, table, T)
    return phi_calc(L, T, simwin, material, table)[0]


def fresnel_t(n1, n2, theta_deg=0.0, exit=False):
    """fresnel.m:1-38: power transmission (Tp, Ts) of a plane interface from
    index n1 into n2 at incidence theta (degrees); with exit=True the same
    surface crossed the other way (the angles and the index ratio swap).
    Real part, as the reference takes it (no transmission past the
    critical angle)."""
    n1 = np.asarray(n1, dtype=np.complex128)
    n2 = np.asarray(n2, dtype=np.complex128)
    ti = np.deg2rad(theta_deg) + 0j
    nr = n1 / n2
    tt = np.arcsin(nr * np.sin(ti))
    if exit:
        ti, tt, nr = tt, ti, 1.0 / nr
    ts = 2 * nr * np.cos(ti) / (nr * np.cos(ti) + np.cos(tt))
    tp = 2 * nr * np.cos(ti) / (nr * np.cos(tt) + np.cos(ti))
    geo = (1.0 / nr) * (np.cos(tt) / np.cos(ti))
    return (tp ** 2 * geo).real, (ts ** 2 * geo).real


def surface_transmission(coating, simwin: SimWindow, material=None, theta_deg=0.0, order=1,
                         table=None, T=20.0, data_dir=None):
    """OpticalSurface.Transmission (OpticalSurface.m:45-82): a number (percent
    if > 1), a table of [T, upper wavelength limit] rows, a callable of the
    wavelengths, "AR" (1), "None" (the bare surface: fresnel's p transmission
    from air into the material, or out of it for the second surface, order
    2), or the name of a transmission data file in `data_dir` (transmission(),
    shifted towards shorter wavelengths at oblique incidence: the curve moves
    by 0.2 lam_max sin^2(theta), lam_max the wavelength of peak transmission,
    in whole grid steps rounded up, OpticalSurface.m:73-82)."""
    lam = simwin.wavelengths
    if isinstance(coating, str):
        if coating == "AR":
            return np.ones(lam.size)
        if coating == "None":
            n = sellmeier_n(lam * 1e6, material, T, table)
            return fresnel_t(1.0, n, theta_deg, exit=(order == 2))[0]
        if data_dir is None:
            raise ValueError(f"coating {coating!r} is a data file: give data_dir")
        import os
        t = transmission(os.path.join(data_dir, coating), lam)[0]
        if theta_deg != 0:
            i = int(np.argmax(t))
            dlam = (lam[i + 1] - lam[i - 1]) / 2
            shift = lam[i] * 0.2 * math.sin(math.radians(theta_deg)) ** 2
            t = np.roll(t, -int(math.ceil(shift / dlam)))
        return t
    if callable(coating):
        return np.asarray(coating(lam), dtype=np.float64)
    c = np.asarray(coating, dtype=np.float64)
    if c.ndim == 0:
        t = np.full(lam.size, float(c))
    else:
        vals, lims = c[:, 0], c[:, 1]
        t = np.ones(lam.size)
        t[lam < lims[0]] = vals[0]
        for k in range(1, len(vals)):
            t[(lam > lims[k - 1]) & (lam < lims[k])] = vals[k]
    return t / 100.0 if np.any(t > 1) else t


def matlab_smooth(y, span):
    """smooth(y, span), 'moving' method (import_interp.m:13 smooths data files
    with span 0.01): span < 1 is a fraction of the points, rounded up; an even
    span drops by one; each point is the mean of the `span` points centred on
    it, and near the ends of the largest centred window that fits (1, 3,
    5, ... points: yy(1) = y(1), yy(2) = mean(y(1:3)), ...)."""
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    if span < 1:
        span = math.ceil(span * n)
    span = min(int(span), n)
    w = span - 1 + span % 2
    if w <= 1:
        return y.copy()
    idx = np.arange(n)
    half = np.minimum(np.minimum(idx, n - 1 - idx), w // 2)
    c = np.concatenate([[0.0], np.cumsum(y)])
    lo, hi = idx - half, idx + half + 1
    return (c[hi] - c[lo]) / (hi - lo)


def makima(x, y, xq, fill=0.0):
    """interp1(x, y, xq, 'makima', fill): the modified Akima piecewise cubic
    (Akima 1970 with the weights w = |d_{i+1} - d_i| + |d_{i+1} + d_i| / 2),
    `fill` outside [x_1, x_n]."""
    from scipy.interpolate import Akima1DInterpolator
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    order = np.argsort(x, kind="stable")
    f = Akima1DInterpolator(x[order], y[order], method="makima")
    out = f(np.asarray(xq, dtype=np.float64), extrapolate=False)
    return np.where(np.isnan(out), fill, out)


def read_data_table(path):
    """The two numeric columns of a data file as readtable gives them
    (header lines and non-numeric rows skipped; comma, tab or space separated)."""
    xs, ys = [], []
    with open(path) as f:
        for line in f:
            parts = line.replace(",", " ").split()
            if len(parts) < 2:
                continue
            try:
                a, b = float(parts[0]), float(parts[1])
            except ValueError:
                continue
            xs.append(a)
            ys.append(b)
    if not xs:
        raise ValueError(f"no numeric rows in {path}")
    return np.array(xs), np.array(ys)


def import_interp(path, x_sim, xfun=None, yfun=None):
    """import_interp.m:1-23: a two-column data file onto x_sim: the y column
    smoothed (smooth, span 0.01), mapped by yfun, padded with zeros at
    min|x_sim| and max(x_sim) when x_sim covers the data on both sides, then
    makima with zero outside."""
    x_dat, y_dat = read_data_table(path)
    if xfun is not None:
        x_dat = xfun(x_dat)
    y_dat = matlab_smooth(y_dat, 0.01)
    if yfun is not None:
        y_dat = yfun(y_dat)
    x_sim = np.asarray(x_sim, dtype=np.float64)
    if np.min(np.abs(x_sim)) < np.min(np.abs(x_dat)) and np.max(np.abs(x_sim)) > np.max(np.abs(x_dat)):
        x_dat = np.concatenate([[np.min(np.abs(x_sim))], x_dat, [np.max(x_sim)]])
        y_dat = np.concatenate([[0.0], y_dat, [0.0]])
    return makima(x_dat, y_dat, x_sim, 0.0)


def transmission(path, l_sim, nsteps=1, nmat=1.0, frnum=0, lim1=300e-9, lim2=None, table=None):
    """transmission.m:1-44: a transmission data file (wavelength in nm, T in
    percent) on the wavelengths l_sim [m] -> (T, T_step, a_step): clipped to
    [0, 1]; divided by the Fresnel factor (4 n / (1 + n)^2)^frnum of the
    material (nmat: indices, or a material name for sellmeier); zero outside
    [lim1, lim2]; renormalised when above 1; per-step loss over nsteps."""
    l_sim = np.asarray(l_sim, dtype=np.float64)
    if lim2 is None:
        lim2 = float(np.max(l_sim))
    T = import_interp(path, l_sim * 1e9, yfun=lambda v: v / 100.0)
    T = np.where(np.isfinite(T), T, 0.0)
    T = np.clip(T, 0.0, 1.0)
    if isinstance(nmat, str):
        nmat = sellmeier_n(l_sim * 1e6, nmat, 20.0, table)
    nmat = np.asarray(nmat)
    with np.errstate(divide="ignore", invalid="ignore"):
        T = np.real(T / ((4.0 * nmat) / (1.0 + nmat) ** 2) ** frnum)
    T = np.where(np.isfinite(T), T, 0.0)
    T = np.where(T < 0, 0.0, T)
    T = np.where((l_sim < lim1) | (l_sim > lim2), 0.0, T)
    if np.max(T) > 1:
        T = T / np.max(T)
    with np.errstate(divide="ignore"):
        a_step = -np.log(T) / nsteps / 2.0
    T_step = np.exp(-2.0 * a_step)
    return T, T_step, a_step


def load_transmission_table(path):
    """The transmission columns of a table in the layout of the reference's
    Materials.csv: {material: (T_file, T_length_m, Fresnel_Reflections)} for
    the rows that name a file (Dielectric.m:47-50,70-78)."""
    import csv
    out = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            name = (row.get("Material") or "").strip()
            tf = (row.get("T_file") or "").strip()
            if not name or not tf:
                continue
            try:
                out[name] = (tf, float(row["T_length_m"]), float(row["Fresnel_Reflections"]))
            except (TypeError, ValueError, KeyError):
                continue
    return out


def bulk_transmission(simwin: "SimWindow", material, path_length_, transmission_table, data_dir,
                      table=None, T=20.0):
    """Dielectric.Transmission (Dielectric.m:68-92): the material's data file
    (T_file, measured over T_length_m with Fresnel_Reflections surfaces, from
    `transmission_table`, see load_transmission_table) through transmission()
    with the material's real index (clamped at 1, Dielectric.m:39-46) within
    the window's spectral limits, scaled to the path length (a number or a
    function of the index, see path_length): T^(path / T_length_m)."""
    import os
    tfile, tlen, frnum = transmission_table[material]
    lam = simwin.wavelengths
    nr = np.real(sellmeier_n(lam * 1e6, material, T, table))
    nr = np.where(nr < 1, 1.0, nr)
    lo, hi = simwin.spectral_limits
    Tm, _, _ = transmission(os.path.join(data_dir, tfile), lam, 1, nr, frnum, lo * 1e-9, hi * 1e-9)
    return Tm ** (path_length(simwin, material, path_length_, table, T) / tlen)


def _movmean(x, k):
    """MATLAB movmean(x, k) for odd k with its default shrinking endpoints."""
    h = k // 2
    c = np.concatenate([[0.0], np.cumsum(x)])
    lo = np.maximum(np.arange(x.size) - h, 0)
    hi = np.minimum(np.arange(x.size) + h + 1, x.size)
    return (c[hi] - c[lo]) / (hi - lo)


def load_gdd(path):
    """A GDD data file as GDDimport2phi reads it: a table with columns
    Wavelength [nm] and GDD [fs^2]."""
    import csv
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    return (np.array([float(r["Wavelength"]) for r in rows]),
            np.array([float(r["GDD"]) for r in rows]))


def gdd_to_phi(simwin: SimWindow, wavelength_nm, gdd_fs2):
    """GDDimport2phi.m:1-33: relative spectral phase of a measured GDD curve
    (a mirror coating).  The GDD is smoothed over 5 samples, spline
    interpolated onto the grid's frequencies (held at the end values outside
    the data), integrated once to the group delay (relative to the reference
    frequency) and once more to the phase.  NaN below 200 nm, where the
    reference's spline extrapolates."""
    from scipy.interpolate import CubicSpline
    w = 2e9 * np.pi * C0 / np.asarray(wavelength_nm, dtype=np.float64)
    gdd = _movmean(np.asarray(gdd_fs2, dtype=np.float64) / 1e30, 5)
    order = np.argsort(w)
    spline = CubicSpline(w[order], gdd[order], bc_type="not-a-knot")
    w_abs, w_rel = simwin.omegas, simwin.rel_omegas
    g = spline(w_abs)
    g[w_abs > w.max()] = gdd[0]
    g[w_abs < w.min()] = gdd[-1]
    dw = np.diff(w_abs)
    dw = np.concatenate([[dw[0]], dw])
    y = g * dw
    gd = np.concatenate([[0.0], np.cumsum(0.5 * (y[1:] + y[:-1]))])
    ref = simwin.reference_index
    gd_rel = gd - gd[ref]
    dphi = gd_rel * np.concatenate([[0.0], np.diff(w_rel)])
    lam = 2 * np.pi * C0 / w_abs
    idx = np.nonzero(lam > 2e-7)[0]
    phi = np.full(simwin.n, np.nan)
    yy = dphi[idx]
    phi[idx] = np.concatenate([[0.0], np.cumsum(0.5 * (yy[1:] + yy[:-1]))])
    return phi - phi[ref]


def _matlab_round(x):
    """MATLAB round: halves away from zero."""
    return int(np.sign(x) * np.floor(abs(x) + 0.5))


def coating_phase(simwin: SimWindow, gdd, theta_deg=0.0):
    """OpticalSurface.Dispersion (OpticalSurface.m:96-114): the relative
    phase of a coating's measured GDD curve `gdd` = (wavelength_nm, gdd_fs2)
    (gdd_to_phi), zero without one.  At oblique incidence the curve moves to
    shorter wavelengths: circshift by -round(0.2 lam0 sin^2(theta) / dlam)
    samples, dlam the (signed) wavelength spacing at the reference index."""
    if gdd is None:
        return np.zeros(simwin.n)
    phi = gdd_to_phi(simwin, *gdd)
    if theta_deg != 0:
        lam = simwin.wavelengths
        ref = simwin.reference_index
        dlam = (lam[ref + 1] - lam[ref - 1]) / 2
        shift = lam[ref] * 0.2 * np.sin(np.deg2rad(theta_deg)) ** 2
        phi = np.roll(phi, -_matlab_round(shift / dlam))
    return phi


def optic_response(simwin: SimWindow, regime="T", s1="None", material="FS", length=1e-3,
                   theta_deg=0.0, s2=None, table=None, T=20.0, gdd=None, gdd2="s1",
                   transmission_table=None, data_dir=None):
    """Optic.simulate (Optic.m:79-99): (transmission, reflection, dispersion
    phase).  Transmission = S1 * bulk * S2 (surface_transmission; s2
    defaults to s1, the second surface crossed on the way out; the bulk's
    data-file transmission over `length` when `transmission_table` and
    `data_dir` are given, see bulk_transmission, else a lossless bulk),
    reflection = 1 - S1's.  The phase is
    S1's coating phase (coating_phase of a measured GDD curve `gdd` =
    (wavelength_nm, gdd_fs2), else 0) and, in transmission, also the bulk's
    relative phase over `length` and S2's coating phase.  S2 carries the
    same curve as S1 unless `gdd2` names another (None: no curve): in the
    reference the second surface defaults to a copy of the first, GDD
    included (Optic.m:57,72-75).  `T` is the bulk's temperature; the bare
    surfaces take fresnel's default 20 C index (fresnel.m:10-16)."""
    s2 = s1 if s2 is None else s2
    t1 = surface_transmission(s1, simwin, material, theta_deg, 1, table, 20.0, data_dir)
    t2 = surface_transmission(s2, simwin, material, theta_deg, 2, table, 20.0, data_dir)
    tb = 1.0
    if transmission_table is not None:
        tb = bulk_transmission(simwin, material, length, transmission_table, data_dir, table, T)
    phi = coating_phase(simwin, gdd, theta_deg)
    if regime == "T":
        g2 = gdd if (isinstance(gdd2, str) and gdd2 == "s1") else gdd2
        phi = phi + bulk_dispersion(simwin, material, length, table, T) + \
            coating_phase(simwin, g2, theta_deg)
    return t1 * tb * t2, 1.0 - t1, phi


def cavity_response(optics):
    """Cavity.simulate (Cavity.m:50-87): the power transmission of a chain
    of optic_response tuples (the reflection of "R" optics, the
    transmission of the others; given as (regime, response)) and the sum of
    their phases."""
    t, phi = 1.0, 0.0
    for regime, (tr, rf, ph_) in optics:
        t = t * (rf if regime == "R" else tr)
        phi = phi + ph_
    return t, phi


def phi2gd(phi, dw):
    """phi2GD.m:1-17 (GD only)."""
    n = phi.size
    phi_r = phi - phi[n // 2 - 1]
    gd = np.diff(phi_r) / dw
    return np.concatenate([[gd[0]], gd])


def poling_from_widths(nsteps, L, widths):
    """QPMcrystal.m:32-51 for given domain widths: walls at their cumulative
    sums, the last domain extended or cut to L, then +-1 per step of
    z = linspace(0, L, nsteps) by the parity of its domain."""
    walls = np.concatenate([[0.0], np.cumsum(np.asarray(widths, dtype=np.float64))])
    if walls[-1] < L:
        walls = np.concatenate([walls, [L]])
    else:
        walls[-1] = L
    z = (np.arange(nsteps, dtype=np.float64) * L) / (nsteps - 1)   # linspace(0, L, nsteps)
    z[-1] = L
    bins = np.searchsorted(walls, z, side="right")        # discretize: edges(k) <= z < edges(k+1)
    bins = np.minimum(bins, walls.size - 1)               # last bin includes the right edge
    q = np.mod(bins, 2)
    return -(2.0 * q - 1.0)


def qpm_poling(nsteps, L, period, duty_off=0.0, uncertainty=0.0, seed=0):
    """QPMcrystal.m:1-57 for one period: +-1 per step.  With a domain-width
    uncertainty the odd domains are drawn from a normal of sigma
    uncertainty/2 around the poled width (QPMcrystal.m:27-29) from a seeded
    PCG64 stream; MATLAB's rng('default') + normrnd stream is not reproduced."""
    n_dom = int(math.floor(2.0 * L / period))
    poled = period * (1.0 - 2.0 * duty_off) / 2.0
    if uncertainty > 0:
        widths = np.random.Generator(np.random.PCG64(seed)).normal(poled, uncertainty / 2.0,
                                                                     n_dom)
    else:
        widths = np.full(n_dom, poled)
    for i in range(1, n_dom - 1, 2):          # domain_widths(2:2:nDs-1), QPMcrystal.m:30
        widths[i] = period - 0.5 * (widths[i - 1] + widths[i + 1])
    return poling_from_widths(nsteps, L, widths)


def _mround(x):
    """MATLAB round: half away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def chirped_grating(z, P1=6.3e-6, P2=2.2e-6, a=0.48, tol=0.1e-6):
    """chirpedgrating.m:1-50: domain widths of a grating whose period runs
    from P1 (+ tol) to P2 as (z / max z)^a, quantised to tol/2 steps; each
    domain's width is read at the position reached by stepping whole
    domains, until the next domain would pass max(z)."""
    z = np.asarray(z, dtype=np.float64)
    zmax = z.max()
    p1 = P1 + tol
    g = p1 + (P2 - p1) * (z / zmax) ** a - (P2 + tol / 2.0)
    dom = _mround(g / tol) * tol / 2.0 + P2 / 2.0
    dz = z[1] - z[0]
    n_dom = int(math.ceil(np.sum(dz / dom)))
    dsteps = _mround(dom / dz).astype(np.int64)
    out, sid, pos = [], 0, 0.0
    for _ in range(n_dom):
        out.append(dom[sid])
        sid = min(sid + int(dsteps[sid]), z.size - 1)
        pos += out[-1]
        if pos + dom[sid] > zmax:
            break
    return np.array(out)


@dataclass
class NonlinearCrystal:
    """Poled crystal (NonlinearCrystal.m, Dielectric.m).

    period, as NonlinearCrystal.pole (NonlinearCrystal.m:83-110): a number
    (one period), two numbers (fan-out: P1 + (P2 - P1) * y / height), more
    numbers (stepped: grating number ceil(y), 1-based), or a callable of the
    step positions returning domain widths (e.g. a chirped_grating)."""
    period: object                 # [m], see above
    length: float                  # [m]
    chi2: float = 2 * 27e-12       # NonlinearCrystal.m:53-54 default for PPLN
    temperature: float = 20.0      # Optic.m:56 default
    duty_off: float = 0.0
    material: str = "PPLN"
    uncertainty: float = 0.0       # domain-width uncertainty [m]
    height: float = 1.0            # fan-out / stepped crystal height (units of y)
    vertical_position: float = 0.0
    seed: int = 0

    def local_period(self):
        """The single period the grating uses, or None for a domain-width callable."""
        if callable(self.period):
            return None
        p = np.atleast_1d(np.asarray(self.period, dtype=np.float64))
        if p.size == 1:
            return float(p[0])
        if p.size == 2:                                   # fanout, NonlinearCrystal.m:131-139
            return float(p[0] + (p[1] - p[0]) / self.height * self.vertical_position)
        return float(p[int(math.ceil(self.vertical_position)) - 1])

    def poling(self, nsteps):
        P = self.local_period()
        if P is None:                                     # QPMcrystal.m:14-15
            z = np.linspace(0.0, self.length, nsteps)
            return poling_from_widths(nsteps, self.length, self.period(z))
        return qpm_poling(nsteps, self.length, P, self.duty_off, self.uncertainty, self.seed)

    def nsteps(self, step):
        return int(math.floor(self.length / step))      # NonlinearCrystal.m:79

    def refractive_index(self, simwin: SimWindow):
        """Dielectric.simulate (Dielectric.m:39-46): real part, clamped at 1."""
        nr = sellmeier_n(simwin.wavelengths * 1e6, self.material, self.temperature)
        return np.where(nr < 1, 1.0, nr)

    def solver_inputs(self, simwin: SimWindow, step=1e-7):
        """Arguments of the solver call (OpticalSim.m:213-226), FFT order."""
        ns = self.nsteps(step)
        P = self.poling(ns)
        nr = self.refractive_index(simwin)
        ref = simwin.reference_index
        n0 = nr[ref]
        w0 = simwin.reference_omega
        G33 = P * self.chi2 * w0 / n0 / 4 / C0
        beta0_abs = 2 * np.pi * nr / simwin.wavelengths
        phi_rel, phi_abs = phi_calc(self.length, self.temperature, simwin, self.material)
        gd = phi2gd(phi_abs, simwin.d_omega)
        beta1_w0 = gd[ref] / self.length
        bdiffw0 = beta0_abs[ref] - beta1_w0 * w0
        T = simwin.in_band().astype(np.float64)          # transmission.m:28 (lossless band)
        alpha = T ** (1.0 / ns)
        hB = phi_rel / self.length * step
        hB = np.where(np.isfinite(hB) & (T > 0), hB, 0.0)
        return dict(G33=G33, beta_op=simwin.fft_order(hB), alpha_w=simwin.fft_order(alpha),
                    bdiffw0=float(bdiffw0), w0=float(w0), h=float(step), nSteps=ns,
                    dt=float(simwin.dt), n_xtal=nr, n0=float(n0))


def sech_pump(simwin: SimWindow, fwhm, energy, area, n_medium=1.0):
    """sech^2-intensity pump with the given energy (Laser.m:66-77).

    E(t) = E0 sech(t/tau), tau = fwhm / (2 ln(1+sqrt 2)) (sechPulse.m:12);
    peak intensity from energy / (sum |E|^2 dt) / area (Laser.m:106-110)."""
    tau = fwhm / (2.0 * math.log(1.0 + math.sqrt(2.0)))
    shape = 1.0 / np.cosh(simwin.times / tau)
    I0 = energy / (np.sum(shape ** 2) * simwin.dt) / area
    E0 = math.sqrt(I0 * 2.0 / n_medium / EPS0 / C0)
    return (E0 * shape).astype(np.complex128)


def quantum_noise(simwin: SimWindow, area, seed, n_medium=1.0):
    """One photon per mode with uniform random phase (new; no reference
    counterpart).  Energy per spectral bin follows I2pow.m:15-18:
    U_k = |Ek|^2 (c n eps0 / 2) A dt / N = hbar w_k."""
    rng = np.random.Generator(np.random.PCG64(seed))
    phase = rng.uniform(0.0, 2 * np.pi, simwin.n)
    w = simwin.omegas
    band = simwin.in_band()
    amp = np.zeros(simwin.n)
    amp[band] = np.sqrt(2 * HBAR * w[band] * simwin.n / (C0 * n_medium * EPS0 * area * simwin.dt))
    Ek = amp * np.exp(1j * phase)
    return SimWindow.k2t(Ek)


def energy_of(simwin: SimWindow, E, area, n_medium=1.0):
    """Pulse energy from the spectral field (OpticalPulse.m:404-418, I2pow.m)."""
    Ek = SimWindow.t2k(E)
    return float(np.sum(np.abs(Ek) ** 2 * (C0 * n_medium * EPS0 / 2) * area * simwin.dt / simwin.n))


def qpm_signal_wavelength(lambda_p, period, temperature=20.0, lo=None, hi=None):
    """Signal wavelength of first-order QPM, k_p - k_s - k_i = 2 pi / period,
    with the Gayer extraordinary index (bisection between 1.05 lambda_p and
    degeneracy, 2 lambda_p, unless lo / hi are given).  Without a root in the
    interval, the end with the smaller phase mismatch."""
    lo = 1.05 * lambda_p if lo is None else lo
    hi = 2.0 * lambda_p if hi is None else hi
    def dk(ls):
        li = 1.0 / (1.0 / lambda_p - 1.0 / ls)
        n = gayer_n(np.array([lambda_p, ls, li]) * 1e6, temperature, 1)
        return 2 * np.pi * (n[0] / lambda_p - n[1] / ls - n[2] / li - 1.0 / period)
    a, b = lo, hi
    fa, fb = dk(a), dk(b)
    if fa * fb > 0:
        # no exact match (e.g. a grating past degeneracy): the best-matched end
        return a if abs(fa) < abs(fb) else b
    for _ in range(80):
        m = 0.5 * (a + b)
        fm = dk(m)
        if fa * fm <= 0:
            b = m
        else:
            a, fa = m, fm
    return 0.5 * (a + b)


def sync_delay(simwin, length, temperature, lambda_s):
    """Cavity delay that cancels the signal's group delay relative to the
    crystal frame (which moves with the pump at the reference wavelength):
    -GD_rel(lambda_s) of one pass (phi_calc + phi2GD)."""
    phi_rel, _ = phi_calc(length, temperature, simwin)
    gd = phi2gd(phi_rel, simwin.d_omega)
    k = int(np.argmin(np.abs(simwin.wavelengths - lambda_s)))
    return -float(gd[k])


def smoothed_top_hat(x, a, b, w):
    """0.5 (tanh((x - a)/w) - tanh((x - b)/w)) (smoothedTopHat.m, tanh form)."""
    return 0.5 * (np.tanh((x - a) / w) - np.tanh((x - b) / w))


@dataclass
class OutputCoupler:
    """Resonator feedback and output coupling.

    With top = smoothed_top_hat(lambda, band, edge), the cavity keeps
      R = (1 - transmission) * top
    of the power, and the output port carries
      T = transmission * top;
    the remaining 1 - top (the depleted pump and the idler) is dumped at the
    band filter every trip.  This is a singly resonant cavity as in the
    reference's OPO setup (toolbox/scripts/simulation_script_template.m:44-48:
    OC coating 0.5*smoothedTopHat(lam, 1.2e-6, 1.8e-6, 30e-9), pump filter
    smoothedTopHat(lam, 0.9e-6, 5.5e-6, 30e-9)), with the coupler's
    transmission as a parameter.  band=None gives a flat coupler: R = 1 - t,
    T = t everywhere (the "R" regime of Optic.m / OpticalPulse.m:163-188)."""
    transmission: float = 0.05
    phase: float = 0.0
    band: tuple | None = (1.2e-6, 1.8e-6)
    edge: float = 30e-9

    def masks(self, simwin: SimWindow):
        lam = simwin.wavelengths
        if self.band is None:
            R = np.full(simwin.n, 1.0 - self.transmission)
        else:
            top = smoothed_top_hat(lam, self.band[0], self.band[1], self.edge)
            top = np.where(simwin.omegas > 0, top, 0.0)
            R = (1.0 - self.transmission) * top
            T = self.transmission * top
            return R, T, np.full(simwin.n, self.phase)
        T = 1.0 - R
        ph = np.full(simwin.n, self.phase)
        return R, T, ph


def cavity_operators(simwin: SimWindow, n_xtal, oc: OutputCoupler, E_pump, E_cav0, n_air=None,
                     chain_before_oc=None, chain_after_oc=None):
    """Crystal-frame spectral operators of one round trip, FFT order.

    The reference trip (OpticalSim.m:230-288) is
      E <- E + E_pump ; refract(xtal) ; solve ; refract(air) ;
      output copy sqrt(T) ; intracavity sqrt(R) ; applyGD(delay).
    All steps after the solver and before the next solve are spectral
    multiplies or a sum, so with X the crystal output spectrum
      next input  = r_in (sqrt(R) e^{-i phi} e^{-i delay dw} r_out X + P_air)
                  = cav * e^{-i delay dw} * X + pump
      output      = sqrt(T) e^{-i phi} r_out X = ocm * X
    with r_in = sqrt(n_air/n_xtal) = 1/r_out (OpticalPulse.m:190-197).
    Other optics (cavity_response of a chain: power transmission, phase)
    multiply by sqrt(T) e^{-i phi}: those before the coupler both paths,
    those after it the fed-back path only.  The delay factor is applied on
    the device (FP64 phase).  Also returns the crystal input of the first
    trip (fftshifted time)."""
    if n_air is None:
        n_air = np.ones(simwin.n)
    R, T, ph = oc.masks(simwin)
    r_in = np.sqrt(n_air / n_xtal)
    r_out = np.sqrt(n_xtal / n_air)
    cav = np.sqrt(R) * np.exp(-1j * ph) * r_out * r_in
    ocm = np.sqrt(T) * np.exp(-1j * ph) * r_out
    for chain, both in ((chain_before_oc, True), (chain_after_oc, False)):
        if chain is not None:
            t_c, phi_c = chain
            f = np.sqrt(t_c) * np.exp(-1j * np.nan_to_num(phi_c))
            cav = cav * f
            if both:
                ocm = ocm * f
    P = r_in * SimWindow.t2k(E_pump)
    X0k = r_in * (SimWindow.t2k(E_cav0) + SimWindow.t2k(E_pump))
    X0 = SimWindow.k2t(X0k)                          # natural time order
    f = SimWindow.fft_order
    return dict(cav=f(cav), ocm=f(ocm), pump=f(P), field0=np.fft.fftshift(X0))


@dataclass
class OPOConfig:
    """One synchronously pumped ring OPO instance (OpticalSim + Cavity + Laser)."""
    n: int = 2 ** 15
    window: float = 12e-12
    lambda_ref: float = 1040e-9
    period: float = 30.5e-6
    length: float = 5e-3
    chi2: float = 2 * 27e-12
    pump_fwhm: float = 200e-15
    pump_energy: float = 2e-9       # 2 W at 1 GHz
    area: float = math.pi * (50e-6) ** 2
    oc_transmission: float = 0.05
    oc_band: tuple | None = (1.2e-6, 1.8e-6)     # fed-back (signal) band; None = flat
    delay: float = 0.3e-12
    # "absolute": delay is the per-trip cavity delay; "sync": delay is the
    # detuning from the synchronous delay of the QPM signal (sync_delay)
    delay_mode: str = "absolute"
    step: float = 1e-7
    error_bounds: tuple = (2.5e-3, 5e-2)   # OpticalSim.m:69 [min, max] percent
    progress_plots: int = 5                 # OpticalSim.m:23
    noise_seed: int = 20261202
    temperature: float = 20.0

    def build(self):
        sw = SimWindow(self.lambda_ref, self.n, self.window)
        xt = NonlinearCrystal(self.period, self.length, self.chi2, self.temperature)
        si = xt.solver_inputs(sw, self.step)
        pump = sech_pump(sw, self.pump_fwhm, self.pump_energy, self.area)
        noise = quantum_noise(sw, self.area, self.noise_seed)
        oc = OutputCoupler(self.oc_transmission, band=self.oc_band)
        ops = cavity_operators(sw, si["n_xtal"], oc, pump, noise)
        sel = math.ceil(si["nSteps"] / (self.progress_plots - 1))   # OpticalSim.m:214
        si.update(maxErr=self.error_bounds[1], minErr=self.error_bounds[0], sel=float(sel),
                  n_cols=self.progress_plots - 1)
        delay = self.delay
        if self.delay_mode == "sync":
            ls = qpm_signal_wavelength(self.lambda_ref, self.period, self.temperature)
            delay = sync_delay(sw, self.length, self.temperature, ls) + self.delay
        elif self.delay_mode != "absolute":
            raise ValueError("delay_mode must be 'absolute' or 'sync'")
        return dict(simwin=sw, crystal=xt, solver=si, pump=pump, noise=noise, ops=ops,
                    delay=delay, oc=oc)
