"""Warm start: band features, bin-level cavity, band-limited handover (CPU)."""
import math

import numpy as np
import pytest

from optifax_amd import physics as ph
from optifax_amd import warmstart as ws


def _cfg(**kw):
    base = dict(n=2 ** 14, window=16e-12, length=8e-3, period=30e-6, pump_energy=4e-9,
                delay=0.0, delay_mode="sync")
    base.update(kw)
    return ph.OPOConfig(**base)


def _pulse_bins(sb, sw, rng, t0=0.4e-12):
    t = sw.times
    fs = 200e12 - ph.C0 / 1040e-9
    E = np.exp(-((t - t0) / 150e-15) ** 2) * np.exp(2j * np.pi * fs * t + 1j * rng.uniform(0, 6))
    E += 0.3 * np.exp(-((t + 0.5e-12) / 100e-15) ** 2) * np.exp(2j * np.pi * (fs + 8e12) * t)
    Xk = ph.SimWindow.fft_order(ph.SimWindow.t2k(E))
    return Xk, sb.bins(Xk)


def test_band_covers_the_resonated_band():
    sw = ph.SimWindow(1040e-9, 2 ** 14, 16e-12)
    sb = ws.band_bins(sw)
    assert sb.D == 2 and sb.nbins == 1024
    assert sb.freqs[0] <= ph.C0 / 1.8e-6 - 10e12 and sb.freqs[-1] >= ph.C0 / 1.2e-6 + 10e12


def test_features_round_trip_and_phase_equivariance():
    sw = ph.SimWindow(1040e-9, 2 ** 14, 16e-12)
    sb = ws.band_bins(sw)
    rng = np.random.default_rng(0)
    _, Xb = _pulse_bins(sb, sw, rng)
    a_ref = 10 * np.linalg.norm(Xb)
    x, th, s = ws.input_features(Xb, 4e-9, 30.5e-6, a_ref)
    assert x.shape == (1, ws.N_IN)
    assert x[0, -2] == pytest.approx(1.0) and x[0, -1] == pytest.approx(0.5)
    assert x[0, -3] == pytest.approx(np.log10(0.1) / 4, rel=1e-6)
    assert np.linalg.norm(x[0, :-3]) == pytest.approx(1.0, rel=1e-5)   # scale-free field part
    # a global phase rotation changes only the canonical angle
    x2, th2, s2 = ws.input_features(Xb * np.exp(1.234j), 4e-9, 30.5e-6, a_ref)
    assert np.allclose(x2, x, atol=2e-6)
    assert math.remainder(th2[0] - th[0] - 1.234, 2 * math.pi) == pytest.approx(0, abs=1e-9)
    # output features: exact round trip through the linear-propagation residual
    lin = np.exp(1j * rng.uniform(0, 6, sb.nbins)) * rng.uniform(0.5, 1.0, sb.nbins)
    F = lin * Xb + 0.1 * Xb * np.abs(Xb) / np.abs(Xb).max()
    y = ws.output_features(F, Xb, lin, th, s)
    Fb = ws.from_output_features(y, Xb, lin, th, s)[0]
    assert np.linalg.norm(Fb - F) / np.linalg.norm(F) < 1e-6
    # the residual of a purely linear pass is zero
    assert np.abs(ws.output_features(lin * Xb, Xb, lin, th, s)).max() < 1e-6


def test_bin_cavity_matches_full_grid_operators():
    cfg = _cfg()
    inst, b = ws.opo_instance(cfg)
    sw = b["simwin"]
    sb = ws.band_bins(sw)
    cav = ws.BandCavity(sb, [inst], [b])
    rng = np.random.default_rng(1)
    Xout = rng.standard_normal(sw.n) + 1j * rng.standard_normal(sw.n)
    dw = 2 * np.pi * np.fft.fftfreq(sw.n, sw.dt)                  # the kernel's dw_of
    nxt = inst["cav"] * np.exp(-1j * inst["delay"] * dw) * Xout + inst["pump"]
    F = sb.bins(inst["cav"] * Xout)
    assert np.allclose(cav.next_input(F[None])[0], sb.bins(nxt), rtol=1e-9, atol=1e-9 * np.abs(nxt).max())
    assert np.allclose(cav.band_out(sb.bins(nxt)[None])[0], F, rtol=1e-7, atol=1e-9 * np.abs(F).max())
    # output-coupled power from the band output
    S = np.abs(sb.bins(inst["ocm"] * Xout)) ** 2
    assert np.allclose(cav.oc_power(F[None])[0], S, rtol=1e-9, atol=1e-12 * S.max())


def test_handover_is_band_limited_interpolation():
    sw = ph.SimWindow(1040e-9, 2 ** 14, 16e-12)
    sb = ws.band_bins(sw)
    Xk, Xb = _pulse_bins(sb, sw, np.random.default_rng(2))
    Xr = sb.to_grid(Xb)
    kc = np.arange(sb.kc[0], sb.kc[-1] + 1)
    kf = (kc - sw.n // 2) % sw.n
    assert np.linalg.norm(Xr[kf] - Xk[kf]) / np.linalg.norm(Xk[kf]) < 1e-6


def test_sync_delay_and_qpm_wavelength():
    ls = ph.qpm_signal_wavelength(1040e-9, 30e-6, 20.0)
    assert 1.45e-6 < ls < 1.55e-6
    # energy conservation closes the QPM condition
    li = 1 / (1 / 1040e-9 - 1 / ls)
    n = ph.gayer_n(np.array([1.04, ls * 1e6, li * 1e6]), 20.0, 1)
    dk = 2 * np.pi * (n[0] / 1040e-9 - n[1] / ls - n[2] / li - 1 / 30e-6)
    assert abs(dk) < 1e-3 * 2 * np.pi / 30e-6
    sw = ph.SimWindow(1040e-9, 2 ** 14, 16e-12)
    d5 = ph.sync_delay(sw, 5e-3, 20.0, ls)
    d8 = ph.sync_delay(sw, 8e-3, 20.0, ls)
    assert 0.4e-12 < d5 < 0.6e-12 and d8 == pytest.approx(d5 * 8 / 5, rel=0.02)
    b = _cfg(delay=20e-15).build()
    assert b["delay"] == pytest.approx(d8 + 20e-15, rel=1e-6)
