"""Host-side input producers (optifax_amd.physics) against the oracle's
independent restatement and against closed-form invariants."""
import math

import numpy as np

import oracle
from optifax_amd import physics as ph


def test_simwindow_matches_reference_definitions():
    sw = ph.SimWindow(1040e-9, 1024, 4e-12)
    assert sw.times[sw.reference_index] == 0.0
    assert abs(sw.dt - 4e-12 / 1024) < 1e-27
    assert abs(sw.omegas[sw.reference_index] - sw.reference_omega) / sw.reference_omega < 1e-15
    assert np.allclose(np.diff(sw.rel_frequencies), 1 / (1024 * sw.dt))


def test_solver_inputs_match_oracle():
    sw = ph.SimWindow(1040e-9, 2048, 6e-12)
    xt = ph.NonlinearCrystal(30e-6, 1e-3, chi2=np.pi * 14e-12)
    si = xt.solver_inputs(sw)
    assert si["nSteps"] == 10000
    pr, pa, gd = oracle.phi_calc(1e-3, 20.0, sw.omegas, sw.rel_omegas, sw.reference_index)
    hb = np.where(np.isfinite(pr), pr / 1e-3 * 1e-7, 0.0)
    band = np.fft.ifftshift(sw.in_band())
    assert np.allclose(si["beta_op"][band], np.fft.ifftshift(hb)[band], rtol=1e-12, atol=1e-15)
    assert np.all(si["beta_op"][~band] == 0) and np.all(si["alpha_w"][~band] == 0)
    assert np.all(si["alpha_w"][band] == 1)
    P = oracle.qpm_poling(10000, 1e-3, 30e-6)
    n0 = ph.gayer_n(1.04, 20.0)
    assert np.allclose(si["G33"], P * np.pi * 14e-12 * sw.reference_omega / n0 / 4 / ph.C0,
                       rtol=1e-14)
    # bdiffw0 = beta0 - beta1*w0 with beta1 from phi2GD's backward difference at w0
    i = sw.reference_index
    beta0 = 2 * np.pi * n0 / 1040e-9
    beta1 = gd[i] / 1e-3
    assert abs(si["bdiffw0"] - (beta0 - beta1 * sw.reference_omega)) < 1e-8 * abs(si["bdiffw0"])


def test_pump_energy_and_noise_photon_number():
    sw = ph.SimWindow(1040e-9, 8192, 10e-12)
    area = 5e-12
    Ep = ph.sech_pump(sw, 200e-15, 1e-9, area)
    assert abs(ph.energy_of(sw, Ep, area) - 1e-9) < 1e-12
    # intensity FWHM of the sech^2 pump
    I = np.abs(Ep) ** 2
    above = sw.times[I >= I.max() / 2]
    assert abs((above[-1] - above[0]) - 200e-15) < 2 * sw.dt
    En = ph.quantum_noise(sw, area, 20261202)
    Ek = ph.SimWindow.t2k(En)
    U = np.abs(Ek) ** 2 * (ph.C0 * ph.EPS0 / 2) * area * sw.dt / sw.n
    band = sw.in_band()
    assert np.allclose(U[band], ph.HBAR * sw.omegas[band], rtol=1e-10)
    assert np.all(U[~band] < 1e-20 * U[band].max())     # FFT round-off only
    # seeded: identical for the same seed, different otherwise
    assert np.array_equal(En, ph.quantum_noise(sw, area, 20261202))
    assert not np.array_equal(En, ph.quantum_noise(sw, area, 1))


def test_cavity_operators_compose_the_reference_trip():
    cfg = ph.OPOConfig(n=512, window=3e-12, length=0.4e-3)
    b = cfg.build()
    ops, sw, si = b["ops"], b["simwin"], b["solver"]
    # with the solver replaced by identity, one trip of the literal sequence
    # (refract out, OC reflect, delay, pump, refract in) equals cav*exp(-i d dw)*X + pump
    X = np.fft.fftshift(ph.SimWindow.k2t(np.random.default_rng(3).standard_normal(512) + 0j))
    Xn = np.fft.ifftshift(X)                        # natural order field in the crystal
    Ek = ph.SimWindow.t2k(Xn) * np.sqrt(si["n_xtal"])            # refract to air (n_air = 1)
    R, _, _ = b["oc"].masks(sw)
    assert R.max() > 0.9 and R.min() < 1e-10        # signal band fed back, the rest dumped
    Ek = Ek * np.sqrt(R)                                         # intracavity OC copy
    Ek = Ek * np.exp(-1j * cfg.delay * sw.rel_omegas)            # applyGD
    Ek = (Ek + ph.SimWindow.t2k(b["pump"])) * np.sqrt(1 / si["n_xtal"])   # pump, refract in
    lit = ph.SimWindow.fft_order(Ek)
    dw = 2 * np.pi * np.fft.fftfreq(512, sw.dt)
    fused = ops["cav"] * np.exp(-1j * cfg.delay * dw) * np.fft.fft(X) + ops["pump"]
    assert np.linalg.norm(fused - lit) / np.linalg.norm(lit) < 1e-12
