"""Shared small workloads for the parity tests (synthetic pulses, MgO:PPLN)."""
import numpy as np

from optifax_amd import physics as ph


def crystal_case(n=256, window=2e-12, L=0.3e-3, energy=2e-10, area=20e-12, period=30e-6,
                 chi2=np.pi * 14e-12, seed=7, plots=5):
    """Single-pass SHG/OPG case: sech^2 pump + one-photon noise, fftshifted."""
    sw = ph.SimWindow(1040e-9, n, window)
    xt = ph.NonlinearCrystal(period, L, chi2=chi2)
    si = xt.solver_inputs(sw)
    E = np.fft.fftshift(ph.sech_pump(sw, 200e-15, energy, area) + ph.quantum_noise(sw, area, seed))
    sel = float(np.ceil(si["nSteps"] / (plots - 1)))
    args = (si["alpha_w"], si["G33"], si["w0"], si["bdiffw0"], si["h"], si["nSteps"], si["dt"],
            si["beta_op"], 5e-2, 2.5e-3, sel)
    return sw, si, E, args


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def opo_case(n=512, window=3e-12, length=0.4e-3, energy=2e-10, area=20e-12, delay=20e-15,
             seed=5, oc_t=0.05):
    cfg = ph.OPOConfig(n=n, window=window, length=length, pump_energy=energy, area=area,
                       delay=delay, noise_seed=seed, oc_transmission=oc_t)
    return cfg, cfg.build()


def oracle_next_input(b, E_cav):
    """Crystal input of the next trip from the oracle's intracavity field."""
    si = b["solver"]
    r_in = np.sqrt(1.0 / si["n_xtal"])
    X = ph.SimWindow.k2t(r_in * (ph.SimWindow.t2k(E_cav) + ph.SimWindow.t2k(b["pump"])))
    return np.fft.fftshift(X)


def oracle_oc_spectra(b, delay, trips):
    """Output-coupled spectrum (FFT order) after each of trips 1..trips, from
    separate oracle runs of the literal OpticalSim.run order."""
    import oracle
    from optifax_amd.physics import SimWindow
    si, sw = b["solver"], b["simwin"]
    R, T, phi = b["oc"].masks(sw)
    out = []
    for t in range(1, trips + 1):
        o = oracle.opo_run(t, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi,
                           delay, sw.rel_omegas, si)
        out.append(SimWindow.fft_order(SimWindow.t2k(o["E_oc"])))
    return out
