"""QPM gain maps (qpmgain.m, sincgain.m): the oracle pinned by closed forms
on the CPU, the device sums against the oracle on the GPU."""
import numpy as np
import pytest

import oracle
from oracle import qpm as oq
from optifax_amd import physics as ph

from cases import rel_l2


# ---------------------------------------------------------------- oracle pins (CPU)
def test_single_domain_crystal_sums_to_the_geometric_series():
    """P = 1 everywhere: qpmgain's sum over z_k = (k-1) dz (the first sample
    with dz_1 = 0) is dz e^{i dk dz} (1 - e^{i dk (n-1) dz}) / (1 - e^{i dk dz})."""
    L = 3e-3
    n, z, dz = oq.z_samples(L, 4096)
    assert n == 3000 and dz[0] == 0.0 and abs(z[-1] - L) < 1e-15
    d = L / (n - 1)
    dk = np.array([0.0, 1.3e3, 2 * np.pi / d, 7.7e4, -2.2e5])
    got = oq.literal_sum(dk, z, dz)
    r = np.exp(1j * dk * d)
    want = np.where(np.abs(1 - r) > 1e-12, d * r * (1 - r ** (n - 1)) / np.where(
        np.abs(1 - r) > 1e-12, 1 - r, 1), (n - 1) * d)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-16)


def test_square_wave_grating_sums_domain_by_domain():
    """A +-1 grating of period Lambda (15-sample domains on the 1 um grid):
    the literal sum equals the sum over domains of each domain's geometric
    series, s_d dz e^{i dk a_d dz} (1 - r^m_d) / (1 - r), r = e^{i dk dz}, at
    any dk; at dk = 2 pi / Lambda its first Fourier component gives
    |sum| = (2/pi) L (100 whole periods)."""
    L, period = 3e-3, 30e-6
    n, z, dz = oq.z_samples(L, 4096)
    NS = int(np.floor(L / 1e-7))
    P = oracle.qpm_poling(NS, L, period)[:: NS // n]
    d = L / (n - 1)
    walls = np.nonzero(np.diff(P))[0] + 1
    starts = np.concatenate([[1], walls])                 # sample 0 has dz = 0
    ends = np.concatenate([walls, [n]])
    for dk in (2 * np.pi / period, 2 * np.pi / period + np.pi / L, 1.234e5, 3.3e5):
        r = np.exp(1j * dk * d)
        want = sum(P[a] * d * np.exp(1j * dk * a * d) * (1 - r ** (b - a)) / (1 - r)
                   for a, b in zip(starts, ends))
        got = oq.literal_sum([dk], z, P * dz)[0]
        assert abs(got - want) < 1e-10 * abs(want), dk
    s = oq.literal_sum([2 * np.pi / period], z, P * dz)[0]
    assert abs(abs(s) / (2 / np.pi * L) - 1) < 1e-2


def test_sinc_gain_law_by_hand():
    """One pump/signal/idler triple and a grating set: at the grating that
    cancels the mismatch the factor is 1 (MATLAB's sinc(0)); a grating that
    leaves dk L / 2 = 1/2 gives sinc(1/2)^2 = (2/pi)^2; the order sum adds
    (sinc(x_3)/3)^2 for m = 3."""
    lam_p, lam_s = 1.04, 1.55
    lam_i = 1.0 / (1.0 / lam_p - 1.0 / lam_s)

    def k(lam):
        return 2 * np.pi * abs(oracle.sellmeier_n(lam, "PPLN", 20.0)[0]) / (lam * 1e-6)

    dk0 = abs(abs(k(lam_p) - k(lam_i)) - k(lam_s))
    L = 1e-3
    P0 = 2 * np.pi / dk0                         # the period that matches exactly
    # a period leaving x = dk L / 2 = 1/2: dk = 1/L
    P1 = 2 * np.pi / (dk0 - 1.0 / L)
    g = oq.sinc_gain(np.array([[lam_p]]), np.array([[lam_s]]), np.array([[lam_i]]), P0, P1, 2, L,
                     "PPLN", 20.0)
    assert g.shape == (1, 1, 2)
    assert abs(g[0, 0, 0] - 1.0) < 1e-9
    assert abs(g[0, 0, 1] - (2 / np.pi) ** 2) < 1e-9
    g3 = oq.sinc_gain(np.array([[lam_p]]), np.array([[lam_s]]), np.array([[lam_i]]), P0, P1, 2, L,
                      "PPLN", 20.0, m=(1.0, 3.0))
    x3 = (dk0 - 3 * dk0) * L / 2
    assert abs(g3[0, 0, 0] - (1.0 + (np.sinc(x3) / 3) ** 2)) < 1e-12


def test_runs_reproduce_the_literal_sum():
    """The product's compression of the sampled grating into runs of equal
    poling (each summed as a geometric series on the device) is exact
    algebra; the host side is checked here: runs cover every nonzero sample
    once, with the sample spacing."""
    from optifax_amd.qpm import runs_of
    L, period = 1e-3, 30e-6
    n, z, dz = oq.z_samples(L, 4096)
    NS = int(np.floor(L / 1e-7))
    P = oracle.qpm_poling(NS, L, period)[:: NS // n]
    z0, ln, wg, d = runs_of(z, P * dz)
    assert abs(d - L / (n - 1)) < 1e-18
    assert ln.sum() == n - 1                        # all but the zero-weight first sample
    assert 60 <= z0.size <= 70                      # ~ L / (period / 2) domains
    # expand the runs back into samples
    zz = np.concatenate([a + d * np.arange(m) for a, m in zip(z0, ln)])
    ww = np.repeat(wg, ln)
    assert np.allclose(zz, z[1:], rtol=0, atol=1e-15)
    assert np.array_equal(ww, (P * dz)[1:])


# ---------------------------------------------------------------- device (GPU)
def _pump_case(n=2 ** 13, window=10e-12, fwhm=300e-15, energy=1e-9, area=np.pi * (30e-6) ** 2):
    sw = ph.SimWindow(1040e-9, n, window)
    E = ph.sech_pump(sw, fwhm, energy, area)
    return sw, E, area


@pytest.mark.gpu
@pytest.mark.parametrize("L,period", [(1e-3, 30e-6), (2e-3, 29.5e-6)])
def test_qpm_gain_map_matches_the_oracle(L, period):
    """qpmgain's map on the device (runs of equal poling summed as geometric
    series) against the oracle's literal sample-by-sample sum, FP64."""
    from optifax_amd import qpm
    sw, E, area = _pump_case()
    pulse = qpm.pulse_descriptors(sw, E, area)
    xt = ph.NonlinearCrystal(period, L)
    got = qpm.qpm_gain(xt, sw, pulse, sigrange_nm=(1400, 1700))
    nr = xt.refractive_index(sw)
    want, pump, signal, idler, dk = oq.qpmgain(L, period, xt.chi2, "PPLN", 20.0, 1e-7,
                                               sw.wavelengths, nr, pulse["abs_spectrum"],
                                               pulse["peak_wavelength"], pulse["wavelength_fwhm"],
                                               pulse["average_intensity"], (1400, 1700))
    assert got["gain"].shape == want.shape and want.size > 100
    assert np.array_equal(got["pump"][0], pump) and np.array_equal(got["signal"][:, 0], signal)
    print(f"L={L}: map {want.shape}, {qpm.zsum.last_runs} runs, {qpm.zsum.last_ms:.3f} ms, "
          f"rel L2 {rel_l2(got['gain'], want):.2e}")
    assert rel_l2(got["gain"], want) < 1e-9
    # the map peaks at the phase-matched signal of the grating
    i = np.unravel_index(np.argmax(want), want.shape)
    ls = ph.qpm_signal_wavelength(1040e-9, period)
    assert abs(signal[i[0]] - ls) < 20e-9


@pytest.mark.gpu
def test_qpm_zsum_literal_samples_and_large_map():
    """Runs of one sample are the literal sum over arbitrary z (here a
    jittered grid); and a full-size map (2000 x 100 outputs over an 8 mm
    crystal) keeps its closed-form value at the phase-matched output."""
    from optifax_amd import qpm
    rng = np.random.default_rng(1)
    z = np.sort(rng.uniform(0, 2e-3, 1500))
    w = rng.standard_normal(1500) * 1e-6
    dk = rng.uniform(-3e5, 3e5, 300)
    got = qpm.zsum(dk, z, w)
    assert qpm.zsum.last_runs == 1500
    assert rel_l2(got, oq.literal_sum(dk, z, w)) < 1e-11
    L, period = 8e-3, 30e-6
    n, z, dz = oq.z_samples(L, 2 ** 15)
    NS = int(np.floor(L / 1e-7))
    P = oracle.qpm_poling(NS, L, period)[:: NS // n]
    dk = np.linspace(1.5e5, 2.7e5, 2000 * 100)
    dk[12345] = 2 * np.pi / period
    got = qpm.zsum(dk, z, P * dz)
    print(f"8 mm, 200000 outputs: {qpm.zsum.last_runs} runs, {qpm.zsum.last_ms:.3f} ms")
    assert abs(abs(got[12345]) / (2 / np.pi * L) - 1) < 1e-2
    pick = rng.choice(dk.size, 200, replace=False)
    assert rel_l2(got[pick], oq.literal_sum(dk[pick], z, P * dz)) < 1e-9


@pytest.mark.gpu
def test_sinc_gain_map_matches_the_oracle():
    from optifax_amd import qpm
    xt = ph.NonlinearCrystal([29e-6, 31e-6], 1e-3)
    lam_p = np.linspace(1.02, 1.06, 20)[:, None]
    lam_s = np.linspace(1.4, 1.7, 300)[None, :]
    lam_i = 1.0 / (1.0 / lam_p - 1.0 / lam_s)
    got, grating_um, conv = qpm.sinc_gain(xt, lam_p, lam_s, lam_i, 51, m=[1, 3])
    want = oq.sinc_gain(lam_p, lam_s, lam_i, 29e-6, 31e-6, 51, 1e-3, "PPLN", 20.0, m=(1.0, 3.0))
    assert got.shape == want.shape == (20, 300, 51)
    assert np.allclose(grating_um, np.linspace(29, 31, 51))
    assert rel_l2(got, want) < 1e-9 and np.all(conv > 0)


def test_host_arrays_match_the_oracle():
    """qpm_gain's host side (wavelength grids, subsampling, mismatches) against
    the oracle's restatement, and the whole map with the oracle's literal
    sum in place of the device's: the same map."""
    from optifax_amd import qpm
    sw, E, area = _pump_case(n=2 ** 12, window=6e-12)
    pulse = qpm.pulse_descriptors(sw, E, area)
    xt = ph.NonlinearCrystal(30e-6, 0.5e-3)
    a = qpm.qpm_arrays(xt, sw, pulse, (1400, 1700))
    nr = xt.refractive_index(sw)
    want, pump, signal, idler, dk = oq.qpmgain(0.5e-3, 30e-6, xt.chi2, "PPLN", 20.0, 1e-7,
                                               sw.wavelengths, nr, pulse["abs_spectrum"],
                                               pulse["peak_wavelength"], pulse["wavelength_fwhm"],
                                               pulse["average_intensity"], (1400, 1700))
    assert np.array_equal(a["pump"][0], pump) and np.array_equal(a["signal"][:, 0], signal)
    assert np.allclose(a["idler"], idler, rtol=1e-14) and np.allclose(a["dk"], dk, rtol=1e-12)
    As = a["g"] * oq.literal_sum(a["dk"], a["z"], a["w"]).reshape(a["dk"].shape)
    gain = a["nS"][:, None] * (np.abs(As) / a["I2A"]) ** 2 / a["Ip"]
    assert rel_l2(gain, want) < 1e-12


def test_pulse_descriptors_of_a_sech_pulse():
    """A sech^2 pulse: DurationCheck is its FWHM and the spectral FWHM its
    transform limit (time-bandwidth product 0.315); the average intensity is
    energy / area / FWHM / 2 (OpticalPulse.m:275-277)."""
    from optifax_amd import qpm
    sw, E, area = _pump_case(fwhm=300e-15, energy=1e-9)
    d = qpm.pulse_descriptors(sw, E, area)
    assert abs(d["average_intensity"] / (1e-9 / area / 300e-15 / 2) - 1) < 2e-3
    lam0 = d["peak_wavelength"]
    dnu = d["wavelength_fwhm"] * ph.C0 / lam0 ** 2
    assert abs(dnu * 300e-15 / 0.3148 - 1) < 2e-2
    assert abs(lam0 - 1040e-9) < 1e-9
