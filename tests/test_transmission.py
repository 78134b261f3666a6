"""Transmission data files (transmission.m, import_interp.m, OpticalSurface.m
coatings, Dielectric.Transmission) on the host: MATLAB's smooth and makima
pinned by hand-worked cases and closed forms, the reference's own H-ZLaF68C
curve (tests/golden/, copied by make_transmission_fixtures.py) through the
whole chain."""
import math
import os

import numpy as np
import pytest

from optifax_amd import physics as ph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CURVE = os.path.join(GOLDEN, "H-ZLaF68C_Transmission_5mm.txt")


def test_smooth_moving_average_by_hand():
    y = np.array([1.0, 2.0, 3.0, 4.0, 10.0, 6.0, 7.0])
    want = [1.0, 2.0, 4.0, 5.0, 6.0, 23.0 / 3.0, 7.0]     # windows 1, 3, 5, 5, 5, 3, 1
    assert np.allclose(ph.matlab_smooth(y, 5), want, rtol=0, atol=1e-14)
    # an even span drops by one; a fraction of the points rounds up (0.3 * 7 -> 3)
    w3 = [1.0, 2.0, 3.0, 17.0 / 3.0, 20.0 / 3.0, 23.0 / 3.0, 7.0]
    assert np.allclose(ph.matlab_smooth(y, 4), w3, rtol=0, atol=1e-14)
    assert np.allclose(ph.matlab_smooth(y, 0.3), w3, rtol=0, atol=1e-14)
    assert np.array_equal(ph.matlab_smooth(y, 0.1), y)   # one point: unchanged


def test_makima_closed_forms():
    x = np.array([0.0, 1.0, 2.5, 3.0, 4.5, 6.0])
    # a straight line is reproduced exactly; nodes are hit; zero outside
    lin = 2.0 * x - 1.0
    xq = np.linspace(-1.0, 7.0, 41)
    got = ph.makima(x, lin, xq)
    inside = (xq >= 0) & (xq <= 6)
    assert np.allclose(got[inside], 2.0 * xq[inside] - 1.0, atol=1e-13)
    assert np.all(got[~inside] == 0.0)
    y = np.array([0.0, 1.0, 0.5, 3.0, 2.0, 2.5])
    assert np.allclose(ph.makima(x, y, x), y, atol=1e-14)
    # the derivative at an interior node is the modified-Akima weighted mean of
    # the neighbouring secant slopes (weights |d_{i+1} - d_i| + |d_{i+1} + d_i| / 2)
    d = np.diff(y) / np.diff(x)
    i = 2
    w1 = abs(d[i + 1] - d[i]) + abs(d[i + 1] + d[i]) / 2
    w2 = abs(d[i - 1] - d[i - 2]) + abs(d[i - 1] + d[i - 2]) / 2
    slope = (w1 * d[i - 1] + w2 * d[i]) / (w1 + w2)
    # (one-sided second-order difference inside the piece to the right of the node)
    h = 1e-5
    f = ph.makima(x, y, [x[i], x[i] + h, x[i] + 2 * h])
    num = (-3 * f[0] + 4 * f[1] - f[2]) / (2 * h)
    assert num == pytest.approx(slope, rel=1e-6)
    # flat data stays flat (no overshoot)
    assert np.allclose(ph.makima(x, np.ones(6), xq[inside]), 1.0, atol=1e-14)


def test_import_interp_pads_and_hits_the_smoothed_nodes():
    x, y = ph.read_data_table(CURVE)
    assert x.size == 30 and x[0] == 330.0
    ys = ph.matlab_smooth(y, 0.01)                        # 0.01 * 30 -> 1 point: unchanged
    assert np.array_equal(ys, y)
    xs = np.concatenate([[100.0], x, [5000.0]])
    got = ph.import_interp(CURVE, xs, yfun=lambda v: v / 100.0)
    assert abs(got[0]) < 1e-15 and abs(got[-1]) < 1e-15   # padded zeros at the far ends
    assert np.allclose(got[1:-1], y / 100.0, atol=1e-14)


def test_transmission_fresnel_limits_and_steps():
    lam = np.linspace(350e-9, 2000e-9, 400)
    T, Ts, a = ph.transmission(CURVE, lam)
    assert np.all((T >= 0) & (T <= 1)) and T.max() > 0.9
    # per-step loss reproduces T over nsteps steps
    T5, Ts5, a5 = ph.transmission(CURVE, lam, nsteps=5)
    ok = T5 > 0
    assert np.allclose(Ts5[ok] ** 5, T5[ok], rtol=1e-12)
    # Fresnel correction of two surfaces of n = 1.5: divide by 0.96^2, renormalise if > 1
    Tf, _, _ = ph.transmission(CURVE, lam, nmat=1.5 * np.ones(lam.size), frnum=2)
    raw = T / 0.96 ** 2
    want = raw / raw.max() if raw.max() > 1 else raw
    assert np.allclose(Tf, want, atol=1e-14)
    # wavelength limits
    Tl, _, _ = ph.transmission(CURVE, lam, lim1=500e-9, lim2=1500e-9)
    assert np.all(Tl[(lam < 500e-9) | (lam > 1500e-9)] == 0)
    assert np.array_equal(Tl[(lam >= 500e-9) & (lam <= 1500e-9)],
                          T[(lam >= 500e-9) & (lam <= 1500e-9)])


def test_coating_file_on_a_surface_and_its_oblique_shift():
    sw = ph.SimWindow(1040e-9, 2 ** 12, 12e-12)
    t0 = ph.surface_transmission("H-ZLaF68C_Transmission_5mm.txt", sw, data_dir=GOLDEN)
    assert np.array_equal(t0, ph.transmission(CURVE, sw.wavelengths)[0])
    t30 = ph.surface_transmission("H-ZLaF68C_Transmission_5mm.txt", sw, theta_deg=30.0,
                                  data_dir=GOLDEN)
    lam = sw.wavelengths
    i = int(np.argmax(t0))
    k = math.ceil(lam[i] * 0.2 * 0.25 / ((lam[i + 1] - lam[i - 1]) / 2))
    assert np.array_equal(t30, np.roll(t0, -k))
    with pytest.raises(ValueError):
        ph.surface_transmission("H-ZLaF68C_Transmission_5mm.txt", sw)


def test_bulk_transmission_scales_with_path_length():
    sw = ph.SimWindow(1040e-9, 2 ** 12, 12e-12)
    ttab = ph.load_transmission_table(os.path.join(GOLDEN, "materials_transmission.csv"))
    gtab = ph.load_material_table(os.path.join(GOLDEN, "materials_transmission.csv"))
    assert ttab["H-ZLaF68C"] == ("H-ZLaF68C_Transmission_5mm.txt", 0.005, 0.0)
    t5 = ph.bulk_transmission(sw, "H-ZLaF68C", 0.005, ttab, GOLDEN, table=gtab)
    lo, hi = sw.spectral_limits
    want = ph.transmission(CURVE, sw.wavelengths, 1, 1.0, 0, lo * 1e-9, hi * 1e-9)[0]
    assert np.allclose(t5, want, atol=1e-15)           # frnum 0: the index drops out
    t10 = ph.bulk_transmission(sw, "H-ZLaF68C", 0.010, ttab, GOLDEN, table=gtab)
    assert np.allclose(t10, t5 ** 2, rtol=1e-12)


def test_optic_with_a_coating_file_and_bulk_absorption():
    """Optic.simulate: T = S1 * bulk * S2 with S2 defaulting to S1."""
    sw = ph.SimWindow(1040e-9, 2 ** 12, 12e-12)
    ttab = ph.load_transmission_table(os.path.join(GOLDEN, "materials_transmission.csv"))
    gtab = ph.load_material_table(os.path.join(GOLDEN, "materials_transmission.csv"))
    f = "H-ZLaF68C_Transmission_5mm.txt"
    t, r, phi = ph.optic_response(sw, "T", f, "H-ZLaF68C", 0.002, table=gtab,
                                  transmission_table=ttab, data_dir=GOLDEN)
    ts = ph.surface_transmission(f, sw, data_dir=GOLDEN)
    tb = ph.bulk_transmission(sw, "H-ZLaF68C", 0.002, ttab, GOLDEN, table=gtab)
    assert np.allclose(t, ts * tb * ts, rtol=1e-14) and np.allclose(r, 1 - ts, rtol=1e-14)


def test_prism_path_follows_the_index():
    """A prism is an optic whose bulk length is a function of its index
    (prismfn_PB.m, Dielectric.PathLength): the path and the bulk phase follow
    n(lambda); a constant function reproduces the plain slab."""
    sw = ph.SimWindow(1040e-9, 2 ** 12, 12e-12)
    # the Pellin-Broca path at the Brewster index (n = tan 54 deg): sin(theta_i) = cos(54 deg)
    nb = math.tan(math.radians(54.0))
    th_i = math.asin(math.sin(math.radians(54.0)) / nb)
    assert th_i == pytest.approx(math.radians(36.0), abs=1e-12)
    a = math.radians(78 + 26 / 60)
    assert ph.prism_pb_path_length(nb) == pytest.approx(2 * 19 * math.sin(a) / math.cos(th_i) / 1e3)
    gtab = ph.load_material_table(os.path.join(GOLDEN, "materials_transmission.csv"))
    mat = "H-ZLaF68C"
    L = ph.path_length(sw, mat, ph.prism_pb_path_length, gtab)
    lam = sw.wavelengths
    band = (lam > 0.7e-6) & (lam < 2.4e-6)
    assert np.all(np.diff(L[band]) != 0) and 0.03 < L[band].min() < L[band].max() < 0.05
    slab = ph.bulk_dispersion(sw, mat, 0.04, gtab)
    same = ph.bulk_dispersion(sw, mat, lambda n: np.full(n.shape, 0.04), gtab)
    assert np.allclose(same[np.isfinite(slab)], slab[np.isfinite(slab)], rtol=0, atol=1e-12)
    prism = ph.bulk_dispersion(sw, mat, ph.prism_pb_path_length, gtab)
    want = ph.phi_calc(L, 20.0, sw, mat, gtab)[0]
    ok = np.isfinite(want)
    assert np.array_equal(prism[ok], want[ok])
