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


def test_crystal_materials_match_oracle():
    lam = np.linspace(0.4, 5.0, 97)
    for mat in ("PPLN", "LN_e", "LN_o", "PPLN_undoped", "OP-GaP", "KTP", "PPKTP"):
        for T in (20.0, 65.0):
            assert np.allclose(ph.sellmeier_n(lam, mat, T), oracle.sellmeier_n(lam, mat, T),
                               rtol=1e-14, atol=0), (mat, T)
    with np.testing.assert_raises(ValueError):
        ph.sellmeier_n(lam, "N-BK7")


def test_solver_inputs_other_material_match_oracle():
    """A KTP crystal: beta from phi_calc's sellmeier branch (20 C whatever the
    crystal temperature), the index at the crystal temperature."""
    sw = ph.SimWindow(1040e-9, 2048, 6e-12)
    xt = ph.NonlinearCrystal(30e-6, 1e-3, material="KTP", temperature=50.0)
    si = xt.solver_inputs(sw)
    pr, pa, gd = oracle.phi_calc_mat(1e-3, 50.0, "KTP", sw.omegas, sw.rel_omegas,
                                     sw.reference_index)
    hb = np.where(np.isfinite(pr), pr / 1e-3 * 1e-7, 0.0)
    band = np.fft.ifftshift(sw.in_band())
    assert np.allclose(si["beta_op"][band], np.fft.ifftshift(hb)[band], rtol=1e-12, atol=1e-15)
    n0 = oracle.sellmeier_n(1.04, "KTP", 50.0)[0]
    assert abs(si["n0"] - n0) < 1e-14
    P = oracle.qpm_poling(10000, 1e-3, 30e-6)
    assert np.allclose(si["G33"], P * xt.chi2 * sw.reference_omega / n0 / 4 / ph.C0, rtol=1e-14)


def test_gratings_match_oracle():
    L, ns = 1e-3, 10000
    z = np.linspace(0.0, L, ns)
    for P1, P2, a in ((6.3e-6, 2.2e-6, 0.48), (28e-6, 32e-6, 1.0), (30e-6, 30e-6, 2.0)):
        w = ph.chirped_grating(z, P1, P2, a, 0.1e-6)
        assert np.array_equal(w, oracle.chirped_grating(z, P1, P2, a, 0.1e-6))
        assert np.array_equal(ph.poling_from_widths(ns, L, w),
                              oracle.qpm_poling_widths(ns, L, w))
    # a crystal whose grating is a callable of the step positions (QPMcrystal.m:14-15)
    xt = ph.NonlinearCrystal(lambda zz: ph.chirped_grating(zz, 31e-6, 29e-6, 1.0, 0.1e-6), L)
    assert np.array_equal(xt.poling(ns), oracle.qpm_poling_widths(
        ns, L, oracle.chirped_grating(z, 31e-6, 29e-6, 1.0, 0.1e-6)))


def test_fanout_and_stepped_periods():
    """NonlinearCrystal.m:88-99,131-139: two periods fan out over the height,
    more periods are stepped gratings picked by ceil(y)."""
    xt = ph.NonlinearCrystal((29e-6, 31e-6), 1e-3, height=10.0, vertical_position=5.0)
    assert abs(xt.local_period() - 30e-6) < 1e-18
    assert np.array_equal(xt.poling(5000), oracle.qpm_poling(5000, 1e-3, 30e-6))
    xs = ph.NonlinearCrystal((28e-6, 29e-6, 30e-6, 31e-6), 1e-3, vertical_position=2.3)
    assert xs.local_period() == 30e-6


def test_domain_width_uncertainty():
    """QPMcrystal.m:27-30 with uncertainty: random odd domains, even domains
    completing each period, so every wall but the random ones stays on the
    period grid; the draw is seeded (MATLAB's stream is not reproduced)."""
    L, period, ns = 1e-3, 30e-6, 20000
    a = ph.qpm_poling(ns, L, period, uncertainty=1e-6, seed=3)
    b = ph.qpm_poling(ns, L, period, uncertainty=1e-6, seed=3)
    c = ph.qpm_poling(ns, L, period, uncertainty=1e-6, seed=4)
    u = ph.qpm_poling(ns, L, period)
    assert np.array_equal(a, b) and not np.array_equal(a, c) and not np.array_equal(a, u)
    assert set(np.unique(a)) == {-1.0, 1.0}
    assert abs(np.count_nonzero(np.diff(a)) - np.count_nonzero(np.diff(u))) <= 1
    assert np.count_nonzero(a != u) < 0.1 * ns      # walls move by ~0.5 um of 15 um domains


def test_glass_table_forms_against_published_indices(tmp_path):
    """Both equation forms of a Materials.csv-layout table (sellmeier.m:22-51)
    against published catalogue indices: N-BK7 (Schott form, Schott
    catalogue coefficients) n_d = 1.5168 at 587.56 nm and 1.5066 at 1064 nm;
    fused silica (Malitson 1965, Schott form) 1.4585 and 1.4496; BK7 in the
    Laurent form of the older Schott catalogue (the CDGM layout) n_d = 1.5168."""
    p = tmp_path / "Materials.csv"
    p.write_text(
        "Material,B1,B2,B3,C1,C2,C3,Equivalent,Equation,T_file\n"
        "N-BK7,1.03961212,0.231792344,1.01046945,0.00600069867,0.0200179144,103.560653,,Schott,x\n"
        "SiO2,0.6961663,0.4079426,0.8974794,0.004679148,0.01351206,97.93400,,Schott,x\n"
        "BK7L,2.2718929,-0.010108077,0.010592509,0.00020816965,-7.6472538e-06,4.9240991e-07,,CDGM,x\n"
        "skip,1,1,1,1,1,1,,Other,x\n")
    t = ph.load_material_table(str(p))
    assert set(t) == {"N-BK7", "SiO2", "BK7L"}
    n = ph.sellmeier_n([0.5875618, 1.064], "N-BK7", table=t)
    assert abs(n[0] - 1.5168) < 1e-4 and abs(n[1] - 1.5066) < 2e-4
    n = ph.sellmeier_n([0.5875618, 1.064], "SiO2", table=t)
    assert abs(n[0] - 1.4585) < 1e-4 and abs(n[1] - 1.4496) < 2e-4
    assert abs(ph.sellmeier_n(0.5875618, "BK7L", table=t) - 1.5168) < 2e-4
    assert abs(ph.sellmeier_n(1.0, "BK7L", table=t) - ph.sellmeier_n(1.0, "N-BK7", table=t)) < 5e-4
    with np.testing.assert_raises(ValueError):
        ph.sellmeier_n(1.0, "SF10", table=t)


def test_glass_bulk_dispersion_gdd(tmp_path):
    """Fused silica's published GVD near 1040 nm is ~18.9 fs^2/mm (Malitson
    index): 10 mm of it gives ~189 fs^2 as the curvature of phi_rel at w0,
    with no group delay left there."""
    p = tmp_path / "Materials.csv"
    p.write_text("Material,B1,B2,B3,C1,C2,C3,Equivalent,Equation\n"
                 "SiO2,0.6961663,0.4079426,0.8974794,0.004679148,0.01351206,97.93400,,Schott\n")
    t = ph.load_material_table(str(p))
    sw = ph.SimWindow(1040e-9, 4096, 10e-12)
    phi = ph.bulk_dispersion(sw, "SiO2", 10e-3, table=t)
    i, dw = sw.reference_index, sw.d_omega
    gdd = (phi[i + 1] - 2 * phi[i] + phi[i - 1]) / dw ** 2
    assert abs(gdd / 189e-30 - 1) < 0.05
    assert abs((phi[i + 1] - phi[i - 1]) / (2 * dw)) < 1e-3 * 10e-3 / ph.C0


def test_fresnel_textbook_cases():
    """Normal incidence into n = 1.5: 1 - ((n-1)/(n+1))^2 = 0.96 for both
    polarisations; p at Brewster's angle: 1; s against 1 - r_s^2 of the
    textbook reflection coefficient; the exit surface at the refracted angle
    transmits what the entry surface does."""
    tp, ts = ph.fresnel_t(1.0, 1.5, 0.0)
    assert abs(tp - 0.96) < 1e-12 and abs(ts - 0.96) < 1e-12
    tb = np.rad2deg(np.arctan(1.5))
    assert abs(ph.fresnel_t(1.0, 1.5, tb)[0] - 1.0) < 1e-12
    th = 40.0
    ti, tt = np.deg2rad(th), np.arcsin(np.sin(np.deg2rad(th)) / 1.5)
    rs = (np.cos(ti) - 1.5 * np.cos(tt)) / (np.cos(ti) + 1.5 * np.cos(tt))
    assert abs(ph.fresnel_t(1.0, 1.5, th)[1] - (1 - rs ** 2)) < 1e-12
    a = ph.fresnel_t(1.0, 1.5, th)
    b = ph.fresnel_t(1.0, 1.5, th, exit=True)
    assert abs(a[0] - b[0]) < 1e-12 and abs(a[1] - b[1]) < 1e-12


def test_surface_transmission_forms():
    sw = ph.SimWindow(1040e-9, 1024, 4e-12)
    assert np.all(ph.surface_transmission("AR", sw) == 1.0)
    assert np.allclose(ph.surface_transmission(95, sw), 0.95)          # percent syntax
    lam = sw.wavelengths
    t = ph.surface_transmission(np.array([[0.2, 1.0e-6], [0.9, 1.2e-6], [0.5, 9e-6]]), sw)
    assert np.all(t[lam < 1.0e-6] == 0.2) and np.all(t[(lam > 1.0e-6) & (lam < 1.2e-6)] == 0.9)
    bare = ph.surface_transmission("None", sw, material="PPLN")
    n = ph.sellmeier_n(lam * 1e6, "PPLN")
    band = (lam > 0.4e-6) & (lam < 4e-6)
    assert np.allclose(bare[band], (1 - ((n - 1) / (n + 1)) ** 2)[band], rtol=1e-12)


def test_gdd_import_constant_gdd_is_a_parabola(tmp_path):
    """A flat GDD curve D gives phi_rel = D (w - w0)^2 / 2 on the grid (the
    two trapezoid integrations are exact for a constant and a linear
    integrand) except at the first bin, where the [0 diff(w_rel)] padding of
    GDDimport2phi.m:26 drops a half trapezoid; movmean matches MATLAB's
    shrinking endpoints."""
    p = tmp_path / "mirror.csv"
    lam = np.linspace(500.0, 2500.0, 81)
    p.write_text("Wavelength,GDD\n" + "".join(f"{l},{-150.0}\n" for l in lam))
    wl, gdd = ph.load_gdd(str(p))
    sw = ph.SimWindow(1040e-9, 2048, 8e-12)
    phi = ph.gdd_to_phi(sw, wl, gdd)
    x = sw.omegas - sw.reference_omega
    ok = np.isfinite(phi)
    ok[0] = False
    assert ok[sw.reference_index] and ok.sum() > 1000
    assert np.allclose(phi[ok], 0.5 * (-150e-30) * x[ok] ** 2, rtol=1e-9, atol=1e-12)
    assert np.allclose(ph._movmean(np.array([1.0, 2, 3, 4, 5, 6]), 5),
                       [2.0, 2.5, 3.0, 4.0, 4.5, 5.0])


def test_optic_chain_into_cavity_operators(tmp_path):
    """Cavity.simulate's products and sums (Cavity.m:50-87): a 50 % mirror
    used in reflection and an AR-coated 1 mm fused-silica window give a
    chain transmission 0.5 x 1 and the window's bulk phase; placed before
    the coupler it scales both spectral operators by sqrt(T) e^{-i phi},
    after it only the fed-back one."""
    p = tmp_path / "Materials.csv"
    p.write_text("Material,B1,B2,B3,C1,C2,C3,Equivalent,Equation\n"
                 "FS,0.6961663,0.4079426,0.8974794,0.004679148,0.01351206,97.93400,,Schott\n")
    t = ph.load_material_table(str(p))
    sw = ph.SimWindow(1040e-9, 1024, 4e-12)
    mirror = ("R", ph.optic_response(sw, "R", 0.5, "FS", 6.35e-3, table=t))
    window = ("T", ph.optic_response(sw, "T", "AR", "FS", 1e-3, table=t))
    assert np.allclose(mirror[1][1], 0.5) and not mirror[1][2].any()
    T_c, phi_c = ph.cavity_response([mirror, window])
    assert np.allclose(T_c, 0.5)
    ok = np.isfinite(phi_c)
    assert np.allclose(phi_c[ok], ph.bulk_dispersion(sw, "FS", 1e-3, table=t)[ok])
    n_x = np.full(sw.n, 2.1)
    oc = ph.OutputCoupler(0.05, band=None)
    E0 = np.zeros(sw.n, complex)
    base = ph.cavity_operators(sw, n_x, oc, E0, E0)
    pre = ph.cavity_operators(sw, n_x, oc, E0, E0, chain_before_oc=(T_c, phi_c))
    post = ph.cavity_operators(sw, n_x, oc, E0, E0, chain_after_oc=(T_c, phi_c))
    f = sw.fft_order(np.sqrt(T_c) * np.exp(-1j * np.nan_to_num(phi_c)))
    assert np.allclose(pre["cav"], base["cav"] * f) and np.allclose(pre["ocm"], base["ocm"] * f)
    assert np.allclose(post["cav"], base["cav"] * f) and np.allclose(post["ocm"], base["ocm"])


def test_material_table_skips_rows_without_coefficients(tmp_path):
    """The reference table's layout (extra trailing columns, an unnamed one)
    with a PPKTP row tagged Schott whose coefficient cells are blank: that
    row is skipped, the Schott and CDGM glasses around it load."""
    p = tmp_path / "Materials.csv"
    p.write_text(
        "Material,B1,B2,B3,C1,C2,C3,Equivalent,Equation,T_file,T_length_m,Fresnel_Reflections,,D1,D2\n"
        "SiO2,0.6961663,0.4079426,0.8974794,0.004679148,0.01351206,97.93400,,Schott,a.txt,0.01,2,,,\n"
        "BK7L,2.2718929,-0.010108077,0.010592509,0.00020816965,-7.6472538e-06,4.9240991e-07,N-BK7,CDGM,a.txt,0.01,2,,,\n"
        "PPKTP,,,,,,,,Schott,b.txt,0.001,2,,,\n")
    t = ph.load_material_table(str(p))
    assert set(t) == {"SiO2", "BK7L"}
    assert abs(ph.sellmeier_n(0.5875618, "SiO2", table=t) - 1.4585) < 1e-4
    assert abs(ph.sellmeier_n(0.5875618, "BK7L", table=t) - 1.5168) < 2e-4


def _flat_gdd_optic(tmp_path, theta=0.0, **kw):
    p = tmp_path / "Materials.csv"
    p.write_text("Material,B1,B2,B3,C1,C2,C3,Equivalent,Equation\n"
                 "FS,0.6961663,0.4079426,0.8974794,0.004679148,0.01351206,97.93400,,Schott\n")
    t = ph.load_material_table(str(p))
    sw = ph.SimWindow(1040e-9, 2048, 8e-12)
    lam = np.linspace(500.0, 2500.0, 81)
    gdd = (lam, np.full(lam.size, -150.0))
    return sw, t, gdd


def test_transmitted_optic_counts_both_coated_surfaces(tmp_path):
    """Optic.simulate in the 'T' regime adds S1's, the bulk's and S2's
    phases (Optic.m:93-97); S2 defaults to a copy of S1, measured GDD
    included (Optic.m:57,72-75).  A flat -150 fs^2 coating on a 1 mm FS
    window therefore gives 2 D (w - w0)^2 / 2 plus the bulk phase; naming
    no curve for S2 leaves one D; in reflection only S1's counts."""
    sw, t, gdd = _flat_gdd_optic(tmp_path)
    x = sw.omegas - sw.reference_omega
    bulk = ph.bulk_dispersion(sw, "FS", 1e-3, table=t)
    _, _, phi2 = ph.optic_response(sw, "T", "AR", "FS", 1e-3, table=t, gdd=gdd)
    _, _, phi1 = ph.optic_response(sw, "T", "AR", "FS", 1e-3, table=t, gdd=gdd, gdd2=None)
    _, _, phir = ph.optic_response(sw, "R", "AR", "FS", 1e-3, table=t, gdd=gdd)
    ok = np.isfinite(phi2) & np.isfinite(bulk)
    ok[0] = False
    D = -150e-30
    assert ok.sum() > 1000
    assert np.allclose(phi2[ok], (2 * D * x ** 2 / 2 + bulk)[ok], rtol=1e-9, atol=1e-9)
    assert np.allclose(phi1[ok], (D * x ** 2 / 2 + bulk)[ok], rtol=1e-9, atol=1e-9)
    assert np.allclose(phir[ok], (D * x ** 2 / 2)[ok], rtol=1e-9, atol=1e-9)


def test_coating_phase_moves_to_shorter_wavelengths_off_normal(tmp_path):
    """OpticalSurface.m:105-110: at incidence theta the GDD phase is
    circshifted by -round(0.2 lam0 sin^2 theta / dlam) samples.  The grid's
    wavelengths fall with the index (dlam < 0), so the curve moves to higher
    indices, i.e. shorter wavelengths."""
    sw, t, gdd = _flat_gdd_optic(tmp_path)
    lam, ref = sw.wavelengths, sw.reference_index
    dlam = (lam[ref + 1] - lam[ref - 1]) / 2
    assert dlam < 0
    th = 30.0
    k = -int(np.sign(0.2 * lam[ref] * 0.25 / dlam) * np.floor(abs(0.2 * lam[ref] * 0.25 / dlam) + 0.5))
    assert k > 0
    p0 = ph.coating_phase(sw, gdd, 0.0)
    p1 = ph.coating_phase(sw, gdd, th)
    assert np.array_equal(np.isnan(p1), np.isnan(np.roll(p0, k)))
    ok = np.isfinite(p1)
    assert np.allclose(p1[ok], np.roll(p0, k)[ok], rtol=0, atol=0)


def test_bare_surface_uses_the_20C_index():
    """fresnel(1, material, theta, lam, exit) calls sellmeier without a
    temperature (fresnel.m:10-16), so a bare 'None' surface's transmission
    does not follow the optic's temperature; the bulk phase does."""
    sw = ph.SimWindow(1040e-9, 1024, 4e-12)
    a = ph.optic_response(sw, "T", "None", "PPLN", 1e-3, T=20.0)
    b = ph.optic_response(sw, "T", "None", "PPLN", 1e-3, T=150.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ok = np.isfinite(a[2]) & np.isfinite(b[2])
    assert not np.allclose(a[2][ok], b[2][ok])
