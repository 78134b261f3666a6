"""Fibre GNLSE (gnlsefib.m): the FP64 oracle's pins, and the kernel body
against it through the CPU emulation (tests/emu/libofx_emu.so) and on the GPU.

The oracle (oracle/gnlse.py) restates gnlsefib.m and MATLAB's ode45
(Dormand-Prince 5(4), Shampine & Reichelt's step control, the continuous
extension at requested points).  No MATLAB output exists to compare with;
the pins are mathematical: the pair's order conditions and its continuous
extension's end values, a linear ODE against its exponential, and GNLSE
cases with closed forms (dispersion alone, self-phase modulation alone, the
fundamental soliton, energy conservation and loss).

The test pulse follows the supercontinuum example distributed with
gnlsefib.m (Dudley, Genty & Coen's test_Dudley: 835 nm, 28.4 fs sech, 10 kW,
gamma 0.11 /W/m, the listed betas, fused-silica Raman), in ps and m, on a
shorter fibre and a smaller grid so that the oracle finishes in seconds.
"""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import gnlse as og  # noqa: E402
from optifax_amd import fibre  # noqa: E402
from optifax_amd import _lib  # noqa: E402

C_NM_PS = 299792458 * 1e9 / 1e12
BETAS = [-11.830e-3, 8.1038e-5, -9.5205e-8, 2.0737e-10, -5.3943e-13, 1.3486e-15, -2.5495e-18,
         3.0524e-21, -1.7140e-24]     # ps^k / m


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def dudley_case(n=512, twidth=12.5, power=10000.0, flength=0.004, raman=True, shock=True):
    T = np.linspace(-twidth / 2, twidth / 2, n)
    A = np.sqrt(power) / np.cosh(T / 0.0284)
    w0 = 2 * np.pi * C_NM_PS / 835.0 if shock else 0.0
    RT = fibre.raman_response(T, 0.0122, 0.032) if raman else 0.0
    return dict(T=T, A=A, w0=w0, gamma=0.11, betas=BETAS, loss=0.0, fr=0.18 if raman else 0.0,
                RT=RT, flength=flength)


def _oracle(c, nsaves=5):
    return og.gnlsefib(c["T"], c["A"], c["w0"], c["gamma"], c["betas"], c["loss"], c["fr"],
                       c["RT"], c["flength"], nsaves)


# ---------------------------------------------------------------- oracle pins
def test_dormand_prince_order_conditions():
    c = np.concatenate([[0.0], og.DP_A])
    b5 = og.DP_B[:, 5]
    b4 = b5 - og.DP_E
    for q in range(5):
        assert np.sum(b5 * c ** q) == pytest.approx(1.0 / (q + 1), abs=1e-14)
    for q in range(4):
        assert np.sum(b4 * c ** q) == pytest.approx(1.0 / (q + 1), abs=1e-14)
    # stage k's coefficients sum to its node
    for k in range(6):
        assert np.sum(og.DP_B[:, k]) == pytest.approx(og.DP_A[k], abs=1e-14)
    # 5th-order condition sum_j b_j sum_k a_jk c_k = 1/6 (b a c)
    Amat = np.zeros((7, 7))
    for k in range(6):
        Amat[k + 1, :] = og.DP_B[:, k]
    assert b5 @ Amat @ c == pytest.approx(1.0 / 6.0, abs=1e-14)
    assert b5 @ Amat @ (c ** 2) == pytest.approx(1.0 / 12.0, abs=1e-14)


def test_continuous_extension_end_values():
    # s = 1 gives the 5th-order solution's weights; d/ds at 0 the first stage
    assert np.allclose(og.DP_BI.sum(axis=1), og.DP_B[:, 5], rtol=0, atol=1e-14)
    assert np.array_equal(og.DP_BI[:, 0], np.eye(7)[0])


def test_ode45_linear_ode_matches_the_exponential():
    lam = -3.0 + 40.0j
    tspan = np.linspace(0.0, 1.0, 11)
    _, y, st = og.ode45(lambda t, y: lam * y, tspan, np.array([1.0 + 0j, 2.0 - 1j]), rtol=1e-6,
                        atol=1e-12)
    want = np.exp(lam * tspan)[:, None] * np.array([1.0 + 0j, 2.0 - 1j])
    assert np.max(np.abs(y - want)) < 1e-5
    assert st["accepted"] > 10 and st["evaluations"] == 1 + 6 * (st["accepted"] + st["rejected"])


def test_dispersion_alone_is_the_closed_form():
    c = dudley_case(n=256, raman=False, shock=False)
    c["gamma"] = 0.0
    Z, AT, AW, W, B, _ = _oracle(c, nsaves=4)
    V, L, Wf, RW, g, Bv, dT = og.operators(c["T"], 0.0, 0.0, c["betas"], 0.0, 0.0, 0.0)
    for i, z in enumerate(Z):
        want = np.fft.fft(np.fft.ifft(c["A"]) * np.exp(L * z))
        assert rel(AT[i], want) < 1e-13


def test_self_phase_modulation_alone():
    c = dudley_case(n=256, raman=False, shock=False, flength=0.01)
    c["betas"] = [0.0]
    Z, AT, _, _, _, _ = _oracle(c, nsaves=3)
    want = c["A"] * np.exp(1j * c["gamma"] * np.abs(c["A"]) ** 2 * Z[-1])
    assert rel(AT[-1], want) < 1e-4
    assert rel(np.abs(AT[-1]), np.abs(c["A"])) < 5e-5     # (RelTol 1e-5)


def test_fundamental_soliton_keeps_its_shape():
    n, T0, gamma, b2 = 512, 0.05, 0.1, -0.01
    T = np.linspace(-4.0, 4.0, n)
    P0 = abs(b2) / (gamma * T0 ** 2)
    A = np.sqrt(P0) / np.cosh(T / T0)
    z0 = math.pi / 2 * T0 ** 2 / abs(b2)        # soliton period
    Z, AT, _, _, _, st = og.gnlsefib(T, A, 0.0, gamma, [b2], 0.0, 0.0, 0.0, z0, 3)
    assert rel(np.abs(AT[-1]), np.abs(A)) < 2e-4
    # phase of the fundamental soliton: gamma P0 z / 2
    ph = np.angle(AT[-1][n // 2] / A[n // 2])
    assert ph == pytest.approx(((gamma * P0 * z0 / 2 + np.pi) % (2 * np.pi)) - np.pi, abs=1e-3)


def test_energy_conserved_without_loss_and_decays_with_loss():
    c = dudley_case(n=256, raman=False, shock=False, power=2000.0, flength=0.02)
    Z, AT, _, _, _, _ = og.gnlsefib(c["T"], c["A"], 0.0, c["gamma"], c["betas"], 0.0, 0.0, 0.0,
                                    c["flength"], 5)
    E = np.sum(np.abs(AT) ** 2, axis=1)
    assert np.max(np.abs(E / E[0] - 1)) < 5e-5       # (RelTol 1e-5; the drift grows with z)
    loss_db = 3.0
    Z, AT, _, _, _, _ = og.gnlsefib(c["T"], c["A"], 0.0, c["gamma"], c["betas"], loss_db, 0.0,
                                    0.0, c["flength"], 5)
    alpha = math.log(10 ** (loss_db / 10))
    E = np.sum(np.abs(AT) ** 2, axis=1)
    assert np.allclose(E / E[0], np.exp(-alpha * Z), rtol=5e-5)


def test_operators_match_the_oracle():
    c = dudley_case(n=256)
    op = fibre.operators(c["T"], c["w0"], c["gamma"], c["betas"], 0.0, c["fr"], c["RT"])
    V, L, W, RW, g, B, dT = og.operators(c["T"], c["w0"], c["gamma"], c["betas"], 0.0, c["fr"],
                                         c["RT"])
    assert np.array_equal(op["L"], L) and np.array_equal(op["W"], W)
    assert np.array_equal(op["RW"], RW) and op["gamma"] == g and op["dT"] == dT


# ---------------------------------------------------------------- emulation
@pytest.fixture(scope="module")
def emu_lib():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "emu")])
    return _lib.load(os.path.join(ROOT, "tests", "emu", "libofx_emu.so"))


MAX_STEPS = 100_000     # every case here takes under 2,000


def _device(c, lib, precision, nsaves=5):
    return fibre.gnlsefib(c["T"], c["A"], c["w0"], c["gamma"], c["betas"], c["loss"], c["fr"],
                          c["RT"], c["flength"], nsaves, precision=precision, lib=lib,
                          max_steps=MAX_STEPS)


def _device_stats(c, lib, precision, nsaves=5):
    op = fibre.operators(c["T"], c["w0"], c["gamma"], c["betas"], c["loss"], c["fr"], c["RT"])
    return fibre.gnlse_batch(c["T"], [c["A"]], [op], c["flength"], nsaves, precision, lib=lib,
                             max_steps=MAX_STEPS)


@pytest.mark.parametrize("raman,shock", [(True, True), (False, True), (False, False)])
def test_emulated_kernel_fp64_matches_the_oracle(emu_lib, raman, shock):
    c = dudley_case(n=256, raman=raman, shock=shock, flength=0.003)
    Zo, ATo, AWo, Wo, Bo, st = _oracle(c)
    Z, AT, AW, W, B = _device(c, emu_lib, "double")
    _, _, stats, _ = _device_stats(c, emu_lib, "double")
    print(f"raman {raman} shock {shock}: oracle {st}, emulation {stats[0].tolist()}, "
          f"AT rel {rel(AT, ATo):.1e}")
    assert stats[0, 0] == st["accepted"] and stats[0, 1] == st["rejected"]
    assert stats[0, 3] == 0
    assert rel(AT, ATo) < 1e-9 and rel(AW, AWo) < 1e-9
    assert np.array_equal(Z, Zo) and np.allclose(W, Wo) and np.allclose(B, Bo)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_emulated_kernel_fp64_small_grids_match_the_oracle(emu_lib, n):
    """Every FFT shape the kernel builds: one pass (n = 4, and the half-length
    Raman transforms at n = 4 and 8), a last radix-2 pass (odd log2 n, and the
    half-length transforms of even ones), the twiddle tables' smallest forms;
    a strongly nonlinear pulse (about 220 steps) with the Raman term."""
    c = dudley_case(n=n, twidth=0.02 * n, power=20000.0, flength=0.02)
    _, ATo, _, _, _, st = _oracle(c)
    _, AT, _, _, _ = _device(c, emu_lib, "double")
    _, _, stats, _ = _device_stats(c, emu_lib, "double")
    assert st["accepted"] > 150
    assert stats[0, 0] == st["accepted"] and stats[0, 1] == st["rejected"]
    assert rel(AT, ATo) < 1e-11


def gain_case(n=512):
    """A weak pulse under strong gain (loss -400 dB/m): the nonlinearity turns
    on as the pulse grows, and ode45 rejects steps (3 of 75 in the oracle)."""
    return dict(dudley_case(n=n, flength=0.05, power=100.0), loss=-400.0)


def test_emulated_kernel_fp64_rejections_match_the_oracle(emu_lib):
    c = gain_case()
    _, ATo, _, _, _, st = _oracle(c)
    _, AT, _, _, _ = _device(c, emu_lib, "double")
    _, _, stats, _ = _device_stats(c, emu_lib, "double")
    print(f"gain: oracle {st}, emulation {stats[0].tolist()}, rel {rel(AT, ATo):.1e}")
    assert st["rejected"] > 0
    assert stats[0, 0] == st["accepted"] and stats[0, 1] == st["rejected"]
    assert stats[0, 2] == st["evaluations"]
    assert rel(AT, ATo) < 1e-9


def test_emulated_kernel_fp32_follows_the_oracle(emu_lib):
    c = dudley_case(n=256, flength=0.003)
    _, ATo, _, _, _, st = _oracle(c)
    _, AT, _, _, _ = _device(c, emu_lib, "single")
    print(f"fp32: rel {rel(AT, ATo):.2e}")
    assert rel(AT, ATo) < 1e-4


def test_argument_checks(emu_lib):
    c = dudley_case(n=256)
    with pytest.raises(ValueError):
        fibre.gnlse_batch(c["T"], [c["A"]], [fibre.operators(c["T"], 0, 0.1, [0.0], 0, 0, 0)],
                          0.01, 2, lib=emu_lib)
    op = fibre.operators(c["T"][:200], 0, 0.1, [0.0], 0, 0, 0)
    with pytest.raises(_lib.OfxError):       # not a power of two
        fibre.gnlse_batch(c["T"][:200], [c["A"][:200]], [op], 0.01, 3, lib=emu_lib)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("raman", [True, False])
def test_gpu_fp64_matches_the_oracle(raman):
    """n = 2^12 (the FP64 limit), the example pulse over 4 mm of fibre."""
    c = dudley_case(n=4096, raman=raman, shock=True, flength=0.004)
    Zo, ATo, AWo, _, _, st = _oracle(c)
    _, AT, AW, _, _ = _device(c, None, "double")
    _, _, stats, ms = _device_stats(c, None, "double")
    print(f"raman {raman}: oracle {st}, device {stats[0].tolist()} in {ms:.1f} ms, "
          f"AT rel {rel(AT, ATo):.1e}")
    assert stats[0, 0] == st["accepted"] and stats[0, 1] == st["rejected"]
    assert rel(AT, ATo) < 1e-9 and rel(AW, AWo) < 1e-9


@pytest.mark.gpu
def test_gpu_fp64_rejections_match_the_oracle():
    c = gain_case(n=4096)
    _, ATo, _, _, _, st = _oracle(c)
    _, AT, _, _, _ = _device(c, None, "double")
    _, _, stats, _ = _device_stats(c, None, "double")
    print(f"gain: oracle {st}, device {stats[0].tolist()}, rel {rel(AT, ATo):.1e}")
    assert st["rejected"] > 0
    assert stats[0, 0] == st["accepted"] and stats[0, 1] == st["rejected"]
    assert rel(AT, ATo) < 1e-9


@pytest.mark.gpu
def test_gpu_fp32_at_2_13_follows_the_oracle():
    c = dudley_case(n=8192, flength=0.004)
    _, ATo, _, _, _, st = _oracle(c)
    _, AT, _, _, _ = _device(c, None, "single")
    print(f"fp32 2^13: oracle {st}, rel {rel(AT, ATo):.2e}")
    assert rel(AT, ATo) < 1e-3


@pytest.mark.gpu
def test_gpu_batch_of_powers_matches_single_runs():
    """A power scan in one launch: each fibre of the batch gives what it gives alone."""
    base = dudley_case(n=1024, flength=0.003)
    fields, ops = [], []
    for p in (2000.0, 5000.0, 10000.0):
        c = dict(base, A=np.sqrt(p) / np.cosh(base["T"] / 0.0284))
        fields.append(c["A"])
        ops.append(fibre.operators(c["T"], c["w0"], c["gamma"], c["betas"], 0.0, c["fr"], c["RT"]))
    Z, AWz, stats, _ = fibre.gnlse_batch(base["T"], fields, ops, base["flength"], 4, "double",
                                         max_steps=MAX_STEPS)
    for i in range(3):
        _, one, st1, _ = fibre.gnlse_batch(base["T"], [fields[i]], [ops[i]], base["flength"], 4,
                                           "double", max_steps=MAX_STEPS)
        assert np.array_equal(stats[i], st1[0]) and np.array_equal(AWz[i], one[0])
