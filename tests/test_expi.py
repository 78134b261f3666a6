"""exp(i x) of the kernels (the dispersion and nonlinear-term phases): the
FP32 form reduces x to a quadrant and |r| <= pi/4 in FP64 (Cody-Waite with
fma) and sums the Taylor polynomials of sin r and cos r in FP32; the FP64
form reduces modulo 2 pi and calls sin / cos.  Pinned against long-double
cos / sin over 4 million arguments up to 1e10 rad (tests/emu/expi_check.cpp,
the kernels' header compiled for the CPU)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_expi_error_bounds():
    emu = os.path.join(ROOT, "tests", "emu")
    subprocess.check_call(["make", "-s", "-C", emu, "expi_check"])
    out = subprocess.check_output([os.path.join(emu, "expi_check")], text=True).split()
    e32, e64 = float(out[0]), float(out[1])
    assert e32 < 1.0e-7            # below one FP32 ulp of 1 (1.19e-7)
    assert e64 < 1.0e-15
