"""Writes tests/golden/fft_golden.json: reference DFT values for the oracle's
FFT at the grid sizes the configs use (2^13 .. 2^16), computed once with
numpy.fft (pocketfft, independent of this repository's code).  Each entry
holds the seed of its input (PCG64, 2N standard normals as interleaved
re/im), 32 output bins spread over the spectrum, and the output's L2 norm.

    python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

SIZES = [1 << 13, 1 << 14, 1 << 15, 1 << 16]


def golden_input(n, seed):
    r = np.random.Generator(np.random.PCG64(seed)).standard_normal(2 * n)
    return r[0::2] + 1j * r[1::2]


def main():
    out = []
    for k, n in enumerate(SIZES):
        seed = 1000 + k
        X = np.fft.fft(golden_input(n, seed))
        bins = sorted({0, 1, n // 2, n - 1, *[int(b) for b in np.linspace(2, n - 2, 28)]})
        out.append(dict(n=n, seed=seed, bins=bins,
                        re=[float(X[b].real) for b in bins], im=[float(X[b].imag) for b in bins],
                        norm=float(np.linalg.norm(X))))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.json")
    with open(path, "w") as f:
        json.dump(dict(source="numpy.fft.fft " + np.__version__, cases=out), f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
