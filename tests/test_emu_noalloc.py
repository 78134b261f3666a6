"""The no-allocation single pass (ofx_crystal_pass_noalloc) through the CPU
emulation of the kernel: its layout in the library's fixed array, its
descriptors and tables give bit for bit what ofx_crystal_pass_host gives, on
consecutive calls of different shapes (the array keeps the previous call's
contents), and sizes beyond the array are refused."""
import ctypes

import numpy as np

from noalloc_cases import OFX_E_ARG, call_host, call_noalloc, host_ptrs, raw_case


def test_noalloc_matches_the_host_entry_on_consecutive_calls(emu_ctx):
    for n, batch, pg, method in ((1024, 2, True, 0), (256, 1, False, 0), (512, 3, False, 1),
                                 (1024, 2, True, 0)):
        c = raw_case(n, batch, pg, seed=n + batch, method=method)
        want = call_host(emu_ctx, c)
        keep = []
        got = call_noalloc(emu_ctx, c, host_ptrs(keep))
        assert got[0] == want[0] == 0
        for a, b in zip(got[1:], want[1:]):
            assert np.array_equal(a, b)


def test_noalloc_refuses_what_does_not_fit(emu_ctx):
    c = raw_case(256, 1, False)
    big = raw_case(256, 1, False)
    big["p"].n = 65536                      # 16 instances of 2^16 points: far beyond the array
    big["batch"] = 17
    rc = emu_ctx.lib.ofx_crystal_pass_noalloc(emu_ctx.h, ctypes.byref(big["p"]), 17, 1, 1, 1, 1,
                                              1, None, None, None)
    assert rc == OFX_E_ARG and b"at most 16" in emu_ctx.lib.ofx_last_error(emu_ctx.h)
    c["p"].nsteps = 3_000_000               # a 300 mm crystal's G table alone is 24 MB
    c["G"] = np.zeros(3_000_000, np.float32)
    rc = emu_ctx.lib.ofx_crystal_pass_noalloc(emu_ctx.h, ctypes.byref(c["p"]), 1,
                                              c["Et"].ctypes.data, c["alpha"].ctypes.data,
                                              c["G"].ctypes.data, 1, c["beta"].ctypes.data,
                                              None, None, None)
    assert rc == OFX_E_ARG and b"bytes of device scratch" in emu_ctx.lib.ofx_last_error(emu_ctx.h)
