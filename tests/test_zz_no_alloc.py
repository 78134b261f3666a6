"""The no-allocation single pass (ofx_crystal_pass_noalloc) on the GPU,
through the C ABI with device arrays, against ofx_crystal_pass_host on the
same inputs: identical fields, ApFT, step multipliers and statuses, on two
consecutive calls, the second smaller (the library's device array keeps the
first call's contents).  The device arrays come from the HIP runtime the
library itself links (ctypes on libamdhip64), test plumbing only."""
import ctypes

import numpy as np
import pytest

from noalloc_cases import call_host, call_noalloc, raw_case

pytestmark = pytest.mark.gpu

_hip = None


def _hip_rt():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class _DevArrays:
    """Device copies of host arrays; fetch() copies one back."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.ptrs = []

    def ptr_of(self, name, a):
        hip = _hip_rt()
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        self.ptrs.append(p)
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0

        def fetch(p=p, shape=a.shape, dtype=a.dtype):
            out = np.empty(shape, dtype)
            assert hip.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
            return out
        return p.value, fetch

    def free(self):
        for p in self.ptrs:
            _hip_rt().hipFree(p)


@pytest.mark.parametrize("shapes", [((4096, 3, True, 0), (1024, 1, False, 0)),
                                    ((2048, 2, False, 1), (512, 1, True, 1))])
def test_noalloc_matches_the_host_entry_twice(gpu_ctx, shapes):
    for n, batch, pg, method in shapes:
        c = raw_case(n, batch, pg, seed=n + batch, method=method)
        print(f"no-alloc pass: N={n} batch={batch} per-instance G={pg} method={method}", flush=True)
        want = call_host(gpu_ctx, c)
        dev = _DevArrays()
        try:
            assert _hip_rt().hipDeviceSynchronize() == 0
            got = call_noalloc(gpu_ctx, c, dev.ptr_of)
        finally:
            dev.free()
        assert got[0] == want[0] == 0
        for a, b in zip(got[1:], want[1:]):
            assert np.array_equal(a, b)
        assert np.all(got[4] == 1)
