"""libofx_hip.so loads without a GPU and exports every function that
include/optifax_hip.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from optifax_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ofx_[a-z0-9_]+|mexFunction)\s*\(", text)))


def test_library_exports_every_declared_function():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libofx_hip.so is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared("optifax_hip.h")
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.EXPORTED)


def test_loads_without_device_and_reports_it():
    L = _lib.load()
    h = ctypes.c_void_p()
    st = L.ofx_create(0, ctypes.byref(h))
    if st == 0:          # a GPU is visible: clean up, nothing else to check here
        L.ofx_destroy(h)
    else:
        assert st in (_lib.OFX_E_NODEVICE, _lib.OFX_E_HIP)
