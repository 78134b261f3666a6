"""ctypes binding of libofx_hip (include/optifax_hip.h).

The library is loaded from optifax_amd/lib/libofx_hip.so (built in-tree by
__graft_entry__.build()).  There is no fallback: if the library is missing,
or no HIP device is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libofx_hip.so")

OFX_OK = 0
OFX_E_ARG = -1
OFX_E_HIP = -2
OFX_E_NOMEM = -3
OFX_E_NONFINITE = -4
OFX_E_TIMEOUT = -5
OFX_E_UNSUPPORTED = -6
OFX_E_NODEVICE = -7

OFX_STATUS_RUNNING = 0
OFX_STATUS_DONE = 1
OFX_STATUS_CONVERGED = 2
OFX_STATUS_NONFINITE = 3
OFX_STATUS_TIMEOUT = 4

OFX_F32 = 0
OFX_F64 = 1
OFX_METHOD_HEUN = 0
OFX_METHOD_RK23 = 1

_ERRNAMES = {
    OFX_E_ARG: "invalid argument",
    OFX_E_HIP: "HIP runtime error",
    OFX_E_NOMEM: "out of memory",
    OFX_E_NONFINITE: "non-finite field",
    OFX_E_TIMEOUT: "cluster barrier timeout",
    OFX_E_UNSUPPORTED: "unsupported",
    OFX_E_NODEVICE: "no HIP device",
}

EXPORTED = [
    "ofx_create", "ofx_destroy", "ofx_last_error", "ofx_device_info",
    "ofx_crystal_pass_host", "ofx_crystal_pass_host_f64", "ofx_crystal_pass_noalloc",
    "ofx_opo_create", "ofx_opo_destroy", "ofx_opo_launch", "ofx_opo_sync",
    "ofx_opo_graph_replay", "ofx_opo_last_ms", "ofx_opo_read", "ofx_opo_write_field",
    "ofx_opo_geometry", "ofx_opo_xcc_ids", "ofx_opo_set_wait_limit",
    "ofx_opo_read_average", "ofx_opo_reset_average", "ofx_opo_keep_history",
    "ofx_opo_read_history",
    "ofx_mlp_create", "ofx_mlp_destroy", "ofx_mlp_last_error", "ofx_mlp_num_params",
    "ofx_mlp_set_params", "ofx_mlp_get_params", "ofx_mlp_forward", "ofx_mlp_gradients",
    "ofx_mlp_step", "ofx_mlp_upload_dataset", "ofx_mlp_train", "ofx_mlp_last_ms",
    "ofx_qpm_zsum", "ofx_sinc_gain",
    "ofx_gnlse_run",
]


class OfxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRNAMES.get(code, 'error')} ({code}): {msg}")
        self.code = code


class PassParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nsteps", ctypes.c_int), ("w0", ctypes.c_double),
                ("bdiffw0", ctypes.c_double), ("h", ctypes.c_double), ("dt", ctypes.c_double),
                ("max_err", ctypes.c_double), ("min_err", ctypes.c_double),
                ("sel", ctypes.c_double), ("n_cols", ctypes.c_int), ("method", ctypes.c_int)]


class Tuning(ctypes.Structure):
    _fields_ = [("cluster", ctypes.c_int), ("n1", ctypes.c_int), ("block", ctypes.c_int),
                ("exchange", ctypes.c_int)]


_DP = ctypes.POINTER(ctypes.c_double)
_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int)
_LLP = ctypes.POINTER(ctypes.c_longlong)


class OpoSpec(ctypes.Structure):
    _fields_ = [("pass_", PassParams), ("delay", ctypes.c_double), ("beta", _DP),
                ("alpha", _DP), ("G", _DP), ("field", _DP), ("pump", _DP), ("cav", _DP),
                ("ocm", _DP), ("cluster", ctypes.c_int), ("n1", ctypes.c_int)]


_lib = None


def load(path: str | None = None):
    """Load libofx_hip; raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"{p} is not built; run __graft_entry__.build() first")
    L = ctypes.CDLL(p)
    vp = ctypes.c_void_p
    L.ofx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.ofx_destroy.argtypes = [vp]
    L.ofx_destroy.restype = None
    L.ofx_last_error.argtypes = [vp]
    L.ofx_last_error.restype = ctypes.c_char_p
    L.ofx_device_info.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, _IP, _IP]
    L.ofx_crystal_pass_host.argtypes = [vp, ctypes.POINTER(PassParams), ctypes.c_int, _FP, _FP, _FP,
                                        ctypes.c_int, _FP, _FP, _FP, _IP, ctypes.POINTER(Tuning)]
    L.ofx_crystal_pass_host_f64.argtypes = [vp, ctypes.POINTER(PassParams), ctypes.c_int, _DP, _DP,
                                            _DP, ctypes.c_int, _DP, _DP, _DP, _IP,
                                            ctypes.POINTER(Tuning)]
    if hasattr(L, "ofx_crystal_pass_noalloc"):   # device pointers (plain addresses)
        L.ofx_crystal_pass_noalloc.argtypes = [vp, ctypes.POINTER(PassParams), ctypes.c_int, vp,
                                               vp, vp, ctypes.c_int, vp, vp, vp, vp]
    L.ofx_opo_create.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OpoSpec),
                                 ctypes.c_int, ctypes.POINTER(Tuning), ctypes.POINTER(vp)]
    L.ofx_opo_destroy.argtypes = [vp]
    L.ofx_opo_destroy.restype = None
    L.ofx_opo_launch.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    L.ofx_opo_sync.argtypes = [vp]
    L.ofx_opo_graph_replay.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    L.ofx_opo_last_ms.argtypes = [vp, _DP]
    L.ofx_opo_read.argtypes = [vp, ctypes.c_int, _DP, _DP, _DP, _IP, _IP, _LLP, _DP, _DP]
    L.ofx_opo_write_field.argtypes = [vp, ctypes.c_int, _DP]
    L.ofx_opo_geometry.argtypes = [vp, ctypes.c_int, _IP, _IP, _IP, _IP, _IP]
    L.ofx_opo_xcc_ids.argtypes = [vp, _IP, ctypes.c_int, _IP]
    L.ofx_opo_set_wait_limit.argtypes = [vp, ctypes.c_double]
    L.ofx_opo_read_average.argtypes = [vp, ctypes.c_int, _DP, _DP, _IP]
    L.ofx_opo_reset_average.argtypes = [vp, ctypes.c_int]
    L.ofx_opo_keep_history.argtypes = [vp, ctypes.c_int]
    L.ofx_opo_read_history.argtypes = [vp, ctypes.c_int, ctypes.c_int, _DP, _DP]
    if hasattr(L, "ofx_mlp_create"):   # the CPU emulation library has no surrogate
        L.ofx_mlp_create.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(vp)]
        L.ofx_mlp_destroy.argtypes = [vp]
        L.ofx_mlp_destroy.restype = None
        L.ofx_mlp_last_error.argtypes = [vp]
        L.ofx_mlp_last_error.restype = ctypes.c_char_p
        L.ofx_mlp_num_params.argtypes = [vp]
        L.ofx_mlp_num_params.restype = ctypes.c_longlong
        L.ofx_mlp_set_params.argtypes = [vp, _FP]
        L.ofx_mlp_get_params.argtypes = [vp, _FP]
        L.ofx_mlp_forward.argtypes = [vp, ctypes.c_int, _FP, _FP]
        L.ofx_mlp_gradients.argtypes = [vp, ctypes.c_int, _FP, _FP, _FP, _FP]
        L.ofx_mlp_step.argtypes = [vp, ctypes.c_int, _FP, _FP, ctypes.c_float, _FP]
        L.ofx_mlp_upload_dataset.argtypes = [vp, ctypes.c_int, _FP, _FP]
        L.ofx_mlp_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, _IP, ctypes.c_float, _FP]
        L.ofx_mlp_last_ms.argtypes = [vp]
        L.ofx_mlp_last_ms.restype = ctypes.c_double
    if hasattr(L, "ofx_qpm_zsum"):     # nor the gain maps
        L.ofx_qpm_zsum.argtypes = [ctypes.c_int, ctypes.c_int, _DP, ctypes.c_int, _DP, _IP, _DP,
                                   ctypes.c_double, _DP, _DP]
        L.ofx_sinc_gain.argtypes = [ctypes.c_int, ctypes.c_int, _DP, ctypes.c_int, _DP,
                                    ctypes.c_int, _DP, ctypes.c_double, _DP, _DP]
    if hasattr(L, "ofx_gnlse_run"):
        L.ofx_gnlse_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _DP, _DP, _DP, _DP, _DP, _DP, _DP, _DP,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_longlong, _DP,
                                    ctypes.POINTER(ctypes.c_longlong), _DP]
    if path is None:
        _lib = L
    return L
