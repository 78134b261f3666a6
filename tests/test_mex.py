"""The compiled MEX gateways (OPOmexBatch = the reference's solver handle,
OPO_mex_BATCH.cpp:19-141; OPOmexCavity = whole round trips) driven through the
mx-API shim exactly as MATLAB would call them."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from optifax_amd import mexshim
from optifax_amd.solver import OPOBatch, OPOmexBatch

from cases import crystal_case, opo_case, rel_l2


def _mex_args(E, args, B=1, cols=4):
    n = E.shape[0]
    alpha, G, w0, bd, h, ns, dt, beta, mx, mn, sel = args
    Et = np.asarray(E, np.complex64).reshape(n, -1)
    return [Et, np.asarray(alpha, np.float32).reshape(1, -1), np.asarray(G, np.float32).reshape(-1, 1),
            np.float32(w0), np.float64(bd), np.float32(h), np.uint32(ns), np.float64(dt),
            np.asarray(beta, np.float32).reshape(1, -1), np.float32(mx), np.float32(mn),
            np.float64(sel), np.zeros((n, cols, B), np.complex64) if B > 1 else
            np.zeros((n, cols), np.complex64), np.zeros((1, ns), np.float32)]


def test_gateways_export_mexfunction():
    for name in ("libOPOmexBatch.so", "libOPOmexCavity.so"):
        L = ctypes.CDLL(os.path.join(mexshim.LIBDIR, name))
        assert hasattr(L, "mexFunction")


def test_wrong_argument_count_uses_reference_error_id():
    with pytest.raises(mexshim.MexError) as e:
        mexshim.call("OPOmexBatch", np.zeros(4, np.complex64))
    assert e.value.identifier == "parallel:gpu:mexGPUExample:InvalidInput"   # OPO_mex_BATCH.cpp:31


def test_shape_errors_before_any_device_work():
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    a = _mex_args(E, args)
    a[2] = a[2][:-1]                      # G one element short
    with pytest.raises(mexshim.MexError) as e:
        mexshim.call("OPOmexBatch", *a)
    assert e.value.identifier == "optifax:OPOmexBatch:size"


def test_empty_batch_returns_copies():
    """numel(Et) = 0: nothing to solve; the gateway returns copies of Et,
    ApFT and stepMods (OPO_mex_BATCH.cpp:42,55-56,120-122) without touching
    a device."""
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    a = _mex_args(E, args)
    a[0] = np.zeros((256, 0), np.complex64)
    a[12] = np.zeros((256, 4, 0), np.complex64)
    a[13] = np.arange(a[13].size, dtype=np.float32).reshape(a[13].shape)
    Et, Ap, sm = mexshim.call("OPOmexBatch", *a)
    assert Et.shape == (256, 0) and Ap.size == 0
    assert np.array_equal(sm, a[13])


@pytest.mark.gpu
def test_opomexbatch_matches_c_api_and_oracle(gpu_ctx):
    sw, si, E, args = crystal_case(n=1024, window=4e-12)
    a = _mex_args(E, args)
    Et, Ap, sm = mexshim.call("OPOmexBatch", *a)
    assert Et.shape == (1024, 1) and Ap.shape == (1024, 4) and sm.shape == (1, args[5])
    # the gateway reads w0, h, maxErr and minErr as single (OPO_mex_BATCH.cpp:45-51)
    same = [float(np.float32(x)) if i in (2, 4, 8, 9) else x for i, x in enumerate(args)]
    E2, Ap2, sm2 = OPOmexBatch(a[0][:, 0], *same, a[12], a[13], ctx=gpu_ctx)
    assert np.array_equal(Et[:, 0], E2.astype(np.complex64))
    assert np.array_equal(sm[0], sm2.reshape(-1))
    Eo, apo, smo, _ = oracle.nee_heun(E, *same, n_cols=4)
    assert np.array_equal(sm[0], smo)
    assert rel_l2(Et[:, 0], Eo) < 2e-4
    mexshim.clear()


@pytest.mark.gpu
def test_opomexcavity_matches_device_batch(gpu_ctx):
    cfg, b = opo_case()
    si, ops = b["solver"], b["ops"]
    n, trips = si["beta_op"].size, 3
    a = [np.asarray(ops["field0"], np.complex64).reshape(n, 1), si["alpha_w"].astype(np.float32),
         si["G33"].astype(np.float32), np.float32(si["w0"]), np.float64(si["bdiffw0"]),
         np.float32(si["h"]), np.uint32(si["nSteps"]), np.float64(si["dt"]),
         si["beta_op"].astype(np.float32), np.float32(si["maxErr"]), np.float32(si["minErr"]),
         np.float64(si["sel"]), ops["pump"], ops["cav"], ops["ocm"], np.float64(cfg.delay),
         np.float64(trips), np.float64(0.0)]
    Et, oc, log, status, esd_oc, esd_x = mexshim.call("OPOmexCavity", *a, nargout=6)
    f32 = lambda x: np.asarray(x, np.float32).astype(np.float64)
    s32 = dict(si, beta_op=f32(si["beta_op"]), alpha_w=f32(si["alpha_w"]), G33=f32(si["G33"]),
               w0=float(f32(si["w0"])), h=float(f32(si["h"])), maxErr=float(f32(si["maxErr"])),
               minErr=float(f32(si["minErr"])))
    run = OPOBatch([dict(solver=s32, field0=a[0][:, 0], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay)], trip_cap=trips, ctx=gpu_ctx)
    run.launch(trips)
    run.sync()
    r = run.read(0)
    assert status[0, 0] == 1
    assert np.array_equal(log[2:, :, 0].T, r["trip_log"][:, 2:])
    assert rel_l2(Et[:, 0], r["field"]) < 1e-6
    e, x, k = run.read_average(0)
    assert k == trips and rel_l2(esd_oc[:, 0], e) < 1e-6 and rel_l2(esd_x[:, 0], x) < 1e-6
    run.close()
    mexshim.clear()
