"""Python host interface over libofx_hip.

`OPOmexBatch` has the argument list and outputs of the reference's solver
handle (OpticalSim.m:237-252; MEX gateway OPO_mex_BATCH.cpp:19-141), so host
code written against `obj.Solver(...)` runs unchanged.  `OPOBatch` is the
device-resident round-trip engine (replaces the OpticalSim.run loop,
OpticalSim.m:230-288).
"""
from __future__ import annotations

import atexit
import ctypes
import threading
import weakref

import numpy as np

from . import _lib
from ._lib import OfxError, OpoSpec, PassParams, Tuning

_ctx_lock = threading.Lock()
_default_ctx = None
_live_batches: "weakref.WeakSet" = weakref.WeakSet()
_live_contexts: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _shutdown():
    # release device objects before the HIP runtime tears itself down
    for b in list(_live_batches):
        b.close()
    for c in list(_live_contexts):
        c.close()


class Context:
    """One HIP device context (ofx_create / ofx_destroy)."""

    def __init__(self, device: int = 0, lib_path: str | None = None):
        self.lib = _lib.load(lib_path)
        h = ctypes.c_void_p()
        st = self.lib.ofx_create(device, ctypes.byref(h))
        if st:
            raise OfxError(st, "ofx_create failed (is a HIP device visible?)")
        self.h = h
        _live_contexts.add(self)

    def check(self, st):
        if st:
            raise OfxError(st, self.lib.ofx_last_error(self.h).decode())

    def info(self):
        name = ctypes.create_string_buffer(128)
        cus = ctypes.c_int()
        khz = ctypes.c_int()
        self.check(self.lib.ofx_device_info(self.h, name, 128, ctypes.byref(cus), ctypes.byref(khz)))
        return dict(name=name.value.decode(), cus=cus.value, wallclock_khz=khz.value)

    def close(self):
        if getattr(self, "h", None):
            for b in list(_live_batches):
                if b.ctx is self:
                    b.close()
            self.lib.ofx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def _tuning(tune):
    if tune is None:
        return None
    ex = tune.get("exchange", 0)
    ex = {"auto": 0, "tags": 1, "barrier": 2}.get(ex, ex)
    t = Tuning(int(tune.get("cluster", 0)), int(tune.get("n1", 0)), int(tune.get("block", 0)),
               int(ex))
    return ctypes.byref(t)


def _method(m):
    if m in (0, "heun"):
        return _lib.OFX_METHOD_HEUN
    if m in (1, "rk23"):
        return _lib.OFX_METHOD_RK23
    raise ValueError(f"method must be 'heun' or 'rk23', got {m!r}")


def _params(n, nsteps, w0, bdiffw0, h, dt, maxErr, minErr, sel, n_cols, method=0):
    return PassParams(int(n), int(nsteps), float(w0), float(bdiffw0), float(h), float(dt),
                      float(maxErr), float(minErr), float(sel), int(n_cols), int(method))


def OPOmexBatch(Et, alpha_w, G, w0, bdiffw0, h, NSteps, dt, beta_op, maxErr, minErr, sel, ApFT,
                stepMods, *, ctx: Context | None = None, precision: str = "single", tune=None,
                return_status: bool = False, method: str = "heun"):
    """[Et, ApFT, stepMods] = OPOmexBatch(Et, alpha_w, G, w0, bdiffw0, h, NSteps, dt,
    beta_op, maxErr, minErr, sel, ApFT, stepMods).

    Et: (N,) or (N, B) complex, fftshifted time (OpticalSim.m:234).
    G: (NSteps,) shared or (NSteps, B) per instance (OPO_CUDA_BATCH_XTAL.cu:222).
    ApFT: (N, P-1) or (N, P-1, B) complex; stepMods: (NSteps,) or (1, NSteps).
    Inputs are not modified; new arrays are returned (OPO_mex_BATCH.cpp:42-56,120-122).
    precision "single" (the MEX contract) or "double" (FP64 fields).
    method "heun" (OPOmexBatch / NEE_CPU.m) or "rk23" (the ode23 variant,
    ode23/OPO_CUDA_ODE23.cu)."""
    ctx = ctx or default_context()
    f64 = precision == "double"
    rdt = np.float64 if f64 else np.float32
    cdt = np.complex128 if f64 else np.complex64
    beta_op = np.asarray(beta_op).reshape(-1)
    N = beta_op.size
    Et_a = np.asarray(Et)
    squeeze = Et_a.ndim == 1
    E2 = Et_a.reshape(N, -1)
    B = E2.shape[1]
    if B == 0:       # an empty batch: nothing to solve, copies returned
        out = (Et_a.copy(), None if ApFT is None else np.array(ApFT, copy=True),
               None if stepMods is None else np.array(stepMods, copy=True))
        return (*out, np.zeros(0, np.int32)) if return_status else out
    Ef = np.ascontiguousarray(E2.T.astype(cdt))          # (B, N) rows
    G_a = np.asarray(G, dtype=rdt)
    nsteps = int(NSteps)
    if G_a.size == nsteps:
        g_batch = 1
        Gf = np.ascontiguousarray(G_a.reshape(-1))
    elif G_a.size == nsteps * B:
        g_batch = B
        Gf = np.ascontiguousarray(G_a.reshape(nsteps, B).T)
    else:
        raise ValueError(f"G must hold NSteps ({nsteps}) or NSteps*B values, got {G_a.size}")
    ap_in = None if ApFT is None else np.asarray(ApFT)
    n_cols = 0
    apf = None
    if ap_in is not None and ap_in.size:
        n_cols = ap_in.size // (N * B)
        if n_cols * N * B != ap_in.size:
            raise ValueError("ApFT size must be N*(P-1)*B")
        apf = np.ascontiguousarray(np.transpose(ap_in.reshape(N, n_cols, B), (2, 1, 0)).astype(cdt))
    sm_in = None if stepMods is None else np.asarray(stepMods, dtype=rdt).reshape(-1)
    smf = None
    if sm_in is not None:
        if sm_in.size != nsteps:
            raise ValueError("stepMods must have NSteps elements")
        smf = np.ascontiguousarray(np.tile(sm_in, (B, 1)))
    p = _params(N, nsteps, w0, bdiffw0, h, dt, maxErr, minErr, sel, n_cols, _method(method))
    status = np.zeros(B, np.int32)
    ptr = (lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) if f64 else \
        (lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    al = np.ascontiguousarray(np.asarray(alpha_w, dtype=rdt).reshape(-1))
    be = np.ascontiguousarray(beta_op.astype(rdt))
    fn = ctx.lib.ofx_crystal_pass_host_f64 if f64 else ctx.lib.ofx_crystal_pass_host
    st = fn(ctx.h, ctypes.byref(p), B, ptr(Ef.view(rdt)), ptr(al), ptr(Gf), g_batch, ptr(be),
            ptr(apf.view(rdt)) if apf is not None else None,
            ptr(smf) if smf is not None else None,
            status.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _tuning(tune))
    if st and st != _lib.OFX_E_NONFINITE:
        ctx.check(st)
    Et_out = Ef.T.copy()
    if squeeze:
        Et_out = Et_out[:, 0]
    ap_out = None
    if apf is not None:
        ap_out = np.transpose(apf, (2, 1, 0)).reshape(ap_in.shape)
    sm_out = None
    if smf is not None:
        sm_out = smf[0].reshape(np.asarray(stepMods).shape)
    if return_status:
        return Et_out, ap_out, sm_out, status
    return Et_out, ap_out, sm_out


class OPOBatch:
    """Device-resident batch of OPO instances (ofx_opo_*).

    Each instance is a dict with keys: solver (the dict of
    physics.NonlinearCrystal.solver_inputs plus maxErr, minErr, sel, n_cols),
    field0 (fftshifted crystal input, complex), and optionally pump, cav, ocm
    (FFT-order complex spectra), delay, and cluster / n1 (this instance's
    launch geometry; 0 or absent = batch tuning or automatic).
    tune: None (built-in heuristic), a dict of cluster / n1 / block for the
    whole batch, or "auto" (timed per grid size by optifax_amd.autotune)."""

    def __init__(self, instances, precision="single", trip_cap=0, tune=None,
                 ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        if isinstance(tune, str):
            if tune != "auto":
                raise ValueError("tune must be None, a dict or 'auto'")
            from .autotune import tune_batch
            instances, block = tune_batch(instances, precision=precision, ctx=self.ctx)
            tune = {"block": block}
        self.tuning = tune
        self.n_inst = len(instances)
        self.prec = _lib.OFX_F64 if precision == "double" else _lib.OFX_F32
        self.trip_cap = int(trip_cap)
        self._keep = []
        specs = (OpoSpec * self.n_inst)()
        self.meta = []
        for i, inst in enumerate(instances):
            s = inst["solver"]
            n = int(np.asarray(s["beta_op"]).size)
            p = _params(n, s["nSteps"], s["w0"], s["bdiffw0"], s["h"], s["dt"], s["maxErr"],
                        s["minErr"], s["sel"], s.get("n_cols", 0),
                        _method(inst.get("method", s.get("method", "heun"))))
            specs[i].pass_ = p
            specs[i].delay = float(inst.get("delay", 0.0))
            specs[i].beta = self._d(s["beta_op"])
            specs[i].alpha = self._d(s["alpha_w"])
            specs[i].G = self._d(s["G33"])
            specs[i].field = self._c(inst["field0"])
            specs[i].pump = self._c(inst["pump"]) if inst.get("pump") is not None else None
            specs[i].cav = self._c(inst["cav"]) if inst.get("cav") is not None else None
            specs[i].ocm = self._c(inst["ocm"]) if inst.get("ocm") is not None else None
            specs[i].cluster = int(inst.get("cluster", 0))
            specs[i].n1 = int(inst.get("n1", 0))
            self.meta.append(dict(n=n, nsteps=int(s["nSteps"]), n_cols=int(s.get("n_cols", 0)),
                                  cavity=inst.get("cav") is not None))
        h = ctypes.c_void_p()
        self.ctx.check(self.lib.ofx_opo_create(self.ctx.h, self.prec, self.n_inst, specs,
                                               self.trip_cap, _tuning(tune), ctypes.byref(h)))
        self.h = h
        self._keep = []
        _live_batches.add(self)

    def _d(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        self._keep.append(a)
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def _c(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128).reshape(-1))
        self._keep.append(a)
        return a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    def launch(self, trips=0, conv_tol=0.0):
        self.ctx.check(self.lib.ofx_opo_launch(self.h, int(trips), float(conv_tol)))

    def graph_replay(self, trips, replays, conv_tol=0.0):
        self.ctx.check(self.lib.ofx_opo_graph_replay(self.h, int(trips), float(conv_tol),
                                                     int(replays)))

    def sync(self, allow_nonfinite=False):
        st = self.lib.ofx_opo_sync(self.h)
        if st and not (allow_nonfinite and st == _lib.OFX_E_NONFINITE):
            self.ctx.check(st)

    def set_wait_limit(self, seconds):
        """Longest wait of one workgroup for its cluster's exchange data before
        the cluster gives up (default 10 s; sync then raises a timeout)."""
        self.ctx.check(self.lib.ofx_opo_set_wait_limit(self.h, float(seconds)))

    def last_ms(self):
        ms = ctypes.c_double()
        self.ctx.check(self.lib.ofx_opo_last_ms(self.h, ctypes.byref(ms)))
        return ms.value

    def read(self, inst):
        m = self.meta[inst]
        n = m["n"]
        field = np.zeros(n, np.complex128)
        oc = np.zeros(n, np.complex128) if m["cavity"] else None
        tl = np.zeros((max(self.trip_cap, 1), 4)) if self.trip_cap else None
        status = ctypes.c_int()
        trips = ctypes.c_int()
        cnt = np.zeros(2, np.int64)
        ap = np.zeros((max(m["n_cols"], 1), n), np.complex128) if m["n_cols"] else None
        sm = np.zeros(m["nsteps"])
        dp = lambda a: None if a is None else a.view(np.float64).ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))
        self.ctx.check(self.lib.ofx_opo_read(
            self.h, inst, dp(field), dp(oc), dp(tl), ctypes.byref(status), ctypes.byref(trips),
            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), dp(ap), dp(sm)))
        return dict(field=field, oc_spec=oc, trip_log=tl, status=status.value,
                    trips_done=trips.value, attempts=int(cnt[0]), accepted=int(cnt[1]),
                    apft=None if ap is None else ap.T.copy(), stepmods=sm)

    def read_average(self, inst):
        """(oc_esd, xout_esd, trips): means of the output-coupled and the
        crystal-output |E(w)|^2 (FFT order) over the trips since creation or
        the last reset_average (OpticalSim.m:302-313)."""
        n = self.meta[inst]["n"]
        oc, xo = np.zeros(n), np.zeros(n)
        trips = ctypes.c_int()
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self.ctx.check(self.lib.ofx_opo_read_average(self.h, inst, dp(oc), dp(xo),
                                                     ctypes.byref(trips)))
        return oc, xo, trips.value

    def reset_average(self, inst):
        self.ctx.check(self.lib.ofx_opo_reset_average(self.h, inst))

    def keep_history(self, trips):
        """Hold the crystal-output spectra of the last `trips` trips of every
        cavity instance on the device (StoredPulses, OpticalSim.m:268-272)."""
        self.ctx.check(self.lib.ofx_opo_keep_history(self.h, int(trips)))

    def read_history(self, inst, trip):
        """(spectrum, stepmods) of cumulative 0-based trip `trip`: the
        crystal-output spectrum (FFT order) and the row of step multipliers
        (StepSizeModifiers(trip, :), OpticalSim.m:172,237)."""
        m = self.meta[inst]
        out = np.zeros(m["n"], np.complex128)
        sm = np.zeros(m["nsteps"])
        dp = lambda a: a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self.ctx.check(self.lib.ofx_opo_read_history(self.h, inst, int(trip), dp(out), dp(sm)))
        return out, sm

    def write_field(self, inst, field):
        a = np.ascontiguousarray(np.asarray(field, dtype=np.complex128).reshape(-1))
        self.ctx.check(self.lib.ofx_opo_write_field(
            self.h, inst, a.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def geometry(self, inst=0):
        v = [ctypes.c_int() for _ in range(5)]
        self.ctx.check(self.lib.ofx_opo_geometry(self.h, inst, *[ctypes.byref(x) for x in v]))
        return dict(zip(["cluster", "n1", "block", "grid", "lds_bytes"], [x.value for x in v]))

    def xcc_ids(self):
        g = self.geometry(0)["grid"]
        ids = np.zeros(g, np.int32)
        cnt = ctypes.c_int()
        self.ctx.check(self.lib.ofx_opo_xcc_ids(self.h, ids.ctypes.data_as(
            ctypes.POINTER(ctypes.c_int)), g, ctypes.byref(cnt)))
        return ids

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.ofx_opo_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
