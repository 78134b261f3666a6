"""Spectral surrogate of one crystal pass.

`SurrogateMLP` drives the device network of optifax_amd/csrc/ofx_mlp.hip
(BF16 MFMA GEMMs with FP32 accumulation, FP32 masters and Adam).  There is no
CPU fallback: without libofx_hip.so or a HIP device every call raises.

Network: n_in -> width -> width -> width -> n_out, GELU after the first three
layers; relative-L2 loss; weights in PyTorch Linear order ([out][in]).
"""
from __future__ import annotations

import ctypes
import weakref
import atexit

import numpy as np

from . import _lib

_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int)
_LIVE: "weakref.WeakSet[SurrogateMLP]" = weakref.WeakSet()


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _fp(a):
    return a.ctypes.data_as(_FP)


def layer_shapes(n_in, width, n_out):
    return [(width, n_in), (width, width), (width, width), (n_out, width)]


def num_params(n_in, width, n_out):
    return sum(o * i + o for o, i in layer_shapes(n_in, width, n_out))


def init_params(n_in, width, n_out, seed=0, out_scale=0.5):
    """He-normal weights, zero biases, FP32 (seeded)."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (o, i) in enumerate(layer_shapes(n_in, width, n_out)):
        sc = np.sqrt(2.0 / i) * (out_scale if l == 3 else 1.0)
        parts.append((rng.standard_normal((o, i)) * sc).astype(np.float32).ravel())
        parts.append(np.zeros(o, np.float32))
    return np.concatenate(parts)


class SurrogateMLP:
    """Device MLP (one object = one set of weights, Adam state and buffers)."""

    def __init__(self, n_in, width=2048, n_out=2048, max_batch=512, device=0, params=None,
                 seed=0):
        self._L = _lib.load()
        self.dims = (int(n_in), int(width), int(n_out))
        self.max_batch = int(max_batch)
        h = ctypes.c_void_p()
        st = self._L.ofx_mlp_create(int(device), int(n_in), int(width), int(n_out),
                                    int(max_batch), ctypes.byref(h))
        if st != _lib.OFX_OK:
            raise _lib.OfxError(st, "ofx_mlp_create failed")
        self._h = h
        self.n_params = int(self._L.ofx_mlp_num_params(h))
        assert self.n_params == num_params(*self.dims)
        self.set_params(init_params(*self.dims, seed=seed) if params is None else params)
        _LIVE.add(self)

    # -- plumbing ---------------------------------------------------------
    def _check(self, st):
        if st != _lib.OFX_OK:
            msg = self._L.ofx_mlp_last_error(self._h) if self._h else b""
            raise _lib.OfxError(st, (msg or b"").decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.ofx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _batch(self, x, t=None):
        n_in, _, n_out = self.dims
        x = _f32(x)
        if x.ndim != 2 or x.shape[1] != n_in:
            raise ValueError(f"x must be [batch][{n_in}]")
        B = x.shape[0]
        if B < 64 or B % 64 or B > self.max_batch:
            raise ValueError(f"batch must be a multiple of 64 in [64, {self.max_batch}]")
        if t is not None:
            t = _f32(t, (B, n_out))
        return x, t, B

    # -- API --------------------------------------------------------------
    def set_params(self, params):
        p = _f32(params, (self.n_params,))
        self._check(self._L.ofx_mlp_set_params(self._h, _fp(p)))

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(self._L.ofx_mlp_get_params(self._h, _fp(p)))
        return p

    def forward(self, x):
        """y [batch][n_out]; any batch size (padded internally to a multiple of 64)."""
        n_in, _, n_out = self.dims
        x = _f32(x)
        if x.ndim != 2 or x.shape[1] != n_in:
            raise ValueError(f"x must be [batch][{n_in}]")
        out = np.empty((x.shape[0], n_out), np.float32)
        for s in range(0, x.shape[0], self.max_batch):
            xs = x[s:s + self.max_batch]
            B = xs.shape[0]
            Bp = max(64, -(-B // 64) * 64)
            if Bp != B:
                xs = np.concatenate([xs, np.zeros((Bp - B, n_in), np.float32)])
            ys = np.empty((Bp, n_out), np.float32)
            self._check(self._L.ofx_mlp_forward(self._h, Bp, _fp(np.ascontiguousarray(xs)),
                                                _fp(ys)))
            out[s:s + B] = ys[:B]
        return out

    def gradients(self, x, t):
        """(loss, gradient in the public parameter layout); no update."""
        x, t, B = self._batch(x, t)
        loss = ctypes.c_float()
        g = np.empty(self.n_params, np.float32)
        self._check(self._L.ofx_mlp_gradients(self._h, B, _fp(x), _fp(t), ctypes.byref(loss),
                                              _fp(g)))
        return float(loss.value), g

    def step(self, x, t, lr):
        """One Adam step on (x, t); returns the loss before the update."""
        x, t, B = self._batch(x, t)
        loss = ctypes.c_float()
        self._check(self._L.ofx_mlp_step(self._h, B, _fp(x), _fp(t), ctypes.c_float(lr),
                                         ctypes.byref(loss)))
        return float(loss.value)

    def upload_dataset(self, X, T):
        n_in, _, n_out = self.dims
        X = _f32(X)
        T = _f32(T)
        if X.ndim != 2 or X.shape[1] != n_in or T.shape != (X.shape[0], n_out):
            raise ValueError("dataset shapes do not match the network")
        self._check(self._L.ofx_mlp_upload_dataset(self._h, X.shape[0], _fp(X), _fp(T)))
        self.dataset_size = X.shape[0]

    def train(self, steps, batch=512, lr=1e-3, seed=0, epochs_order=True):
        """`steps` Adam steps on minibatches of the uploaded dataset.

        Minibatches walk seeded permutations of the dataset (a fresh one per
        epoch).  Returns (losses [steps], device milliseconds)."""
        n = self.dataset_size
        rng = np.random.default_rng(seed)
        idx = []
        need = steps * batch
        while sum(len(a) for a in idx) < need:
            idx.append(rng.permutation(n) if epochs_order else rng.integers(0, n, n))
        idx = np.ascontiguousarray(np.concatenate(idx)[:need], dtype=np.int32)
        losses = np.empty(steps, np.float32)
        self._check(self._L.ofx_mlp_train(self._h, int(batch), int(steps), idx.ctypes.data_as(_IP),
                                          ctypes.c_float(lr), _fp(losses)))
        return losses, float(self._L.ofx_mlp_last_ms(self._h))

    def last_ms(self):
        return float(self._L.ofx_mlp_last_ms(self._h))


@atexit.register
def _close_live():
    for m in list(_LIVE):
        m.close()


# ---------------------------------------------------------------------------
# Spectral features: 1024 bins of amplitude and unwrapped phase
# ---------------------------------------------------------------------------
class SpectralBins:
    """A decimated spectral window of a SimWindow grid.

    Bin j is the grid frequency with SimWindow (centred) index
    k0 + D*j, j = 0..nbins-1, so the bins span nbins*D grid points.  D is a
    power of two: sampling a spectrum every D points is exact for a field
    confined to |t| < T/(2D) (T = window), and `to_grid` inverts it by
    band-limited interpolation (periodised time field, one period kept).

    Features of a spectrum X (FFT order, as the solver holds it):
      amplitude  sqrt(|X_j| / a_ref)    (compressed: 4th root of power)
      phase      unwrap(angle(X_j)) / pi (MATLAB unwrap along the bins,
                 as OpticalPulse.m:318-322 computes SpectralPhase)
    """

    def __init__(self, simwin, f_lo_hz, nbins=1024, D=None, a_ref=1.0):
        n = simwin.n
        if D is None:
            D = max(1, n // (4 * nbins))        # 16 ps window, 2^14 grid: D = 4
        if D & (D - 1) or nbins * D > n:
            raise ValueError("D must be a power of two with nbins*D <= n")
        f = simwin.rel_frequencies + C0_HZ / simwin.reference_wave
        k0 = int(np.searchsorted(f, f_lo_hz))
        k0 = min(max(k0, 0), n - nbins * D)
        self.n, self.nbins, self.D, self.k0 = n, int(nbins), int(D), k0
        self.kc = k0 + D * np.arange(nbins)                  # centred indices
        self.kf = (self.kc - n // 2) % n                     # FFT-order indices
        self.freqs = f[self.kc]
        self.a_ref = float(a_ref)
        self.window = simwin.temporal_range

    def bins(self, Xk):
        """FFT-order spectra [..., n] -> complex bins [..., nbins]."""
        return np.asarray(Xk)[..., self.kf]

    def features(self, Xb):
        """complex bins -> (amplitude, phase) feature arrays."""
        amp = np.sqrt(np.abs(Xb) / self.a_ref)
        ph = np.unwrap(np.angle(Xb), axis=-1) / np.pi
        return amp, ph

    def from_features(self, amp, ph):
        return (self.a_ref * np.asarray(amp, np.float64) ** 2) * np.exp(1j * np.pi * np.asarray(ph))

    def to_grid(self, Xb):
        """complex bins -> FFT-order spectrum on the full grid (zero outside
        the window band), band-limited: the time field of the bins is periodic
        with period T/D; one period centred on t = 0 is kept."""
        n, D = self.n, self.D
        Xc = np.zeros(np.shape(Xb)[:-1] + (n,), np.complex128)
        Xc[..., self.kc] = Xb
        Xf = np.fft.ifftshift(Xc, axes=-1)                   # FFT order
        e = np.fft.ifft(Xf, axis=-1)                         # time, FFT order (t = 0 first)
        m = np.fft.fftfreq(n) * n                            # sample offsets from t = 0
        keep = np.abs(m) < n / (2 * D)
        edge = np.abs(m) == n / (2 * D)
        w = keep * 1.0 + edge * 0.5
        return np.fft.fft(e * (D * w), axis=-1)


C0_HZ = 299792458.0
