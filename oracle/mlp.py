"""FP64 restatement of the surrogate MLP — test infrastructure, not product.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this
module, as the checker of the device kernels in optifax_amd/csrc/ofx_mlp.hip.

The reference has no learned component, so nothing here follows a reference
file: the network is the one the surrogate row of the scope table names
(4 weight layers, GELU, relative-L2 loss, Adam), written out in numpy.

Network (weights [out][in]):
    z1 = x W0^T + b0, h1 = gelu(z1), ..., y = h3 W3^T + b3
    loss = mean_r ||y_r - t_r|| / ||t_r||
Public parameter vector: W0, b0, W1, b1, W2, b2, W3, b3 (row-major, FP32 on
the device).

``bf16=True`` rounds (round-to-nearest-even) at exactly the points where the
device keeps BF16: the network input, the weights, each GELU output and each
output-side gradient dL/dz that feeds a GEMM (dL/dy, dL/dz3, dL/dz2, dL/dz1);
the arithmetic between those points stays FP64.  That is the sharp
comparison for the device; ``bf16=False`` is the exact FP64 network.

Parity pins (tests/test_surrogate.py):
  * gradients vs central finite differences of the loss, FP64, on a small
    network (a dropped term, a wrong transpose or a missing GELU' fails);
  * dL/dy closed form vs finite differences of the loss alone;
  * GELU vs its tabulated values (gelu(1) = 0.8413447460685429,
    gelu(-1) = -0.15865525393145707) and GELU' vs finite differences;
  * Adam vs its defining recurrences evaluated by hand for two steps;
  * bf16_round vs hand-worked bit patterns (ties to even).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def bf16_round(x):
    """float -> bfloat16 (round to nearest, ties to even) -> float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = u + 0x7FFF + ((u >> 16) & 1)
    b = (r & 0xFFFF0000).astype(np.uint32)
    return b.view(np.float32).astype(np.float64).reshape(np.shape(x))


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def layer_shapes(n_in, width, n_out):
    return [(width, n_in), (width, width), (width, width), (n_out, width)]


def num_params(n_in, width, n_out):
    return sum(o * i + o for o, i in layer_shapes(n_in, width, n_out))


def unpack(params, n_in, width, n_out):
    """Views (W, b) per layer into the public parameter vector."""
    p = np.asarray(params)
    out, off = [], 0
    for o, i in layer_shapes(n_in, width, n_out):
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    assert off == p.size, "parameter vector size does not match the network"
    return out


def forward(params, x, dims, bf16=False):
    """y and the cache of (h_l, z_{l+1}) for the backward pass."""
    n_in, width, n_out = dims
    rnd = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float64), n_in, width, n_out)
    h = rnd(np.asarray(x, dtype=np.float64))
    hs, zs = [h], []
    for l, (W, b) in enumerate(layers):
        z = h @ rnd(W).T + b
        zs.append(z)
        if l < 3:
            h = rnd(gelu(z))
            hs.append(h)
    return zs[-1], (hs, zs, layers)


def rel_l2_loss(y, t):
    """mean over rows of ||y - t|| / ||t|| and dL/dy."""
    d = y - t
    ne = np.sqrt((d * d).sum(axis=1))
    nt = np.sqrt((t * t).sum(axis=1))
    nt = np.where(nt > 0, nt, 1.0)
    B = y.shape[0]
    loss = float(np.mean(ne / nt))
    sc = np.where(ne > 0, 1.0 / (B * np.where(ne > 0, ne, 1.0) * nt), 0.0)
    return loss, d * sc[:, None]


def loss_and_grads(params, x, t, dims, bf16=False):
    """(loss, gradient in the public layout, y)."""
    n_in, width, n_out = dims
    rnd = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    y, (hs, zs, layers) = forward(params, x, dims, bf16)
    loss, dy = rel_l2_loss(y, np.asarray(t, dtype=np.float64))
    dz = rnd(dy)
    grads = [None] * 8
    for l in range(3, -1, -1):
        W, _ = layers[l]
        grads[2 * l] = dz.T @ hs[l]
        grads[2 * l + 1] = dz.sum(axis=0)
        if l > 0:
            dh = dz @ rnd(W)
            dz = rnd(dh * gelu_d(zs[l - 1]))
    return loss, np.concatenate([g.ravel() for g in grads]), y


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
