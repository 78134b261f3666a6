"""Surrogate MLP: oracle pins (CPU) and device kernels vs the oracle (GPU)."""
import numpy as np
import pytest

from oracle import mlp as om
from optifax_amd import surrogate as sg


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _problem(dims, batch, seed, scale=1.0):
    n_in, width, n_out = dims
    rng = np.random.default_rng(seed)
    p = sg.init_params(n_in, width, n_out, seed=seed).astype(np.float64)
    # non-zero biases so their gradients and the bias path are exercised
    for W, b in om.unpack(p, *dims):
        b[:] = 0.1 * rng.standard_normal(b.size)
    x = rng.standard_normal((batch, n_in)) * scale
    t = rng.standard_normal((batch, n_out))
    return p, x, t


# ---------------------------------------------------------------- oracle pins
def test_bf16_round_bit_patterns():
    e = 2.0 ** -8
    vals = np.array([1.0, 1 + e, 1 + 3 * e, 1 + e + 2.0 ** -20, -(1 + 3 * e), 3.0e38, 0.0])
    want = np.array([1.0, 1.0, 1 + 4 * e, 1 + 2 * e, -(1 + 4 * e), 2.9980560e38, 0.0])
    got = om.bf16_round(vals)
    assert np.allclose(got[:5], want[:5], rtol=0, atol=0)
    assert got[-1] == 0.0
    # 3e38 rounds to a representable bf16 value: 8 significant bits
    assert np.float32(got[5]).view(np.uint32) & 0xFFFF == 0


def test_gelu_tabulated_and_derivative():
    assert om.gelu(np.array(1.0)) == pytest.approx(0.8413447460685429, abs=1e-15)
    assert om.gelu(np.array(-1.0)) == pytest.approx(-0.15865525393145707, abs=1e-15)
    assert om.gelu(np.array(0.0)) == 0.0
    z = np.linspace(-6, 6, 97)
    h = 1e-6
    fd = (om.gelu(z + h) - om.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - om.gelu_d(z))) < 1e-8


def test_loss_gradient_closed_form():
    rng = np.random.default_rng(3)
    y = rng.standard_normal((4, 7))
    t = rng.standard_normal((4, 7))
    loss, dy = om.rel_l2_loss(y, t)
    want = np.mean(np.linalg.norm(y - t, axis=1) / np.linalg.norm(t, axis=1))
    assert loss == pytest.approx(want, rel=1e-14)
    h = 1e-6
    for (r, c) in [(0, 0), (1, 3), (3, 6), (2, 2)]:
        yp = y.copy(); yp[r, c] += h
        ym = y.copy(); ym[r, c] -= h
        fd = (om.rel_l2_loss(yp, t)[0] - om.rel_l2_loss(ym, t)[0]) / (2 * h)
        assert dy[r, c] == pytest.approx(fd, rel=1e-6, abs=1e-10)


def test_oracle_gradients_match_finite_differences():
    dims = (5, 8, 6)
    p, x, t = _problem(dims, 3, seed=11)
    loss, g, _ = om.loss_and_grads(p, x, t, dims)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd[i] = (om.loss_and_grads(pp, x, t, dims)[0] - om.loss_and_grads(pm, x, t, dims)[0]) / (2 * h)
    assert _rel(g, fd) < 1e-7
    # every layer's block individually (a transposed or dropped block fails here)
    off = 0
    for o, i in om.layer_shapes(*dims):
        for n in (o * i, o):
            assert _rel(g[off:off + n], fd[off:off + n]) < 1e-6
            off += n


def test_adam_by_hand_two_steps():
    p = np.array([1.0, -2.0, 0.5])
    g1 = np.array([0.1, -0.2, 0.0])
    g2 = np.array([0.3, 0.1, -0.4])
    lr = 0.01
    p1, m, v = om.adam(p, g1, np.zeros(3), np.zeros(3), 1, lr)
    # step 1: m = 0.1 g, v = 0.001 g^2, mh = g, vh = g^2 -> p - lr * g / (|g| + eps)
    want1 = p - lr * g1 / (np.abs(g1) + 1e-8)
    assert np.allclose(p1, want1, rtol=0, atol=1e-15)
    p2, m2, v2 = om.adam(p1, g2, m, v, 2, lr)
    m_ = 0.9 * 0.1 * g1 + 0.1 * g2
    v_ = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    want2 = p1 - lr * (m_ / (1 - 0.81)) / (np.sqrt(v_ / (1 - 0.999 ** 2)) + 1e-8)
    assert np.allclose(p2, want2, rtol=0, atol=1e-15)


def test_bf16_emulation_stays_close_to_fp64():
    dims = (40, 64, 64)
    p, x, t = _problem(dims, 16, seed=5)
    l64, g64, y64 = om.loss_and_grads(p, x, t, dims)
    l16, g16, y16 = om.loss_and_grads(p, x, t, dims, bf16=True)
    assert _rel(y16, y64) < 2e-2
    assert _rel(g16, g64) < 5e-2
    assert abs(l16 - l64) / l64 < 1e-2


def test_product_param_layout_matches_oracle():
    dims = (70, 128, 64)
    assert sg.num_params(*dims) == om.num_params(*dims)
    assert [s for s in sg.layer_shapes(*dims)] == om.layer_shapes(*dims)


# ---------------------------------------------------------------- device
SMALL = (100, 128, 128)       # n_in not a multiple of 64: exercises the padded input path


def _per_layer(g, gref, dims):
    out, off = [], 0
    for o, i in om.layer_shapes(*dims):
        for n in (o * i, o):
            out.append(_rel(g[off:off + n], gref[off:off + n]))
            off += n
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("batch", [64, 128])
def test_device_forward_matches_oracle(batch):
    p, x, t = _problem(SMALL, batch, seed=21)
    net = sg.SurrogateMLP(*SMALL, max_batch=128, params=p.astype(np.float32))
    y = net.forward(x)
    p32 = p.astype(np.float32).astype(np.float64)
    y16, _ = om.forward(p32, x, SMALL, bf16=True)
    y64, _ = om.forward(p32, x, SMALL)
    print(f"forward batch {batch}: rel vs bf16 oracle {_rel(y, y16):.2e}, vs fp64 {_rel(y, y64):.2e}")
    assert _rel(y, y16) < 3e-3
    assert _rel(y, y64) < 3e-2
    net.close()


@pytest.mark.gpu
def test_device_gradients_match_oracle():
    batch = 128
    p, x, t = _problem(SMALL, batch, seed=22)
    net = sg.SurrogateMLP(*SMALL, max_batch=128, params=p.astype(np.float32))
    loss, g = net.gradients(x, t)
    p32 = p.astype(np.float32).astype(np.float64)
    l16, g16, _ = om.loss_and_grads(p32, x, t, SMALL, bf16=True)
    l64, g64, _ = om.loss_and_grads(p32, x, t, SMALL)
    errs = _per_layer(g, g16, SMALL)
    print("per-tensor rel err vs bf16 oracle:", " ".join(f"{e:.1e}" for e in errs))
    print(f"loss {loss:.6f} oracle {l16:.6f} (fp64 {l64:.6f}); all grads vs fp64 {_rel(g, g64):.2e}")
    assert abs(loss - l16) / l16 < 1e-4
    assert max(errs) < 1e-2
    assert _rel(g, g16) < 5e-3
    assert _rel(g, g64) < 5e-2
    # same call twice: deterministic
    loss2, g2 = net.gradients(x, t)
    assert loss2 == loss and np.array_equal(g, g2)
    net.close()


@pytest.mark.gpu
def test_device_adam_step_matches_oracle():
    batch = 64
    p, x, t = _problem(SMALL, batch, seed=23)
    p0 = p.astype(np.float32)
    net = sg.SurrogateMLP(*SMALL, max_batch=128, params=p0)
    _, g = net.gradients(x, t)
    lr = 1e-3
    net.step(x, t, lr)
    p1 = net.get_params().astype(np.float64)
    want, m, v = om.adam(p0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, lr)
    assert _rel(p1 - p0, want - p0) < 1e-4
    # second step with the same batch: moments carried over
    _, g2 = net.gradients(x, t)
    net.step(x, t, lr)
    p2 = net.get_params().astype(np.float64)
    want2, _, _ = om.adam(p1, g2.astype(np.float64), m, v, 2, lr)
    assert _rel(p2 - p1, want2 - p1) < 1e-3
    net.close()


@pytest.mark.gpu
def test_device_training_reduces_loss():
    # a smooth nonlinear map the network can fit
    dims = (100, 128, 128)
    rng = np.random.default_rng(7)
    A = rng.standard_normal((dims[0], dims[2])) / np.sqrt(dims[0])
    X = rng.standard_normal((2048, dims[0])).astype(np.float32)
    T = (np.tanh(X @ A) + 0.5).astype(np.float32)
    net = sg.SurrogateMLP(*dims, max_batch=256, seed=1)
    net.upload_dataset(X, T)
    losses, ms = net.train(200, batch=256, lr=1e-3, seed=2)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f} in {ms:.1f} ms")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]
    assert np.mean(losses[-20:]) < np.mean(losses[:20])
    net.close()


@pytest.mark.gpu
def test_device_full_size_gradients():
    """The surrogate's own shape: 2050 -> 2048 x 3 -> 2048, batch 512."""
    dims = (2050, 2048, 2048)
    batch = 512
    p, x, t = _problem(dims, batch, seed=24)
    net = sg.SurrogateMLP(*dims, max_batch=batch, params=p.astype(np.float32))
    loss, g = net.gradients(x, t)
    p32 = p.astype(np.float32).astype(np.float64)
    l16, g16, _ = om.loss_and_grads(p32, x, t, dims, bf16=True)
    errs = _per_layer(g, g16, dims)
    print("full size per-tensor rel err:", " ".join(f"{e:.1e}" for e in errs))
    assert abs(loss - l16) / l16 < 1e-4
    assert _rel(g, g16) < 5e-3
    net.close()
