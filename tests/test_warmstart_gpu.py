"""Surrogate warm start on the device: what the trained network predicts
better than linear propagation alone, and the gate's fallback."""
import math

import numpy as np
import pytest

from optifax_amd import physics as ph
from optifax_amd import warmstart as ws
from optifax_amd.solver import OPOBatch

pytestmark = pytest.mark.gpu


def _cfgs(n_inst, seed, detuning=(50e-15, 110e-15)):
    rng = np.random.default_rng(seed)
    return [ph.OPOConfig(n=2 ** 13, window=16e-12, length=8e-3,
                         period=float(rng.uniform(29.5e-6, 30.5e-6)),
                         pump_energy=float(rng.uniform(3.5e-9, 4.5e-9)), pump_fwhm=1e-12,
                         area=math.pi * (50e-6) ** 2, delay=float(rng.uniform(*detuning)),
                         delay_mode="sync", error_bounds=(0.0, 1e30), noise_seed=100 + i)
            for i in range(n_inst)]


def _rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def test_trained_surrogate_beats_linear_propagation(gpu_ctx):
    """Train on pairs from device trajectories of 12 instances; on pairs held
    out at random from the same trajectories, the one-pass prediction must be
    closer to the solver than the crystal's linear propagation, the
    prediction that needs no network.  (With this little data the network
    does not generalise to other instances; that needs the full data set.)"""
    X, F, pe, pr, lin, sb = ws.generate_pairs(_cfgs(12, 1), 40, log=None, ctx=gpu_ctx)
    rng = np.random.default_rng(0)
    perm = rng.permutation(X.shape[0])
    nv = 128
    vi, ti = perm[:nv], perm[nv:]
    x, y, a_ref = ws.make_dataset(X[ti], F[ti], pe[ti], pr[ti], lin[ti])
    net, losses, val, _ = ws.train_surrogate(x, y, 3000, batch=64, width=512, log=None)
    sur = ws.BandSurrogate(net, a_ref)
    e_net = _rel(sur(X[vi], pe[vi], pr[vi], lin[vi]), F[vi])
    e_lin = _rel(lin[vi] * X[vi], F[vi])
    print(f"held-out one-pass error: network median {np.median(e_net):.3f}, "
          f"linear propagation median {np.median(e_lin):.3f}; training loss "
          f"{np.mean(losses[-50:]):.3f}, validation {val:.3f}")
    assert np.all(np.isfinite(losses))
    assert np.median(e_net) < np.median(e_lin)
    net.close()


def test_gate_falls_back_to_the_plain_solve(gpu_ctx):
    """An untrained network is rejected by the gate: every instance restarts
    from its own noise field and then follows the plain solve trip for trip."""
    from optifax_amd.surrogate import SurrogateMLP
    cfgs = _cfgs(2, 3)
    pairs = [ws.opo_instance(c) for c in cfgs]
    insts = [p[0] for p in pairs]
    builds = [p[1] for p in pairs]
    net = SurrogateMLP(ws.N_IN, 256, ws.N_OUT, max_batch=64, seed=7)
    sur = ws.BandSurrogate(net, 1.0)
    warm = ws.solve_warm(cfgs, insts, builds, sur, conv_tol=1e-4, budget=6, chunk=5, log=None,
                         ctx=gpu_ctx)
    assert warm["fallback"].all()
    plain = OPOBatch(insts, trip_cap=8, ctx=gpu_ctx)
    plain.launch(5, 1e-4)
    plain.sync()
    for i in range(2):
        # the warm batch ran 1 gated trip, then 5 from the noise field
        assert np.array_equal(warm["fields"][i], plain.read(i)["field"])
    plain.close()
    net.close()
