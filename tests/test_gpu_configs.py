"""The HIP path on the workloads at their stated shapes, against the FP64
oracle where it finishes in seconds and through exact step bookkeeping and
status where it does not.

* The single pass as stated: 2^13 points, 10 ps, 1 mm MgO:PPLN at 30 um,
  1 nJ sech^2 in 5 um^2, seeded noise, P = 5.  The field grows without bound
  and the FP64 oracle itself turns non-finite; the device must take the same
  step sequence up to that point and report it at the same attempt.
* The uneven batch: 256 instances drawn from seed 3 (period 28-32 um, 1-10
  mm, 2^14-2^16 points, +-50 fs from synchronism) in one persistent launch for
  one round trip, FP64 fields; the shortest crystal of each grid size against
  the oracle's literal round trip.
* Repeated OPOmexBatch calls (one per round trip under OpticalSim.run) reuse
  the device run and give what fresh calls give.
"""
import concurrent.futures as cf
import os
import sys

import numpy as np
import pytest

import oracle
from optifax_amd.physics import SimWindow
from optifax_amd.solver import Context, OPOBatch, OPOmexBatch

from cases import crystal_case, oracle_next_input, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

pytestmark = pytest.mark.gpu


def _stated_single_pass(method):
    sw, si, E, args = crystal_case(n=8192, window=10e-12, L=1e-3, energy=1e-9, area=5e-12,
                                   period=30e-6)
    args = list(args)
    if method == "rk23":
        args[8], args[9] = 1e-4, 5e-6           # tolerance 1e-6 (maxErr, minErr in percent)
    return sw, si, E, args


@pytest.mark.parametrize("method", ["rk23", "heun"])
def test_stated_single_pass_goes_nonfinite_at_the_oracles_step(gpu_ctx, method):
    """FP64 fields: identical accepted-step sequence up to the blow-up and
    the NONFINITE status.  RK23 finds NaN in the attempt's stage field, in
    the oracle and on the device alike, so the attempt counts agree; Heun
    finds it at the start of the next attempt, which the device counts and
    the oracle (NEE_CPU.m:44-46 leaves before the attempt) does not."""
    sw, si, E, args = _stated_single_pass(method)
    solve = oracle.rk23 if method == "rk23" else oracle.nee_heun
    Eo, apo, smo, cnt = solve(E, *args, n_cols=4, allow_nonfinite=True)
    s = dict(si, maxErr=args[8], minErr=args[9], sel=args[10], n_cols=4)
    run = OPOBatch([dict(solver=s, field0=E, method=method)], precision="double", ctx=gpu_ctx)
    run.launch(0)
    run.sync(allow_nonfinite=True)
    r = run.read(0)
    run.close()
    print(f"{method}: oracle {int(cnt[0])} attempts / {int(cnt[1])} accepted, device "
          f"{r['attempts']} / {r['accepted']}, status {r['status']}")
    assert r["status"] == 3                                  # OFX_STATUS_NONFINITE
    assert np.array_equal(r["stepmods"], smo)
    assert r["accepted"] == int(cnt[1])
    assert r["attempts"] == int(cnt[0]) + (1 if method == "heun" else 0)
    if method == "rk23":
        # the field keeps the last accepted step (finite)
        assert np.all(np.isfinite(r["field"])) and rel_l2(r["field"], Eo) < 1e-6


_BATCH = {}


def _batch():
    if not _BATCH:
        import bench
        insts, meta = bench.batch_workload()
        _BATCH["v"] = (insts, meta)
    return _BATCH["v"]


def _oracle_trip(inst):
    """The oracle's literal round trip from the instance's noise start."""
    b = inst["_build"]
    si, sw = b["solver"], b["simwin"]
    R, T, phi = b["oc"].masks(sw)
    o = oracle.opo_run(1, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi,
                       b["delay"], sw.rel_omegas, si)
    return oracle_next_input(b, o["E_cav"]), SimWindow.fft_order(SimWindow.t2k(o["E_oc"])), \
        o["steps"][0]


def test_uneven_batch_of_256_one_trip_fp64(gpu_ctx):
    from optifax_amd import physics as ph
    insts, meta = _batch()
    run = OPOBatch(insts, precision="double", trip_cap=2, ctx=gpu_ctx)
    run.launch(1)
    run.sync()
    rs = [run.read(i) for i in range(len(insts))]
    run.close()
    assert all(r["status"] == 1 for r in rs)                 # OFX_STATUS_DONE
    assert all(r["trips_done"] == 1 for r in rs)
    assert all(np.all(np.isfinite(r["field"])) and np.all(np.isfinite(r["oc_spec"])) for r in rs)
    # the shortest crystal of each grid size against the oracle
    picks = []
    for n in (2 ** 14, 2 ** 15, 2 ** 16):
        picks.append(min((m[1], i) for i, m in enumerate(meta) if m[0] == n)[1])
    rng = np.random.default_rng(3)                           # the batch's own draw, replayed
    cfgs = []
    for i in range(len(insts)):
        lg = int(rng.integers(14, 17))
        cfg = ph.OPOConfig(n=2 ** lg, window=12e-12 * 2.0 ** (lg - 15), progress_plots=2,
                           period=float(rng.uniform(28e-6, 32e-6)),
                           length=float(rng.uniform(1e-3, 10e-3)),
                           delay=float(rng.uniform(-50e-15, 50e-15)), delay_mode="sync",
                           noise_seed=int(rng.integers(1 << 30)))
        cfgs.append(cfg)
    checks = [dict(insts[i], _build=cfgs[i].build()) for i in picks]
    for c, i in zip(checks, picks):
        assert np.array_equal(c["_build"]["ops"]["field0"], insts[i]["field0"])
    with cf.ThreadPoolExecutor(3) as ex:
        refs = list(ex.map(_oracle_trip, checks))
    for i, (X, oc, steps) in zip(picks, refs):
        r = rs[i]
        print(f"instance {i} (N={meta[i][0]}, {meta[i][1] * 1e3:.2f} mm): field rel L2 "
              f"{rel_l2(r['field'], X):.2e}, steps {r['trip_log'][0, 2:]} vs {steps}")
        assert np.array_equal(r["trip_log"][0, 2:].astype(int), steps)
        assert rel_l2(r["field"], X) < 1e-9
        assert rel_l2(r["oc_spec"], oc) < 1e-9


def test_repeated_mex_calls_reuse_the_run(gpu_ctx):
    """OpticalSim.run calls the solver once per trip with the same shapes:
    the context keeps the device run between such calls.  A sequence of
    calls (new data with the same shapes, another grid, the first shapes
    again) returns what a fresh context returns for each call, bit for bit."""
    cases = []
    for seed, n in ((1, 1024), (2, 1024), (3, 2048), (4, 1024)):
        sw, si, E, args = crystal_case(n=n, window=4e-12 * n / 1024, seed=seed)
        G = args[1] * (1.0 + 0.1 * seed)
        cases.append((E, (args[0], G) + tuple(args[2:]), n))
    got = [OPOmexBatch(E, *a, np.zeros((n, 4), complex), np.zeros(a[5]), ctx=gpu_ctx)
           for E, a, n in cases]
    for (E, a, n), g in zip(cases, got):
        fresh = Context(0)
        want = OPOmexBatch(E, *a, np.zeros((n, 4), complex), np.zeros(a[5]), ctx=fresh)
        fresh.close()
        for x, y in zip(g, want):
            assert np.array_equal(x, y)
