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
        # the field keeps the last accepted step: finite, ~1e142 V/m after
        # the runaway growth, which amplifies the last-digit differences of
        # the two FP64 computations (measured 1.5e-4 relative L2)
        assert np.all(np.isfinite(r["field"])) and rel_l2(r["field"], Eo) < 1e-3


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


def _xtal_pass_case(xt, n=4096, window=8e-12, seed=3):
    from optifax_amd import physics as ph
    sw = ph.SimWindow(1040e-9, n, window)
    si = xt.solver_inputs(sw)
    E = np.fft.fftshift(ph.sech_pump(sw, 200e-15, 2e-10, 20e-12) + ph.quantum_noise(sw, 20e-12, seed))
    args = (si["alpha_w"], si["G33"], si["w0"], si["bdiffw0"], si["h"], si["nSteps"], si["dt"],
            si["beta_op"], 5e-2, 2.5e-3, float(np.ceil(si["nSteps"] / 4)))
    return si, E, args


def _crystals():
    from optifax_amd import physics as ph
    return {
        "KTP, chirped 40->36 um": ph.NonlinearCrystal(
            lambda z: ph.chirped_grating(z, 40e-6, 36e-6, 1.0, 0.1e-6), 0.4e-3,
            chi2=np.pi * 14e-12, material="KTP"),
        "OP-GaP": ph.NonlinearCrystal(60e-6, 0.3e-3, chi2=np.pi * 14e-12, material="OP-GaP"),
        "PPLN fan-out 29-31 um at y=0.5": ph.NonlinearCrystal(
            [29e-6, 31e-6], 0.4e-3, chi2=np.pi * 14e-12, vertical_position=0.5),
        "PPLN stepped, grating 3": ph.NonlinearCrystal(
            [28e-6, 29e-6, 30.5e-6, 32e-6], 0.4e-3, chi2=np.pi * 14e-12, vertical_position=2.5),
        "PPLN 30 um, 1 um domain uncertainty": ph.NonlinearCrystal(
            30e-6, 0.4e-3, chi2=np.pi * 14e-12, uncertainty=1e-6, seed=7),
    }


@pytest.mark.parametrize("name", list(_crystals()))
def test_pass_other_crystals_fp64_and_fp32(gpu_ctx, name):
    """Solver inputs of the other materials and grating forms (sellmeier.m,
    NonlinearCrystal.m:84-139, chirpedgrating.m, QPMcrystal.m) through the
    HIP kernel at 2^12 points: FP64 fields reproduce the oracle's step
    sequence and field; FP32 follows the same steps."""
    si, E, args = _xtal_pass_case(_crystals()[name])
    Eo, apo, smo, cnt = oracle.nee_heun(E, *args, n_cols=4)
    Eg, apg, smg = OPOmexBatch(E, *args, np.zeros((E.size, 4), complex), np.zeros(args[5]),
                               ctx=gpu_ctx, precision="double")
    assert np.array_equal(smg, smo)
    assert rel_l2(Eg, Eo) < 1e-9 and rel_l2(apg, apo) < 1e-9
    Ef, _, smf = OPOmexBatch(E, *args, np.zeros((E.size, 4), complex), np.zeros(args[5]),
                             ctx=gpu_ctx)
    print(f"{name}: {int(cnt[0])} attempts, FP32 rel L2 {rel_l2(Ef, Eo):.2e}")
    assert np.array_equal(smf, smo) and rel_l2(Ef, Eo) < 2e-4


def test_round_trips_with_an_optic_chain_fp64(gpu_ctx):
    """A cavity with an optic chain (Cavity.m:50-87): a 1 mm undoped LN
    window with a measured -150 fs^2 coating on both faces and 10 % loss
    before the coupler, a 99 % mirror after it, folded into the device
    operators, against the literal OpticalSim.run sequence with the same
    optics in the coupler masks; 2^13 points, 0.4 mm PPLN, three trips."""
    from optifax_amd import physics as ph
    cfg = ph.OPOConfig(n=2 ** 13, window=6e-12, length=0.4e-3, pump_energy=2e-10,
                       area=20e-12, delay=20e-15, noise_seed=5)
    b = cfg.build()
    si, sw = b["solver"], b["simwin"]
    lam = np.linspace(500.0, 2500.0, 81)
    gdd = (lam, np.full(lam.size, -150.0))
    t_w, _, phi_w = ph.optic_response(sw, "T", 0.9 ** 0.5, "PPLN_undoped", 1e-3, gdd=gdd)
    window = ("T", (t_w, None, np.nan_to_num(phi_w)))
    mirror = ("R", ph.optic_response(sw, "R", 0.01, "PPLN_undoped", 1e-3))
    T_b, phi_b = ph.cavity_response([window])
    T_a, phi_a = ph.cavity_response([mirror])
    ops = ph.cavity_operators(sw, si["n_xtal"], b["oc"], b["pump"], b["noise"],
                              chain_before_oc=(T_b, phi_b), chain_after_oc=(T_a, phi_a))
    R, T, phi = b["oc"].masks(sw)
    trips = 3
    # literal order: the optics before the coupler act on both coupler
    # outputs, the mirror after it on the fed-back field only (a mirror in
    # reflection adds no phase, so one coupler phase serves both outputs)
    assert not np.any(phi_a)
    o = oracle.opo_run(trips, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R * T_b * T_a,
                       T * T_b, phi + phi_b, cfg.delay, sw.rel_omegas, si)
    run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay)], precision="double", trip_cap=trips,
                   ctx=gpu_ctx)
    run.launch(trips)
    run.sync()
    r = run.read(0)
    run.close()
    assert np.array_equal(r["trip_log"][:, 2:].astype(int), o["steps"])
    assert rel_l2(r["field"], oracle_next_input(b, o["E_cav"])) < 1e-9
    assert rel_l2(r["oc_spec"], SimWindow.fft_order(SimWindow.t2k(o["E_oc"]))) < 1e-9


def test_config5_cavity_from_a_saturated_field_fp64(gpu_ctx):
    """Config 5's cavity at its grid (2^16 points, 16 ps), band coupler (5 %
    on 1.2-1.8 um), 4 nJ 200 fs pump and -50 fs detuning from synchronism,
    PPLN at the fan-out's local period of 30 um, adaptive Heun with the
    default error bounds, from a saturated start (pump plus a signal pulse at
    30 % of the pump's peak field): two FP64 round trips against the literal
    OpticalSim.run sequence of the oracle.  The crystal is shortened to
    0.5 mm so that the oracle finishes in seconds; the 8 mm crystal changes
    only the number of steps."""
    from optifax_amd import physics as ph
    from optifax_amd import warmstart as ws
    from cases import oracle_next_input
    cfg = ph.OPOConfig(n=2 ** 16, window=16e-12, length=0.5e-3, period=30e-6, pump_energy=4e-9,
                       delay=-50e-15, delay_mode="sync", noise_seed=1000)
    b = cfg.build()
    si, sw = b["solver"], b["simwin"]
    seed = np.fft.ifftshift(ws.smooth_seed(sw, cfg, 0.3))      # natural time order
    E0 = b["noise"] + seed
    R, T, phi = b["oc"].masks(sw)
    trips = 2
    o = oracle.opo_run(trips, E0, b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi, b["delay"],
                       sw.rel_omegas, si)
    ops = b["ops"]
    run = OPOBatch([dict(solver=si, field0=oracle_next_input(b, E0), pump=ops["pump"],
                         cav=ops["cav"], ocm=ops["ocm"], delay=b["delay"])],
                   precision="double", trip_cap=trips, ctx=gpu_ctx)
    run.launch(trips)
    run.sync()
    r = run.read(0)
    run.close()
    print(f"config 5 cavity, 0.5 mm: steps per trip {o['steps'].tolist()}, "
          f"field rel L2 {rel_l2(r['field'], oracle_next_input(b, o['E_cav'])):.1e}")
    assert np.array_equal(r["trip_log"][:, 2:].astype(int), o["steps"])
    assert rel_l2(r["field"], oracle_next_input(b, o["E_cav"])) < 1e-9
    assert rel_l2(r["oc_spec"], SimWindow.fft_order(SimWindow.t2k(o["E_oc"]))) < 1e-9
