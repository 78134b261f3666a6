"""The persistent pass kernel's body (optifax_amd/csrc/ofx_pass_impl.h),
compiled for the CPU (tests/emu/libofx_emu.so, one thread per workgroup), run
through the unchanged runtime layer and compared with the FP64 oracle.  This
checks the kernel's indexing, four-step FFT, halo handling, step control and
cavity stage on a machine without a GPU; the GPU tests repeat the same checks
on the device."""
import numpy as np
import pytest

import oracle
from optifax_amd import OfxError
from optifax_amd.solver import OPOBatch, OPOmexBatch

from cases import crystal_case, opo_case, oracle_next_input, oracle_oc_spectra, rel_l2


def _run(ctx, E, args, n, precision, C, apft_cols=4, exchange="auto"):
    ns = args[5]
    return OPOmexBatch(E, *args, np.zeros((n, apft_cols), complex), np.zeros(ns), ctx=ctx,
                       precision=precision, tune={"cluster": C, "exchange": exchange})


# both ways a workgroup learns that its cluster's exchange data are complete
EXCHANGES = ("tags", "barrier")


@pytest.mark.parametrize("exchange", EXCHANGES)
@pytest.mark.parametrize("n,C,window", [(256, 1, 2e-12), (256, 2, 2e-12), (256, 4, 2e-12),
                                        (1024, 2, 4e-12), (1024, 8, 4e-12)])
def test_pass_fp64_matches_oracle(emu_ctx, n, C, window, exchange):
    sw, si, E, args = crystal_case(n=n, window=window)
    Eo, apo, smo, cnt = oracle.nee_heun(E, *args, n_cols=4)
    Eg, apg, smg = _run(emu_ctx, E, args, n, "double", C, exchange=exchange)
    assert np.array_equal(smg, smo)                 # identical adaptive step sequence
    assert rel_l2(Eg, Eo) < 1e-10
    assert rel_l2(apg, apo) < 1e-10


def test_pass_fp32_tracks_oracle(emu_ctx):
    sw, si, E, args = crystal_case(n=512, window=3e-12)
    Eo, apo, smo, _ = oracle.nee_heun(E, *args, n_cols=4)
    Eg, apg, smg = _run(emu_ctx, E, args, 512, "single", 2)
    assert np.array_equal(smg, smo)
    # FP32 fields over ~750 high-gain steps (measured 4e-5; FP32 rounding per step)
    assert rel_l2(Eg, Eo) < 2e-4


@pytest.mark.parametrize("exchange", EXCHANGES)
def test_round_trips_match_literal_opticalsim(emu_ctx, exchange):
    cfg, b = opo_case()
    si, ops = b["solver"], b["ops"]
    sw = b["simwin"]
    trips = 3
    R, T, phi = b["oc"].masks(sw)
    o = oracle.opo_run(trips, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi,
                       cfg.delay, sw.rel_omegas, si)
    for C in (1, 2):
        run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                             ocm=ops["ocm"], delay=cfg.delay)], precision="double", trip_cap=trips,
                       ctx=emu_ctx, tune={"cluster": C, "exchange": exchange})
        run.launch(trips)
        run.sync()
        r = run.read(0)
        assert r["trips_done"] == trips
        assert np.array_equal(r["trip_log"][:, 2:].astype(int), o["steps"])
        assert rel_l2(r["field"], oracle_next_input(b, o["E_cav"])) < 1e-10
        from optifax_amd.physics import SimWindow
        assert rel_l2(r["oc_spec"], SimWindow.fft_order(SimWindow.t2k(o["E_oc"]))) < 1e-10
        assert np.allclose(r["trip_log"][1:, 0], o["conv"][1:], rtol=1e-9)
        assert np.allclose(r["trip_log"][:, 1], o["oc_energy"], rtol=1e-10)
        run.close()


def test_batch_invariance_and_per_instance_G(emu_ctx):
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    n, ns = 256, args[5]
    B = 3
    Eb = np.stack([E] * B, axis=1)
    Eg, apg, smg = OPOmexBatch(Eb, *args, np.zeros((n, 4, B), complex), np.zeros(ns), ctx=emu_ctx,
                               precision="double", tune={"cluster": 2})
    E1, _, _ = _run(emu_ctx, E, args, n, "double", 2)
    for b in range(B):
        assert np.array_equal(Eg[:, b], E1)
    # per-instance poling (G is NSteps x B, OPO_CUDA_BATCH_XTAL.cu:222): flip one sign pattern
    G = np.stack([args[1], -args[1], args[1]], axis=1)
    a2 = list(args)
    a2[1] = G
    Eg2, _, _ = OPOmexBatch(Eb, *a2, np.zeros((n, 4, B), complex), np.zeros(ns), ctx=emu_ctx,
                            precision="double")
    a3 = list(args)
    a3[1] = -args[1]
    Eo, _, _, _ = oracle.nee_heun(E, *a3, n_cols=4)
    assert rel_l2(Eg2[:, 1], Eo) < 1e-10
    assert np.array_equal(Eg2[:, 0], Eg2[:, 2])


def test_nonfinite_input_reports_status(emu_ctx):
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    E2 = np.stack([E, E], axis=1)
    E2[5, 1] = np.nan
    Eg, _, _, st = OPOmexBatch(E2, *args, np.zeros((256, 4, 2), complex), np.zeros(args[5]),
                               ctx=emu_ctx, precision="double", return_status=True)
    assert list(st) == [1, 3]             # DONE, NONFINITE
    assert np.all(np.isfinite(Eg[:, 0]))


def test_forced_rejection_terminates(emu_ctx):
    # maxErr = 0: every attempt above stepmod 1 is rejected; the stepmod > 1
    # guard (NEE_CPU.m:68) makes the pass finish with unit steps.
    sw, si, E, args = crystal_case(n=256, window=2e-12, L=0.02e-3)
    a = list(args)
    a[8], a[9] = 0.0, -1.0
    Eo, _, smo, cnt = oracle.nee_heun(E, *a, n_cols=4)
    Eg, _, smg = _run(emu_ctx, E, a, 256, "double", 2)
    assert np.array_equal(smg, smo) and set(np.unique(smo[smo > 0])) == {1.0}
    assert rel_l2(Eg, Eo) < 1e-10


def test_empty_batch_is_a_no_op(emu_ctx):
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    sm_in = np.arange(args[5], dtype=np.float32)
    Eg, apg, smg, st = OPOmexBatch(np.zeros((256, 0), complex), *args, None, sm_in, ctx=emu_ctx,
                                   return_status=True)
    assert Eg.shape == (256, 0) and apg is None and st.size == 0
    assert np.array_equal(smg, sm_in)


def test_argument_errors(emu_ctx):
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    with pytest.raises(ValueError):
        OPOmexBatch(E, args[0], args[1][:-1], *args[2:], None, None, ctx=emu_ctx)
    a = list(args)
    a[10] = 1.0                      # sel must exceed 1
    with pytest.raises(OfxError):
        OPOmexBatch(E, *a, None, None, ctx=emu_ctx)
    with pytest.raises(OfxError):
        OPOmexBatch(E[:100], args[0][:100], args[1], *args[2:7], args[7][:100], *args[8:], None,
                    None, ctx=emu_ctx)      # N not a power of two


@pytest.mark.parametrize("exchange", EXCHANGES)
@pytest.mark.parametrize("n,C,window", [(256, 1, 2e-12), (256, 2, 2e-12), (1024, 4, 4e-12),
                                        (1024, 8, 4e-12)])
def test_rk23_pass_fp64_matches_oracle(emu_ctx, n, C, window, exchange):
    """RK23 (ode23 variant): same accept/reject sequence and FP64 fields as orc_rk23,
    including the K4 boundary exchange between workgroups (C > 1)."""
    sw, si, E, args = crystal_case(n=n, window=window)
    Eo, apo, smo, cnt = oracle.rk23(E, *args, n_cols=4)
    assert cnt[0] > cnt[1]                          # the case exercises rejections
    Eg, apg, smg = OPOmexBatch(E, *args, np.zeros((n, 4), complex), np.zeros(args[5]),
                               ctx=emu_ctx, precision="double", tune={"cluster": C, "exchange": exchange},
                               method="rk23")
    assert np.array_equal(smg, smo)
    assert rel_l2(Eg, Eo) < 1e-10
    assert rel_l2(apg, apo) < 1e-10


def test_rk23_round_trips_match_literal_opticalsim(emu_ctx):
    cfg, b = opo_case()
    si, ops, sw = b["solver"], b["ops"], b["simwin"]
    trips = 2
    R, T, phi = b["oc"].masks(sw)
    o = oracle.opo_run(trips, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi,
                       cfg.delay, sw.rel_omegas, si, method=1)
    run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay, method="rk23")], precision="double",
                   trip_cap=trips, ctx=emu_ctx, tune={"cluster": 2})
    run.keep_history(trips)
    run.launch(trips)
    run.sync()
    r = run.read(0)
    assert np.array_equal(r["trip_log"][:, 2:].astype(int), o["steps"])
    assert rel_l2(r["field"], oracle_next_input(b, o["E_cav"])) < 1e-10
    # the first trip's row of step multipliers is the oracle RK23 pass's
    _, _, sm_o, _ = oracle.rk23(ops["field0"], si["alpha_w"], si["G33"], si["w0"], si["bdiffw0"],
                                si["h"], si["nSteps"], si["dt"], si["beta_op"], si["maxErr"],
                                si["minErr"], si["sel"], n_cols=0, want_stepmods=True)
    assert np.array_equal(run.read_history(0, 0)[1], sm_o)
    run.close()


def test_rk23_mixed_with_heun_in_one_launch(emu_ctx):
    """Heun and RK23 instances share one launch (separate clusters, own halos)."""
    sw, si, E, args = crystal_case(n=256, window=2e-12)
    inst = lambda m: dict(solver=dict(si, maxErr=args[8], minErr=args[9], sel=args[10]),
                          field0=E, method=m)
    run = OPOBatch([inst("heun"), inst("rk23"), inst("heun")], precision="double", ctx=emu_ctx,
                   tune={"cluster": 2})
    run.launch(0)
    run.sync()
    Eh, _, _, _ = oracle.nee_heun(E, *args)
    Er, _, _, _ = oracle.rk23(E, *args)
    assert rel_l2(run.read(0)["field"], Eh) < 1e-10
    assert rel_l2(run.read(1)["field"], Er) < 1e-10
    assert rel_l2(run.read(2)["field"], Eh) < 1e-10
    run.close()


def test_converged_instances_are_skipped_until_rewritten(emu_ctx):
    cfg, b = opo_case()
    si, ops = b["solver"], b["ops"]
    run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay)], precision="double", trip_cap=8,
                   ctx=emu_ctx, tune={"cluster": 2})
    run.launch(4, 1e30)                 # any finite change counts as converged (trip 2)
    run.sync()
    r = run.read(0)
    assert r["status"] == 2 and r["trips_done"] == 2
    run.launch(3, 1e30)                 # converged earlier: skipped
    run.sync()
    r2 = run.read(0)
    assert r2["trips_done"] == 2 and np.array_equal(r2["field"], r["field"])
    run.write_field(0, r["field"])      # rewritten: runs again, first trip has no previous
    run.launch(3, 1e30)
    run.sync()
    r3 = run.read(0)
    assert r3["trips_done"] == 4 and np.isinf(r3["trip_log"][2, 0])
    run.close()


@pytest.mark.parametrize("exchange", EXCHANGES)
def test_wait_limit_gives_up_and_the_next_launch_recovers(emu_ctx, exchange):
    """A zero wait limit makes the cluster give up at its first wait (every
    member leaves through its own bounded poll); the field stays as it was,
    and after the limit is restored the same launch reproduces a fresh run
    exactly (the tags of the abandoned launch are never mistaken for new
    data)."""
    cfg, b = opo_case()
    si, ops = b["solver"], b["ops"]
    inst = dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                ocm=ops["ocm"], delay=cfg.delay)
    tune = {"cluster": 4, "exchange": exchange}
    ref = OPOBatch([inst], precision="double", trip_cap=4, ctx=emu_ctx, tune=tune)
    ref.launch(2)
    ref.sync()
    want = ref.read(0)
    ref.close()
    run = OPOBatch([inst], precision="double", trip_cap=8, ctx=emu_ctx, tune=tune)
    f0 = run.read(0)["field"]
    run.set_wait_limit(0.0)
    with pytest.raises(OfxError, match="timed out"):
        run.launch(2)
        run.sync()
    assert np.array_equal(run.read(0)["field"], f0)
    run.set_wait_limit(10.0)
    run.write_field(0, f0)
    run.launch(2)
    run.sync()
    got = run.read(0)
    assert np.array_equal(got["field"], want["field"])
    assert np.array_equal(got["oc_spec"], want["oc_spec"])
    with pytest.raises(OfxError):
        run.set_wait_limit(-1.0)
    run.close()


@pytest.mark.parametrize("exchange", EXCHANGES)
def test_run_averages_and_trip_history(emu_ctx, exchange):
    """Device-side run averages (ESDOutAverage and the crystal-output ESD
    behind ESDPumpDepAverage / ESDIdlerICAverage, OpticalSim.m:302-313) and
    the trip history (StoredPulses, OpticalSim.m:268-272, and the rows of
    StepSizeModifiers, OpticalSim.m:172,237) against the oracle: the held
    crystal output times the output-coupling transfer must give the oracle's
    output spectrum of that trip, and the held step multipliers the oracle
    pass's on that trip's crystal input."""
    cfg, b = opo_case()
    si, ops = b["solver"], b["ops"]
    oc = oracle_oc_spectra(b, cfg.delay, 6)
    run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay)], precision="double", trip_cap=8,
                   ctx=emu_ctx, tune={"cluster": 2, "exchange": exchange})
    e, x, k = run.read_average(0)
    assert k == 0 and not e.any() and not x.any()
    run.keep_history(3)
    run.launch(1)
    run.launch(2)
    run.sync()
    e, x, k = run.read_average(0)
    assert k == 3 and rel_l2(e, np.mean([abs(s) ** 2 for s in oc[:3]], axis=0)) < 1e-10
    R, T, phi = b["oc"].masks(b["simwin"])
    sw = b["simwin"]
    inputs = [ops["field0"]] + [
        oracle_next_input(b, oracle.opo_run(t, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"],
                                            R, T, phi, cfg.delay, sw.rel_omegas, si)["E_cav"])
        for t in (1, 2)]
    for t in range(3):
        h, sm = run.read_history(0, t)
        assert rel_l2(h * ops["ocm"], oc[t]) < 1e-10
        _, _, sm_o, _ = oracle.nee_heun(inputs[t], si["alpha_w"], si["G33"], si["w0"],
                                        si["bdiffw0"], si["h"], si["nSteps"], si["dt"],
                                        si["beta_op"], si["maxErr"], si["minErr"], si["sel"],
                                        n_cols=0, want_stepmods=True)
        assert np.array_equal(sm, sm_o)
    with pytest.raises(OfxError):
        run.read_history(0, 3)                 # not run yet
    with pytest.raises(OfxError):
        run.keep_history(2)                    # once per run
    run.reset_average(0)
    run.launch(3)                              # trips 3..5; the ring of 3 now holds 3..5
    run.sync()
    e, x, k = run.read_average(0)
    assert k == 3 and rel_l2(e, np.mean([abs(s) ** 2 for s in oc[3:6]], axis=0)) < 1e-10
    hs = [run.read_history(0, t)[0] for t in (3, 4, 5)]
    for t, h in zip((3, 4, 5), hs):
        assert rel_l2(h * ops["ocm"], oc[t]) < 1e-10
    assert rel_l2(x, np.mean([abs(h) ** 2 for h in hs], axis=0)) < 1e-12
    with pytest.raises(OfxError):
        run.read_history(0, 2)                 # overwritten by trip 5
    run.close()


def test_round_trips_with_an_optic_chain_match_literal_opticalsim(emu_ctx):
    """A lossy, dispersive optic before the coupler (power transmission 0.9,
    the phase of 1 mm of dispersive bulk) folded into the device operators
    gives the trips of the literal sequence whose coupler masks carry that
    optic (R T_c, T T_c, phi + phi_c)."""
    from optifax_amd import physics as ph
    cfg, b = opo_case()
    si, sw = b["solver"], b["simwin"]
    T_c = np.full(sw.n, 0.9)
    phi_c = np.nan_to_num(ph.bulk_dispersion(sw, "LN_o", 1e-3))
    ops = ph.cavity_operators(sw, si["n_xtal"], b["oc"], b["pump"], b["noise"],
                              chain_before_oc=(T_c, phi_c))
    R, T, phi = b["oc"].masks(sw)
    trips = 3
    o = oracle.opo_run(trips, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R * T_c,
                       T * T_c, phi + phi_c, cfg.delay, sw.rel_omegas, si)
    run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay)], precision="double", trip_cap=trips,
                   ctx=emu_ctx, tune={"cluster": 2})
    run.launch(trips)
    run.sync()
    r = run.read(0)
    assert np.array_equal(r["trip_log"][:, 2:].astype(int), o["steps"])
    assert rel_l2(r["field"], oracle_next_input(b, o["E_cav"])) < 1e-10
    from optifax_amd.physics import SimWindow
    assert rel_l2(r["oc_spec"], SimWindow.fft_order(SimWindow.t2k(o["E_oc"]))) < 1e-10
    run.close()


def test_pass_with_ktp_and_a_chirped_grating_matches_oracle(emu_ctx):
    """Solver inputs of another material and a chirped grating (G(z) no
    longer periodic) through the emulated kernel: the FP64 pass reproduces
    the oracle's step sequence and field."""
    from optifax_amd import physics as ph
    sw = ph.SimWindow(1040e-9, 512, 3e-12)
    xt = ph.NonlinearCrystal(lambda z: ph.chirped_grating(z, 40e-6, 36e-6, 1.0, 0.1e-6), 0.3e-3,
                             chi2=np.pi * 14e-12, material="KTP")
    si = xt.solver_inputs(sw)
    assert len(np.unique(np.diff(np.nonzero(np.diff(si["G33"]))[0]))) > 1   # walls not periodic
    E = np.fft.fftshift(ph.sech_pump(sw, 200e-15, 2e-10, 20e-12) + ph.quantum_noise(sw, 20e-12, 3))
    args = (si["alpha_w"], si["G33"], si["w0"], si["bdiffw0"], si["h"], si["nSteps"], si["dt"],
            si["beta_op"], 5e-2, 2.5e-3, float(np.ceil(si["nSteps"] / 4)))
    Eo, apo, smo, cnt = oracle.nee_heun(E, *args, n_cols=4)
    Eg, apg, smg = _run(emu_ctx, E, args, 512, "double", 2)
    assert np.array_equal(smg, smo)
    assert rel_l2(Eg, Eo) < 1e-10 and rel_l2(apg, apo) < 1e-10


def test_plain_pass_is_not_a_round_trip(emu_ctx):
    """A plain crystal pass (launch(0)) on a cavity run counts as no trip:
    the trip count, the run averages and the history ring are left to the
    round trips, so a later read_average divides by the trips actually
    averaged and read_history holds no stale row for the pass."""
    cfg, b = opo_case()
    si, ops = b["solver"], b["ops"]
    run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                         ocm=ops["ocm"], delay=cfg.delay)], precision="double", trip_cap=4,
                   ctx=emu_ctx, tune={"cluster": 2})
    run.keep_history(4)
    run.launch(1)
    run.sync()
    e1, x1, k1 = run.read_average(0)
    h0, sm0 = run.read_history(0, 0)
    log1 = run.read(0)["trip_log"].copy()
    run.launch(0)
    run.sync()
    r = run.read(0)
    assert r["trips_done"] == 1 and k1 == 1
    e, x, k = run.read_average(0)
    assert k == 1 and np.array_equal(e, e1) and np.array_equal(x, x1)
    with pytest.raises(OfxError):
        run.read_history(0, 1)
    h, sm = run.read_history(0, 0)
    assert np.array_equal(h, h0) and np.array_equal(sm, sm0)
    assert np.array_equal(r["trip_log"], log1)
    run.launch(1)
    run.sync()
    assert run.read(0)["trips_done"] == 2
    assert run.read_average(0)[2] == 2
    h1, sm1 = run.read_history(0, 1)
    assert sm1.any() and np.isfinite(h1).all()
    run.close()


def test_repeated_pass_calls_reuse_the_run(emu_ctx):
    """OpticalSim.run calls the solver once per trip with the same shapes;
    the context keeps the device run between such calls.  Results of a
    sequence of calls (same shapes and new data, then another grid, then the
    first shapes again) equal those of a fresh context per call."""
    from optifax_amd.solver import Context
    import os
    ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cases = []
    for seed, n in ((1, 256), (2, 256), (3, 512), (4, 256)):
        sw, si, E, args = crystal_case(n=n, window=2e-12 * n / 256, seed=seed)
        G = args[1] * (1.0 + 0.1 * seed)      # new coupling per call
        cases.append((E, (args[0], G) + args[2:], n))
    got = []
    for E, args, n in cases:
        got.append(_run(emu_ctx, E, args, n, "double", 2))
    for (E, args, n), g in zip(cases, got):
        fresh = Context(lib_path=os.path.join(ROOT, "tests", "emu", "libofx_emu.so"))
        want = _run(fresh, E, args, n, "double", 2)
        fresh.close()
        for a, b in zip(g, want):
            assert np.array_equal(a, b)
