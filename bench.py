"""Benchmark: OPO round trips per second on one MI355X.

Workload (one "step" = one round trip): synchronously pumped ring OPO,
N = 2^15 points over a 12 ps window at 1040 nm, 5 mm MgO:PPLN with a
30.5 um grating, 200 fs sech^2 pump of 2 W at 1 GHz, 5 % output coupler,
0.3 ps cavity detuning, started from seeded one-photon-per-mode noise.
Adaptive Heun split-step with the reference's default error bounds
([2.5e-3, 5e-2] %, h = 0.1 um).  Synthetic inputs, resident in HBM before
the timed region; FP32 complex fields with FP64 phase arguments and norms.

Prints one JSON line.  `roofline` is for the dominant kernel (the persistent
pass kernel) measured with HIP events on its own stream; `cpu_baseline` is
the FP64 C restatement of the reference's CPU solver (oracle/, plotting and
pause removed) timed on a bounded sample on one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "OPO round trips/s; time-to-steady-state warm-start vs plain @1e-4 err"


def algorithmic_bytes(n, accepted, rejected, trips):
    """Per instance: 24 N B per accepted adaptive step (16 B field r/w + 8 B
    beta/alpha), 16 N B per rejected step, 40 N B per round trip (cavity)."""
    return 24.0 * n * accepted + 16.0 * n * rejected + 40.0 * n * trips


def workload():
    """The ring OPO: the output coupler is Optic("R", 0.05) over the whole band
    (intracavity sqrt(0.95), output sqrt(0.05)), the delay 0.3 ps per trip."""
    from optifax_amd import physics as ph
    cfg = ph.OPOConfig(oc_band=None)
    return cfg, cfg.build()


def measured_traffic(key):
    """HBM bytes per launch of the pass kernel from the committed rocprofv3
    --pmc summary of this same command (profiles/pmc_traffic.json:
    FETCH_SIZE + WRITE_SIZE averaged over the run's dispatches, each divided
    by what the counter reads for a device copy of known size); None without
    a summary."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)[key]
    except (OSError, ValueError, KeyError):
        return None
    return rec["traffic_bytes_per_launch"]


def batch_workload(n_inst=256, seed=3):
    """256 uneven OPO instances drawn from a seed: poling period 28-32 um,
    crystal length 1-10 mm, grid 2^14-2^16 at a fixed dt of 12 ps / 2^15
    (windows 6 / 12 / 24 ps), detuning +-50 fs around the synchronous delay,
    one progress chunk per pass (sel = NSteps)."""
    from optifax_amd import physics as ph
    rng = np.random.default_rng(seed)
    insts, meta = [], []
    for i in range(n_inst):
        lg = int(rng.integers(14, 17))
        cfg = ph.OPOConfig(n=2 ** lg, window=12e-12 * 2.0 ** (lg - 15), progress_plots=2,
                           period=float(rng.uniform(28e-6, 32e-6)),
                           length=float(rng.uniform(1e-3, 10e-3)),
                           delay=float(rng.uniform(-50e-15, 50e-15)), delay_mode="sync",
                           noise_seed=int(rng.integers(1 << 30)))
        b = cfg.build()
        ops = b["ops"]
        insts.append(dict(solver=b["solver"], field0=ops["field0"], pump=ops["pump"],
                          cav=ops["cav"], ocm=ops["ocm"], delay=b["delay"]))
        meta.append((2 ** lg, cfg.length))
    return insts, meta


def warmstart_configs(n_inst=32, seed=5, n=2 ** 14):
    """Detuned ring cavities: cavity detunings (+50 .. +110 fs from synchronism) of an
    8 mm fan-out PPLN (local period 29-31 um, one position per instance),
    4 W at 1 GHz of 1 ps sech^2 pump in pi (50 um)^2, 16 ps window on n
    points (2^14 by default; the specified grid is 2^16), Heun with fixed
    0.4 um steps (error control off: maxErr = inf, minErr = 0)."""
    from optifax_amd import physics as ph
    rng = np.random.default_rng(seed)
    periods = rng.permutation(np.linspace(29e-6, 31e-6, n_inst))
    dets = np.linspace(50e-15, 110e-15, n_inst)
    return [ph.OPOConfig(n=n, window=16e-12, length=8e-3, period=float(p),
                         pump_energy=4e-9, pump_fwhm=1e-12, area=np.pi * (50e-6) ** 2,
                         delay=float(d), delay_mode="sync", error_bounds=(0.0, 1e30),
                         noise_seed=1000 + i)
            for i, (p, d) in enumerate(zip(periods, dets))]


def training_configs(n_inst=64, seed=11):
    """Trajectories for the surrogate's data, drawn apart from the solve's
    instances: period 29-31 um, detuning +30 .. +110 fs, pump 3.5-4.5 nJ."""
    from optifax_amd import physics as ph
    rng = np.random.default_rng(seed)
    return [ph.OPOConfig(n=2 ** 14, window=16e-12, length=8e-3,
                         period=float(rng.uniform(29e-6, 31e-6)),
                         pump_energy=float(rng.uniform(3.5e-9, 4.5e-9)), pump_fwhm=1e-12,
                         area=np.pi * (50e-6) ** 2, delay=float(rng.uniform(30e-15, 110e-15)),
                         delay_mode="sync", error_bounds=(0.0, 1e30), noise_seed=5000 + i)
            for i in range(n_inst)]


def main_warmstart(args):
    """Time to steady state (trip-to-trip change of the output-coupled
    spectrum < 1e-4, 1000-trip budget): plain FP32 round trips from noise
    (phase "plain") and the surrogate warm start (phase "warm": trajectory
    data, training, surrogate trips, gated plain trips); "all" runs both."""
    from optifax_amd import warmstart as ws
    log = lambda m: print(m, file=sys.stderr, flush=True)
    cfgs = warmstart_configs(args.ws_instances, n=args.ws_grid)
    pairs = [ws.opo_instance(c) for c in cfgs]
    insts = [p[0] for p in pairs]
    builds = [p[1] for p in pairs]
    log(f"warmstart: {len(insts)} instances, phase {args.ws_phase}")
    cfg_out = {
        "workload": f"detuned ring OPOs at 2^{int(np.log2(args.ws_grid))} points: {len(insts)} detunings "
                    "+50..+110 fs, 8 mm fan-out PPLN "
                    "29-31 um, 4 W @ 1 GHz 1 ps sech^2 in pi (50 um)^2, 5% OC, fixed 0.4 um "
                    "steps; steady state = change between trips < 1e-4, budget 1000 trips",
        "instances": len(insts),
    }
    plain = warm = None
    if args.ws_phase in ("all", "plain"):
        plain = ws.solve_plain(insts, conv_tol=1e-4, budget=1000, log=log,
                               max_seconds=args.ws_max_seconds, precision=args.precision)
        cfg_out.update(plain_seconds=plain["seconds"],
                       plain_conv_seconds=[None if not np.isfinite(x) else round(float(x), 2)
                                           for x in plain["conv_seconds"]],
                       plain_trips_median=float(np.median(plain["trips"])),
                       plain_trips_max=int(np.max(plain["trips"])),
                       converged_plain=int((plain["status"] == 2).sum()))
    if args.ws_phase in ("all", "warm"):
        t0 = time.perf_counter()
        X, F, pe, pr, lin, sb = ws.generate_pairs(training_configs(args.ws_train_instances),
                                                  args.ws_train_trips, log=log)
        t_data = time.perf_counter() - t0
        x, y, a_ref = ws.make_dataset(X, F, pe, pr, lin)
        t0 = time.perf_counter()
        net, losses, val, ms = ws.train_surrogate(x, y, args.ws_train_steps, log=log)
        t_train = time.perf_counter() - t0
        sur = ws.BandSurrogate(net, a_ref, max_norm=float(np.max(np.linalg.norm(X - 0, axis=1))))
        warm = ws.solve_warm(cfgs, insts, builds, sur, conv_tol=1e-4, budget=1000, log=log,
                             max_seconds=args.ws_max_seconds)
        cfg_out.update(warm_seconds=warm["seconds"], surrogate_seconds=warm["surrogate_seconds"],
                       warm_conv_seconds=[None if not np.isfinite(x) else round(float(x), 2)
                                          for x in warm["conv_seconds"]],
                       warm_trips_median=float(np.median(warm["trips"])),
                       warm_trips_max=int(np.max(warm["trips"])),
                       surrogate_trips_median=float(np.median(warm["surrogate_trips"])),
                       converged_warm=int((warm["status"] == 2).sum()),
                       fallbacks=int(warm["fallback"].sum()),
                       gate_err_median=float(np.median(warm["gate_err"])),
                       train_pairs=int(x.shape[0]), data_seconds=t_data, train_seconds=t_train,
                       train_loss_last=float(np.mean(losses[-50:])), val_loss=val,
                       mlp_train_device_ms=ms)
        net.close()
    value, unit = None, None
    if plain is not None and warm is not None:
        value = plain["seconds"] / warm["seconds"]
        unit = "x (time to steady state, plain / warm start)"
        errs = []
        for i in np.where((plain["status"] == 2) & (warm["status"] == 2))[0]:
            Sp = np.abs(plain["oc"][i]) ** 2
            Sw = np.abs(warm["oc"][i]) ** 2
            errs.append(float(np.linalg.norm(Sw - Sp) / np.linalg.norm(Sp)))
        cfg_out["steady_state_rel_l2_warm_vs_plain_median"] = \
            float(np.median(errs)) if errs else None
    elif plain is not None:
        value, unit = plain["seconds"], "s to steady state (plain)"
    else:
        value, unit = warm["seconds"], "s to steady state (warm start, incl. surrogate trips)"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": unit,
        "n_gpus": 1,
        "steps": int(max(np.max(r["trips"]) for r in (plain, warm) if r is not None)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": plain is not None and warm is not None,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32" if args.precision == "single" else "f64") + " fields (plain), f32 fields + bf16 surrogate (warm)",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise; surrogate "
                "trained on solver trajectories)",
        "config": cfg_out,
    }
    if args.ws_out:
        arrs = {}
        for tag, r in (("plain", plain), ("warm", warm)):
            if r is not None:
                arrs[f"{tag}_trips"] = r["trips"]
                arrs[f"{tag}_status"] = r["status"]
                arrs[f"{tag}_oc"] = np.stack(r["oc"])
        if warm is not None:
            arrs.update(sur_trips=warm["surrogate_trips"], gate_err=warm["gate_err"])
        np.savez(args.ws_out, **arrs)
    print(json.dumps(out))


MFMA_BF16_PEAK_TFLOPS = 2500.0


def mlp_step_flops(batch, n_in, width, n_out):
    """Algorithmic flops of one training step: forward, weight gradients, and
    the data gradients of layers 1-3 (the input needs none).  n_in is the
    network's input width (the device pads it to 128 columns; the padding is
    not counted)."""
    fwd = 2.0 * batch * (n_in * width + 2 * width * width + width * n_out)
    dgrad = 2.0 * batch * (2 * width * width + width * n_out)
    return 2.0 * fwd + dgrad


def main_steady(args):
    """Accuracy of the FP32 steady state: eight detuned 8 mm cavities run to
    steady state (change between trips < 1e-4) with FP32 fields and with
    FP64 fields (the FP64 device mode reproduces the FP64 oracle's step
    sequence exactly and its fields to rounding); value = median relative L2
    distance of their output-coupled power spectra."""
    from optifax_amd import warmstart as ws
    log = lambda m: print(m, file=sys.stderr, flush=True)
    pick = [3, 30, 7, 18, 24, 28, 23, 21]          # the eight fastest to settle in FP32
    cfgs = [warmstart_configs(32)[i] for i in pick]
    insts = [ws.opo_instance(c)[0] for c in cfgs]
    res = {}
    for prec in ("single", "double"):
        res[prec] = ws.solve_plain(insts, conv_tol=1e-4, budget=args.steady_budget, chunk=10,
                                   log=log, precision=prec)
    s32 = [np.abs(o) ** 2 for o in res["single"]["oc"]]
    s64 = [np.abs(o) ** 2 for o in res["double"]["oc"]]
    both = np.isfinite(res["single"]["conv_seconds"]) & np.isfinite(res["double"]["conv_seconds"])
    rel = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(s32, s64)])
    out = {
        "metric": METRIC,
        "value": float(np.median(rel[both])) if both.any() else None,
        "unit": "relative L2, FP32 vs FP64 steady-state output power spectrum (median)",
        "n_gpus": 1,
        "steps": args.steady_budget,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 vs f64",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise)",
        "config": {
            "workload": "detuned ring OPOs (8 mm fan-out PPLN) at 2^14 points, 8 of the 32 detunings (+50..+110 fs), each run "
                        "to a change between trips < 1e-4 in FP32 and in FP64",
            "instances": len(insts),
            "converged_both": int(both.sum()),
            "rel_l2_per_instance": [float(x) for x in rel],
            "trips_f32": res["single"]["trips"].tolist(),
            "trips_f64": res["double"]["trips"].tolist(),
            "seconds_f32": res["single"]["seconds"],
            "seconds_f64": res["double"]["seconds"],
        },
    }
    print(json.dumps(out))


def main_esd(args):
    """The ring OPO over 300 round trips from noise with FP32 fields and
    with FP64 fields (the FP64 device mode reproduces the FP64 oracle's step
    sequence exactly and its fields to rounding; the oracle itself needs ~55 s
    per trip here).  value = relative L2 distance of the run-averaged
    output-coupled energy spectral densities, mean over all trips of
    |E_oc(w)|^2 (OpticalSim.averageSimPulses, OpticalSim.m:302-307: the sum of
    each trip's output ESD divided by the number of trips), accumulated on
    the device."""
    from optifax_amd.solver import Context, OPOBatch
    cfg, b = workload()
    si, ops = b["solver"], b["ops"]
    ctx = Context(0)
    trips = args.esd_trips
    res = {}
    for prec in ("single", "double"):
        run = OPOBatch([dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                             ocm=ops["ocm"], delay=b["delay"])],
                       precision=prec, trip_cap=trips + 1, ctx=ctx)
        t0 = time.perf_counter()
        done = 0
        while done < trips:
            k = min(25, trips - done)
            run.launch(k)
            run.sync(allow_nonfinite=True)
            done += k
            print(f"esd: {prec} {done}/{trips} trips {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        r = run.read(0)
        avg, _, n_avg = run.read_average(0)       # summed on the device as each trip ends
        res[prec] = dict(seconds=wall, last=np.abs(r["oc_spec"]) ** 2, avg=avg, n_avg=n_avg,
                         status=r["status"], energy=r["trip_log"][:trips, 1].copy())
        run.close()
    ctx.close()
    rl2 = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))
    s32, s64 = res["single"], res["double"]
    en32, en64 = s32["energy"], s64["energy"]
    marks = [t for t in (1, 10, 25, 50, 100, 150, 200, 250, 300) if t <= trips]
    out = {
        "metric": METRIC,
        "value": rl2(s32["avg"], s64["avg"]),
        "unit": f"relative L2, FP32 vs FP64 run-averaged output-coupled ESD over {trips} trips",
        "n_gpus": 1,
        "steps": trips,
        "warmup": 0,
        "ms_per_step": 1e3 * s32["seconds"] / trips,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 vs f64",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise)",
        "config": {
            "workload": "synchronously pumped ring OPO, N=2^15, 12 ps window, 1040 nm; 5 mm "
                        "MgO:PPLN 30.5 um; 2 W @ 1 GHz 200 fs sech^2; 5% OC; 0.3 ps delay; "
                        f"{trips} trips from noise in FP32 and in FP64",
            "trips_averaged": [s32["n_avg"], s64["n_avg"]],
            "status": [s32["status"], s64["status"]],
            "last_trip_esd_rel_l2": rl2(s32["last"], s64["last"]),
            "avg_power_rel_diff": float(abs(en32.mean() - en64.mean()) / en64.mean()),
            "seconds_f32": s32["seconds"],
            "seconds_f64": s64["seconds"],
            "oc_energy_f64": {str(t): float(en64[t - 1]) for t in marks},
            "oc_energy_rel_diff": {str(t): float(abs(en32[t - 1] - en64[t - 1]) /
                                                 max(abs(en64[t - 1]), 1e-300))
                                   for t in marks},
        },
    }
    print(json.dumps(out))


def config5_configs(n_det=8, fixed_step=None, pump_fwhm=200e-15):
    """Config 5 as specified: 2^16 points, 16 ps, 8 mm fan-out PPLN at the
    local period 30 um, 4 W at 1 GHz (4 nJ sech^2 of pump_fwhm; config 5
    does not state the length, the ring workload's is 200 fs), 5 % coupler on the 1.2-1.8 um band, n_det of the 64 detunings
    linspace(-50, 50) fs (every 64/n_det-th, ends included by linspace over
    n_det), adaptive Heun with the default error bounds (fixed_step: error
    control off and base step fixed_step metres; the pass then keeps its
    initial step multiplier of 4, NEE_CPU.m:10, so it takes steps of
    4 fixed_step)."""
    from optifax_amd import physics as ph
    kw = {}
    if fixed_step:
        kw = dict(step=float(fixed_step), error_bounds=(0.0, 1e30))
    return [ph.OPOConfig(n=2 ** 16, window=16e-12, length=8e-3, period=30e-6,
                         pump_energy=4e-9, pump_fwhm=float(pump_fwhm), delay=float(d),
                         delay_mode="sync", noise_seed=1000 + i, **kw)
            for i, d in enumerate(np.linspace(-50e-15, 50e-15, n_det))]


def main_config5(args):
    """Whether config 5's cavities reach a steady state: FP64 plain round
    trips from one-photon noise for n detunings at once (one persistent
    launch per 10 trips), convergence test as specified (change of the
    output-coupled spectrum between trips < 1e-4) until all converged, the
    trip budget (--c5-trips) or --ws-max-seconds.  value = detunings
    converged; per detuning: the change between trips (median over the last
    50 trips), the output energy's rms fluctuation over the last 50 trips
    relative to its mean, and the adaptive attempts per trip."""
    from optifax_amd import warmstart as ws
    from optifax_amd.solver import OPOBatch
    log = lambda m: print(m, file=sys.stderr, flush=True)
    cfgs = config5_configs(args.c5_detunings, args.c5_fixed_step, args.c5_pump_fwhm)
    insts = [ws.opo_instance(c)[0] for c in cfgs]
    budget = args.c5_trips
    run = OPOBatch(insts, trip_cap=budget, precision=args.precision)
    t0 = time.perf_counter()
    done = 0
    while done < budget:
        k = min(10, budget - done)
        run.launch(k, 1e-4)
        run.sync(allow_nonfinite=True)
        done += k
        st = [run.read(i)["status"] for i in range(len(insts))]
        now = time.perf_counter() - t0
        log(f"config5: {done} trips, {sum(x == 2 for x in st)}/{len(st)} converged, {now:.0f} s")
        if all(x in (2, 3) for x in st) or now > args.ws_max_seconds:
            break
    wall = time.perf_counter() - t0
    rs = [run.read(i) for i in range(len(insts))]
    run.close()
    per = []
    for c, r in zip(cfgs, rs):
        n = r["trips_done"]
        tl = r["trip_log"][:n]
        w = tl[-50:]
        en = w[:, 1]
        per.append({"detuning_fs": round(c.delay * 1e15, 2), "status": r["status"],
                    "trips": int(n),
                    "change_median_last50": float(np.median(w[:, 0])),
                    "oc_energy_rms_rel_last50": float(en.std() / en.mean()) if en.mean() > 0 else None,
                    "oc_energy_mean_last50": float(en.mean()),
                    "attempts_per_trip_median": float(np.median(tl[:, 2]))})
    out = {
        "metric": METRIC,
        "value": int(sum(p["status"] == 2 for p in per)),
        "unit": f"of {len(per)} detunings at steady state (change between trips < 1e-4)",
        "n_gpus": 1,
        "steps": done,
        "warmup": 0,
        "ms_per_step": 1e3 * wall / max(done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "double" else "f32",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise)",
        "config": {
            "workload": f"config 5 cavities: 2^16 points, 16 ps, 8 mm PPLN at 30 um, 4 nJ "
                        f"{args.c5_pump_fwhm * 1e15:g} fs sech^2 at 1 GHz, 5% OC on 1.2-1.8 um, "
                        f"{len(per)} detunings "
                        "linspace(-50, 50) fs, "
                        + (f"error control off, fixed {4 * args.c5_fixed_step * 1e6:g} um steps"
                           if args.c5_fixed_step
                           else "adaptive Heun (default error bounds)")
                        + f", plain trips from noise, budget {budget}",
            "seconds": wall,
            "per_detuning": per,
        },
    }
    print(json.dumps(out))


def fibre_example(n=2 ** 13):
    """The supercontinuum example distributed with gnlsefib.m (Dudley, Genty &
    Coen's test_Dudley, in ps and m): 835 nm, 10 kW 28.4 fs sech in 15 cm of
    PCF (gamma 0.11 /W/m, betas to 10th order), fused-silica Raman (fr 0.18,
    12.2 / 32 fs) and self-steepening, 12.5 ps window on n points, 200 saves."""
    from optifax_amd import fibre
    c_nm_ps = 299792458 * 1e9 / 1e12
    T = np.linspace(-12.5 / 2, 12.5 / 2, n)
    betas = [-11.830e-3, 8.1038e-5, -9.5205e-8, 2.0737e-10, -5.3943e-13, 1.3486e-15,
             -2.5495e-18, 3.0524e-21, -1.7140e-24]
    return dict(T=T, A=np.sqrt(10000.0) / np.cosh(T / 0.0284), w0=2 * np.pi * c_nm_ps / 835.0,
                gamma=0.11, betas=betas, loss=0.0, fr=0.18,
                RT=fibre.raman_response(T, 0.0122, 0.032), flength=0.15, nsaves=200)


def main_fibre(args):
    """Fibre GNLSE (gnlsefib.m) on the device: the example pulse at 2^13
    points with FP32 fields, one fibre (latency) and a batch of fibres with
    peak powers 5-15 kW in one launch, one workgroup each (throughput;
    value = fibre solves/s of the batch); FP64 fields at 2^12 beside it.
    Parity: the single FP32 solve's final temporal field against the FP64
    oracle (oracle/gnlse.py, numpy, one core), whose wall time is the CPU
    baseline."""
    from optifax_amd import fibre
    c = fibre_example()
    op = fibre.operators(c["T"], c["w0"], c["gamma"], c["betas"], c["loss"], c["fr"], c["RT"])
    run = lambda fields, ops, prec, T=c["T"]: fibre.gnlse_batch(T, fields, ops, c["flength"],
                                                                 c["nsaves"], prec)
    run([c["A"]], [op], "single")                       # warm-up (module load, first launch)
    t0 = time.perf_counter()
    Z, AWz, st1, ms1 = run([c["A"]], [op], "single")
    wall1 = time.perf_counter() - t0
    AT1, _ = fibre.lab_frame(Z, AWz[0], op, c["T"].size, op["dT"])
    B = args.fibre_batch
    powers = np.linspace(5000.0, 15000.0, B)
    fields = [np.sqrt(p) / np.cosh(c["T"] / 0.0284) for p in powers]
    t0 = time.perf_counter()
    _, _, stB, msB = run(fields, [op] * B, "single")
    wallB = time.perf_counter() - t0
    c64 = fibre_example(2 ** 12)
    op64 = fibre.operators(c64["T"], c64["w0"], c64["gamma"], c64["betas"], 0.0, c64["fr"],
                           c64["RT"])
    _, _, st64, ms64 = fibre.gnlse_batch(c64["T"], [c64["A"]], [op64], c64["flength"],
                                         c64["nsaves"], "double")
    out = {
        "metric": METRIC,
        "value": B / (msB * 1e-3),
        "unit": f"fibre solves/s (batch of {B}, device time)",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": msB,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 step control",
        "data": "synthetic (the gnlsefib.m example pulse; the batch scans its peak power)",
        "config": {
            "workload": "fibre GNLSE (gnlsefib.m): 2^13 points, 12.5 ps, 835 nm 28.4 fs sech, "
                        "15 cm PCF, gamma 0.11, 10 betas, Raman fr 0.18, self-steepening, "
                        "ode45 (RelTol 1e-5, NormControl), 200 saves",
            "single_device_ms": ms1, "single_wall_s": wall1,
            "single_steps": [int(st1[0, 0]), int(st1[0, 1])],
            "batch": B, "batch_device_ms": msB, "batch_wall_s": wallB,
            "batch_steps_median": float(np.median(stB[:, 0])),
            "batch_status_ok": int((stB[:, 3] == 0).sum()),
            "fp64_2e12_device_ms": ms64, "fp64_2e12_steps": [int(st64[0, 0]), int(st64[0, 1])],
        },
    }
    if not args.no_cpu_baseline:
        import oracle.gnlse as og
        t0 = time.perf_counter()
        Zo, ATo, _, _, _, sto = og.gnlsefib(c["T"], c["A"], c["w0"], c["gamma"], c["betas"],
                                            c["loss"], c["fr"], c["RT"], c["flength"],
                                            c["nsaves"])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": 1.0 / dt, "unit": "fibre solves/s", "cores": 1, "kind": "port",
            "sample": f"the single example solve in the FP64 oracle (numpy FFTs, one core): "
                      f"{dt:.1f} s, {sto['accepted']} steps"}
        out["parity"] = {
            "against": "FP64 oracle, final temporal field",
            "rel_l2": float(np.linalg.norm(AT1[-1] - ATo[-1]) / np.linalg.norm(ATo[-1])),
            "oracle_steps": [sto["accepted"], sto["rejected"]]}
    print(json.dumps(out))


def main_qpm(args):
    """QPM gain map (qpmgain.m): an 8 mm MgO:PPLN at 30 um and a 300 fs, 1 nJ
    sech^2 pump at 1040 nm on a 2^13-point window; signals 1100-1700 nm (at
    most 2000) x pump wavelengths within +-3 FWHM (at most 100), the crystal
    sampled as qpmgain.m samples it.  value = mismatch sums per second of the
    device z-sum (runs of equal poling summed in closed form); parity and the
    CPU baseline from the oracle's literal sample-by-sample sum on 200 of the
    outputs (numpy, one core)."""
    from optifax_amd import physics as ph, qpm
    sw = ph.SimWindow(1040e-9, 2 ** 13, 10e-12)
    area = np.pi * (30e-6) ** 2
    E = ph.sech_pump(sw, 300e-15, 1e-9, area)
    pulse = qpm.pulse_descriptors(sw, E, area)
    xt = ph.NonlinearCrystal(30e-6, 8e-3)
    a = qpm.qpm_arrays(xt, sw, pulse, (1100.0, 1700.0), step=1e-7)
    dk = a["dk"].reshape(-1)
    for _ in range(max(args.warmup, 1)):
        qpm.zsum(dk, a["z"], a["w"])
    ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got = qpm.zsum(dk, a["z"], a["w"])
        ms.append(qpm.zsum.last_ms)
    wall = time.perf_counter() - t0
    dev_ms = float(np.mean(ms))
    out = {
        "metric": METRIC,
        "value": dk.size / (dev_ms * 1e-3),
        "unit": "QPM mismatch sums/s (device time)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (sech^2 pump spectrum through an 8 mm PPLN grating)",
        "config": {
            "workload": "QPM gain map (qpmgain.m): 8 mm MgO:PPLN 30 um, 300 fs 1 nJ sech^2 pump "
                        "at 1040 nm, 2^13 points over 10 ps, signals 1100-1700 nm",
            "map": list(a["dk"].shape), "z_samples": int(a["z"].size),
            "runs": int(qpm.zsum.last_runs), "device_ms": dev_ms,
        },
    }
    if not args.no_cpu_baseline:
        import oracle.qpm as oq
        pick = np.random.default_rng(0).choice(dk.size, 200, replace=False)
        t0 = time.perf_counter()
        want = oq.literal_sum(dk[pick], a["z"], a["w"])
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": pick.size / dt, "unit": "QPM mismatch sums/s", "cores": 1, "kind": "port",
            "sample": f"the oracle's literal sum (one exponential per z sample, numpy) for 200 "
                      f"of the {dk.size} outputs: {dt:.2f} s"}
        out["parity"] = {
            "against": "FP64 oracle, literal sum over the z samples, 200 outputs",
            "rel_l2": float(np.linalg.norm(got[pick] - want) / np.linalg.norm(want))}
    print(json.dumps(out))


def main_pass(args):
    """A single SHG/OPG crystal pass: 2^13 points, 10 ps window, 1040 nm;
    sech^2 200 fs pump in 5 um^2 with seeded one-photon noise; 30 um PPLN,
    d_eff 14 pm/V; RK23 at tolerance 1e-6 (maxErr 1e-4 %, minErr 5e-6 %);
    P = 5.  As stated (1 nJ over 1 mm) the FP64 oracle itself diverges to
    NaN by 0.8 mm, so the pass runs 0.2 mm at 0.1 nJ (the parity case).
    value = single-pass solves/s from the device time of each solve, its
    input rewritten in HBM before each one."""
    from optifax_amd import physics as ph
    from optifax_amd.solver import Context, OPOBatch
    L, energy, area = 0.2e-3, 1e-10, 5e-12
    sw = ph.SimWindow(1040e-9, 8192, 10e-12)
    si = ph.NonlinearCrystal(30e-6, L, chi2=np.pi * 14e-12).solver_inputs(sw)
    E = np.fft.fftshift(ph.sech_pump(sw, 200e-15, energy, area) + ph.quantum_noise(sw, area, 11))
    si = dict(si, maxErr=1e-4, minErr=5e-6, sel=float(np.ceil(si["nSteps"] / 4)), n_cols=4)
    ctx = Context(0)
    run = OPOBatch([dict(solver=si, field0=E, method="rk23")], precision=args.precision, ctx=ctx)
    for _ in range(args.warmup):
        run.write_field(0, E)
        run.launch(0)
        run.sync()
    dev_ms, t0 = [], time.perf_counter()
    for _ in range(args.steps):
        run.write_field(0, E)
        run.launch(0)
        run.sync()
        dev_ms.append(run.last_ms())
    wall = time.perf_counter() - t0
    r = run.read(0)
    att = r["attempts"] // (args.warmup + args.steps)
    acc = r["accepted"] // (args.warmup + args.steps)
    gbs = algorithmic_bytes(8192, acc, att - acc, 0) / (np.mean(dev_ms) * 1e-3) / 1e9
    run.close()
    ctx.close()
    out = {
        "metric": METRIC,
        "value": 1e3 / float(np.mean(dev_ms)),
        "unit": "single-pass solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(dev_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "single" else "f64",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise)",
        "config": {
            "workload": "single SHG/OPG pass, N=2^13, 10 ps, 1040 nm; 0.2 mm "
                        "MgO:PPLN 30 um, d_eff 14 pm/V; 0.1 nJ 200 fs sech^2 in 5 um^2; RK23 "
                        "tol 1e-6; P = 5",
            "adaptive_attempts_per_solve": att,
            "accepted_per_solve": acc,
            "solves_per_s_host_clock": args.steps / wall,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        import oracle
        t0 = time.perf_counter()
        _, _, _, cnt = oracle.rk23(E, si["alpha_w"], si["G33"], si["w0"], si["bdiffw0"], si["h"],
                                   si["nSteps"], si["dt"], si["beta_op"], si["maxErr"],
                                   si["minErr"], si["sel"], n_cols=4, want_stepmods=False)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "single-pass solves/s", "cores": 1,
                               "kind": "port",
                               "sample": f"one whole solve ({int(cnt[0])} adaptive attempts, "
                                         f"{dt:.1f} s)"}
    print(json.dumps(out))


def main_mex(args):
    """The MEX drop-in as OpticalSim.run drives it: one OPOmexBatch call per
    round trip (host arrays in and out, OpticalSim.m:237-252) on the ring
    workload's crystal pass (N=2^15, 5 mm, FP32, ApFT with 4 columns,
    stepMods).  The context keeps the device run between calls with the same
    shapes, so only the first call plans and allocates.  Reported: the first
    call's wall time, the later calls' mean wall time, the device time of
    the same pass (HIP events around one launch of the kept run) and their
    difference, the per-call host overhead (uploads, readbacks, conversions)."""
    from optifax_amd.solver import Context, OPOBatch, OPOmexBatch
    cfg, b = workload()
    si = b["solver"]
    n = si["beta_op"].size
    E = b["ops"]["field0"]
    args_ = (si["alpha_w"], si["G33"], si["w0"], si["bdiffw0"], si["h"], si["nSteps"], si["dt"],
             si["beta_op"], si["maxErr"], si["minErr"], si["sel"])
    ctx = Context(0)
    walls = []
    for it in range(args.warmup + args.steps + 1):
        t0 = time.perf_counter()
        E, _, _ = OPOmexBatch(E, *args_, np.zeros((n, 4), complex), np.zeros(si["nSteps"]),
                              ctx=ctx)
        walls.append(time.perf_counter() - t0)
    # device time of the same pass: one launch of a run kept on the device
    run = OPOBatch([dict(solver=dict(si, n_cols=4), field0=E)], ctx=ctx)
    run.launch(0)
    run.sync()
    dev = []
    for _ in range(3):
        run.write_field(0, E)
        run.launch(0)
        run.sync()
        dev.append(run.last_ms())
    run.close()
    ctx.close()
    warm = walls[1 + args.warmup:]
    out = {
        "metric": "OPOmexBatch calls/s (one call per round trip)",
        "value": 1.0 / float(np.mean(warm)),
        "unit": "calls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.mean(warm)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (the ring workload's noise start, each call's output the next input)",
        "config": {"workload": "OPOmexBatch crystal pass of the ring workload (N=2^15, 5 mm PPLN, "
                               "ApFT 4 columns, stepMods), host arrays",
                   "first_call_ms": 1e3 * walls[0],
                   "device_pass_ms": float(np.mean(dev)),
                   "host_overhead_ms_per_call": 1e3 * float(np.mean(warm)) - float(np.mean(dev))},
    }
    print(json.dumps(out))


def main_surrogate(args):
    """The surrogate's shape: 2050 -> 2048 x3 -> 2048 MLP, batch 512, Adam, on a
    synthetic device-resident set of 8192 pairs; steps = Adam steps."""
    from optifax_amd.surrogate import SurrogateMLP
    n_in, width, n_out, batch = 2050, 2048, 2048, 512
    rng = np.random.default_rng(0)
    X = rng.standard_normal((8192, n_in)).astype(np.float32)
    T = rng.standard_normal((8192, n_out)).astype(np.float32)
    net = SurrogateMLP(n_in, width, n_out, max_batch=batch, seed=0)
    net.upload_dataset(X, T)
    if args.warmup > 0:
        net.train(args.warmup, batch=batch)
    t0 = time.perf_counter()
    losses, ms = net.train(args.steps, batch=batch)
    wall = time.perf_counter() - t0
    flops = mlp_step_flops(batch, n_in, width, n_out)
    tflops = flops / (ms / args.steps * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "surrogate training steps/s (batch 512)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "surrogate MLP 2050-2048-2048-2048-2048, GELU, batch 512, Adam, "
                               "relative-L2 loss",
                   "device_ms_per_step": ms / args.steps,
                   "gflop_per_step": flops / 1e9},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": MFMA_BF16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
                     "traffic": None},
    }
    net.close()
    print(json.dumps(out))


def cavity_state(b, X):
    """The oracle's intracavity field E_cav (natural time order, before the
    pump is added) whose next crystal input is the device field X (fftshifted
    time): inverts pump addition and refraction into the crystal."""
    from optifax_amd.physics import SimWindow
    si = b["solver"]
    r_in = np.sqrt(1.0 / si["n_xtal"])
    return SimWindow.k2t(SimWindow.t2k(np.fft.ifftshift(X)) / r_in - SimWindow.t2k(b["pump"]))


def cpu_baseline(b, X, method="heun"):
    """The FP64 oracle (one core) runs one whole round trip (OpticalSim.run's
    order: pump, refraction, the adaptive crystal pass, output coupling,
    delay) from the crystal input X of the first timed trip.  Returns the
    baseline record, the oracle's output-coupled spectrum (FFT order) of
    that trip (the parity reference) and its (attempted, accepted) steps."""
    import oracle
    from optifax_amd.physics import SimWindow
    si, sw = b["solver"], b["simwin"]
    R, T, phi = b["oc"].masks(sw)
    E_cav = cavity_state(b, X)
    t0 = time.perf_counter()
    o = oracle.opo_run(1, E_cav, b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi, b["delay"],
                       sw.rel_omegas, si, method=0 if method == "heun" else 1)
    dt = time.perf_counter() - t0
    att, acc = (int(v) for v in o["steps"][0])
    rec = {
        "value": 1.0 / dt,
        "unit": "round trips/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one whole round trip from the first timed trip's crystal input "
                   f"({att} adaptive steps, {acc} accepted, N={sw.n}), {dt:.1f} s"),
    }
    return rec, SimWindow.fft_order(SimWindow.t2k(o["E_oc"])), (att, acc)


def cpu_baseline_batch(insts, meta, attempts_per_trip, sample_steps=3000):
    """Oracle (FP64) on the host cores for the uneven batch: the instances
    are independent, so the CPU path runs one instance per core.  Sample: the
    first `sample_steps` minimal steps of one instance per core (spread over
    the batch and its grid sizes), all cores at once, giving a cost per
    adaptive step for each grid size under full load; the batch's round trip
    is then the sum over instances of (GPU-measured adaptive steps) x (that
    cost), divided by the cores.  The single-core figure repeats the sample
    with one instance of each grid size on an otherwise idle host."""
    import concurrent.futures as cf
    import oracle
    oracle.lib()
    try:
        cores = min(16, len(os.sched_getaffinity(0)))
    except AttributeError:
        cores = min(16, os.cpu_count() or 1)
    sizes = sorted({n for n, _ in meta})
    order = sorted(range(len(insts)), key=lambda i: (sizes.index(meta[i][0]), i))
    # every grid size once, then the other cores spread over the batch
    pick = [next(i for i in order if meta[i][0] == n) for n in sizes]
    rest = [i for i in order if i not in pick]
    extra = max(cores - len(pick), 0) if rest else 0
    pick += [rest[(k * len(rest)) // extra] for k in range(extra)]

    def one(i):
        si = insts[i]["solver"]
        ns = min(sample_steps, si["nSteps"])
        t0 = time.perf_counter()
        _, _, _, cnt = oracle.nee_heun(insts[i]["field0"], si["alpha_w"], si["G33"][:ns],
                                       si["w0"], si["bdiffw0"], si["h"], ns, si["dt"],
                                       si["beta_op"], si["maxErr"], si["minErr"], float(ns),
                                       n_cols=0, want_stepmods=False)
        return meta[i][0], time.perf_counter() - t0, int(cnt[0])

    def per_size(rows):
        out = {}
        for n in sizes:
            t = sum(r[1] for r in rows if r[0] == n)
            c = sum(r[2] for r in rows if r[0] == n)
            out[n] = t / max(c, 1)
        return out

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(cores) as ex:
        loaded = list(ex.map(one, pick))
    t_par = time.perf_counter() - t0
    t0 = time.perf_counter()
    alone = [one(next(i for i in pick if meta[i][0] == n)) for n in sizes]
    t_one = time.perf_counter() - t0
    c_par, c_one = per_size(loaded), per_size(alone)
    trip_par = sum(a * c_par[n] for a, (n, _) in zip(attempts_per_trip, meta)) / cores
    trip_one = sum(a * c_one[n] for a, (n, _) in zip(attempts_per_trip, meta))
    return {
        "value": len(insts) / trip_par,
        "unit": "instance round trips/s",
        "cores": cores,
        "kind": "port",
        "single_core_value": len(insts) / trip_one,
        "sample": (f"first {sample_steps} minimal steps of {cores} instances at once "
                   f"({sum(r[2] for r in loaded)} adaptive steps, {t_par:.1f} s) and of one "
                   f"instance per grid size alone ({t_one:.1f} s); seconds per adaptive step "
                   + ", ".join(f"2^{int(np.log2(n))}: {c_par[n] * 1e3:.2f} ms loaded / "
                               f"{c_one[n] * 1e3:.2f} ms alone" for n in sizes)
                   + "; scaled by each instance's GPU-measured adaptive steps per trip"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed round trips")
    ap.add_argument("--warmup", type=int, default=2, help="untimed round trips")
    ap.add_argument("--precision", default="single", choices=["single", "double"])
    ap.add_argument("--method", default="heun", choices=["heun", "rk23"],
                    help="crystal solver of the ring workload (rk23: the ode23 variant)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ws-instances", type=int, default=32)
    ap.add_argument("--ws-train-instances", type=int, default=32)
    ap.add_argument("--ws-grid", type=int, default=2 ** 14,
                    help="grid points of the warmstart workload's instances (specified: 65536)")
    ap.add_argument("--ws-max-seconds", type=float, default=900.0,
                    help="stop each solve after this many seconds (results so far are reported)")
    ap.add_argument("--ws-train-trips", type=int, default=250)
    ap.add_argument("--ws-train-steps", type=int, default=20000)
    ap.add_argument("--ws-out", default="")
    ap.add_argument("--steady-budget", type=int, default=600,
                    help="trip budget of each run of the steady workload")
    ap.add_argument("--ws-phase", default="all", choices=["all", "plain", "warm"])
    ap.add_argument("--esd-trips", type=int, default=300,
                    help="round trips of each run of the esd workload")
    ap.add_argument("--workload", default="ring",
                    choices=["ring", "batch", "warmstart", "surrogate", "steady", "esd", "pass",
                             "mex", "config5", "fibre", "qpm"],
                    help="ring: one synchronously pumped ring OPO (the headline line); batch: 256 uneven "
                         "instances in one persistent launch")
    ap.add_argument("--fibre-batch", type=int, default=256)
    ap.add_argument("--c5-detunings", type=int, default=8)
    ap.add_argument("--c5-trips", type=int, default=200,
                    help="trip budget of the config5 workload (specified: 1000)")
    ap.add_argument("--c5-pump-fwhm", type=float, default=200e-15)
    ap.add_argument("--c5-fixed-step", type=float, default=0.0,
                    help="config5 with error control off and this fixed step (m); 0 = adaptive")
    ap.add_argument("--tune", default="heuristic", choices=["heuristic", "auto"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 256 MiB three times (the known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: one GPU only; --gpus {args.gpus} is not supported (independent "
              "instances would shard across GPUs, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        return main_batch(args)
    if args.workload == "warmstart":
        return main_warmstart(args)
    if args.workload == "surrogate":
        return main_surrogate(args)
    if args.workload == "steady":
        return main_steady(args)
    if args.workload == "esd":
        return main_esd(args)
    if args.workload == "pass":
        return main_pass(args)
    if args.workload == "mex":
        return main_mex(args)
    if args.workload == "config5":
        return main_config5(args)
    if args.workload == "fibre":
        return main_fibre(args)
    if args.workload == "qpm":
        return main_qpm(args)
    from optifax_amd.solver import Context, OPOBatch
    cfg, b = workload()
    si, ops = b["solver"], b["ops"]
    n = si["beta_op"].size
    ctx = Context(0)
    trip_cap = args.warmup + args.steps + 1
    inst = dict(solver=si, field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                ocm=ops["ocm"], delay=b["delay"], method=args.method)
    run = OPOBatch([inst], precision=args.precision, trip_cap=trip_cap, ctx=ctx,
                   tune="auto" if args.tune == "auto" else None)
    geo = run.geometry(0)
    # warm-up trips (also builds the one-trip graph)
    if args.warmup > 0:
        run.graph_replay(1, args.warmup)
        run.sync()
    X0 = run.read(0)["field"]                  # crystal input of the first timed trip
    t0 = time.perf_counter()
    run.graph_replay(1, args.steps)
    run.sync()
    wall = time.perf_counter() - t0
    ev_ms = run.last_ms()                      # HIP events around the replays
    r = run.read(0)
    log = r["trip_log"][args.warmup:args.warmup + args.steps]
    att = int(log[:, 2].sum())
    acc = int(log[:, 3].sum())
    rej = att - acc
    # the first timed trip again from the same input (the kernel is
    # deterministic), for its output-coupled spectrum
    run.write_field(0, X0)
    run.launch(1)
    run.sync()
    oc_run = run.read(0)["oc_spec"]
    run.close()
    # the other field precision from the same input: its first trip's output
    # and its rate over the same number of graph-replayed trips
    other = "double" if args.precision == "single" else "single"
    run2 = OPOBatch([dict(inst, field0=X0)], precision=other, trip_cap=args.steps + 3, ctx=ctx)
    run2.launch(1)
    run2.sync()
    oc_other = run2.read(0)["oc_spec"]
    run2.graph_replay(1, 1)
    run2.sync()
    t1 = time.perf_counter()
    run2.graph_replay(1, args.steps)
    run2.sync()
    wall2 = time.perf_counter() - t1
    run2.close()
    bytes_total = algorithmic_bytes(n, acc, rej, args.steps)
    per_launch_ms = ev_ms / args.steps
    achieved = bytes_total / args.steps / (per_launch_ms * 1e-3) / 1e9
    value = args.steps / wall
    rates = {args.precision: value, other: args.steps / wall2}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "single" else "f64",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise)",
        "config": {
            "workload": "synchronously pumped ring OPO, N=2^15, 12 ps window, 1040 nm; 5 mm "
                        "MgO:PPLN 30.5 um; 2 W @ 1 GHz 200 fs sech^2; 5% OC (flat R 0.95); "
                        "0.3 ps delay" + ("" if args.method == "heun" else
                                          "; RK23 (ode23 variant) crystal solver"),
            "instances": 1,
            "adaptive_steps_per_trip": att / args.steps,
            "rejected_per_trip": rej / args.steps,
            "cluster_workgroups": geo["cluster"],
            "block": geo["block"],
            "n1": geo["n1"],
            "tune": args.tune,
        },
        "fp32_round_trips_per_s": rates["single"],
        "fp64_round_trips_per_s": rates["double"],
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(("ring_n32768_f32" if args.precision == "single"
                                         else "ring_n32768_f64") +
                                        ("" if args.method == "heun" else "_rk23")),
        },
    }
    if not args.no_cpu_baseline:
        base, oc_oracle, (o_att, o_acc) = cpu_baseline(b, X0, args.method)
        out["cpu_baseline"] = base
        f32 = oc_run if args.precision == "single" else oc_other
        f64 = oc_other if args.precision == "single" else oc_run
        rel = lambda a, ref: float(np.linalg.norm(a - ref) / np.linalg.norm(ref))
        out["parity"] = {
            "against": "FP64 CPU oracle (oracle.opo_run): one round trip from the first timed "
                       "trip's crystal input; relative L2 of the output-coupled spectrum",
            "rel_l2_fp32": rel(f32, oc_oracle),
            "rel_l2_fp64": rel(f64, oc_oracle),
            "rel_l2_fp32_power_spectrum": rel(np.abs(f32) ** 2, np.abs(oc_oracle) ** 2),
            "oracle_adaptive_steps": o_att,
            "oracle_accepted_steps": o_acc,
        }
    ctx.close()
    print(json.dumps(out))


def _geometry_report(run, meta, tune):
    rep = {}
    if tune == "auto":
        from optifax_amd.autotune import tune_batch
        for n, best in getattr(tune_batch, "last_report", {}).items():
            rep[str(n)] = {k: best[k] for k in ("cluster", "n1", "block", "us_per_step",
                                                 "candidates", "tuning_s")}
        return rep
    for i, (n, _) in enumerate(meta):
        if str(n) not in rep:
            g = run.geometry(i)
            rep[str(n)] = {"cluster": g["cluster"], "n1": g["n1"], "block": g["block"]}
    return rep


def main_batch(args):
    """The uneven batch: all instances in one persistent launch per round trip."""
    from optifax_amd.solver import Context, OPOBatch
    ctx = Context(0)
    t_build = time.perf_counter()
    insts, meta = batch_workload()
    t_build = time.perf_counter() - t_build
    trip_cap = args.warmup + args.steps + 1
    tune = "auto" if args.tune == "auto" else None
    t_tune = time.perf_counter()
    run = OPOBatch(insts, precision=args.precision, trip_cap=trip_cap, ctx=ctx, tune=tune)
    t_tune = time.perf_counter() - t_tune
    print(f"batch: {len(insts)} instances built in {t_build:.1f} s, batch created in "
          f"{t_tune:.1f} s, grid {run.geometry(0)['grid']}", file=sys.stderr, flush=True)
    if args.warmup > 0:
        run.launch(args.warmup)
        run.sync(allow_nonfinite=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        run.launch(1)
        run.sync(allow_nonfinite=True)
        print(f"batch: timed trip {k + 1}/{args.steps} {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    att = acc = 0
    bytes_total = 0.0
    att_inst = []
    for i, (n, L) in enumerate(meta):
        log = run.read(i)["trip_log"][args.warmup:args.warmup + args.steps]
        a_i, c_i = int(log[:, 2].sum()), int(log[:, 3].sum())
        att += a_i
        acc += c_i
        att_inst.append(a_i / args.steps)
        bytes_total += algorithmic_bytes(n, c_i, a_i - c_i, args.steps)
    inst_trips = len(insts) * args.steps
    achieved = bytes_total / wall / 1e9
    out = {
        "metric": METRIC,
        "value": inst_trips / wall,
        "unit": "instance round trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "single" else "f64",
        "data": "synthetic (sech^2 pump + seeded one-photon-per-mode noise)",
        "config": {
            "workload": "256 uneven OPO instances in one persistent launch: period 28-32 um, "
                        "length 1-10 mm, grid 2^14-2^16 at dt = 12 ps/2^15, detuning +-50 fs "
                        "(seed 3)",
            "instances": len(insts),
            "adaptive_steps_per_instance_trip": att / inst_trips,
            "tune": args.tune,
            "grid_workgroups": run.geometry(0)["grid"],
            "geometry_per_grid": _geometry_report(run, meta, args.tune),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    run.close()
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(insts, meta, att_inst)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
