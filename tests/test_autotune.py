"""Launch-geometry autotuner: candidate space, choice and per-instance use."""
import numpy as np
import pytest

from optifax_amd import autotune
from optifax_amd.solver import OPOBatch
from tests.cases import opo_case, rel_l2


def _inst(n=256, window=2e-12, length=0.2e-3):
    cfg, b = opo_case(n=n, window=window, length=length)
    ops = b["ops"]
    return dict(solver=b["solver"], field0=ops["field0"], pump=ops["pump"], cav=ops["cav"],
                ocm=ops["ocm"], delay=cfg.delay)


def test_candidates_cover_valid_splits():
    c = autotune.candidates(1 << 15)
    assert all((1 << 15) % n1 == 0 and C <= min(n1, (1 << 15) // n1) for C, n1, _ in c)
    assert {n1 for _, n1, _ in c} == {128, 256, 512}
    assert {b for _, _, b in c} == {256, 512}
    assert max(C for C, _, _ in c) == 128
    with pytest.raises(ValueError):
        autotune.candidates(1000)


def test_tune_picks_a_launchable_geometry_and_caches(emu_ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("OFX_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._MEM.clear()
    inst = _inst()
    best = autotune.tune(inst, batch=2, precision="double", steps=8, ctx=emu_ctx)
    assert (best["cluster"], best["n1"], best["block"]) in autotune.candidates(256)
    assert best["candidates"] > 4 and best["us_per_step"] >= 0   # the emulation reports no device time
    # second call: served from the cache (memory, then disk)
    again = autotune.tune(inst, batch=2, precision="double", steps=8, ctx=emu_ctx)
    assert again == best
    autotune._MEM.clear()
    assert autotune.tune(inst, batch=2, precision="double", steps=8, ctx=emu_ctx) == best


def test_tuned_uneven_batch_matches_default_geometry(emu_ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("OFX_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._MEM.clear()
    insts = [_inst(256, 2e-12), _inst(512, 3e-12), _inst(256, 2e-12)]
    tuned, block = autotune.tune_batch(insts, precision="double", steps=8, ctx=emu_ctx)
    assert all("cluster" in t and "n1" in t for t in tuned)
    assert tuned[0]["cluster"] == tuned[2]["cluster"]
    a = OPOBatch(tuned, precision="double", trip_cap=2, tune={"block": block}, ctx=emu_ctx)
    b = OPOBatch(insts, precision="double", trip_cap=2, ctx=emu_ctx)
    for run in (a, b):
        run.launch(2)
        run.sync()
    for i in range(3):
        ra, rb = a.read(i), b.read(i)
        assert ra["trips_done"] == rb["trips_done"] == 2
        # a different FFT split only reorders FP64 rounding
        assert rel_l2(ra["field"], rb["field"]) < 1e-9
    ga = a.geometry(1)
    assert (ga["cluster"], ga["n1"]) == (tuned[1]["cluster"], tuned[1]["n1"])
    a.close()
    b.close()


def test_measurement_pass_takes_steps(emu_ctx):
    """The truncated pass the tuner times must run (a chunk longer than the
    truncated crystal would leave nothing to time)."""
    inst = _inst()
    proto = autotune._truncate(inst, 12)
    assert proto["solver"]["nSteps"] == 12 and proto["solver"]["sel"] <= 12
    from optifax_amd.solver import OPOBatch
    run = OPOBatch([dict(proto, cluster=2)], precision="double", ctx=emu_ctx)
    run.launch(0)
    run.sync()
    assert run.read(0)["attempts"] >= 3
    run.close()


def test_measure_counts_each_launch_once(emu_ctx, monkeypatch):
    """The run's attempt counter is cumulative over its launches: every
    repeat must be divided by its own attempts, not the running total."""
    from optifax_amd import solver
    monkeypatch.setattr(solver.OPOBatch, "last_ms", lambda self: 1.0)
    inst = _inst()
    t1 = autotune.measure(inst, (2, 16, 256), 1, "double", steps=12, ctx=emu_ctx, repeats=1)
    t3 = autotune.measure(inst, (2, 16, 256), 1, "double", steps=12, ctx=emu_ctx, repeats=3)
    assert t1 is not None and t3 == t1
