"""bench.py's contract pieces that run without a GPU."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

from cases import opo_case  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and "one GPU only" in p.stderr


def test_algorithmic_bytes_model():
    # 24 N per accepted step, 16 N per rejected step, 40 N per round trip
    assert bench.algorithmic_bytes(1024, 10, 3, 2) == 1024 * (24 * 10 + 16 * 3 + 40 * 2)


def test_batch_cpu_baseline_runs_one_instance_per_core():
    insts, meta = [], []
    for n in (256, 512, 256):
        cfg, b = opo_case(n=n, window=2e-12 * n / 256, length=0.1e-3)
        ops = b["ops"]
        insts.append(dict(solver=b["solver"], field0=ops["field0"]))
        meta.append((n, cfg.length))
    r = bench.cpu_baseline_batch(insts, meta, [100.0, 120.0, 90.0], sample_steps=50)
    assert r["kind"] == "port" and r["cores"] >= 1
    assert r["value"] > 0 and r["single_core_value"] > 0
    assert "2^8" in r["sample"] and "2^9" in r["sample"]


def test_batch_cpu_baseline_on_fewer_cores_than_grid_sizes(monkeypatch):
    insts, meta = [], []
    for n in (256, 512, 1024):
        cfg, b = opo_case(n=n, window=2e-12 * n / 256, length=0.05e-3)
        insts.append(dict(solver=b["solver"], field0=b["ops"]["field0"]))
        meta.append((n, cfg.length))
    monkeypatch.setattr(bench.os, "sched_getaffinity", lambda pid: {0}, raising=False)
    r = bench.cpu_baseline_batch(insts, meta, [50.0, 50.0, 50.0], sample_steps=20)
    assert r["cores"] == 1 and r["value"] > 0


def test_ring_cpu_baseline_is_one_oracle_round_trip_from_the_device_state():
    """cavity_state inverts the pump addition and refraction, so the oracle
    round trip it starts takes the device's crystal input; the baseline's
    output spectrum equals the oracle's own for that trip."""
    import oracle
    from cases import oracle_next_input, rel_l2
    from optifax_amd.physics import SimWindow
    cfg, b = opo_case(n=256, window=2e-12, length=0.1e-3)
    sw, si = b["simwin"], b["solver"]
    R, T, phi = b["oc"].masks(sw)
    o1 = oracle.opo_run(1, b["noise"], b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi,
                        b["delay"], sw.rel_omegas, si)
    X = oracle_next_input(b, o1["E_cav"])          # a device-style crystal input
    assert rel_l2(oracle_next_input(b, bench.cavity_state(b, X)), X) < 1e-13
    rec, oc, (att, acc) = bench.cpu_baseline(b, X)
    o2 = oracle.opo_run(1, o1["E_cav"], b["pump"], np.ones(sw.n), si["n_xtal"], R, T, phi,
                        b["delay"], sw.rel_omegas, si)
    assert rel_l2(oc, SimWindow.fft_order(SimWindow.t2k(o2["E_oc"]))) < 1e-12
    assert rec["kind"] == "port" and rec["cores"] == 1 and rec["value"] > 0
    assert (att, acc) == tuple(int(v) for v in o2["steps"][0])
