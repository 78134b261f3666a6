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
