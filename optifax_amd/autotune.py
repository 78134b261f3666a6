"""Launch-geometry autotuner for the persistent crystal-pass kernel.

For a grid size N, a precision and the number of instances B of that size
that will share the GPU, the tunable schedule of the kernel is

  cluster  C   workgroups per instance (the field is split over C LDS tiles)
  n1           four-step split N = n1 * n2 (row FFTs of n1 and n2 points,
               radix-4 Stockham passes with a radix-2 tail when log2 is odd)
  block        threads per workgroup

Every candidate that fits the LDS limit is timed on the caller's own
instance, truncated to its first `steps` minimal steps, with k copies in one
launch, k = min(B, slots // C) (the number that run at once).  The makespan
of B instances is then ceil(B / k) * t_step(C, n1, block, k) per step.  The
lowest makespan wins.  Results are cached per (device, N, precision, B) in
memory and in a JSON file (OFX_TUNE_CACHE, default ~/.cache/optifax_amd/
tune.json).
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib

_MEM: dict = {}


def _cache_path():
    return os.environ.get("OFX_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "optifax_amd",
                                       "tune.json"))


def _load_cache():
    p = _cache_path()
    try:
        with open(p) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key, value):
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        d = _load_cache()
        d[key] = value
        tmp = p + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def candidates(n, max_cluster=128):
    """(cluster, n1, block) triples to try for grid size n."""
    lg = int(round(math.log2(n)))
    if 1 << lg != n:
        raise ValueError("n must be a power of two")
    splits = sorted({(lg + 1) // 2, (lg + 1) // 2 + 1, lg // 2})
    out = []
    for s1 in splits:
        n1 = 1 << s1
        n2 = n // n1
        if n1 < 4 or n2 < 4:
            continue
        C = 1
        while C <= min(n1, n2, max_cluster):
            for block in (256, 512):
                out.append((C, n1, block))
            C *= 2
    return out


def _truncate(inst, steps):
    s = dict(inst["solver"])
    ns = int(s["nSteps"])
    k = max(1, min(ns, int(steps)))
    s["nSteps"] = k
    s["G33"] = np.asarray(s["G33"])[:k]
    # one chunk covering the truncated pass (a longer chunk would leave no
    # complete chunk to run), no snapshots; alpha_w keeps the full crystal's
    # per-step loss
    s["sel"] = float(min(float(s["sel"]), k))
    s["n_cols"] = 0
    return dict(solver=s, field0=inst["field0"], delay=inst.get("delay", 0.0),
                method=inst.get("method", s.get("method", "heun")))


def _time_pass(inst, cand, k, precision, steps, ctx, repeats):
    """(best seconds of one launch, adaptive attempts per copy in it) for k
    concurrent copies of the pass truncated to `steps` minimal steps, or None
    when the candidate cannot launch."""
    from .solver import OPOBatch
    C, n1, block = cand
    proto = _truncate(inst, steps)
    proto.update(cluster=C, n1=n1)
    try:
        run = OPOBatch([proto] * k, precision=precision, tune={"block": block}, ctx=ctx)
    except _lib.OfxError:
        return None
    try:
        best = None
        before = 0
        for _ in range(repeats):
            run.write_field(0, proto["field0"])
            for i in range(1, k):
                run.write_field(i, proto["field0"])
            run.launch(0)
            run.sync(allow_nonfinite=True)
            ms = run.last_ms()
            total = run.read(0)["attempts"]       # cumulative over the run's launches
            att, before = total - before, total
            if att < 1:
                return None
            best = (ms * 1e-3, att) if best is None or ms * 1e-3 < best[0] else best
        return best
    finally:
        run.close()


def measure(inst, cand, k, precision="single", steps=1500, ctx=None, repeats=2):
    """Seconds per adaptive step of one pass with k concurrent copies, or None
    when the candidate cannot launch.  The cost per step is the marginal one
    between passes truncated to `steps` and 3 `steps` minimal steps, so the
    launch's fixed costs (table loads into LDS, the first exchange) do not
    count: a long crystal pass amortises them, a short timing run does not."""
    a = _time_pass(inst, cand, k, precision, steps, ctx, repeats)
    if a is None:
        return None
    b = _time_pass(inst, cand, k, precision, 3 * steps, ctx, repeats)
    if b is None:
        return None
    (t1, n1), (t2, n2) = a, b
    if n2 > n1 and t2 > t1:
        return (t2 - t1) / (n2 - n1)
    return t2 / n2


def tune(inst, batch=1, precision="single", steps=1500, ctx=None, use_cache=True, verbose=False,
         slots=None):
    """Best (cluster, n1, block) for `batch` instances shaped like `inst` on
    `slots` workgroup slots (default: one per CU of the device)."""
    from .solver import default_context
    ctx = ctx or default_context()
    info = ctx.info()
    n = int(np.asarray(inst["solver"]["beta_op"]).size)
    slots = int(slots) if slots else int(info["cus"])
    key = f"{info['name']}|n={n}|{precision}|B={int(batch)}|slots={slots}"
    if use_cache:
        if key in _MEM:
            return dict(_MEM[key])
        disk = _load_cache()
        if key in disk:
            _MEM[key] = disk[key]
            return dict(disk[key])
    results = []
    t0 = time.time()
    for cand in candidates(n):
        C = cand[0]
        k = max(1, min(int(batch), slots // C))
        t = measure(inst, cand, k, precision, steps, ctx)
        if t is None:
            continue
        waves = math.ceil(batch / k)
        results.append((waves * t, t, k, cand))
        if verbose:
            print(f"  C={C:3d} n1={cand[1]:5d} block={cand[2]} k={k:3d}: "
                  f"{t * 1e6:8.2f} us/step, makespan {waves * t * 1e6:8.2f} us/step", flush=True)
    if not results:
        raise RuntimeError(f"no launchable geometry for N={n}")
    results.sort(key=lambda r: r[0])
    span, t, k, (C, n1, block) = results[0]
    best = dict(cluster=C, n1=n1, block=block, us_per_step=t * 1e6, concurrent=k,
                makespan_us_per_step=span * 1e6, candidates=len(results),
                tuning_s=time.time() - t0)
    _MEM[key] = best
    if use_cache:
        _store_cache(key, best)
    return dict(best)


def tune_batch(instances, precision="single", steps=1500, ctx=None, verbose=False):
    """Per-instance geometry for an uneven batch: instances are grouped by N,
    each group tuned for its own size and count (first instance of the group
    as the prototype) on its share of the CUs.  The groups run side by side
    in one launch, so a group is given the share of the workgroup slots that
    its work takes (sum over its instances of N log2 N times the minimal
    steps): tuned against the whole device, a group would pick clusters that
    are fastest alone but crowd out the others.  Returns (instances with
    cluster/n1 set, block)."""
    from .solver import default_context
    groups: dict = {}
    for i, inst in enumerate(instances):
        n = int(np.asarray(inst["solver"]["beta_op"]).size)
        groups.setdefault(n, []).append(i)
    cus = int((ctx or default_context()).info()["cus"])
    work = {n: sum(n * math.log2(n) * float(instances[i]["solver"]["nSteps"]) for i in ids)
            for n, ids in groups.items()}
    total = sum(work.values())
    share = {n: max(1, int(round(cus * w / total))) for n, w in work.items()}
    out = [dict(inst) for inst in instances]
    blocks = {}
    for n, ids in groups.items():
        best = tune(instances[ids[0]], batch=len(ids), precision=precision, steps=steps, ctx=ctx,
                    verbose=verbose, slots=share[n])
        for i in ids:
            out[i]["cluster"] = best["cluster"]
            out[i]["n1"] = best["n1"]
        blocks[n] = (len(ids), best["block"])
    # one block size per launch: the one chosen for the most instances
    block = max(blocks.values())[1]
    tune_batch.last_report = {n: dict(tune(instances[ids[0]], batch=len(ids), precision=precision,
                                           steps=steps, ctx=ctx, slots=share[n]), slots=share[n])
                              for n, ids in groups.items()}
    return out, block
