"""Autotuner for the batched plan's launch options.

Picks, per length bucket, the vector-load width (bytes in flight per lane per
step of the streaming passes), the streaming blocks per CU (up to the
resident count equal chunks, above it chunks shrinking towards the end of
the pass), the longest task run whole in one block (k_small) and the
candidates held in LDS by the finishing kernel.  Results do not depend on these options, only speed does.

The search is coordinate descent: one knob at a time over its candidates,
keeping the best, timed with HIP events over a few executions of the plan on
the caller's data.  tuning.json (next to this file) maps a bucket key to the
chosen options; Plan applies the entry of its bucket when one exists.
"""
import json
import math
import os

from . import _lib

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")

SPACE = (
    ("vload_bytes", (16, 32, 64)),
    ("blocks_per_cu", (2, 4, 8, 16)),
    ("small_cap", (1024, 2048, 4096, 8192)),
    ("refine_cap", (2048, 4096, 8192, 16384)),
)
DEFAULT = {"vload_bytes": 32, "blocks_per_cu": 8, "small_cap": 0, "refine_cap": 0}


def bucket_key(task_lens, ks):
    """Bucket of a plan: log2 of the mean task length, log4 of the task count,
    log4 of the largest k."""
    lens = [int(x) for x in task_lens] or [0]
    mean = max(1.0, sum(lens) / len(lens))
    kmax = max([int(k) for k in ks] or [1])
    return (f"len2^{int(round(math.log2(mean)))}_tasks4^{int(round(math.log(max(1, len(lens)), 4)))}"
            f"_k4^{int(round(math.log(max(1, kmax), 4)))}")


def load_table(path=TABLE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_table(table, path=TABLE):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def apply(plan, table=None):
    """Sets the plan's options from the table entry of its bucket, if any."""
    table = load_table() if table is None else table
    entry = table.get(bucket_key(plan.task_lens, plan.ks))
    if entry:
        plan.set_options(**entry["options"])
        return entry["options"]
    return None


def _time(radik, plan, vo, io, d_x, d_i, steps):
    from .device import Event
    for _ in range(max(3, steps)):  # (the first executions of a staging run slower)
        plan.execute(vo, io, d_x, d_i)
    radik.synchronize()
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(steps):
        plan.execute(vo, io, d_x, d_i)
    e1.record()
    e1.synchronize()
    return e0.elapsed_ms(e1) / steps


def tune(plan, vo, io, d_x, d_i=None, steps=5, space=SPACE, log=None):
    """Coordinate descent over `space`; returns (best options, ms, trials)."""
    import radik
    best = dict(DEFAULT)
    plan.set_options(**best)
    best_ms = _time(radik, plan, vo, io, d_x, d_i, steps)
    trials = [(dict(best), best_ms)]
    for name, values in space:
        for v in values:
            if v == best[name]:
                continue
            cand = dict(best)
            cand[name] = v
            try:
                plan.set_options(**cand)
                ms = _time(radik, plan, vo, io, d_x, d_i, steps)
            except _lib.RadiKError:
                continue
            trials.append((cand, ms))
            if log:
                log(f"tune {cand} {ms:.4f} ms")
            if ms < best_ms:
                best, best_ms = cand, ms
    plan.set_options(**best)
    return best, best_ms, trials


def record(table, plan, options, ms):
    table[bucket_key(plan.task_lens, plan.ks)] = {"options": options, "ms": round(ms, 4)}
    return table
