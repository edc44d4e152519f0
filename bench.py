"""Benchmark of the radix top-k path on one MI355X.

Default workload: the uneven batch (4096 tasks, lengths log-uniform in
[2^10, 2^20], k = 256, per-task sigma log-uniform in [1e-3, 1e3], 30 %
duplicates of 8 fixed values in a quarter of the tasks, a 1e-4 fraction of
+-0 / +-min subnormal / +-Inf) through the persistent batched form (one plan,
hipGraph replay).  One step = one top-k pass over the whole batch with the
input resident in HBM.

The JSON line also carries, for the dominant kernel, its roofline
(algorithmic bytes per launch / its mean event-timed duration) and the CPU
baseline (the oracle's restatement of the reference schedule on host cores,
on a bounded sample of the same workload), plus secondary lines for the
uniform batch, the 2^28 single task and the per-task plain form.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-secondary]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "radik.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

SPECIALS = np.array([0.0, -0.0, 1.4e-45, -1.4e-45, np.inf, -np.inf], dtype=np.float32)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def uneven_batch(T=4096, seed=4096):
    rng = np.random.default_rng(seed)
    lens = np.round(2.0 ** (10 + 10 * rng.random(T))).astype(np.int64)
    sigma = (10.0 ** (-3 + 6 * rng.random(T))).astype(np.float32)
    dup_task = rng.random(T) < 0.25
    total = int(lens.sum())
    x = rng.standard_normal(total, dtype=np.float32)
    tid = np.repeat(np.arange(T, dtype=np.int32), lens)
    x *= sigma[tid]
    fixed = np.arange(1, 9, dtype=np.float32) * np.float32(0.5)
    m = dup_task[tid] & (rng.random(total, dtype=np.float32) < 0.3)
    x[m] = fixed[rng.integers(0, 8, int(m.sum()))] * sigma[tid[m]]
    del m
    sp = np.nonzero(rng.random(total, dtype=np.float32) < 1e-4)[0]
    x[sp] = SPECIALS[rng.integers(0, 6, sp.size)]
    return x, lens


def dominant_roofline(plan, phase_ms, steps, N, out_pairs, idx_bytes):
    """Roofline of the slower of the two streaming kernels (each reads the
    input once: 4 B/element; compaction also writes the winner pairs)."""
    hist_ms = phase_ms["hist"] / steps
    comp_ms = phase_ms["compact"] / steps
    if comp_ms >= hist_ms:
        name, t_ms, byts = "k_compact", comp_ms, 4.0 * N + (4 + idx_bytes) * out_pairs
    else:
        name, t_ms, byts = "k_hist", hist_ms, 4.0 * N
    achieved = byts / (t_ms * 1e-3)
    return name, {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                  "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic_from_profiles(name)}


def traffic_from_profiles(kernel, workload="uneven"):
    """HBM bytes per launch of `kernel` on `workload`, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries (scaled by the copy-kernel
    calibration), if any."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None


def time_plan(radik, plan, vo, io, d_x, steps, warmup):
    from radik import Event
    for _ in range(warmup):
        plan.execute(vo, io, d_x)
    radik.synchronize()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        plan.execute(vo, io, d_x)
    e1.record()
    e1.synchronize()
    radik.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_ms(e1) / steps, wall / steps * 1e3


def phase_times(radik, plan, vo, io, d_x, steps, d_i=None):
    plan.set_timing(True)
    for _ in range(steps):
        plan.execute(vo, io, d_x, d_i)
    radik.synchronize()
    ms, n = plan.timing()
    plan.set_timing(False)
    return ms, n


def cpu_baseline(x, lens, k, flags, max_tasks=256):
    """Oracle's restatement of the reference schedule on host cores, on the
    first max_tasks tasks of the workload (threads = min(16, cores))."""
    import oracle as O
    sub = lens[:max_tasks]
    n = int(sub.sum())
    xs = x[:n]
    threads = max(1, min(16, os.cpu_count() or 1))
    O.topk_schedule(xs[:100000], [100000], k, flags, threads=threads)  # warm
    times = []
    t_end = time.perf_counter() + 20.0
    while len(times) < 3 or (time.perf_counter() < t_end and len(times) < 10):
        t0 = time.perf_counter()
        O.topk_schedule(xs, sub, k, flags, threads=threads)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end:
            break
    t = float(np.median(times))
    return {"value": round(n / t / 1e9, 4), "unit": "Gelem/s", "cores": threads, "kind": "port",
            "sample": f"first {len(sub)} tasks of the uneven batch ({n} elements), k={k}, "
                      f"reference 3-pass schedule restated in C (oracle/radik_oracle.c), OpenMP over tasks, "
                      f"median of {len(times)} runs"}


def secondary(radik, O, warmup, steps):
    from radik import DeviceArray, Event
    out = {}
    # uniform batch 1000 x 1000, k = 100 (and the 64-level tie variant)
    rng = np.random.default_rng(1001)
    u = rng.random(1000 * 1000, dtype=np.float32)
    for name, data in (("uniform_1000x1000_k100", u), ("ties64_1000x1000_k100", np.floor(u * 64) / 64)):
        d = DeviceArray.from_host(data.astype(np.float32))
        p = radik.Plan([1000] * 1000, 100)
        vo = DeviceArray.zeros((100, 1000), np.float32)
        io = DeviceArray.zeros((100, 1000), np.int32)
        ms, _ = time_plan(radik, p, vo, io, d, max(steps, 20), warmup)
        out[name] = {"ms": round(ms, 4), "Gelem_s": round(1e6 / (ms * 1e-3) / 1e9, 2)}
    # single task 2^28, Int32 custom indices (a permutation), k grid
    n = 1 << 28
    log("secondary: 2^28 single task")
    rng = np.random.default_rng(2828)
    x = rng.standard_normal(n, dtype=np.float32)
    perm = np.random.default_rng(2829).permutation(n).astype(np.int32)
    d = DeviceArray.from_host(x)
    di = DeviceArray.from_host(perm)
    del perm
    res = {}
    for k in (16, 1024, 4096, 16384):
        p = radik.Plan([n], k, indices_dtype=np.int32)
        vo = DeviceArray.zeros((k, 1), np.float32)
        io = DeviceArray.zeros((k, 1), np.int32)
        for _ in range(warmup):
            p.execute(vo, io, d, di)
        radik.synchronize()
        e0, e1 = Event(), Event()
        e0.record()
        for _ in range(steps):
            p.execute(vo, io, d, di)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        ph, cnt = phase_times(radik, p, vo, io, d, steps, di)
        B = 8.0 * n + 8.0 * k
        r = {"ms": round(ms, 4), "Gelem_s": round(n / (ms * 1e-3) / 1e9, 1),
             "path_hbm_frac": round(B / (ms * 1e-3) / HBM_PEAK, 4)}
        if ph:
            r["phase_ms"] = {kk: round(v / cnt, 4) for kk, v in ph.items()}
        res[f"k{k}"] = r
        # spot check against the oracle at k = 1024
        if k == 1024:
            v = vo.to_host()[:, 0]
            ev, _ = O.topk_exact(x, [n], k, O.LARGEST | O.ASCEND)
            r["matches_oracle"] = bool(np.array_equal(v.view(np.uint32), ev[0].view(np.uint32)))
    out["single_2p28"] = res
    return out


def tune_all(radik, out_path):
    """Autotunes the plan options on the benchmark workloads; writes the table."""
    from radik import DeviceArray, autotune
    table = autotune.load_table()
    x, lens = uneven_batch()
    work = [("uneven", x, lens, 256, None)]
    rng = np.random.default_rng(2828)
    n = 1 << 28
    x3 = rng.standard_normal(n, dtype=np.float32)
    for k in (16, 1024, 4096, 16384):
        work.append((f"single2^28_k{k}", x3, np.array([n]), k, None))
    u = np.random.default_rng(1001).random(1000 * 1000, dtype=np.float32)
    work.append(("uniform1000x1000", u, np.array([1000] * 1000), 100, None))
    rows = np.random.default_rng(5150).integers(1, 257, 8)
    for r in (1, 16, 128):
        xs = (np.random.default_rng(int(r)).standard_normal(r * 128256) * 3).astype(np.float32)
        work.append((f"rows{r}x128256", xs, np.array([128256] * r), 50, None))
    for name, xx, ll, k, _ in work:
        d = DeviceArray.from_host(xx)
        T = len(ll)
        p = radik.Plan(ll, k, tuned=False)
        vo = DeviceArray.zeros((k, T), np.float32)
        io = DeviceArray.zeros((k, T), np.int32)
        best, ms, trials = autotune.tune(p, vo, io, d, steps=5, log=log)
        autotune.record(table, p, best, ms)
        log(f"tuned {name}: {best} {ms:.4f} ms (default {trials[0][1]:.4f} ms)")
        del d
    autotune.save_table(table, out_path)
    autotune.save_table(table)
    print(json.dumps(table), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--only-batched", action="store_true",
                    help="time the batched plan only (profiling runs)")
    ap.add_argument("--tune", default="",
                    help="run the autotuner on the bench workloads and write the table to this path")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="also run the plain copy kernel over this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.only_batched:
        args.no_secondary = args.no_cpu = args.no_plain = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the radix top-k path runs on a single MI355X here "
              "(multi-GPU sharding is a design note, not a measured path).", file=sys.stderr)
        sys.exit(2)

    import oracle as O
    import radik
    from radik import DeviceArray

    info = radik.device_info()
    log("device", info, radik.lds_caps())
    if args.copy_bytes:
        from radik import _lib
        nb = args.copy_bytes // 16 * 16
        src = DeviceArray.zeros(nb // 4, np.float32)
        dst = DeviceArray.zeros(nb // 4, np.float32)
        for _ in range(5):
            _lib.check(_lib.load().radik_diag_copy(dst.ptr, src.ptr, nb, None))
        del src, dst
        log(f"copy kernel: 5 x {nb} bytes")
    if args.tune:
        tune_all(radik, args.tune)
        return
    t0 = time.perf_counter()
    x, lens = uneven_batch()
    N = int(lens.sum())
    T = len(lens)
    k = 256
    log(f"uneven batch: T={T} N={N} generated in {time.perf_counter() - t0:.1f}s")
    d_x = DeviceArray.from_host(x)
    vo = DeviceArray.zeros((k, T), np.float32)
    io = DeviceArray.zeros((k, T), np.int32)
    plan = radik.Plan(lens, k, largest=True, rev=False)
    ms, wall_ms = time_plan(radik, plan, vo, io, d_x, args.steps, args.warmup)
    value = N / (ms * 1e-3) / 1e9
    log(f"batched: {ms:.3f} ms/step ({value:.1f} Gelem/s)")

    # parity spot check of the timed output (a few tasks) against the oracle
    vh, ih = vo.to_host(), io.to_host()
    off = np.concatenate([[0], np.cumsum(lens)])
    check_ok = True
    for t in (0, 1, T // 2, T - 1):
        ev, ei = O.topk_exact(x[off[t]:off[t + 1]], [lens[t]], k, O.LARGEST | O.ASCEND)
        check_ok &= bool(np.array_equal(vh[:, t].view(np.uint32), ev[0].view(np.uint32)) and
                         np.array_equal(ih[:, t], ei[0]))

    phase, cnt = phase_times(radik, plan, vo, io, d_x, args.steps)
    kname, roof = dominant_roofline(plan, phase, cnt, N, k * T, 4)
    path_frac = (8.0 * N + 8.0 * k * T) / (ms * 1e-3) / HBM_PEAK

    extra = {"phase_ms": {kk: round(v / cnt, 4) for kk, v in phase.items()}, "dominant_kernel": kname,
             "path_hbm_frac": round(path_frac, 4), "wall_ms_per_step": round(wall_ms, 4),
             "oracle_spot_check": check_ok}

    if not args.no_plain:
        # plain form: one drop-in call per task (synchronous, as the reference)
        ws = radik.RadiKWorkspace(int(lens.max()), 1)
        t1 = time.perf_counter()
        ntask = 0
        for t in range(T):
            xt = d_x.view(int(off[t]), int(lens[t]))
            radik.topk_radix_select_(vo.column(t), io.column(t), ws, xt, None, [int(lens[t])], k)
            ntask += 1
            if time.perf_counter() - t1 > 30:
                break
        plain_s = time.perf_counter() - t1
        plain_rate = float(off[ntask]) / plain_s / 1e9
        extra["plain_per_task"] = {"Gelem_s": round(plain_rate, 2), "tasks_timed": ntask}
        extra["batched_over_plain"] = round(value / plain_rate, 2)
        log(f"plain per-task: {plain_rate:.2f} Gelem/s over {ntask} tasks")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, lens, k, O.LARGEST | O.ASCEND)
        log("cpu baseline", cpu)
    del x

    if not args.no_secondary:
        extra["secondary"] = secondary(radik, O, args.warmup, args.steps)

    line = {
        "metric": "Gelem/s (batched top-k, uneven batch) + % HBM peak",
        "value": round(value, 3),
        "unit": "Gelem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy, generated on host, resident in HBM before timing)",
        "config": {"workload": "uneven batch: 4096 tasks, len log-uniform [2^10,2^20] (seed 4096), k=256, "
                               "largest, rev=false, Int32 indices, persistent batched plan (hipGraph)",
                   "total_elements": N, "num_tasks": T, "k": k, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
