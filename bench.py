"""Benchmark of the radix top-k path on one MI355X.

Default workload: the uneven batch (4096 tasks, lengths log-uniform in
[2^10, 2^20], k = 256, per-task sigma log-uniform in [1e-3, 1e3], 30 %
duplicates of 8 fixed values in a quarter of the tasks, a 1e-4 fraction of
+-0 / +-min subnormal / +-Inf) through the persistent batched form (one plan,
staged once, its kernels launched eagerly each step).  One step = one top-k
pass over the whole batch with the input resident in HBM.

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


def dominant_roofline(phase_ms, steps, info, out_pairs, idx_bytes, workload):
    """Roofline of the dominant kernel, pass A (k_hist): it reads every
    element of the long tasks once (4 B/element; the short and medium tasks
    are read by k_small / k_medium on the side stream).  When the re-read
    pass k_compact takes longer it is the dominant one instead (4 B/element
    of the long tasks plus the winner pairs it writes)."""
    n_long = info["long_elements"]
    hist_ms = phase_ms["hist"] / steps
    comp_ms = phase_ms["compact"] / steps
    if comp_ms >= hist_ms:
        name, t_ms, byts = "k_compact", comp_ms, 4.0 * n_long + (4 + idx_bytes) * out_pairs
    else:
        name, t_ms, byts = "k_hist", hist_ms, 4.0 * n_long
    achieved = byts / (t_ms * 1e-3)
    return name, {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                  "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                  "traffic": traffic_from_profiles(name, workload)}, byts


def path_fracs(ms, N, k, T, idx_bytes=4):
    """Whole-path HBM fractions: under the two-read model of a radix select
    with candidate spill (8 B/element + the k output pairs per task), and
    with the one read of the input this pipeline makes (4 B/element + the
    output pairs).  The second is the one that cannot exceed 1."""
    out = (4 + idx_bytes) * k * T
    t = ms * 1e-3
    return {"path_hbm_frac": round((8.0 * N + out) / t / HBM_PEAK, 4),
            "path_hbm_frac_one_read": round((4.0 * N + out) / t / HBM_PEAK, 4)}


def traffic_from_profiles(kernel, workload="uneven"):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE passes (each scaled by what the counter
    reads for the plain copy kernel), or None.  profiles/traffic.json names
    the library source hash it was measured on; a figure measured on other
    kernel sources is not reported."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernels_sha256_16") != kernels_hash():
        return None
    return t.get(workload, {}).get(kernel)


def kernels_hash():
    """First 16 hex digits of the SHA-256 of the kernel sources."""
    import hashlib
    h = hashlib.sha256()
    for name in ("radik_kernels.hip", "radik_pipeline.hip", "radik_device.h", "radik_host.h"):
        with open(os.path.join(ROOT, "radik.jl_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


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
    for _ in range(2):  # (the first phase-timed executions run slower)
        plan.execute(vo, io, d_x, d_i)
    radik.synchronize()
    plan.set_timing(True)  # (resets the sums)
    for _ in range(steps):
        plan.execute(vo, io, d_x, d_i)
    radik.synchronize()
    ms, n = plan.timing()
    plan.set_timing(False)
    return ms, n


def _median_time(fn, budget_s, min_runs=5, max_runs=2000):
    fn()
    fn()
    fn()  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < min_runs or (time.perf_counter() < t_end and len(times) < max_runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times)


def host_threads():
    """Host threads for the CPU baseline: the CPU share the environment gives
    this process (OMP_NUM_THREADS when set), else every core."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return os.cpu_count() or 1


def cpu_baseline(x4, lens4, k4, flags):
    """The reference's 3-pass schedule restated in C (oracle/radik_oracle.c,
    OpenMP) on the host: the headline line is the plain single-task
    configuration (n = 1e6 N(0,1), seed 42, k = 100, largest, automatic
    indices) on every thread the process may use, the element range split
    across threads.  Beside it: the textbook line (quickselect + sort of
    (key, position) pairs, one thread: nth_element + sort) on the same task,
    and the schedule over the first 256 tasks of the uneven batch (threads
    over tasks)."""
    import oracle as O
    threads = host_threads()
    x1 = np.random.default_rng(42).standard_normal(1_000_000, dtype=np.float32)
    t_sched, r1 = _median_time(lambda: O.topk_schedule(x1, [x1.size], 100, O.LARGEST | O.ASCEND,
                                                          threads=threads), 8.0)
    t_text, r2 = _median_time(lambda: O.topk_exact(x1, [x1.size], 100, O.LARGEST | O.ASCEND, threads=1), 4.0)
    sub = lens4[:256]
    n = int(sub.sum())
    xs = x4[:n]
    t_c4, r3 = _median_time(lambda: O.topk_schedule(xs, sub, k4, flags, threads=threads), 8.0, min_runs=3,
                            max_runs=10)
    return {"value": round(x1.size / t_sched / 1e9, 4), "unit": "Gelem/s", "cores": threads, "kind": "port",
            "sample": (f"n=1e6 N(0,1) seed 42, k=100, largest, rev=false (the plain single-task configuration): "
                       f"reference 3-pass schedule restated in C (oracle/radik_oracle.c), OpenMP over the "
                       f"element range, {threads} threads, median of {r1} runs, {t_sched * 1e3:.3f} ms"),
            "textbook_nth_element_sort": {"value": round(x1.size / t_text / 1e9, 4), "unit": "Gelem/s",
                                          "cores": 1, "ms": round(t_text * 1e3, 3), "runs": r2},
            "uneven_first_256_tasks": {"value": round(n / t_c4 / 1e9, 4), "unit": "Gelem/s", "cores": threads,
                                       "elements": n, "runs": r3}}


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
        # the drop-in topk! into the same pre-allocated buffers (synchronous)
        t_drop, _ = _median_time(lambda: radik.topk_(vo, io, d, [1000] * 1000, 100), 1.0, max_runs=200)
        out[name] = {"ms": round(ms, 4), "Gelem_s": round(1e6 / (ms * 1e-3) / 1e9, 2),
                     "topk_drop_in_ms": round(t_drop * 1e3, 4)}
    # the plain single task: n = 1e6 N(0,1) (seed 42), k = 100, drop-in topk
    x1 = np.random.default_rng(42).standard_normal(1_000_000, dtype=np.float32)
    d1 = DeviceArray.from_host(x1)
    t1, _ = _median_time(lambda: radik.topk(d1, 100), 1.0, max_runs=500)
    v1, i1 = radik.topk(d1, 100)
    e1v, e1i = O.topk_exact(x1, [x1.size], 100, O.LARGEST | O.ASCEND)
    out["single_1e6_k100_topk"] = {"ms": round(t1 * 1e3, 4), "Gelem_s": round(1e6 / t1 / 1e9, 3),
                                   "matches_oracle": bool(np.array_equal(v1.to_host().view(np.uint32),
                                                                         e1v[0].view(np.uint32)) and
                                                          np.array_equal(i1.to_host(), e1i[0]))}
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
        r = {"ms": round(ms, 4), "Gelem_s": round(n / (ms * 1e-3) / 1e9, 1)}
        r.update(path_fracs(ms, n, k, 1))
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
        best, ms, trials = autotune.tune(p, vo, io, d, steps=10, log=log)
        autotune.record(table, p, best, ms)
        log(f"tuned {name}: {best} {ms:.4f} ms (default {trials[0][1]:.4f} ms)")
        del d
    autotune.save_table(table, out_path)
    autotune.save_table(table)
    print(json.dumps(table), flush=True)


def spiky_rows(rng, rows, cols):
    """N(0, 3) rows with 0-4 entries per row set to exactly +20."""
    m = (rng.standard_normal(rows * cols, dtype=np.float32) * np.float32(3)).astype(np.float32)
    nsp = rng.integers(0, 5, rows)
    r = np.repeat(np.arange(rows), nsp)
    m[r * cols + rng.integers(0, cols, r.size)] = np.float32(20)
    return m


def serving(radik, seconds=3.0, seed=5150, pool_size=12, budget=16 << 30, inflight=4):
    """Poisson arrivals of rows x 128256 requests (rows U{1..256}, k in {1, 50,
    1024}), served (a) by one synchronous drop-in call per request in arrival
    order and (b) by the continuous-batching server, at the same frozen rates:
    lambda = 0.7 / (mean plain service time) and a saturating 1.5 / (mean
    plain service time).  Latency = arrival -> results complete; p50/p99 and
    rows/s for both forms at both loads."""
    from radik import DeviceArray
    cols = 128256
    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(pool_size):
        rows = int(rng.integers(1, 257))
        k = int(rng.choice([1, 50, 1024]))
        x = DeviceArray.from_host(spiky_rows(rng, rows, cols))
        pool.append((rows, k, x, DeviceArray.zeros((k, rows), np.float32), DeviceArray.zeros((k, rows), np.int32)))
    ws = radik.RadiKWorkspace(cols, 256)

    def plain(req):
        rows, k, x, vo, io = req
        radik.topk_radix_select_(vo, io, ws, x, None, [cols] * rows, k, LARGEST=True, ASCEND=False)

    for req in pool:  # warm and calibrate
        plain(req)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        for req in pool:
            plain(req)
    s_mean = (time.perf_counter() - t0) / (reps * len(pool))

    def wait_until(t):
        while time.perf_counter() < t:
            pass

    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 3)
    srv = radik.Server(budget_bytes=budget, max_inflight=inflight, largest=True, rev=True)
    for req in pool:  # warm the slots' buffers
        srv.wait(srv.submit(req[2], req[0], cols, req[1], req[3], req[4]))
    out = {"plain_service_ms": round(s_mean * 1e3, 4), "budget_bytes": budget, "slots": inflight}
    for name, load in (("load70", 0.7), ("load150", 1.5)):
        lam = load / s_mean
        n_req = max(50, int(seconds * lam))
        arr = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / lam, n_req))
        which = np.random.default_rng(seed + 2).integers(0, len(pool), n_req)
        rows_total = int(sum(pool[w][0] for w in which))
        lat_plain = np.empty(n_req)
        t0 = time.perf_counter()
        for i in range(n_req):
            wait_until(t0 + arr[i])
            plain(pool[which[i]])
            lat_plain[i] = time.perf_counter() - (t0 + arr[i])
        span_plain = time.perf_counter() - t0
        st0 = srv.stats()
        tickets = np.empty(n_req, dtype=np.int64)
        lag = np.empty(n_req)
        t0 = time.perf_counter()
        for i in range(n_req):
            wait_until(t0 + arr[i])
            lag[i] = time.perf_counter() - (t0 + arr[i])  # submit delay behind the arrival
            rows, k, x, vo, io = pool[which[i]]
            tickets[i] = srv.submit(x, rows, cols, k, vo, io)
        lat_srv = np.array([srv.wait(int(t)) for t in tickets]) * 1e-3 + lag
        span_srv = time.perf_counter() - t0
        st = srv.stats()
        out[name] = {
            "lambda_req_s": round(lam, 1), "requests": n_req,
            "plain": {"p50_ms": pct(lat_plain, 50), "p99_ms": pct(lat_plain, 99),
                      "rows_s": round(rows_total / span_plain, 1)},
            "served": {"p50_ms": pct(lat_srv, 50), "p99_ms": pct(lat_srv, 99),
                       "rows_s": round(rows_total / span_srv, 1), "launches": st["launches"] - st0["launches"],
                       "max_coalesced": st["max_coalesced"]},
        }
        log(name, out[name])
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--only-batched", action="store_true",
                    help="time the batched plan only (profiling runs)")
    ap.add_argument("--only-serving", action="store_true", help="run the serving comparison only")
    ap.add_argument("--workload", default="uneven", choices=("uneven", "single"),
                    help="batched-plan workload: the uneven batch, or one 2^28 task (profiling runs)")
    ap.add_argument("--k", type=int, default=256)
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
    if args.only_serving:
        print(json.dumps(serving(radik)), flush=True)
        return
    t0 = time.perf_counter()
    if args.workload == "single":
        x = np.random.default_rng(2828).standard_normal(1 << 28, dtype=np.float32)
        lens = np.array([1 << 28])
    else:
        x, lens = uneven_batch()
    N = int(lens.sum())
    T = len(lens)
    k = args.k
    log(f"{args.workload}: T={T} N={N} generated in {time.perf_counter() - t0:.1f}s")
    d_x = DeviceArray.from_host(x)
    vo = DeviceArray.zeros((k, T), np.float32)
    io = DeviceArray.zeros((k, T), np.int32)
    plan = radik.Plan(lens, k, largest=True, rev=False)
    ms, wall_ms = time_plan(radik, plan, vo, io, d_x, args.steps, args.warmup)
    value = N / (ms * 1e-3) / 1e9
    log(f"batched: {ms:.3f} ms/step ({value:.1f} Gelem/s)")

    # parity of the whole timed output against the oracle (values and
    # indices bit-exact); a mismatch voids the measurement
    vh, ih = vo.to_host(), io.to_host()
    off = np.concatenate([[0], np.cumsum(lens)])
    t_chk = time.perf_counter()
    ev, ei = O.topk_exact(x, lens, k, O.LARGEST | O.ASCEND, threads=host_threads())
    bad_v = int(np.count_nonzero(vh.view(np.uint32) != ev.T.view(np.uint32)))
    bad_i = int(np.count_nonzero(ih.astype(np.int64) != ei.T.astype(np.int64)))
    check_ok = bad_v == 0 and bad_i == 0
    del ev, ei
    log(f"oracle check of all {T} tasks: {'ok' if check_ok else 'MISMATCH'} "
        f"({bad_v} values, {bad_i} indices differ; {time.perf_counter() - t_chk:.1f}s)")

    info = plan.info()
    phase, cnt = phase_times(radik, plan, vo, io, d_x, args.steps)
    kname, roof, kbytes = dominant_roofline(phase, cnt, info, k * T, 4, args.workload)

    extra = {"phase_ms": {kk: round(v / cnt, 4) for kk, v in phase.items()}, "dominant_kernel": kname,
             "dominant_kernel_bytes": int(kbytes), "plan_info": info,
             **path_fracs(ms, N, k, T), "wall_ms_per_step": round(wall_ms, 4),
             "oracle_check": {"tasks": T, "ok": check_ok, "values_differ": bad_v, "indices_differ": bad_i}}
    if roof["traffic"]:
        roof_ratio = roof["traffic"] / kbytes
        extra["dominant_traffic_over_algorithmic"] = round(roof_ratio, 4)

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
        log("serving")
        extra["secondary"]["serving_rows128256"] = serving(radik)
        log(extra["secondary"]["serving_rows128256"])

    line = {
        "metric": "Gelem/s (batched top-k, uneven batch) + % HBM peak",
        "value": round(value, 3) if check_ok else None,
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
        "config": {"workload": ("uneven batch: 4096 tasks, len log-uniform [2^10,2^20] (seed 4096), "
                                f"k={k}, largest, rev=false, Int32 indices, persistent batched plan (eager launches)"
                                if args.workload == "uneven" else
                                f"single task: 2^28 N(0,1) (seed 2828), k={k}, largest, rev=false, "
                                "persistent batched plan (eager launches)"),
                   "total_elements": N, "num_tasks": T, "k": k, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
