"""Parity of the HIP path (through the C-ABI) with the oracle: bit-exact
values, bit-exact indices (threshold ties to the lowest position).

Cases: the committed golden fixtures, the reference's own end-to-end shapes
(test/test_algorithm.jl) on seeded data, pass-level intermediates
(test/test_count_bin.jl ramps, test/test_select_bin.jl), edge cases, and
size-independent properties at full configuration sizes."""
import sys

import numpy as np
import pytest

import oracle as O
import radik
from radik import DeviceArray

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_topk(x, lens, k, flags, idx=None, idx_dtype=np.int32):
    v, i = O.topk_exact(x, lens, k, flags, idx_in=idx, idx_dtype=idx_dtype)
    return v.T, i.T  # [k, T] like the device outputs


def run_ws(x, lens, k, flags, idx=None, idx_dtype=np.int32, vload=32, ws=None):
    T = len(lens)
    d_x = DeviceArray.from_host(np.asarray(x, np.float32))
    d_i = DeviceArray.from_host(np.asarray(idx, idx_dtype)) if idx is not None else None
    vo = DeviceArray.zeros((k, T), np.float32)
    io = DeviceArray.zeros((k, T), idx_dtype)
    if ws is None:
        ws = radik.RadiKWorkspace(max(max(lens), 1), T)
    radik.topk_radix_select_(vo, io, ws, d_x, d_i, lens, k, LARGEST=bool(flags & O.LARGEST),
                             ASCEND=bool(flags & O.ASCEND), WITHSCALE=bool(flags & O.WITHSCALE),
                             WITHIDXIN=idx is not None, VLOAD_SIZE_BYTES=vload)
    return vo.to_host(), io.to_host()


def assert_same(got, exp, name):
    gv, gi = got
    ev, ei = exp
    assert gv.shape == ev.shape, name
    bad = np.nonzero(gv.view(np.uint32) != ev.view(np.uint32))
    assert bad[0].size == 0, f"{name}: {bad[0].size} value mismatches, first at {bad[0][0]},{bad[1][0]}"
    bad = np.nonzero(gi.astype(np.int64) != ei.astype(np.int64))
    assert bad[0].size == 0, f"{name}: {bad[0].size} index mismatches, first at {bad[0][0]},{bad[1][0]}"


def test_device_is_gfx950():
    info = radik.device_info()
    say("device", info, radik.lds_caps())
    assert info["arch"].startswith("gfx950")


def test_wave_and_block_scan():
    # counterpart of test/test_utils.jl:53-106 (@warpreduce) on wave64
    import ctypes
    from radik import _lib
    for x in (np.arange(1, 257, dtype=np.uint32),
              np.random.default_rng(0).integers(0, 1000, 4096).astype(np.uint32)):
        n = x.size
        d = DeviceArray.from_host(x)
        w = DeviceArray.zeros(n, np.uint32)
        b = DeviceArray.zeros(n, np.uint32)
        _lib.check(_lib.load().radik_diag_scan(d.ptr, w.ptr, b.ptr, n, None))
        ew = np.concatenate([np.cumsum(c) for c in x.reshape(-1, 64)]).astype(np.uint32)
        eb = np.concatenate([np.cumsum(c) for c in x.reshape(-1, 256)]).astype(np.uint32)
        assert np.array_equal(w.to_host(), ew)
        assert np.array_equal(b.to_host(), eb)
        if x[0] == 1:
            assert w.to_host()[:64].tolist() == np.cumsum(np.arange(1, 65)).tolist()


def _diag_pass1(x, lens, k, flags):
    from radik import _lib
    T = len(lens)
    d = DeviceArray.from_host(np.asarray(x, np.float32))
    h = DeviceArray.zeros(T * 4096, np.uint32)
    s = DeviceArray.zeros(T * 4, np.uint32)
    _lib.check(_lib.load().radik_diag_pass1(d.ptr, _lib.i64_array(lens), T, k, flags, h.ptr, s.ptr, None))
    return h.to_host().reshape(T, 4096), s.to_host().reshape(T, 4)


@pytest.mark.parametrize("flags", [O.LARGEST | O.WITHSCALE, O.LARGEST, 0, O.WITHSCALE])
def test_pass1_histogram_ramps(known_answers, flags):
    # test/test_count_bin.jl:11-61: the three ramps; exact integer histograms
    ramps = known_answers["count_bin_ramps"]["ramps"]
    parts = [np.arange(lo, hi + 1, dtype=np.float32) for lo, hi in ramps]
    lens = [p.size for p in parts]
    say("pass1 ramps", flags)
    h, sel = _diag_pass1(np.concatenate(parts), lens, 100, flags)
    for t, p in enumerate(parts):
        exp = O.rank_histogram(p, flags, 20, 12)
        assert np.array_equal(h[t].astype(np.int64), exp)
        # bin selection on the same histogram (test/test_select_bin.jl rule)
        b, krem, cnt = O.select_bin(exp, 100, False)
        b -= 1
        if krem == cnt:
            assert sel[t, 0] == b + 1 and sel[t, 1] == 0xFFFFFFFF
        else:
            assert (sel[t, 0], sel[t, 1], sel[t, 2], sel[t, 3]) == (b, b, krem, cnt)


def test_pass1_histogram_random_and_ties():
    rng = np.random.default_rng(9)
    lens = [1, 4097, 100000, 5, 300000, 0, 77777]
    x = rng.standard_normal(sum(lens)).astype(np.float32)
    x[:50000] = np.round(x[:50000] * 4) / 4
    say("pass1 random")
    h, _ = _diag_pass1(x, lens, 64, O.LARGEST)
    off = 0
    for t, n in enumerate(lens):
        assert np.array_equal(h[t].astype(np.int64), O.rank_histogram(x[off:off + n], O.LARGEST, 20, 12))
        off += n


def test_golden_cases_workspace(golden_cases):
    for c in golden_cases:
        say("golden ws", c["name"])
        got = run_ws(c["x"], list(c["lens"]), c["k"], c["flags"], idx=c["idx"])
        assert_same(got, (c["vals"].T, c["inds"].T), c["name"])


def test_golden_cases_plan(golden_cases):
    for c in golden_cases:
        if c["idx"] is not None:
            continue
        for use_graph in (True, False):
            say("golden plan", c["name"], use_graph)
            lens = list(c["lens"])
            T, k = len(lens), c["k"]
            fl = c["flags"]
            p = radik.Plan(lens, k, largest=bool(fl & O.LARGEST), rev=not (fl & O.ASCEND),
                           withscale=bool(fl & O.WITHSCALE), use_graph=use_graph)
            d_x = DeviceArray.from_host(c["x"])
            vo = DeviceArray.zeros((k, T), np.float32)
            io = DeviceArray.zeros((k, T), np.int32)
            for rep in range(3):  # graph replays must all match
                vo.fill(np.float32(-7))
                p.execute(vo, io, d_x)
                radik.synchronize()
                assert_same((vo.to_host(), io.to_host()), (c["vals"].T, c["inds"].T), f"{c['name']} rep{rep}")


@pytest.mark.parametrize("largest", [True, False])
@pytest.mark.parametrize("rev", [True, False])
def test_reference_basic_grid(largest, rev):
    # test/test_algorithm.jl:6-40: n = 2^14, K in 16..4096, largest x rev
    rng = np.random.default_rng(42 + 2 * largest + rev)
    n = 1 << 14
    for k in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096):
        x = rng.random(n, dtype=np.float32)
        say("basic", largest, rev, k)
        d = DeviceArray.from_host(x)
        v, i = radik.topk(d, k, largest=largest, rev=rev)
        v, i = v.to_host(), i.to_host()
        flags = (O.LARGEST if largest else 0) | (0 if rev else O.ASCEND)
        ev, ei = oracle_topk(x, [n], k, flags)
        assert np.array_equal(v, ev[:, 0]) and np.array_equal(i, ei[:, 0])
        # the reference's own assertions
        exp = np.sort(np.partition(x, n - k)[n - k:] if largest else np.partition(x, k - 1)[:k])
        assert np.array_equal(v, exp[::-1] if rev else exp)
        assert np.all((1 <= i) & (i <= n)) and len(np.unique(i)) == k
        assert np.array_equal(x[i - 1], v)


def test_reference_multitask_workspace_reuse():
    # test/test_algorithm.jl:42-96: reused workspace, global idx_in
    lens = [1000, 2000, 1500, 500, 5000]
    k = 64
    ws = radik.RadiKWorkspace(5000, 5, np.int32)
    rng = np.random.default_rng(4)
    for rep in range(2):
        x = rng.random(sum(lens), dtype=np.float32)
        idx = np.arange(1, sum(lens) + 1, dtype=np.int32)
        say("multitask", rep)
        got = run_ws(x, lens, k, O.LARGEST | O.ASCEND, idx=idx, ws=ws)
        assert_same(got, oracle_topk(x, lens, k, O.LARGEST | O.ASCEND, idx=idx), "multitask")
        off = 0
        for t, n in enumerate(lens):
            part = x[off:off + n]
            assert np.array_equal(got[0][:, t], np.sort(np.partition(part, n - k)[n - k:]))
            assert np.array_equal(x[got[1][:, t] - 1], got[0][:, t])
            off += n


def test_reference_overloads_and_batch():
    # test/test_algorithm.jl:98-150
    rng = np.random.default_rng(5)
    x = rng.random(1000, dtype=np.float32)
    d = DeviceArray.from_host(x)
    v1, i1 = radik.topk(d, 100, largest=True, rev=False)
    ind = DeviceArray.from_host(np.arange(1, 1001, dtype=np.int32))
    v2, i2 = radik.topk(d, 100, indices=ind, largest=True, rev=False)
    assert np.array_equal(v1.to_host(), v2.to_host()) and np.array_equal(i1.to_host(), i2.to_host())
    lens = [1000, 1500, 800]
    xb = rng.random(sum(lens), dtype=np.float32)
    db = DeviceArray.from_host(xb)
    va, ia = radik.topk(db, lens, 50)
    seq = DeviceArray.from_host(np.concatenate([np.arange(1, n + 1, dtype=np.int32) for n in lens]))
    vb, ib = radik.topk(db, lens, 50, indices=seq)
    assert np.array_equal(va.to_host(), vb.to_host()) and np.array_equal(ia.to_host(), ib.to_host())
    assert_same((va.to_host(), ia.to_host()), oracle_topk(xb, lens, 50, O.LARGEST | O.ASCEND), "batch")


def test_reference_errors_on_gpu():
    # test/test_algorithm.jl:259-269
    d = DeviceArray.from_host(np.random.default_rng(0).standard_normal(1000).astype(np.float32))
    with pytest.raises(radik.RadiKError):
        radik.topk_(DeviceArray.zeros((50, 1), np.float32), DeviceArray.zeros((50, 1), np.int32), d, [1000], 100)
    with pytest.raises(radik.RadiKError):
        radik.topk(d, 100, indices=DeviceArray.from_host(np.arange(1, 501, dtype=np.int32)))


def test_edge_cases():
    rng = np.random.default_rng(6)
    cases = [
        ("empty_and_tiny", rng.standard_normal(3).astype(np.float32), [0, 1, 2, 0], 2),
        ("k_equals_n", rng.standard_normal(5000).astype(np.float32), [5000], 5000),
        ("k_one_ties", np.full(9000, 1.5, np.float32), [9000], 1),
        ("all_equal_large", np.full(300000, -2.0, np.float32), [300000], 1000),
        ("signed_zeros", np.tile(np.array([0.0, -0.0], np.float32), 6000), [12000], 6001),
        ("int64_idx", rng.standard_normal(20000).astype(np.float32), [20000], 100),
    ]
    for name, x, lens, k in cases:
        for flags in (O.LARGEST | O.ASCEND, 0):
            say("edge", name, flags)
            idt = np.int64 if name == "int64_idx" else np.int32
            got = run_ws(x, lens, k, flags, idx_dtype=idt)
            assert_same(got, oracle_topk(x, lens, k, flags, idx_dtype=idt), name)


def test_vload_widths_and_determinism():
    rng = np.random.default_rng(8)
    lens = [70001, 123457, 9999]
    x = (np.round(rng.standard_normal(sum(lens)) * 8) / 8).astype(np.float32)
    exp = oracle_topk(x, lens, 300, O.LARGEST | O.ASCEND)
    for vload in (4, 16, 32, 64):
        say("vload", vload)
        assert_same(run_ws(x, lens, 300, O.LARGEST | O.ASCEND, vload=vload), exp, f"vload{vload}")


def test_large_single_task_properties():
    # n = 2^24 (scaled from the 2^28 configuration): exact against the oracle
    n = 1 << 24
    rng = np.random.default_rng(2828)
    x = rng.standard_normal(n, dtype=np.float32)
    perm = rng.permutation(n).astype(np.int32)
    for k in (16, 1024, 4096):
        for flags in (O.LARGEST, O.LARGEST | O.ASCEND, 0):
            say("large", k, flags)
            got = run_ws(x, [n], k, flags, idx=perm)
            assert_same(got, oracle_topk(x, [n], k, flags, idx=perm), f"2^24 k{k} f{flags}")


def test_predicted_pass_covers_ties_and_specials():
    # long tasks (>= 2^17, k <= len/64) take the sampled single-read path
    rng = np.random.default_rng(31)
    n = 1 << 18
    x = rng.standard_normal(4 * n).astype(np.float32)
    x[n:2 * n][rng.random(n) < 0.3] = np.float32(3.5)          # heavy tie near the threshold
    x[2 * n:3 * n] = np.round(x[2 * n:3 * n] * 2) / 2            # few levels
    sp = rng.integers(3 * n, 4 * n, 500)
    x[sp] = np.array([0.0, -0.0, np.inf, -np.inf, 1.4e-45], np.float32)[rng.integers(0, 5, 500)]
    lens = [n, n, n, n]
    for k in (1, 256, 4096):
        for flags in (O.LARGEST | O.ASCEND, 0, O.LARGEST | O.WITHSCALE):
            say("predicted", k, flags)
            assert_same(run_ws(x, lens, k, flags), oracle_topk(x, lens, k, flags), f"pred k{k} f{flags}")


def test_prediction_miss_forces_reread():
    # The sampled positions (group j of G starts at j * len / G) hold large
    # values that are absent elsewhere, so the sample overstates how many
    # large keys the task has and the predicted bound is too tight; the
    # fallback re-read must still give the exact result.
    n = 1 << 20
    k = n // 64
    G = 64
    x = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    for j in range(G):
        st = j * n // G
        x[st:st + 256] = 100.0 + np.arange(j * 256, j * 256 + 256, dtype=np.float32)
    say("prediction miss")
    assert_same(run_ws(x, [n], k, O.LARGEST | O.ASCEND), oracle_topk(x, [n], k, O.LARGEST | O.ASCEND), "miss")
    # An exact tie seen only by the sample (4 entries at the start of each
    # sample group): the bound narrowed to a position covers 16 of the 256
    # ties, the spill falls short of m and the task is re-read.
    x = np.random.default_rng(4).standard_normal(n).astype(np.float32)
    for j in range(G):
        st = j * n // G
        x[st:st + 4] = np.float32(5.0)
    say("position-bound miss")
    for k in (256, 200):
        assert_same(run_ws(x, [n], k, O.LARGEST), oracle_topk(x, [n], k, O.LARGEST), f"position miss k{k}")


def test_heavy_ties_at_the_threshold():
    # Tens of thousands of exact ties in the threshold bin: the candidate set
    # exceeds the finishing block's LDS, so the krem-th candidate is bracketed
    # from a sample of the spill (ties resolve by lowest position).  Also a
    # constant task and one batch mixing tie-heavy and plain tasks.
    rng = np.random.default_rng(4242)
    n = 1 << 20
    x = rng.standard_normal(n).astype(np.float32)
    x[rng.random(n) < 0.05] = np.float32(3.5)
    for k in (64, 256, 3000):
        for flags in (O.LARGEST, O.LARGEST | O.ASCEND, O.LARGEST | O.WITHSCALE):
            say("heavy ties", k, flags)
            assert_same(run_ws(x, [n], k, flags), oracle_topk(x, [n], k, flags), f"ties k{k} f{flags}")
    c = np.full(1 << 18, 1.25, np.float32)
    say("constant task")
    assert_same(run_ws(c, [c.size], 100, O.LARGEST), oracle_topk(c, [c.size], 100, O.LARGEST), "constant")
    lens = [1 << 19, 1 << 17, 1 << 18, 70000]
    xs = rng.standard_normal(sum(lens)).astype(np.float32)
    xs[: 1 << 19][rng.random(1 << 19) < 0.04] = np.float32(3.0)
    xs[(1 << 19) + (1 << 17):][rng.random((1 << 18) + 70000) < 0.03] = np.float32(2.75)
    say("tie mix")
    assert_same(run_ws(xs, lens, 256, O.LARGEST), oracle_topk(xs, lens, 256, O.LARGEST), "tie mix")
    # many tasks: the finishing blocks keep fewer candidates in LDS
    lens = [1 << 16] * 300
    xs = rng.standard_normal(sum(lens)).astype(np.float32)
    xs[rng.random(xs.size) < 0.05] = np.float32(3.25)
    say("tie batch")
    assert_same(run_ws(xs, lens, 256, O.LARGEST), oracle_topk(xs, lens, 256, O.LARGEST), "tie batch")


def test_heavy_ties_reread_tasks():
    # Tie-heavy tasks that are read a second time inside one block (medium
    # tasks beside the passes, and long unsampled ones: k > len / 64): more
    # candidates than the block's LDS, bounded by a sample of the candidates
    # the re-read left in the spill region.
    rng = np.random.default_rng(4343)
    lens = [40000, 50000, 65000, 3000, 30000]
    xs = rng.standard_normal(sum(lens)).astype(np.float32)
    xs[rng.random(xs.size) < 0.08] = np.float32(3.0)
    for k, flags in ((256, O.LARGEST), (100, O.LARGEST | O.ASCEND), (256, O.LARGEST | O.WITHSCALE), (1000, 0)):
        say("medium ties", k, flags)
        assert_same(run_ws(xs, lens, k, flags), oracle_topk(xs, lens, k, flags), f"medium ties k{k} f{flags}")
    n = 200000
    x = rng.standard_normal(n).astype(np.float32)
    x[rng.random(n) < 0.10] = np.float32(2.2)  # ~20000 ties, beyond a single task's 16384 LDS entries
    for k in (4000, 9000):
        say("unsampled ties", k)
        assert_same(run_ws(x, [n], k, O.LARGEST), oracle_topk(x, [n], k, O.LARGEST), f"unsampled ties k{k}")


def test_uneven_batch_with_duplicates():
    # The uneven configuration's data at reduced scale (200 tasks of 2^16..2^20
    # elements, a quarter with 30 % of their entries on 8 fixed values, +-0 /
    # subnormal / +-Inf specials) through the batched plan: exact ties across
    # the threshold get a spill bound with a position from the sample.
    rng = np.random.default_rng(2020)
    T = 200
    lens = np.round(2.0 ** (16 + 4 * rng.random(T))).astype(np.int64)
    sigma = (10.0 ** (-3 + 6 * rng.random(T))).astype(np.float32)
    dup = rng.random(T) < 0.25
    x = rng.standard_normal(int(lens.sum()), dtype=np.float32)
    tid = np.repeat(np.arange(T), lens)
    x *= sigma[tid]
    fixed = np.arange(1, 9, dtype=np.float32) * np.float32(0.5)
    m = dup[tid] & (rng.random(x.size, dtype=np.float32) < 0.3)
    x[m] = fixed[rng.integers(0, 8, int(m.sum()))] * sigma[tid[m]]
    sp = np.nonzero(rng.random(x.size) < 1e-4)[0]
    x[sp] = np.array([0.0, -0.0, 1.4e-45, -1.4e-45, np.inf, -np.inf], np.float32)[rng.integers(0, 6, sp.size)]
    d = DeviceArray.from_host(x)
    for k, largest, rev in ((256, True, False), (1, True, True), (1000, True, False), (256, False, True)):
        flags = (O.LARGEST if largest else 0) | (0 if rev else O.ASCEND)
        say("uneven dups", k, largest, rev)
        p = radik.Plan(list(lens), k, largest=largest, rev=rev)
        vo = DeviceArray.zeros((k, T), np.float32)
        io = DeviceArray.zeros((k, T), np.int32)
        for rep in range(2):
            p.execute(vo, io, d)
            radik.synchronize()
            assert_same((vo.to_host(), io.to_host()), oracle_topk(x, list(lens), k, flags),
                        f"uneven dups k{k} largest{largest} rev{rev} rep{rep}")


def test_large_k_multi_block_sort():
    # k above 4096 sorts in several blocks (runs + ranks); ties and specials
    rng = np.random.default_rng(16384)
    n = 1 << 20
    x = rng.standard_normal(n).astype(np.float32)
    x[rng.integers(0, n, 20000)] = np.float32(2.5)  # a heavy tie inside the top-k
    x[rng.integers(0, n, 300)] = np.array([np.inf, -0.0, 0.0], np.float32)[rng.integers(0, 3, 300)]
    for k in (4097, 16384, 70000):
        for flags in (O.LARGEST | O.ASCEND, O.LARGEST, 0):
            say("large k", k, flags)
            assert_same(run_ws(x, [n], k, flags), oracle_topk(x, [n], k, flags), f"k{k} f{flags}")
    # one launch mixing result sizes below and above 4096 (per-task k)
    lens = [n, 300000, 70000, 9000]
    ks = [20000, 100, 5000, 9000]
    xs = rng.standard_normal(sum(lens)).astype(np.float32)
    p = radik.Plan(lens, ks, largest=True, rev=False)
    d = DeviceArray.from_host(xs)
    vo = DeviceArray.zeros(sum(ks), np.float32)
    io = DeviceArray.zeros(sum(ks), np.int32)
    p.execute(vo, io, d)
    radik.synchronize()
    v, i = vo.to_host(), io.to_host()
    off = oo = 0
    for L, k in zip(lens, ks):
        ev, ei = O.topk_exact(xs[off:off + L], [L], k, O.LARGEST | O.ASCEND)
        assert np.array_equal(v[oo:oo + k].view(np.uint32), ev[0].view(np.uint32)), (L, k)
        assert np.array_equal(i[oo:oo + k], ei[0]), (L, k)
        off += L
        oo += k


def test_nan_inputs_follow_total_order():
    # NaN policy: a NaN is ranked by its key (IEEE totalOrder), so +NaN is the
    # largest and -NaN the smallest value; payload NaNs keep their bits.
    # Short, medium, long (sampled) tasks; largest, smallest and WITHSCALE
    # (a NaN first element scales by 0).
    rng = np.random.default_rng(2718)
    nans = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0xFF800001, 0x7FFFFFFF, 0xFFFFFFFF, 0x7FC12345],
                    dtype=np.uint32).view(np.float32)
    lens = [1000, 30000, 1 << 18, 1 << 20, 3000]
    x = rng.standard_normal(sum(lens)).astype(np.float32)
    for frac, (a, b) in zip((0.001, 0.01, 0.0001, 0.00002, 0.2),
                            np.stack([np.cumsum([0] + lens[:-1]), np.cumsum(lens)], 1)):
        m = np.nonzero(rng.random(b - a) < frac)[0] + a
        x[m] = nans[rng.integers(0, 7, m.size)]
    x[lens[0]] = np.float32(np.nan)  # the second task's first element (its WITHSCALE scaler)
    for k in (1, 100, 2000):
        for flags in (O.LARGEST | O.ASCEND, O.LARGEST, 0, O.ASCEND, O.LARGEST | O.WITHSCALE):
            say("nan", k, flags)
            assert_same(run_ws(x, lens, k, flags), oracle_topk(x, lens, k, flags), f"nan k{k} f{flags}")
    p = radik.Plan(lens, 100, largest=True, rev=False)
    d = DeviceArray.from_host(x)
    vo = DeviceArray.zeros((100, len(lens)), np.float32)
    io = DeviceArray.zeros((100, len(lens)), np.int32)
    p.execute(vo, io, d)
    radik.synchronize()
    assert_same((vo.to_host(), io.to_host()), oracle_topk(x, lens, 100, O.LARGEST | O.ASCEND), "nan plan")
