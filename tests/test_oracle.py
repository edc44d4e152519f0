"""Pins the CPU oracle (oracle/radik_oracle.c) before it checks anything.

Each function is checked against the reference's own known answers
(tests/golden/known_answers.json, transcribed from WilliBee/RadiK.jl test/*.jl),
against library routines (numpy.lexsort / numpy.cumsum / numpy.histogram) or a
closed form, and the two independent top-k restatements against each other.
The mutation tests build oracle variants with one planted mistake each and
require the pins to fail for every one.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def f32(v):
    if isinstance(v, str):
        return np.float32(float(v))
    return np.float32(v)


# ---------------------------------------------------------------- key / bins
def test_get_bin_id_known_answer(known_answers):
    ka = known_answers["get_bin_id_order"]
    bins = [O.bin_id(f32(v), ka["left"], ka["right"]) for v in ka["vals"]]
    assert all(0 <= b < 256 for b in bins)
    assert bins[0] == ka["first"] and bins[-1] == ka["last"]
    assert all(np.diff(np.array(bins, dtype=np.int64)) < 0)
    # IEEE-754 closed form of the same bins
    assert bins == [255, 194, 192, 191, 128, 64, 63, 61, 0]


def _ieee_specials():
    u = np.array([0x00000000, 0x80000000, 0x00000001, 0x80000001, 0x007FFFFF, 0x807FFFFF, 0x00800000,
                  0x80800000, 0x7F7FFFFF, 0xFF7FFFFF, 0x7F800000, 0xFF800000, 0x3F800000, 0xBF800000],
                 dtype=np.uint32)
    return u.view(np.float32)


def test_key_closed_form_and_order():
    vals = _ieee_specials()
    k = np.array([O.key(v) for v in vals], dtype=np.uint32)
    assert np.array_equal(k, O.keys(vals))
    # order: numeric order, with -0 just below +0
    order = sorted(range(len(vals)), key=lambda i: (float(vals[i]), not np.signbit(vals[i])))
    ks = k[order]
    assert np.all(np.diff(ks.astype(np.int64)) > 0)
    # inverse
    for v in vals:
        assert np.float32(O.lib().orc_key_inv(O.key(v))).view(np.uint32) == np.float32(v).view(np.uint32)


def test_bin_id_digits_match_key_bits():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(200).astype(np.float32)
    for left, right in ((0, 20), (12, 20), (24, 24), (0, 30)):
        exp = ((O.keys(x).astype(np.uint64) << np.uint64(left)) & np.uint64(0xFFFFFFFF)) >> np.uint64(right)
        got = np.array([O.bin_id(v, left, right) for v in x], dtype=np.uint64)
        assert np.array_equal(got, exp)


# ---------------------------------------------------------------- scans
def test_warp_scan_known_answers(known_answers):
    ka = known_answers["warp_scan"]
    assert O.warp_scan(np.arange(1, 33), 32, True).tolist() == ka["up_1_32"]
    assert O.warp_scan(np.arange(1, 33), 32, False).tolist() == ka["down_1_32"]
    assert O.warp_scan(np.arange(1, 65), 32, True).tolist() == ka["up_1_64_ws32"]
    # wave64 (AMD): one scan over all 64 lanes
    assert O.warp_scan(np.arange(1, 65), 64, True).tolist() == np.cumsum(np.arange(1, 65)).tolist()
    r = np.random.default_rng(0).integers(0, 100, 256)
    got = O.warp_scan(r, 64, False).reshape(4, 64)
    assert np.array_equal(got, np.cumsum(r.reshape(4, 64)[:, ::-1], axis=1)[:, ::-1])


def test_cross_warp_reduction_known_answers(known_answers):
    for case in known_answers["cross_warp_reduction"]:
        got = O.cross_warp_reduction(case["input"], case["block"], case["warp"], case["largest"])
        assert got.tolist() == case["expected"], case["source"]


def test_find_bin_and_write_known_answers(known_answers):
    for c in known_answers["find_bin_and_write"]:
        got = O.find_bin_and_write(c["counts"], c["old_k"], c["neighbor"], c["sum"], c["thread_x"], c["unroll"],
                                   c["largest"])
        exp = None if c["expected"] is None else tuple(c["expected"])
        assert got == exp, c["source"]


def test_select_bin_known_answers(known_answers):
    for c in known_answers["select_bin"]:
        for t, kk in enumerate(c["k"]):
            h = c["hist"][str(t)]
            exp = tuple(c["expected"][t] if len(c["expected"]) > 1 else c["expected"][0])
            assert O.select_bin(h, kk, c["largest"]) == exp, c["source"]
            for block in c["blocks"]:
                for warp in (32, 64):
                    if block % warp:
                        continue
                    assert O.select_bin_blocked(h, block, warp, kk, c["largest"]) == exp, (c["source"], block)


def test_select_bin_brute_force():
    rng = np.random.default_rng(7)
    for _ in range(200):
        h = rng.integers(0, 5, 256) * (rng.random(256) < 0.3)
        tot = int(h.sum())
        if tot == 0:
            continue
        k = int(rng.integers(1, tot + 1))
        for largest in (True, False):
            hh = h[::-1] if largest else h
            c = np.cumsum(hh)
            j = int(np.searchsorted(c, k))  # first index with cumsum >= k
            b = (255 - j) if largest else j
            before = int(c[j] - hh[j])
            exp = (b + 1, k - before, int(h[b]))
            assert O.select_bin(h, k, largest) == exp
            assert O.select_bin_blocked(h, 256, 64, k, largest) == exp
        assert O.select_bin(h, tot + 1, True) is None  # k beyond the total: nothing


# ---------------------------------------------------------------- histograms
def test_count_bin_ramps(known_answers):
    ka = known_answers["count_bin_ramps"]
    for lo, hi in ka["ramps"]:
        data = np.arange(lo, hi + 1, dtype=np.float32)
        for withscale in (False, True):
            got = O.count_bin_ex(data, ka["left"], ka["right"], ka["histlen"], withscale, True)
            s = data[0] if withscale else np.float32(0)
            vals = (data - s).astype(np.float32) if withscale else data
            exp = np.bincount(O.keys(vals) >> 30, minlength=4)
            assert np.array_equal(got, exp)
            assert got.sum() == data.size


def test_rank_histogram_is_reversed_key_histogram():
    x = np.random.default_rng(1).standard_normal(5000).astype(np.float32)
    hk = np.bincount(O.keys(x) >> 20, minlength=4096)
    assert np.array_equal(O.rank_histogram(x, O.LARGEST, 20, 12), hk[::-1])
    assert np.array_equal(O.rank_histogram(x, 0, 20, 12), hk)


# ---------------------------------------------------------------- filter
def _filter_data(spec):
    if spec == "1:1000":
        return np.arange(1, 1001, dtype=np.float32)
    return np.array(spec, dtype=np.float32)


def test_filter_known_answers(known_answers):
    for c in known_answers["filter"]:
        data = _filter_data(c["data"])
        flags = (O.LARGEST if c["largest"] else 0) | (O.WITHSCALE if c["withscale"] else 0)
        if "kth_scaled_of" in c:
            kth_val = np.float32(c["kth_scaled_of"]) - data[0]
        else:
            kth_val = np.float32(c["kth"])
        kth_rank = (~O.key(kth_val)) & 0xFFFFFFFF if c["largest"] else O.key(kth_val)
        vals, pos = O.filter_task(data, c["k"], kth_rank, c["boundary"], flags)
        assert len(vals) == c["expected_count"], c["source"]
        exp = c["expected_set"]
        if exp == "1:1000":
            exp = list(range(1, 1001))
        assert sorted(vals.tolist()) == sorted(float(v) for v in exp), c["source"]
        if "expected_indices" in c:
            idx = np.array(c["indices"])[pos]
            assert sorted(idx.tolist()) == sorted(c["expected_indices"])


def test_filter_multitask_known_answer(known_answers):
    c = known_answers["filter_multitask"]
    for task, exp in zip(c["tasks"], c["expected_sets"]):
        data = np.array(task, dtype=np.float32)
        kth = data[c["kth_index"] - 1]
        vals, _ = O.filter_task(data, c["k"], (~O.key(kth)) & 0xFFFFFFFF, c["boundary"], O.LARGEST)
        assert sorted(vals.tolist()) == sorted(exp)


# ---------------------------------------------------------------- top-k
def lexsort_topk(x, lens, k, flags, idx_in=None):
    """Library-routine statement of the contract: numpy.lexsort."""
    out_v = np.zeros((len(lens), k), np.float32)
    out_i = np.full((len(lens), k), 0, np.int64)
    off = 0
    for t, n in enumerate(lens):
        v = x[off:off + n]
        s = np.float32(0)
        if flags & O.WITHSCALE and n:
            s = v[0] if np.isfinite(v[0]) else np.float32(0)
        with np.errstate(over="ignore", invalid="ignore"):
            sv = np.where(np.isnan(v), v, (v - s).astype(np.float32)) if flags & O.WITHSCALE else v
        r = O.keys(sv)
        if flags & O.LARGEST:
            r = ~r
        pos = np.arange(n)
        sel = np.lexsort((pos, r))[:min(k, n)]
        ok = O.keys(v[sel])
        if not flags & O.ASCEND:
            ok = ~ok
        order = sel[np.lexsort((sel, ok))]
        m = len(order)
        out_v[t, :m] = v[order]
        if idx_in is not None:
            out_i[t, :m] = idx_in[off + order]
        else:
            out_i[t, :m] = order + 1
        out_i[t, m:] = -1
        off += n
    return out_v, out_i


def _random_case(rng):
    T = int(rng.integers(1, 5))
    lens = rng.integers(0, 400, T)
    k = int(rng.integers(0, 120))
    flags = int(rng.integers(0, 8))
    kind = int(rng.integers(0, 5))
    n = int(lens.sum())
    if kind == 0:
        x = rng.standard_normal(n)
    elif kind == 1:
        x = np.round(rng.standard_normal(n) * 2) / 2
    elif kind == 2:
        x = _ieee_specials()[rng.integers(0, 14, n)]
    elif kind == 3:
        x = 1.0 + rng.integers(0, 30, n) * 2.0 ** -21
    else:
        x = _with_nans(rng, np.round(rng.standard_normal(n) * 4) / 4)
    return np.asarray(x, np.float32), lens, k, flags


def _nans():
    """Quiet and signalling NaNs of both signs, with payloads."""
    return np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0xFF800001, 0x7FFFFFFF, 0xFFFFFFFF, 0x7FC12345],
                    dtype=np.uint32).view(np.float32)


def _with_nans(rng, x):
    x = np.asarray(x, np.float32).copy()
    if x.size:
        m = rng.random(x.size) < 0.1
        x[m] = _nans()[rng.integers(0, 7, int(m.sum()))]
    return x


def test_exact_matches_lexsort():
    rng = np.random.default_rng(11)
    for _ in range(300):
        x, lens, k, flags = _random_case(rng)
        v, i = O.topk_exact(x, lens, k, flags, idx_dtype=np.int64)
        ev, ei = lexsort_topk(x, lens, k, flags)
        assert np.array_equal(v.view(np.uint32), ev.view(np.uint32))
        assert np.array_equal(i, ei)


def test_nan_policy_is_ieee_total_order():
    """NaN policy of the contract: a NaN is ranked by its key, i.e. by IEEE
    754 totalOrder: -NaN < -Inf < ... < -0 < +0 < ... < +Inf < +NaN, larger
    payloads further out.  (The reference skips NaN in its first histogram,
    src/kernels/count_bin.jl:111, but maps it to floatmin/floatmax when
    scaling, src/utils.jl:69; the two disagree, so the contract fixes one.)"""
    vals = np.array([0x7FC00000, 0xFFC00000, 0x7F800000, 0xFF800000, 0x00000000, 0x80000000, 0x3F800000,
                     0xBF800000, 0x7FFFFFFF, 0xFFFFFFFF, 0x7F800001, 0xFF800001], np.uint32).view(np.float32)
    # totalOrder stated on the bits: negatives by descending magnitude bits,
    # then positives by ascending magnitude bits
    u = vals.view(np.uint32).astype(np.int64)
    tot = np.where(u >> 31 == 1, -(u & 0x7FFFFFFF) - 1, u)
    order = np.argsort(tot, kind="stable")
    n = vals.size
    v, i = O.topk_exact(vals, [n], n, O.ASCEND)          # smallest first, ascending
    assert np.array_equal(v[0].view(np.uint32), vals[order].view(np.uint32))
    v, i = O.topk_exact(vals, [n], 3, O.LARGEST)         # 3 largest, descending
    assert np.array_equal(v[0].view(np.uint32), vals[order[::-1][:3]].view(np.uint32))
    assert v[0].view(np.uint32)[0] == 0xFFFFFFFF or np.isnan(v[0][0])
    # WITHSCALE: a NaN first element scales by 0 and NaNs keep their bits
    x = np.array([np.nan, 1.0, -np.nan, 3.0, 2.0], np.float32)
    v, i = O.topk_exact(x, [5], 5, O.LARGEST | O.WITHSCALE | O.ASCEND)
    ev, ei = lexsort_topk(x, [5], 5, O.LARGEST | O.WITHSCALE | O.ASCEND)
    assert np.array_equal(v.view(np.uint32), ev.view(np.uint32)) and np.array_equal(i, ei)
    for fl in (O.LARGEST, 0, O.LARGEST | O.WITHSCALE, O.WITHSCALE | O.ASCEND):
        y = _with_nans(np.random.default_rng(fl), np.random.default_rng(fl + 9).standard_normal(2000))
        for k in (1, 7, 300):
            v, i = O.topk_exact(y, [y.size], k, fl, idx_dtype=np.int64)
            ev, ei = lexsort_topk(y, [y.size], k, fl)
            assert np.array_equal(v.view(np.uint32), ev.view(np.uint32)) and np.array_equal(i, ei)
            a = O.topk_schedule(y, [y.size], k, fl)
            assert np.array_equal(a[0].view(np.uint32), v.view(np.uint32))


def test_exact_brute_force_subsets():
    """n <= 10: the chosen set is the unique size-min(k,n) subset that
    dominates its complement in (rank, position) order."""
    import itertools
    rng = np.random.default_rng(5)
    for _ in range(60):
        n = int(rng.integers(1, 11))
        k = int(rng.integers(1, n + 1))
        x = (np.round(rng.standard_normal(n)) / 2).astype(np.float32)
        flags = int(rng.integers(0, 4))
        r = O.keys(x)
        if flags & O.LARGEST:
            r = ~r
        best = None
        for sub in itertools.combinations(range(n), k):
            rest = [j for j in range(n) if j not in sub]
            if all((int(r[a]), a) < (int(r[b]), b) for a in sub for b in rest):
                assert best is None
                best = set(sub)
        _, i = O.topk_exact(x, [n], k, flags)
        assert set((i[0] - 1).tolist()) == best


def test_schedule_matches_exact():
    rng = np.random.default_rng(13)
    for _ in range(300):
        x, lens, k, flags = _random_case(rng)
        a = O.topk_exact(x, lens, k, flags)
        b = O.topk_schedule(x, lens, k, flags)
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
        assert np.array_equal(a[1], b[1])


def test_reorder_turns_one_output_order_into_the_other():
    # top-k in selection order re-sorted by key equals the oracle's other
    # output order (the 2^28 parity test derives its expectations this way)
    rng = np.random.default_rng(5)
    x = (np.round(rng.standard_normal(50000) * 64) / 64).astype(np.float32)
    x[rng.integers(0, x.size, 300)] = np.array([0.0, -0.0, np.inf], np.float32)[rng.integers(0, 3, 300)]
    for largest in (True, False):
        sel = O.LARGEST if largest else O.ASCEND
        sv, si = O.topk_exact(x, [x.size], 3000, sel)
        other = (O.LARGEST | O.ASCEND) if largest else 0
        ov, oi = O.topk_exact(x, [x.size], 3000, other)
        for k in (1, 17, 3000):
            ev, ei = O.topk_exact(x, [x.size], k, other)
            o = O.reorder(sv[0][:k], descending=not largest)
            assert np.array_equal(sv[0][:k][o].view(np.uint32), ev[0].view(np.uint32))
            assert np.array_equal(si[0][:k][o], ei[0])


def test_schedule_single_task_threads_match_serial():
    # one task on several threads (the C1 CPU baseline): same outputs and
    # pass statistics as the serial schedule, ties in position order
    rng = np.random.default_rng(17)
    cases = [rng.standard_normal(100000), np.round(rng.standard_normal(100000) * 3) / 3,
             np.full(50000, 2.5), _with_nans(rng, rng.standard_normal(80000)),
             _ieee_specials()[rng.integers(0, 14, 60000)]]
    for x in cases:
        x = np.asarray(x, np.float32)
        for k in (1, 100, 4000):
            for flags in (O.LARGEST | O.ASCEND, 0, O.LARGEST | O.WITHSCALE):
                a = O.topk_schedule(x, [x.size], k, flags, threads=1, stats=True)
                b = O.topk_schedule(x, [x.size], k, flags, threads=5, stats=True)
                assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
                assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
                e = O.topk_exact(x, [x.size], k, flags)
                assert np.array_equal(e[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(e[1], b[1])


def test_schedule_pass_counts():
    # all-equal data runs all three passes; distinct values stop after pass 1/2
    x = np.full(5000, 3.0, np.float32)
    _, _, st = O.topk_schedule(x, [5000], 10, O.LARGEST, stats=True)
    assert st[0, 9] == 3 and st[0, 1] == 10 and st[0, 2] == 5000
    x = np.arange(5000, dtype=np.float32)
    _, _, st = O.topk_schedule(x, [5000], 10, O.LARGEST, stats=True)
    assert st[0, 9] >= 1


def test_golden_cases_reproduce(golden_cases):
    for c in golden_cases:
        v, i = O.topk_exact(c["x"], c["lens"], c["k"], c["flags"], idx_in=c["idx"])
        assert np.array_equal(v.view(np.uint32), c["vals"].view(np.uint32)), c["name"]
        assert np.array_equal(i, c["inds"]), c["name"]
        ev, ei = lexsort_topk(c["x"], c["lens"], c["k"], c["flags"], idx_in=c["idx"])
        assert np.array_equal(ev.view(np.uint32), c["vals"].view(np.uint32)), c["name"]
        assert np.array_equal(ei, c["inds"]), c["name"]


# ---------------------------------------------------------------- mutations
def _pins_pass(lib_path):
    """Runs a compact subset of the pins against an oracle build; True if all hold."""
    L = ctypes.CDLL(lib_path)
    old = O._lib
    O._lib = None
    saved = O.LIB_PATH
    try:
        O.LIB_PATH = lib_path
        O.lib()
        ok = True
        try:
            bins = [O.bin_id(f32(v), 0, 24) for v in ["inf", 100.0, 2.0, 1.0, 0.0, -1.0, -2.0, -100.0, "-inf"]]
            ok &= bins == [255, 194, 192, 191, 128, 64, 63, 61, 0]
            ok &= O.select_bin([2, 8, 14, 3] + [0] * 252, 5, True) == (3, 2, 14)
            rng = np.random.default_rng(2)
            for _ in range(40):
                x, lens, k, flags = _random_case(rng)
                v, i = O.topk_exact(x, lens, k, flags, idx_dtype=np.int64)
                ev, ei = lexsort_topk(x, lens, k, flags)
                ok &= np.array_equal(v.view(np.uint32), ev.view(np.uint32)) and np.array_equal(i, ei)
                a = O.topk_schedule(x, lens, k, flags)
                ok &= np.array_equal(a[0].view(np.uint32), v.view(np.uint32))
            # WITHSCALE with an infinite first element
            x = np.array([np.inf, 1, 2, 3, 4], np.float32)
            v, _ = O.topk_exact(x, [5], 2, O.LARGEST | O.WITHSCALE | O.ASCEND)
            ev, _ = lexsort_topk(x, [5], 2, O.LARGEST | O.WITHSCALE | O.ASCEND)
            ok &= np.array_equal(v, ev)
        except Exception:
            ok = False
        return bool(ok)
    finally:
        O._lib = old
        O.LIB_PATH = saved
        del L


@pytest.mark.parametrize("plant", [0, 1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(plant, tmp_path):
    out = str(tmp_path / f"liboracle_p{plant}.so")
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    f"-DORC_PLANT={plant}", "-o", out, os.path.join(ROOT, "oracle", "radik_oracle.c"), "-lm"],
                   check=True)
    assert _pins_pass(out) == (plant == 0)


# ---------------------------------------------------------------- other key types
def _numeric_topk(vals_f, lens, k, flags):
    """numpy statement of the contract for data without NaN: order by numeric
    value (numpy's own compare; -0 below +0), ties by position."""
    out = []
    off = 0
    for n in lens:
        v = np.asarray(vals_f[off:off + n])
        neg0 = np.signbit(v) & (v == 0)
        pos = np.arange(n)
        # rank: largest first (descending value, +0 before -0) or smallest first
        if flags & O.LARGEST:
            sel = np.lexsort((pos, neg0, -v.astype(np.float64)))[:min(k, n)]
        else:
            sel = np.lexsort((pos, ~neg0, v.astype(np.float64)))[:min(k, n)]
        sv = v[sel].astype(np.float64)
        s0 = neg0[sel]
        if flags & O.ASCEND:
            order = sel[np.lexsort((sel, ~s0, sv))]
        else:
            order = sel[np.lexsort((sel, s0, -sv))]
        out.append(order)
        off += n
    return out


def _typed_cases(rng, kt):
    n = int(rng.integers(1, 3000))
    if kt == O.KEY_F16:
        x = (rng.standard_normal(n) * 8).astype(np.float16)
        x[rng.random(n) < 0.2] = np.float16(1.5)            # ties
        x[rng.random(n) < 0.01] = np.array([np.inf, -np.inf, 0.0, -0.0, 6e-8], np.float16)[rng.integers(0, 5)]
        return x, x.astype(np.float32)
    if kt == O.KEY_BF16:
        f = (rng.standard_normal(n) * 100).astype(np.float32)
        f[rng.random(n) < 0.2] = np.float32(3.0)
        b = O.f32_to_bf16(f)
        b[rng.random(n) < 0.01] = np.array([0x7F80, 0xFF80, 0x0000, 0x8000, 0x0001], np.uint16)[rng.integers(0, 5)]
        return b, O.bf16_to_f32(b)
    x = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    x[rng.random(n) < 0.2] = 7
    x[:3] = [np.iinfo(np.int32).min, np.iinfo(np.int32).max, 0]
    return x, x


@pytest.mark.parametrize("kt", [1, 2, 3])
def test_typed_exact_matches_numpy_order(kt):
    rng = np.random.default_rng(100 + kt)
    for _ in range(40):
        x, xf = _typed_cases(rng, kt)
        lens = [x.size // 3, x.size - x.size // 3]
        k = int(rng.integers(1, 400))
        for flags in (O.LARGEST | O.ASCEND, O.LARGEST, 0, O.ASCEND):
            v, i = O.topk_exact_typed(kt, x, lens, k, flags)
            exp = _numeric_topk(xf, lens, k, flags)
            off = 0
            for t, n in enumerate(lens):
                m = len(exp[t])
                assert np.array_equal(i[t, :m], exp[t] + 1)
                assert np.array_equal(v[t, :m].view(np.uint16 if kt in (1, 2) else np.uint32),
                                      x[off + exp[t]].view(np.uint16 if kt in (1, 2) else np.uint32))
                assert np.all(i[t, m:] == -1)
                off += n


def test_typed_keys_closed_form():
    # float16 / bfloat16: the key orders every pattern by IEEE totalOrder
    h = np.arange(1 << 16, dtype=np.uint32)
    k16 = np.array([O.key_typed(int(b), O.KEY_F16) for b in h[::97]], np.int64)
    sign = h[::97] >> 15
    tot = np.where(sign == 1, -(h[::97] & 0x7FFF).astype(np.int64) - 1, h[::97].astype(np.int64))
    assert np.array_equal(np.argsort(k16, kind="stable"), np.argsort(tot, kind="stable"))
    assert O.key_typed(0x8000, O.KEY_F16) < O.key_typed(0x0000, O.KEY_F16)  # -0 < +0
    # int32: two's complement order
    vals = np.array([-2**31, -5, -1, 0, 1, 2**31 - 1], np.int64)
    ks = [O.key_typed(int(np.uint32(v & 0xFFFFFFFF)), O.KEY_I32) for v in vals]
    assert ks == sorted(ks) and len(set(ks)) == len(ks)
    # float32 through the typed entry equals the Float32 oracle
    rng = np.random.default_rng(3)
    x = _with_nans(rng, rng.standard_normal(5000))
    for flags in (O.LARGEST | O.ASCEND, 0):
        a = O.topk_exact(x, [2000, 3000], 150, flags)
        b = O.topk_exact_typed(O.KEY_F32, x, [2000, 3000], 150, flags)
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])


def test_f64_keys_are_ieee_total_order():
    # totalOrder as a signed integer: negative patterns -(magnitude bits) - 1
    rng = np.random.default_rng(64)
    b = rng.integers(0, 2**63, 4000, dtype=np.int64).astype(np.uint64)
    b[::2] |= np.uint64(1 << 63)
    b[:8] = np.array([0, 1 << 63, 0x7FF0000000000000, 0xFFF0000000000000, 0x7FF8000000000000,
                      0xFFF8000000000001, 1, (1 << 63) | 1], np.uint64)
    mag = (b & np.uint64(0x7FFFFFFFFFFFFFFF)).astype(np.int64)
    tot = np.where(b >> np.uint64(63) != 0, -mag - 1, mag)
    k = O.keys64(b.view(np.float64))
    assert np.array_equal(np.argsort(k, kind="stable"), np.argsort(tot, kind="stable"))
    assert O.keys64(np.array([-0.0]))[0] < O.keys64(np.array([0.0]))[0]


def test_f64_exact_matches_the_float32_oracle_on_widened_data():
    # float32 -> float64 is exact and keeps the order of every non-NaN
    # value and of quiet NaNs, so the C restatement on the float32 data
    # fixes the float64 restatement's selection, ties and output order
    rng = np.random.default_rng(65)
    for rep in range(30):
        n = int(rng.integers(1, 4000))
        x = (rng.standard_normal(n) * 10).astype(np.float32)
        x[rng.random(n) < 0.2] = np.float32(2.5)
        sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 1e-45, -1e-45], np.float32)
        x[rng.random(n) < 0.02] = sp[rng.integers(0, sp.size)]
        lens = [n // 3, n - n // 3]
        k = int(rng.integers(1, 500))
        idx = rng.permutation(n).astype(np.int32) if rep % 2 else None
        for flags in (O.LARGEST | O.ASCEND, O.LARGEST, 0, O.ASCEND):
            ev, ei = O.topk_exact(x, lens, k, flags, idx_in=idx)
            gv, gi = O.topk_exact_f64(x.astype(np.float64), lens, k, flags, idx_in=idx)
            assert np.array_equal(gi, ei)
            assert np.array_equal(gv.astype(np.float32).view(np.uint32), ev.view(np.uint32))


def test_f64_exact_matches_numpy_sort_on_distinct_values():
    rng = np.random.default_rng(66)
    x = rng.standard_normal(10000) * 1e200
    v, i = O.topk_exact_f64(x, [10000], 300, O.LARGEST)
    assert np.array_equal(v[0], np.sort(x)[::-1][:300])
    assert np.array_equal(i[0], np.argsort(-x)[:300] + 1)
    v, i = O.topk_exact_f64(x, [10000], 300, O.ASCEND)
    assert np.array_equal(v[0], np.sort(x)[:300])


def test_f32_to_bf16_rounding():
    # round to nearest even on the dropped 16 bits
    f = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.4e38, 1e-40], np.float32)
    b = O.f32_to_bf16(f)
    assert b[0] == 0x3F80 and b[1] == 0x3F80 and b[2] == 0x3F81  # ties to even, above half rounds up
    assert O.bf16_to_f32(b)[3] == -2.5


# ---------------------------------------------------------------- sampling
def test_sampling_oracle_closed_forms():
    # weights 3, 2, 1 (T = 1): c = 3, 5, 6
    v = np.log(np.array([[3.0, 2.0, 1.0]]))
    for u, r in ((0.0, 0), (0.49, 0), (0.51, 1), (0.82, 1), (0.84, 2), (0.999, 2)):
        assert O.sample_topp(v, [3], 1.0, 1.0, [u])[0][0] == r
    # nucleus top_p = 0.5: c_i >= 3 at i = 1, only the first survives
    assert O.sample_topp(v, [3], 1.0, 0.5, [0.99])[0][0] == 0
    # top_p = 0.8: c_i >= 4.8 at i = 2; u * c_2 = 0.7 * 5 = 3.5 -> rank 1
    assert O.sample_topp(v, [3], 1.0, 0.8, [0.7])[0][0] == 1
    # greedy and equal values (uniform over the tie)
    assert O.sample_topp(v, [3], 0.0, 0.9, [0.99])[0][0] == 0
    eq = np.full((1, 8), 2.5)
    for j in range(8):
        assert O.sample_topp(eq, [8], 0.7, 1.0, [(j + 0.5) / 8])[0][0] == j
    # temperature scales log-weights: T = 2 halves them (weights 3^.5, 2^.5, 1)
    w = np.sqrt([3.0, 2.0, 1.0])
    c = np.cumsum(w)
    u = (c[0] + 0.01) / c[-1]
    assert O.sample_topp(v, [3], 2.0, 1.0, [u])[0][0] == 1
    # empty rows
    assert O.sample_topp(np.zeros((1, 4)), [0], 1.0, 1.0, [0.3])[0][0] == -1
