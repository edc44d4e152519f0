"""Key types beyond the reference's Float32 (Float16, BFloat16, Int32,
Float64), through the C-ABI, against the oracle's typed restatements:
bit-exact value patterns and indices (ties to the lowest position).  Short,
medium and long (sampled, and over 2^20 elements) tasks, ties and special
values (+-0, +-Inf, NaN, subnormals, the integer extremes), both selection
directions and output orders, the drop-in and the batched plan, and a result
set beyond 4096 (the multi-block sort).  Float64 runs its own path (drop-in
and workspace): exact ties across the threshold, custom Int32 and Int64
indices, k up to its limit of 8192, and the rejections (k above it, plans,
WITHSCALE)."""
import sys

import numpy as np
import pytest

import oracle as O
import radik
from radik import DeviceArray

pytestmark = pytest.mark.gpu

KEYS = {O.KEY_F16: "f16", O.KEY_BF16: "bf16", O.KEY_I32: "i32"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_data(kt, n, rng):
    if kt == O.KEY_F16:
        x = (rng.standard_normal(n) * 4).astype(np.float16)
        x[rng.random(n) < 0.03] = np.float16(6.5)  # ties near the top
        sp = rng.integers(0, n, max(1, n // 5000))
        x[sp] = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 6e-8, -6e-8], np.float16)[rng.integers(0, 7, sp.size)]
        return x
    if kt == O.KEY_BF16:
        f = (rng.standard_normal(n) * 1e3).astype(np.float32)
        f[rng.random(n) < 0.03] = np.float32(3000.0)
        b = O.f32_to_bf16(f)
        sp = rng.integers(0, n, max(1, n // 5000))
        b[sp] = np.array([0x0000, 0x8000, 0x7F80, 0xFF80, 0x7FC0, 0xFFC1, 0x0001], np.uint16)[rng.integers(0, 7, sp.size)]
        return b
    x = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    x[rng.random(n) < 0.03] = 2**31 - 77
    sp = rng.integers(0, n, max(1, n // 5000))
    x[sp] = np.array([-2**31, 2**31 - 1, 0, -1], np.int32)[rng.integers(0, 4, sp.size)]
    return x


def bits(a, kt):
    return np.asarray(a).view(np.uint16 if kt in (O.KEY_F16, O.KEY_BF16) else np.uint32)


def check(kt, got_v, got_i, ev, ei, name):
    gv = bits(got_v, kt)
    xv = bits(ev, kt)
    bad = np.nonzero(gv != xv)
    assert bad[0].size == 0, f"{name}: {bad[0].size} value mismatches, first at {tuple(b[0] for b in bad)}"
    bad = np.nonzero(got_i.astype(np.int64) != ei.astype(np.int64))
    assert bad[0].size == 0, f"{name}: {bad[0].size} index mismatches, first at {tuple(b[0] for b in bad)}"


@pytest.mark.parametrize("kt", [O.KEY_F16, O.KEY_BF16, O.KEY_I32])
def test_keytype_batch_drop_in_and_plan(kt):
    rng = np.random.default_rng(300 + kt)
    lens = [700, 2048, 5000, 40000, 65536, 200000, 1 << 20, 3 << 20, 0, 1500]
    x = make_data(kt, sum(lens), rng)
    d = DeviceArray.from_host(x)
    key = KEYS[kt]
    for k in (1, 100, 1000):
        for largest, rev in ((True, False), (False, True), (True, True), (False, False)):
            flags = (O.LARGEST if largest else 0) | (0 if rev else O.ASCEND)
            ev, ei = O.topk_exact_typed(kt, x, lens, k, flags, threads=O.max_threads())
            say(key, "drop-in", k, largest, rev)
            v, i = radik.topk(d, lens, k, largest=largest, rev=rev, key=key)
            check(kt, v.to_host(), i.to_host(), ev.T, ei.T, f"{key} topk k{k} l{largest} r{rev}")
            if k == 100:
                say(key, "plan", k, largest, rev)
                p = radik.Plan(lens, k, largest=largest, rev=rev, key=key)
                vo = DeviceArray.zeros((k, len(lens)), x.dtype)
                io = DeviceArray.zeros((k, len(lens)), np.int32)
                for rep in range(2):
                    p.execute(vo, io, d)
                    radik.synchronize()
                    check(kt, vo.to_host(), io.to_host(), ev.T, ei.T, f"{key} plan k{k} l{largest} r{rev} {rep}")


@pytest.mark.parametrize("kt", [O.KEY_F16, O.KEY_BF16, O.KEY_I32])
def test_keytype_large_task_indices_and_big_k(kt):
    rng = np.random.default_rng(400 + kt)
    n = 1 << 22
    x = make_data(kt, n, rng)
    idx = rng.permutation(n).astype(np.int32)
    d = DeviceArray.from_host(x)
    di = DeviceArray.from_host(idx)
    key = KEYS[kt]
    for k in (64, 5000):
        for flags in (O.LARGEST | O.ASCEND, 0):
            say(key, "large", k, flags)
            ev, ei = O.topk_exact_typed(kt, x, [n], k, flags, idx_in=idx)
            ws = radik.RadiKWorkspace(n, 1)
            vo = DeviceArray.zeros((k, 1), x.dtype)
            io = DeviceArray.zeros((k, 1), np.int32)
            radik.topk_radix_select_(vo, io, ws, d, di, [n], k, LARGEST=bool(flags & O.LARGEST),
                                     ASCEND=bool(flags & O.ASCEND), WITHIDXIN=True, key=key)
            check(kt, vo.to_host(), io.to_host(), ev.T, ei.T, f"{key} large k{k} f{flags}")


def test_keytype_rejects_withscale_and_bad_dtypes():
    d = DeviceArray.from_host(np.zeros(100, np.float16))
    vo = DeviceArray.zeros((10, 1), np.float16)
    io = DeviceArray.zeros((10, 1), np.int32)
    ws = radik.RadiKWorkspace(100, 1)
    with pytest.raises(radik.RadiKError):
        radik.topk_radix_select_(vo, io, ws, d, None, [100], 10, WITHSCALE=True)
    with pytest.raises(radik.RadiKError):
        radik.topk(DeviceArray.from_host(np.zeros(100, np.int64)), 10)


# ------------------------------------------------------------------ Float64
def make_f64(n, rng, levels=0):
    if levels:
        x = rng.integers(0, levels, n).astype(np.float64) * 0.125 - 3.0
    else:
        x = rng.standard_normal(n) * 1e3
        x[rng.random(n) < 0.03] = 2718.25  # ties near the top
    sp = rng.integers(0, n, max(1, n // 5000))
    x[sp] = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 5e-324, -5e-324])[rng.integers(0, 8, sp.size)]
    return x


def check64(got_v, got_i, ev, ei, name):
    bad = np.nonzero(np.asarray(got_v).view(np.uint64) != np.asarray(ev).view(np.uint64))
    assert bad[0].size == 0, f"{name}: {bad[0].size} value mismatches, first at {tuple(b[0] for b in bad)}"
    bad = np.nonzero(got_i.astype(np.int64) != ei.astype(np.int64))
    assert bad[0].size == 0, f"{name}: {bad[0].size} index mismatches, first at {tuple(b[0] for b in bad)}"


def test_f64_batch_drop_in():
    rng = np.random.default_rng(500)
    lens = [700, 2048, 5000, 40000, 65536, 200000, 1 << 20, 0, 1500]
    x = make_f64(sum(lens), rng)
    d = DeviceArray.from_host(x)
    for k in (1, 100, 1000):
        for largest, rev in ((True, False), (False, True), (True, True), (False, False)):
            flags = (O.LARGEST if largest else 0) | (0 if rev else O.ASCEND)
            say("f64 drop-in", k, largest, rev)
            ev, ei = O.topk_exact_f64(x, lens, k, flags)
            v, i = radik.topk(d, lens, k, largest=largest, rev=rev)
            check64(v.to_host(), i.to_host(), ev.T, ei.T, f"f64 topk k{k} l{largest} r{rev}")


def test_f64_large_task_ties_indices_and_max_k():
    rng = np.random.default_rng(501)
    n = 1 << 22
    idx = rng.permutation(n).astype(np.int32)
    di = DeviceArray.from_host(idx)
    ws = radik.RadiKWorkspace(n, 1)
    for levels in (0, 64):  # 64 levels: exact ties across the threshold
        x = make_f64(n, rng, levels)
        d = DeviceArray.from_host(x)
        for k in (64, 5000, O_MAX_K):
            for flags in (O.LARGEST | O.ASCEND, 0):
                say("f64 large", levels, k, flags)
                ev, ei = O.topk_exact_f64(x, [n], k, flags, idx_in=idx)
                vo = DeviceArray.zeros((k, 1), np.float64)
                io = DeviceArray.zeros((k, 1), np.int32)
                radik.topk_radix_select_(vo, io, ws, d, di, [n], k, LARGEST=bool(flags & O.LARGEST),
                                         ASCEND=bool(flags & O.ASCEND), WITHIDXIN=True)
                check64(vo.to_host(), io.to_host(), ev.T, ei.T, f"f64 large L{levels} k{k} f{flags}")


O_MAX_K = 8192


def test_f64_limits():
    x = DeviceArray.from_host(np.zeros(10000, np.float64))
    with pytest.raises(radik.RadiKError):
        radik.topk(x, O_MAX_K + 1)
    with pytest.raises(radik.RadiKError):
        radik.Plan([10000], 10, key="f64")
    ws = radik.RadiKWorkspace(10000, 1)
    vo = DeviceArray.zeros((10, 1), np.float64)
    io = DeviceArray.zeros((10, 1), np.int32)
    with pytest.raises(radik.RadiKError):
        radik.topk_radix_select_(vo, io, ws, x, None, [10000], 10, WITHSCALE=True)
    # Int64 indices through the drop-in
    rng = np.random.default_rng(502)
    y = make_f64(30000, rng)
    ii = rng.permutation(30000).astype(np.int64)
    ev, ei = O.topk_exact_f64(y, [30000], 500, O.LARGEST, idx_in=ii, idx_dtype=np.int64)
    v, i = radik.topk(DeviceArray.from_host(y), 500, indices=DeviceArray.from_host(ii), rev=True)
    check64(v.to_host(), i.to_host(), ev[0], ei[0], "f64 int64 indices")
