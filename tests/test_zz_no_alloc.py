"""The allocation-free entry point (radik_topk_noalloc) against the drop-in
it varies (radik_topk), through the C-ABI: identical values and indices on
two calls in a row, the second with a smaller input, and the limits
rejected with RADIK_ELIMIT."""
import sys

import numpy as np
import pytest

import radik
from radik import DeviceArray, _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def call(fn, x, lens, k, largest, rev, idx=None, idx_dtype=np.int32):
    T = len(lens)
    d = DeviceArray.from_host(x)
    di = DeviceArray.from_host(idx) if idx is not None else None
    vo = DeviceArray.zeros((k, T), np.float32).fill(np.float32(-7))
    io = DeviceArray.zeros((k, T), idx_dtype)
    L = _lib.load()
    rc = getattr(L, fn)(vo.ptr, io.ptr, np.dtype(idx_dtype).itemsize, k, T, d.ptr,
                        di.ptr if di is not None else None, 0 if idx is None else idx.size,
                        _lib.i64_array(lens), T, k, 1 if largest else 0, 1 if rev else 0, None)
    _lib.check(rc)
    return vo.to_host(), io.to_host()


def test_noalloc_matches_topk_twice_then_smaller():
    rng = np.random.default_rng(9090)
    big = [100, 3000, 70000, 1 << 20, 2048, 40000, 300000, 0, 1 << 18, 5000]
    small = [900, 65536, 200000, 17]
    for lens, k, largest, rev, with_idx in ((big, 256, True, False, True), (small, 100, False, True, False),
                                            (big, 1000, True, True, False), (small, 3000, True, False, True)):
        n = sum(lens)
        x = rng.standard_normal(n).astype(np.float32)
        x[rng.random(n) < 0.02] = np.float32(1.75)
        idx = rng.permutation(n).astype(np.int32) if with_idx else None
        say("noalloc", len(lens), n, k, largest, rev, with_idx)
        a = call("radik_topk", x, lens, k, largest, rev, idx)
        b = call("radik_topk_noalloc", x, lens, k, largest, rev, idx)
        assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
        assert np.array_equal(a[1], b[1])


def test_noalloc_int64_indices_and_limits():
    rng = np.random.default_rng(9191)
    lens = [1 << 21, 1 << 21]  # exactly the element limit, two long tasks
    x = rng.standard_normal(sum(lens)).astype(np.float32)
    idx = np.arange(sum(lens), dtype=np.int64) * 3
    say("noalloc at the element limit, int64 indices")
    a = call("radik_topk", x, lens, 64, True, False, idx, np.int64)
    b = call("radik_topk_noalloc", x, lens, 64, True, False, idx, np.int64)
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and np.array_equal(a[1], b[1])
    with pytest.raises(radik.RadiKError) as e:
        call("radik_topk_noalloc", np.zeros((1 << 22) + 1, np.float32), [(1 << 22) + 1], 8, True, False)
    assert e.value.code == _lib.RADIK_ELIMIT
    with pytest.raises(radik.RadiKError) as e:
        call("radik_topk_noalloc", np.zeros(513 * 2049, np.float32), [2049] * 513, 8, True, False)
    assert e.value.code == _lib.RADIK_ELIMIT
