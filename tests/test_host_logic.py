"""Host-side logic of the library, on the CPU (no device call): pass A's
chunk table, the routing of tasks to the short / medium / streaming paths,
and the serving engine's batch formation and admission footprint.  The
reference sizes one grid from the shortest and longest task
(src/algorithm.jl:205-219); these are the rules that replace it."""
import ctypes

import numpy as np
import pytest

from radik import _lib

L = _lib.load()


def chunks(total, blocks, spread, guided, gran, min_chunk):
    n = ctypes.c_int64()
    _lib.check(L.radik_diag_plan_chunks(total, blocks, spread, guided, gran, min_chunk, None, 0, ctypes.byref(n)))
    starts = (ctypes.c_int64 * (n.value + 1))()
    _lib.check(L.radik_diag_plan_chunks(total, blocks, spread, guided, gran, min_chunk, starts, n.value + 1,
                                        ctypes.byref(n)))
    return np.array(starts[:], dtype=np.int64)


@pytest.mark.parametrize("guided", [0, 1])
def test_chunk_table_covers_the_element_space(guided):
    gran = 256 * 4 * 2
    for total in (1, 2047, 2048, 2049, 10**6, 622540295, 1 << 28):
        for blocks in (1, 256, 1024, 2048):
            st = chunks(total, blocks, 1 if not guided else 2, guided, gran, (8 if guided else 4) * gran)
            sizes = np.diff(st)
            assert st[0] == 0 and st[-1] == total
            assert np.all(sizes > 0)
            # every chunk but the last starts and ends on the granularity
            assert np.all(st[:-1] % gran == 0)
            if not guided:
                # equal chunks (the last one takes the remainder)
                assert np.all(sizes[:-1] == sizes[0])
                assert len(sizes) <= max(blocks, 1) or sizes[0] == (8 if guided else 4) * gran
            else:
                # shrinking towards the end, never below the minimum except the merged tail
                assert np.all(np.diff(sizes[:-1]) <= 0)
                assert np.all(sizes[:-1] >= 8 * gran)


def test_chunk_table_guided_starts_large_and_ends_small():
    gran, blocks = 2048, 1024
    st = chunks(622540295, blocks, 1, 1, gran, 8 * gran)
    sizes = np.diff(st)
    assert sizes[0] >= 622540295 // blocks - gran
    assert sizes[-2] <= 16 * gran
    assert len(sizes) > blocks


def route(lens, ks, small_cap=2048, predict=1, medium=1):
    n = len(lens)
    cls = (ctypes.c_int32 * n)()
    groups = (ctypes.c_int64 * n)()
    _lib.check(L.radik_diag_route(_lib.i64_array(lens), _lib.i64_array(ks), n, small_cap, predict, medium, cls,
                                  groups))
    return list(cls), list(groups)


def test_routing_rules():
    cases = [  # (len, k, class, groups)
        (0, 10, 0, 0), (2048, 100, 0, 0),           # short: run whole in one block
        (2049, 100, 1, 0), (65535, 2048, 1, 0),     # medium: one block beside the passes
        (65535, 2049, 3, 0),                        # m above the medium cap: streaming, unsampled
        (32767, 256, 1, 0),                         # below the sampled length: medium
        (32768, 256, 2, 16),                        # from 2^15 on sampled (at least 16 groups)
        (65536, 1024, 2, 16),                       # k <= len / 64: sampled (at least 16 groups)
        (65536, 1025, 1, 0),                        # not sampled -> medium
        (70000, 2000, 3, 0),                        # too long for medium, k > len / 64
        (1 << 20, 256, 2, 64), (1 << 22, 256, 2, 256), (1 << 28, 16, 2, 1024),  # 1/64 of the task, <= 1024
        (1 << 28, 1 << 22, 2, 1024), (1 << 28, (1 << 22) + 1, 3, 0),
    ]
    cls, groups = route([c[0] for c in cases], [c[1] for c in cases])
    for (ln, k, c, g), got_c, got_g in zip(cases, cls, groups):
        assert (got_c, got_g) == (c, g), (ln, k)
    # without prediction nothing is sampled; without medium routing those
    # tasks stream
    cls, groups = route([1 << 20, 30000], [256, 256], predict=0)
    assert cls == [3, 1] and groups == [0, 0]
    cls, _ = route([1 << 20, 30000], [256, 256], medium=0)
    assert cls == [2, 3]
    cls, _ = route([5000], [10], small_cap=8192)
    assert cls == [0]


def take(bytes_, share, used, budget):
    t = ctypes.c_int64()
    _lib.check(L.radik_diag_batch(_lib.i64_array(bytes_), len(bytes_), share, used, budget, ctypes.byref(t)))
    return t.value


def test_batch_formation_and_admission():
    G = 1 << 30
    share, budget = 4 * G, 16 * G
    assert take([], share, 0, budget) == 0
    assert take([G, G, G, G, G], share, 0, budget) == 4            # bounded by the slot share
    assert take([G, 2 * G, 2 * G], share, 0, budget) == 2          # oldest first, no reordering
    assert take([10 * G, G], share, 0, budget) == 1                # above the share: alone
    assert take([10 * G], share, 7 * G, budget) == 0               # no room in the pool yet
    assert take([10 * G], share, 6 * G, budget) == 1
    assert take([G, G, G], share, 14 * G, budget) == 2             # bounded by the pool's room
    assert take([16 * G], share, 0, budget) == 1                   # the whole budget when idle


def test_footprint_counts_spill_and_histograms():
    def fp(r, c):
        b = ctypes.c_int64()
        _lib.check(L.radik_diag_footprint(r, c, ctypes.byref(b)))
        return b.value
    assert fp(0, 128256) == 0
    one = fp(1, 128256)
    assert one >= 8 * 128256 + 4 * 4096  # 8-byte spill entries per element + the 4096-bin histogram
    assert fp(256, 128256) == 256 * one
    assert fp(1, 2 * 128256) - one == 8 * 128256
