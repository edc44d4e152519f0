"""Parity at the benchmark configurations' real shapes, through the C-ABI,
against the oracle: bit-exact values and indices (ties to the lowest
position).

* the uneven batch (4096 tasks, lengths log-uniform in [2^10, 2^20], k = 256,
  duplicates and +-0 / subnormal / +-Inf specials), whole output, through the
  graph-captured plan (short and medium tasks on the forked side stream) and
  eagerly;
* a small plan mixing short, medium and long tasks, so the side-stream fork
  inside a captured graph runs in every test run, and the same kind of plan
  under every launch knob (where the side stream forks and joins, the medium
  tasks' LDS share, the sampling threshold), whose results must not change;
* the 2^28-element task with Int32 custom indices (a seeded permutation),
  k in {16, 1024, 4096, 16384} x largest x rev (the grid of the reference's
  test/test_algorithm.jl:6-40 at the benchmark size);
* long tasks re-read by pass B over several blocks whose winners and
  candidates are then gathered in one block's LDS (prediction off, and a tie
  pattern that the sampled bound covers only in part).
"""
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


def same(gv, gi, ev, ei, name):
    bad = np.nonzero(gv.view(np.uint32) != ev.view(np.uint32))
    assert bad[0].size == 0, f"{name}: {bad[0].size} value mismatches, first at {tuple(b[0] for b in bad)}"
    bad = np.nonzero(gi.astype(np.int64) != ei.astype(np.int64))
    assert bad[0].size == 0, f"{name}: {bad[0].size} index mismatches, first at {tuple(b[0] for b in bad)}"


# ----------------------------------------------------------------- uneven batch
@pytest.fixture(scope="module")
def uneven():
    from bench import uneven_batch
    say("uneven batch: generating")
    x, lens = uneven_batch()
    ev, ei = O.topk_exact(x, lens, 256, O.LARGEST | O.ASCEND, threads=O.max_threads())
    say("uneven batch: oracle done", int(lens.sum()))
    return x, lens, ev.T, ei.T


@pytest.mark.parametrize("use_graph", [True, False])
def test_uneven_batch_full(uneven, use_graph):
    x, lens, ev, ei = uneven
    T, k = len(lens), 256
    say("uneven batch full, graph", use_graph)
    d = DeviceArray.from_host(x)
    p = radik.Plan(list(lens), k, largest=True, rev=False, use_graph=use_graph)
    vo = DeviceArray.zeros((k, T), np.float32)
    io = DeviceArray.zeros((k, T), np.int32)
    for rep in range(2):
        vo.fill(np.float32(-7))
        p.execute(vo, io, d)
        radik.synchronize()
        same(vo.to_host(), io.to_host(), ev, ei, f"uneven graph={use_graph} rep{rep}")
    info = p.info()
    say("plan info", info)
    assert info["n_short"] > 0 and info["n_medium"] > 0 and info["n_long"] > 0
    assert info["long_elements"] + info["medium_elements"] + info["short_elements"] == int(lens.sum())


def test_mixed_plan_side_stream():
    # short (k_small), medium (k_medium) and long sampled tasks in one plan
    rng = np.random.default_rng(77)
    lens = [500, 2048, 3000, 40000, 65536, 70000, 1 << 18, 1500, 30000, 1 << 20, 900, 300000, 0, 5000]
    x = rng.standard_normal(sum(lens)).astype(np.float32)
    x[rng.random(x.size) < 0.02] = np.float32(2.0)  # ties near the threshold
    d = DeviceArray.from_host(x)
    for k, largest, rev in ((256, True, False), (64, False, True), (1000, True, True)):
        flags = (O.LARGEST if largest else 0) | (0 if rev else O.ASCEND)
        ev, ei = O.topk_exact(x, lens, k, flags)
        for use_graph in (True, False):
            say("mixed plan", k, largest, rev, use_graph)
            p = radik.Plan(lens, k, largest=largest, rev=rev, use_graph=use_graph)
            vo = DeviceArray.zeros((k, len(lens)), np.float32)
            io = DeviceArray.zeros((k, len(lens)), np.int32)
            for rep in range(3):
                vo.fill(np.float32(-7))
                p.execute(vo, io, d)
                radik.synchronize()
                same(vo.to_host(), io.to_host(), ev.T, ei.T, f"mixed k{k} graph={use_graph} rep{rep}")
            info = p.info()
            assert info["n_short"] > 0 and info["n_medium"] > 0 and info["n_long"] > 0, info


def test_mixed_plan_launch_options():
    # The plan's launch knobs change the schedule, never the results: the
    # side stream joined before k_finish or at the end, forked at the start,
    # after the sampling or after the prediction, other LDS shares for the
    # medium tasks, other sampling thresholds; eager and graph replay.
    rng = np.random.default_rng(78)
    lens = [700, 2500, 40000, 65536, 1 << 18, 1 << 20, 3000, 33000, 0, 150000]
    x = rng.standard_normal(sum(lens)).astype(np.float32)
    x[rng.random(x.size) < 0.01] = np.float32(1.5)
    d = DeviceArray.from_host(x)
    k = 128
    ev, ei = O.topk_exact(x, lens, k, O.LARGEST | O.ASCEND)
    sets = [{"join_early": 0}, {"join_early": 1}, {"fork_late": 1}, {"fork_late": 2, "join_early": 0},
            {"longest_first": 0}, {"longest_first": 1},
            {"medium_per_cu": 1}, {"medium_per_cu": 8}, {"sample_min_len": 4096}, {"sample_min_len": 65536}]
    for opts in sets:
        for use_graph in (False, True):
            say("mixed plan options", opts, "graph", use_graph)
            p = radik.Plan(lens, k, largest=True, rev=False, use_graph=use_graph)
            for name, val in opts.items():
                p.set_option(name, val)
            vo = DeviceArray.zeros((k, len(lens)), np.float32)
            io = DeviceArray.zeros((k, len(lens)), np.int32)
            for rep in range(2):
                vo.fill(np.float32(-7))
                p.execute(vo, io, d)
                radik.synchronize()
                same(vo.to_host(), io.to_host(), ev.T, ei.T, f"options {opts} graph={use_graph} rep{rep}")
    p = radik.Plan(lens, k, largest=True, rev=False)
    for name, bad in (("join_early", 2), ("fork_late", 3), ("medium_per_cu", 0), ("sample_min_len", 100),
                      ("longest_first", 2),
                      ("no_such_option", 1)):
        with pytest.raises(radik.RadiKError):
            p.set_option(name, bad)


# ----------------------------------------------------------------- 2^28 task
@pytest.fixture(scope="module")
def single_2p28():
    n = 1 << 28
    say("2^28: generating input and permutation")
    x = np.random.default_rng(2828).standard_normal(n, dtype=np.float32)
    perm = np.random.default_rng(2829).permutation(n).astype(np.int32)
    # The oracle's top-16384 for largest (rev = true: descending value, ties
    # by position) and smallest (rev = false), each in selection order; the
    # top-k for smaller k is its first k entries, and the other output order
    # a stable re-sort of those by key.
    ref = {}
    for largest in (True, False):
        flags = O.LARGEST if largest else O.ASCEND
        v, i = O.topk_exact(x, [n], 16384, flags, idx_in=perm)
        ref[largest] = (v[0], i[0])
    say("2^28: oracle done")
    return x, perm, ref


@pytest.mark.parametrize("k", [16, 1024, 4096, 16384])
def test_single_2p28_grid(single_2p28, k):
    x, perm, ref = single_2p28
    n = x.size
    d = DeviceArray.from_host(x)
    di = DeviceArray.from_host(perm)
    for largest in (True, False):
        sv, si = ref[largest][0][:k], ref[largest][1][:k]  # selection order
        for rev in (False, True):
            # selection order is rev=true for largest, rev=false for smallest
            if largest == rev:
                ev, ei = sv, si
            else:
                o = O.reorder(sv, descending=not largest)
                ev, ei = sv[o], si[o]
            say("2^28", k, "largest", largest, "rev", rev)
            vo = DeviceArray.zeros((k, 1), np.float32)
            io = DeviceArray.zeros((k, 1), np.int32)
            ws = radik.RadiKWorkspace(n, 1)
            radik.topk_radix_select_(vo, io, ws, d, di, [n], k, LARGEST=largest, ASCEND=not rev, WITHIDXIN=True)
            same(vo.to_host()[:, 0], io.to_host()[:, 0], ev, ei, f"2^28 k{k} largest{largest} rev{rev} ws")
            del ws
            # the batched plan (hipGraph replay) on the same task
            p = radik.Plan([n], k, largest=largest, rev=rev, indices_dtype=np.int32)
            for rep in range(2):
                vo.fill(np.float32(-7))
                p.execute(vo, io, d, di)
                radik.synchronize()
                same(vo.to_host()[:, 0], io.to_host()[:, 0], ev, ei, f"2^28 k{k} largest{largest} rev{rev} plan")


# ----------------------------------------------------------------- pass-B re-read
def test_reread_over_several_blocks_gathered_in_lds():
    # Tasks over 2^20 elements whose spill does not hold their selection are
    # re-read by pass B over several blocks; with few winners + candidates
    # (they fit the finishing block's LDS) both are listed winners first.
    n = 1 << 21
    rng = np.random.default_rng(2121)
    perm = rng.permutation(n).astype(np.int32)
    x = rng.standard_normal(n).astype(np.float32)
    d = DeviceArray.from_host(x)
    di = DeviceArray.from_host(perm)
    for k in (200, 256):
        for flags in (O.LARGEST | O.ASCEND, O.ASCEND, O.LARGEST):
            ev, ei = O.topk_exact(x, [n], k, flags, idx_in=perm)
            say("reread prediction off", k, flags)
            p = radik.Plan([n], k, largest=bool(flags & O.LARGEST), rev=not (flags & O.ASCEND),
                           indices_dtype=np.int32)
            p.set_options(predict=False)
            vo = DeviceArray.zeros((k, 1), np.float32)
            io = DeviceArray.zeros((k, 1), np.int32)
            for use_graph in (True, False):
                p.set_graph(use_graph)
                vo.fill(np.float32(-7))
                p.execute(vo, io, d, di)
                radik.synchronize()
                same(vo.to_host(), io.to_host(), ev.T, ei.T, f"reread k{k} f{flags} graph={use_graph}")
    # The sampled groups (group j of G = 128 starts at j * n / G) hold large
    # distinct values found nowhere else: the sample overstates them, the
    # predicted bound spills too few entries and pass B re-reads the task.
    G = 128
    x = rng.standard_normal(n).astype(np.float32)
    for j in range(G):
        st = j * n // G
        x[st:st + 256] = 100.0 + np.arange(j * 256, j * 256 + 256, dtype=np.float32)
    d = DeviceArray.from_host(x)
    for k in (200, 256):
        for flags in (O.LARGEST, O.LARGEST | O.ASCEND):
            say("reread prediction miss", k, flags)
            ev, ei = O.topk_exact(x, [n], k, flags, idx_in=perm)
            vo = DeviceArray.zeros((k, 1), np.float32)
            io = DeviceArray.zeros((k, 1), np.int32)
            ws = radik.RadiKWorkspace(n, 1)
            radik.topk_radix_select_(vo, io, ws, d, di, [n], k, LARGEST=bool(flags & O.LARGEST),
                                     ASCEND=bool(flags & O.ASCEND), WITHIDXIN=True)
            same(vo.to_host(), io.to_host(), ev.T, ei.T, f"prediction miss k{k} f{flags}")


# ----------------------------------------------------------------- rank tail
@pytest.mark.parametrize("k", [100, 1000, 20000])
def test_rank_tail_second_digit_cut(k):
    # Tasks over 2^20 elements whose spill holds their m best finish in the
    # rank tail.  When the m-th best lies in the predicted first digit, the
    # tail lists only the candidates up to the second-digit bin holding it.
    # Cases: smooth data; an exact tie straddling the m-th best (the cut
    # takes its whole bin); a tie of 5 % of the task in the threshold digit
    # (the list exceeds the tail, which falls back); two big tasks at once.
    n = 1 << 22
    rng = np.random.default_rng(4242 + k)
    base = rng.standard_normal(n).astype(np.float32)
    cases = {"smooth": (base, [n])}
    x = base.copy()
    v = np.partition(base, n - k)[n - k]  # the k-th largest
    x[rng.integers(0, n, 3000)] = v
    cases["tie_at_threshold"] = (x, [n])
    x = base.copy()
    x[rng.random(n) < 0.05] = np.float32(2.75)
    cases["tie_heavy"] = (x, [n])
    cases["two_tasks"] = (base, [(1 << 21) + 777, (1 << 21) - 999])
    for name, (x, lens) in cases.items():
        T = len(lens)
        lens = [int(a) for a in lens]
        xx = x[:sum(lens)]
        d = DeviceArray.from_host(xx)
        perm = rng.permutation(sum(lens)).astype(np.int32)
        di = DeviceArray.from_host(perm)
        for flags in (O.LARGEST | O.ASCEND, O.LARGEST, O.ASCEND):
            largest, rev = bool(flags & O.LARGEST), not (flags & O.ASCEND)
            ev, ei = O.topk_exact(xx, lens, k, flags, idx_in=perm, threads=O.max_threads())
            say("rank tail", name, k, flags)
            vo = DeviceArray.zeros((k, T), np.float32)
            io = DeviceArray.zeros((k, T), np.int32)
            ws = radik.RadiKWorkspace(sum(lens), T)
            radik.topk_radix_select_(vo, io, ws, d, di, lens, k, LARGEST=largest, ASCEND=not rev, WITHIDXIN=True)
            same(vo.to_host(), io.to_host(), ev.T, ei.T, f"rank tail {name} k{k} f{flags} ws")
            del ws
            p = radik.Plan(lens, k, largest=largest, rev=rev, indices_dtype=np.int32)
            for rep in range(2):
                vo.fill(np.float32(-7))
                p.execute(vo, io, d, di)
                radik.synchronize()
                same(vo.to_host(), io.to_host(), ev.T, ei.T, f"rank tail {name} k{k} f{flags} plan rep{rep}")
