"""Serving engine and batched plans: every result is identical alone and in
the mix (bit-exact), requests are coalesced, admission is enforced."""
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


def spiky(rng, rows, cols):
    """N(0, 3) rows with 0-4 entries per row set to exactly +20."""
    m = (rng.standard_normal(rows * cols) * 3).astype(np.float32)
    for r in range(rows):
        c = rng.integers(0, cols, int(rng.integers(0, 5)))
        m[r * cols + c] = 20.0
    return m


def alone(x, rows, cols, k, flags):
    ws = radik.RadiKWorkspace(cols, rows)
    d = DeviceArray.from_host(x)
    vo = DeviceArray.zeros((k, rows), np.float32)
    io = DeviceArray.zeros((k, rows), np.int32)
    radik.topk_radix_select_(vo, io, ws, d, None, [cols] * rows, k, LARGEST=bool(flags & O.LARGEST),
                             ASCEND=bool(flags & O.ASCEND))
    return vo.to_host(), io.to_host()


def test_server_same_alone_and_in_mix():
    rng = np.random.default_rng(5150)
    shapes = [(int(rng.integers(1, 9)), int(c), int(k))
              for c, k in zip(rng.choice([128256, 40000, 3000, 700], 24), rng.choice([1, 50, 1024], 24))]
    reqs = []
    for rows, cols, k in shapes:
        x = spiky(rng, rows, cols)
        reqs.append((x, rows, cols, k, DeviceArray.from_host(x), DeviceArray.zeros((k, rows), np.float32),
                     DeviceArray.zeros((k, rows), np.int32)))
    srv = radik.Server(budget_bytes=4 << 30, max_inflight=2, largest=True, rev=True)
    say("submitting", len(reqs))
    tickets = [srv.submit(d, rows, cols, k, vo, io) for (x, rows, cols, k, d, vo, io) in reqs]
    lat = [srv.wait(t) for t in tickets]
    st = srv.stats()
    say("stats", st, "latency ms p50", np.median(lat))
    srv.close()
    assert st["rows"] == sum(r[1] for r in reqs)
    assert st["max_coalesced"] >= 2, "requests that arrive together must share a launch"
    for j, (x, rows, cols, k, d, vo, io) in enumerate(reqs):
        mix = (vo.to_host(), io.to_host())
        solo = alone(x, rows, cols, k, O.LARGEST)
        assert np.array_equal(mix[0].view(np.uint32), solo[0].view(np.uint32)), j
        assert np.array_equal(mix[1], solo[1]), j
        if j < 6:
            ev, ei = O.topk_exact(x, [cols] * rows, k, O.LARGEST)
            assert np.array_equal(mix[0].view(np.uint32), ev.T.view(np.uint32)), j
            assert np.array_equal(mix[1], ei.T), j


def test_server_admission_budget():
    srv = radik.Server(budget_bytes=64 << 20, max_inflight=1)
    d = DeviceArray.zeros(256 * 128256, np.float32)
    vo = DeviceArray.zeros((1, 256), np.float32)
    io = DeviceArray.zeros((1, 256), np.int32)
    with pytest.raises(radik.RadiKError) as e:
        srv.submit(d, 256, 128256, 1, vo, io)
    assert e.value.code == radik._lib.RADIK_EADMIT
    small = DeviceArray.from_host(np.arange(1000, dtype=np.float32))
    vo1 = DeviceArray.zeros((3, 1), np.float32)
    io1 = DeviceArray.zeros((3, 1), np.int32)
    t = srv.submit(small, 1, 1000, 3, vo1, io1)
    srv.wait(t)
    assert vo1.to_host()[:, 0].tolist() == [999.0, 998.0, 997.0]
    assert io1.to_host()[:, 0].tolist() == [1000, 999, 998]
    srv.close()


def test_plan_per_task_k_matches_separate_calls():
    rng = np.random.default_rng(77)
    lens = [1 << 18, 5000, 300000, 1 << 17, 70000]
    ks = [16, 100, 2048, 1, 700]
    x = rng.standard_normal(sum(lens)).astype(np.float32)
    p = radik.Plan(lens, ks, largest=False, rev=False)
    d = DeviceArray.from_host(x)
    vo = DeviceArray.zeros(sum(ks), np.float32)
    io = DeviceArray.zeros(sum(ks), np.int32)
    for _ in range(3):
        p.execute(vo, io, d)
    radik.synchronize()
    v, i = vo.to_host(), io.to_host()
    off, oo = 0, 0
    for n, k in zip(lens, ks):
        ev, ei = O.topk_exact(x[off:off + n], [n], k, O.ASCEND)
        assert np.array_equal(v[oo:oo + k].view(np.uint32), ev[0].view(np.uint32))
        assert np.array_equal(i[oo:oo + k], ei[0])
        off += n
        oo += k


def test_server_large_request_runs_alone_within_the_pool():
    # budget 1 GiB over 4 slots: a request above the slot share (256 MiB)
    # but within the budget is admitted and runs in a batch of its own
    rng = np.random.default_rng(606)
    rows, cols, k = 160, 200000, 64
    x = spiky(rng, rows, cols)
    d = DeviceArray.from_host(x)
    vo = DeviceArray.zeros((k, rows), np.float32)
    io = DeviceArray.zeros((k, rows), np.int32)
    small = []
    for _ in range(6):
        xs = spiky(rng, 2, 5000)  # (the device copies stay referenced until the results are read)
        small.append((xs, DeviceArray.from_host(xs), DeviceArray.zeros((10, 2), np.float32),
                      DeviceArray.zeros((10, 2), np.int32)))
    srv = radik.Server(budget_bytes=1 << 30, max_inflight=4, largest=True, rev=True)
    say("large request beside small ones")
    tickets = [srv.submit(ds, 2, 5000, 10, v, i) for xs, ds, v, i in small[:3]]
    tickets.append(srv.submit(d, rows, cols, k, vo, io))
    tickets += [srv.submit(ds, 2, 5000, 10, v, i) for xs, ds, v, i in small[3:]]
    for t in tickets:
        srv.wait(t)
    srv.close()
    ev, ei = O.topk_exact(x, [cols] * rows, k, O.LARGEST)
    assert np.array_equal(vo.to_host().view(np.uint32), ev.T.view(np.uint32))
    assert np.array_equal(io.to_host(), ei.T)
    for xs, ds, v, i in small:
        ev, ei = O.topk_exact(xs, [5000] * 2, 10, O.LARGEST)
        assert np.array_equal(v.to_host().view(np.uint32), ev.T.view(np.uint32)) and np.array_equal(i.to_host(), ei.T)


def test_server_close_with_a_backlog():
    # more requests than slots, closed without waiting: close drains them
    rng = np.random.default_rng(707)
    srv = radik.Server(budget_bytes=256 << 20, max_inflight=1, largest=True, rev=True)
    reqs = []
    for _ in range(24):
        x = spiky(rng, 3, 20000)
        reqs.append((x, DeviceArray.from_host(x), DeviceArray.zeros((5, 3), np.float32),
                     DeviceArray.zeros((5, 3), np.int32)))
    say("backlog of", len(reqs))
    for x, d, vo, io in reqs:
        srv.submit(d, 3, 20000, 5, vo, io)
    srv.close()
    radik.synchronize()
    for x, d, vo, io in reqs:
        ev, ei = O.topk_exact(x, [20000] * 3, 5, O.LARGEST)
        assert np.array_equal(vo.to_host().view(np.uint32), ev.T.view(np.uint32))
        assert np.array_equal(io.to_host(), ei.T)
