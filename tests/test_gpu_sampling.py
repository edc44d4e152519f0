"""Temperature / top-p sampling fused after the top-k (radik_plan_sample),
through the C-ABI, against the oracle's float64 restatement.

The device evaluates the weights and their running sums in float32; a
decision may legitimately fall on the other side of a boundary that lies
within float32 rounding of it.  Tolerance: the sampled rank must equal the
oracle's, except in rows whose float64 decision lies within 1e-5 (relative
to the row's total weight) of a boundary, where a neighbouring rank is
accepted.  The token must be the index at the sampled rank."""
import sys

import numpy as np
import pytest

import oracle as O
import radik
from radik import DeviceArray

pytestmark = pytest.mark.gpu

TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def logits(rng, rows, cols):
    x = (rng.standard_normal(rows * cols) * 3).astype(np.float32)
    nsp = rng.integers(0, 5, rows)
    r = np.repeat(np.arange(rows), nsp)
    x[r * cols + rng.integers(0, cols, r.size)] = np.float32(20)  # exact ties at the top
    return x


def check_rows(vals, idx, ranks, tokens, m, temperature, top_p, u):
    er, slack = O.sample_topp(vals.T, m, temperature, top_p, u)
    for t in range(len(m)):
        if m[t] == 0:
            assert ranks[t] == -1 and tokens[t] == -1
            continue
        if ranks[t] != er[t]:
            assert slack[t] < TOL and abs(int(ranks[t]) - int(er[t])) <= 1, (t, ranks[t], er[t], slack[t])
        assert tokens[t] == idx[ranks[t], t]


@pytest.mark.parametrize("key", ["f32", "bf16"])
def test_sampling_after_topk_rows(key):
    rng = np.random.default_rng(5151)
    rows, cols = 96, 32000
    x = logits(rng, rows, cols)
    data = x if key == "f32" else O.f32_to_bf16(x)
    d = DeviceArray.from_host(data)
    lens = [cols] * rows
    lens[5] = 0  # an empty row
    lens[7] = 20  # a row shorter than k
    n_in = sum(lens)
    d = DeviceArray.from_host(data[:n_in])
    for k in (1, 50, 1024):
        p = radik.Plan(lens, k, largest=True, rev=True, key=key)
        vo = DeviceArray.zeros((k, rows), data.dtype)
        io = DeviceArray.zeros((k, rows), np.int32)
        p.execute(vo, io, d)
        vals = vo.to_host()
        vals_f = vals if key == "f32" else O.bf16_to_f32(vals)
        idx = io.to_host()
        m = [min(k, L) for L in lens]
        for temperature, top_p in ((0.0, 0.9), (0.7, 0.9), (1.0, 1.0), (1.5, 0.5), (0.3, 0.0)):
            u = rng.random(rows).astype(np.float32)
            du = DeviceArray.from_host(u)
            say("sampling", key, k, temperature, top_p)
            tok, rk = p.sample(vo, io, temperature, top_p, du)
            radik.synchronize()
            check_rows(vals_f, idx, rk.to_host(), tok.to_host(), m, temperature, top_p, u.astype(np.float64))


def test_sampling_rejects_ascending_plans():
    p = radik.Plan([100], 10, largest=True, rev=False)
    vo = DeviceArray.zeros((10, 1), np.float32)
    io = DeviceArray.zeros((10, 1), np.int32)
    with pytest.raises(radik.RadiKError):
        p.sample(vo, io, 1.0, 0.9, DeviceArray.zeros(1, np.float32))
