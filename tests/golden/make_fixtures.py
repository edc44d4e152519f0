"""Writes the golden fixtures under tests/golden/.

known_answers.json  inputs and expected outputs transcribed from the
                    reference's own tests (WilliBee/RadiK.jl test/*.jl, cited
                    per case).  Data only: the reference code is not run.
cases.npz           seeded inputs with the oracle's outputs (values, indices)
                    for the GPU parity tests: configuration shapes scaled down,
                    tie-heavy and special-value data.  Inputs are stored as raw
                    float32, so no RNG or libm drift matters.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

F32_MIN_SUB = np.float32(1.4e-45)
F32_MAX_SUB = np.uint32(0x007FFFFF).view(np.float32)


def known_answers():
    ka = {}
    # test/test_utils.jl:9-23 -- 8-bit bins of sorted specials: first 255, last 0,
    # strictly decreasing.
    ka["get_bin_id_order"] = {
        "source": "test/test_utils.jl:9-23",
        "vals": ["inf", 100.0, 2.0, 1.0, 0.0, -1.0, -2.0, -100.0, "-inf"],
        "left": 0, "right": 24, "first": 255, "last": 0, "strictly_decreasing": True,
    }
    # test/test_utils.jl:53-106 -- Up scan gives cumsum, Down gives reverse cumsum.
    ka["warp_scan"] = {
        "source": "test/test_utils.jl:78-105",
        "up_1_32": np.cumsum(np.arange(1, 33)).tolist(),
        "down_1_32": np.cumsum(np.arange(1, 33)[::-1])[::-1].tolist(),
        "up_1_64_ws32": np.cumsum(np.arange(1, 33)).tolist() + np.cumsum(np.arange(33, 65)).tolist(),
    }
    # test/test_select_bin.jl:13-55
    ka["cross_warp_reduction"] = [
        {"source": "test/test_select_bin.jl:32-42", "largest": True, "block": 12, "warp": 4,
         "input": list(range(12, 0, -1)),
         "expected": (np.arange(12, 0, -1) + np.array([12] * 4 + [4] * 4 + [0] * 4)).tolist()},
        {"source": "test/test_select_bin.jl:44-54", "largest": False, "block": 12, "warp": 4,
         "input": list(range(1, 13)),
         "expected": (np.arange(1, 13) + np.array([0] * 4 + [4] * 4 + [12] * 4)).tolist()},
    ]
    # test/test_select_bin.jl:60-74
    ka["compute_neighbor"] = {
        "source": "test/test_select_bin.jl:60-74",
        "largest": {"prefix": [9, 5, 2], "expected": [5, 2, 0]},
        "smallest": {"prefix": [2, 5, 9], "expected": [0, 2, 5]},
    }
    # test/test_select_bin.jl:79-188 (bin ids are 1-based)
    ka["find_bin_and_write"] = [
        {"source": "test/test_select_bin.jl:84-104", "counts": [3, 5, 2, 8], "old_k": 10, "neighbor": 5,
         "sum": 23, "thread_x": 2, "unroll": 4, "largest": True, "expected": [8, 5, 8]},
        {"source": "test/test_select_bin.jl:106-119", "counts": [3, 5, 2, 8], "old_k": 10, "neighbor": 5,
         "sum": 23, "thread_x": 2, "unroll": 4, "largest": False, "expected": [6, 2, 5]},
        {"source": "test/test_select_bin.jl:122-144", "counts": [3, 5, 2, 8], "old_k": 24, "neighbor": 10,
         "sum": 23, "thread_x": 2, "unroll": 4, "largest": True, "expected": None},
        {"source": "test/test_select_bin.jl:151-168", "counts": [5, 3, 2], "old_k": 1, "neighbor": 0,
         "sum": 10, "thread_x": 1, "unroll": 3, "largest": False, "expected": [1, 1, 5]},
        {"source": "test/test_select_bin.jl:170-187", "counts": [5, 3, 2], "old_k": 10, "neighbor": 0,
         "sum": 10, "thread_x": 1, "unroll": 3, "largest": False, "expected": [3, 2, 2]},
    ]
    # test/test_select_bin.jl:194-352 -- select_bin_kernel! over 256 bins.
    sb = []
    h = [0] * 256
    h[0:4] = [2, 8, 14, 3]
    sb.append({"source": "test/test_select_bin.jl:194-222", "hist": {"0": h}, "k": [5], "largest": True,
               "blocks": [256], "expected": [[3, 2, 14]]})
    sb.append({"source": "test/test_select_bin.jl:224-252", "hist": {"0": h}, "k": [5], "largest": False,
               "blocks": [256], "expected": [[2, 3, 8]]})
    h10 = [1] * 10 + [0] * 246
    sb.append({"source": "test/test_select_bin.jl:254-282", "hist": {"0": h10}, "k": [5], "largest": True,
               "blocks": [32, 64, 128, 256], "expected": [[6, 1, 1]]})
    h2 = [0] * 256
    h2[100:110] = [2] * 10
    h3 = [0] * 256
    h3[200:206] = [3] * 6
    sb.append({"source": "test/test_select_bin.jl:284-326", "hist": {"0": h10, "1": h2, "2": h3},
               "k": [5, 7, 8], "largest": True, "blocks": [256],
               "expected": [[6, 1, 1], [107, 1, 2], [204, 2, 3]]})
    h5 = [5] + [0] * 255
    sb.append({"source": "test/test_select_bin.jl:328-352", "hist": {"0": h5}, "k": [3], "largest": True,
               "blocks": [256], "expected": [[1, 3, 5]]})
    ka["select_bin"] = sb
    # test/test_count_bin.jl:11-24 -- three ramps, 4-bin digit (LEFT 0, RIGHT 30).
    ka["count_bin_ramps"] = {
        "source": "test/test_count_bin.jl:11-61",
        "ramps": [[1, 200000], [500, 800], [1000, 2000]],
        "left": 0, "right": 30, "histlen": 4, "pack": 4,
    }
    # test/test_filter.jl known-answer sets.
    ka["filter"] = [
        {"source": "test/test_filter.jl:8-49", "data": "1:1000", "k": 1000, "withscale": False,
         "largest": True, "kth": 0.0, "boundary": 0, "expected_count": 1000, "expected_set": "1:1000"},
        {"source": "test/test_filter.jl:51-104 (scaled, boundary 0)", "data": "1:1000", "k": 5, "withscale": True,
         "largest": True, "kth_scaled_of": 996, "boundary": 0, "expected_count": 4,
         "expected_set": [1000, 999, 998, 997]},
        {"source": "test/test_filter.jl:106-121 (scaled, boundary 1)", "data": "1:1000", "k": 5, "withscale": True,
         "largest": True, "kth_scaled_of": 996, "boundary": 1, "expected_count": 5,
         "expected_set": [1000, 999, 998, 997, 996]},
        {"source": "test/test_filter.jl:123-138 (unscaled, kth = 995)", "data": "1:1000", "k": 5,
         "withscale": False, "largest": True, "kth": 995.0, "boundary": 0, "expected_count": 5,
         "expected_set": [1000, 999, 998, 997, 996]},
        {"source": "test/test_filter.jl:705-745 (WITHIDXIN)", "data": [100, 85, 92, 78, 95, 88, 90, 82],
         "indices": [10, 11, 12, 13, 14, 15, 16, 17], "k": 3, "withscale": False, "largest": True, "kth": 92.0,
         "boundary": 3, "expected_count": 3, "expected_set": [100, 95, 92], "expected_indices": [10, 14, 12]},
        {"source": "test/test_filter.jl:783-826 (empty result)", "data": [1, 2, 3], "k": 2, "withscale": False,
         "largest": True, "kth": 100.0, "boundary": 0, "expected_count": 0, "expected_set": []},
        {"source": "test/test_filter.jl:829-891 (all equal)", "data": [50, 50, 50, 50, 50], "k": 3,
         "withscale": False, "largest": True, "kth": 50.0, "boundary": 3, "expected_count": 3,
         "expected_set": [50, 50, 50]},
    ]
    ka["filter_multitask"] = {
        "source": "test/test_filter.jl:254-315",
        "tasks": [[1, 2, 3, 4, 5], [10, 20, 30, 40, 50], [100, 200, 300, 400, 500]],
        "k": 2, "kth_index": 4, "boundary": 2,
        "expected_sets": [[5, 4], [50, 40], [500, 400]],
    }
    # test/test_algorithm.jl -- end-to-end shapes (data there is unseeded GPU
    # rand!, so only the shapes, k grid and properties are transcribed).
    ka["algorithm"] = {
        "source": "test/test_algorithm.jl:6-257",
        "basic": {"n": 16384, "k": [16, 32, 64, 128, 256, 512, 1024, 2048, 4096]},
        "multitask": {"task_lens": [1000, 2000, 1500, 500, 5000], "k": 64, "ws_n": 5000},
        "overloads": {"n": 1000, "k": 100},
        "batch": {"task_lens": [1000, 1500, 800], "k": 50},
        "errors": {"n": 1000, "k": 100, "wrong_out": [50, 1], "wrong_indices": 500},
    }
    return ka


def specials():
    return np.array([0.0, -0.0, F32_MIN_SUB, -F32_MIN_SUB, F32_MAX_SUB, -F32_MAX_SUB,
                     np.inf, -np.inf, 1.17549435e-38, -1.17549435e-38, 3.4028235e38, -3.4028235e38],
                    dtype=np.float32)


def make_cases():
    """Seeded small cases; each entry: x, lens, k, flags, idx_in (or empty), vals, idx."""
    rng = np.random.default_rng(20251203)
    cases = []

    shared = {}

    def add(name, x, lens, k, flags, idx=None):
        x = np.asarray(x, dtype=np.float32)
        v, i = O.topk_exact(x, lens, k, flags, idx_in=idx)
        # inputs are stored once and referenced by id
        xid = shared.setdefault(("x", x.tobytes()), (f"x{len(shared)}", x))[0]
        iid = ""
        if idx is not None:
            idx = np.asarray(idx, np.int32)
            iid = shared.setdefault(("i", idx.tobytes()), (f"i{len(shared)}", idx))[0]
        cases.append((name, xid, np.asarray(lens, dtype=np.int64), k, flags, iid, v, i))

    L, A, S = O.LARGEST, O.ASCEND, O.WITHSCALE
    # configs[0]-shaped, scaled down: N(0,1), k=100, largest, rev=false
    add("normal_5e4_k100", rng.standard_normal(50_000), [50_000], 100, L | A)
    # configs[1]-shaped: uniform batch, k=100; and the 64-level tie variant
    u = rng.random(100 * 1000, dtype=np.float32)
    add("uniform_100x1000_k100", u, [1000] * 100, 100, L | A)
    add("ties64_100x1000_k100", np.floor(u * 64) / 64, [1000] * 100, 100, L | A)
    # configs[2]-shaped, scaled down: custom Int32 indices (permutation), all four modes
    n = 1 << 16
    x = rng.standard_normal(n).astype(np.float32)
    perm = rng.permutation(n).astype(np.int32)
    for k in (16, 1024, 4096):
        for fl in (L, L | A, 0, A):
            add(f"normal_2p16_k{k}_f{fl}", x, [n], k, fl, perm)
    # configs[3]-shaped, scaled down: uneven lengths, sigma spread, duplicates, specials
    T = 32
    lens = np.round(2.0 ** (10 + 4 * rng.random(T))).astype(np.int64)
    parts = []
    for t in range(T):
        sig = 10.0 ** (-3 + 6 * rng.random())
        v = (rng.standard_normal(lens[t]) * sig).astype(np.float32)
        if t % 4 == 0:
            m = rng.random(lens[t]) < 0.3
            fixed = (np.arange(1, 9, dtype=np.float32) * 0.5 * np.float32(sig)).astype(np.float32)
            v[m] = fixed[rng.integers(0, 8, m.sum())]
        sp = rng.random(lens[t]) < 2e-3
        v[sp] = specials()[rng.integers(0, 8, sp.sum())]
        parts.append(v)
    add("uneven32_k256", np.concatenate(parts), lens, 256, L | A)
    # configs[4]-shaped rows (narrow): spikes of +20 -> ties at k=1
    rows, cols = 4, 12000
    m = (rng.standard_normal(rows * cols) * 3).astype(np.float32)
    for r in range(rows):
        m[r * cols + rng.integers(0, cols, 4)] = 20.0
    for k in (1, 50, 1024):
        add(f"spikes_4x12000_k{k}", m, [cols] * rows, k, L)
    # WITHSCALE with clustered values and subnormal differences
    c = (np.float32(1.0) + rng.integers(0, 50, 30000).astype(np.float32) * np.float32(2 ** -20)).astype(np.float32)
    add("withscale_clustered_k100", c, [30000], 100, L | A | S)
    add("withscale_subnormal_k64", (rng.integers(-40, 40, 5000) * F32_MIN_SUB).astype(np.float32), [5000], 64,
        L | A | S)
    # all equal, n < k, empty tasks, k = 1
    add("all_equal_k300", np.full(5000, 7.0, np.float32), [5000], 300, L | A)
    add("ragged_small", rng.standard_normal(2100), [0, 5, 1000, 1, 1094], 64, L | A)
    add("k_gt_n", rng.standard_normal(300), [100, 200], 250, 0)
    add("specials_only", np.tile(specials(), 400), [len(specials()) * 400], 500, L)
    return cases, {key: arr for key, arr in shared.values()}


def main():
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(known_answers(), f, indent=1)
    cases, shared = make_cases()
    arrays = {"in_" + key: arr for key, arr in shared.items()}
    for j, (name, xid, lens, k, flags, iid, v, i) in enumerate(cases):
        pre = f"c{j:02d}_"
        arrays[pre + "name"] = np.array(name)
        arrays[pre + "x"] = np.array(xid)
        arrays[pre + "lens"] = lens
        arrays[pre + "k"] = np.array(k, np.int64)
        arrays[pre + "flags"] = np.array(flags, np.int64)
        arrays[pre + "idx"] = np.array(iid)
        arrays[pre + "vals"] = v
        arrays[pre + "inds"] = i
    np.savez_compressed(os.path.join(HERE, "cases.npz"), **arrays)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
