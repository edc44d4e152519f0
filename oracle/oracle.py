"""ctypes front of the CPU oracle (oracle/radik_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg, always as the checker.  The product never imports it.

Flag bits match the C-ABI of the HIP library: LARGEST=1, ASCEND=2,
WITHSCALE=4, WITHIDXIN=8 (the Val switches of topk_radix_select!,
RadiK.jl src/algorithm.jl:170-175).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

LARGEST, ASCEND, WITHSCALE, WITHIDXIN = 1, 2, 4, 8

_lib = None


def build():
    """Compile liboracle.so from radik_oracle.c (gcc, no fast-math)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.orc_key.argtypes = [ctypes.c_float]
        L.orc_key.restype = ctypes.c_uint32
        L.orc_key_inv.argtypes = [ctypes.c_uint32]
        L.orc_key_inv.restype = ctypes.c_float
        L.orc_bin_id.argtypes = [ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.orc_bin_id.restype = ctypes.c_uint32
        L.orc_sample_scaler.argtypes = [ctypes.c_float, ctypes.c_int]
        L.orc_sample_scaler.restype = ctypes.c_float
        L.orc_scaled.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.orc_scaled.restype = ctypes.c_float
        L.orc_apply_scaling_ref.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.orc_apply_scaling_ref.restype = ctypes.c_float
        L.orc_warp_scan.argtypes = [i64p, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.orc_warp_scan.restype = None
        L.orc_cross_warp_reduction.argtypes = [i64p, i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_cross_warp_reduction.restype = None
        L.orc_find_bin_and_write.argtypes = [i64p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p]
        L.orc_find_bin_and_write.restype = ctypes.c_int
        L.orc_select_bin_blocked.argtypes = [i64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int64, ctypes.c_int, i64p]
        L.orc_select_bin_blocked.restype = ctypes.c_int
        L.orc_select_bin.argtypes = [i64p, ctypes.c_int, ctypes.c_int64, ctypes.c_int, i64p]
        L.orc_select_bin.restype = ctypes.c_int
        L.orc_count_bin_ex.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p]
        L.orc_count_bin_ex.restype = None
        L.orc_rank_histogram.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int,
                                         ctypes.c_int, i64p]
        L.orc_rank_histogram.restype = None
        L.orc_filter.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32,
                                 ctypes.c_int64, ctypes.c_uint32, ctypes.c_void_p, i64p]
        L.orc_filter.restype = ctypes.c_int64
        for name in ("orc_topk_exact",):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, i64p, ctypes.c_int64,
                          ctypes.c_int64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
            f.restype = ctypes.c_int
        L.orc_topk_schedule.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, i64p, ctypes.c_int64,
                                        ctypes.c_int64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                        i64p, ctypes.c_int]
        L.orc_topk_schedule.restype = ctypes.c_int
        L.orc_key_typed.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.orc_key_typed.restype = ctypes.c_uint32
        L.orc_topk_exact_typed.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, i64p,
                                           ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int]
        L.orc_topk_exact_typed.restype = ctypes.c_int
        L.orc_nstat.restype = ctypes.c_int
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def key(v):
    return lib().orc_key(float(np.float32(v)))


def keys(arr):
    """Vectorised orc_key (same closed form, numpy)."""
    b = np.ascontiguousarray(arr, dtype=np.float32).view(np.uint32)
    return np.where(b >> 31 != 0, b ^ np.uint32(0xFFFFFFFF), b ^ np.uint32(0x80000000)).astype(np.uint32)


def bin_id(v, left, right):
    return lib().orc_bin_id(float(np.float32(v)), left, right)


def warp_scan(v, ws, up):
    a, p = _i64(np.array(v).copy())
    lib().orc_warp_scan(p, a.size, ws, 1 if up else 0)
    return a


def cross_warp_reduction(inp, block, warp, largest):
    ps, pp = _i64(np.array(inp).copy())
    sm, sp = _i64(np.array(inp).copy())
    lib().orc_cross_warp_reduction(pp, sp, block, warp, 1 if largest else 0)
    return sm


def find_bin_and_write(counts, old_k, neighbor, total, thread_x, unroll, largest):
    c, cp = _i64(counts)
    out, op = _i64(np.zeros(3))
    ok = lib().orc_find_bin_and_write(cp, old_k, neighbor, total, thread_x, unroll, 1 if largest else 0, op)
    return (int(out[0]), int(out[1]), int(out[2])) if ok else None


def select_bin_blocked(hist, block, warp, k, largest):
    h, hp = _i64(hist)
    out, op = _i64(np.zeros(3))
    ok = lib().orc_select_bin_blocked(hp, h.size, block, warp, k, 1 if largest else 0, op)
    if ok < 0:
        raise ValueError("histlen must be a multiple of block and block of warp")
    return (int(out[0]), int(out[1]), int(out[2])) if ok else None


def select_bin(hist, k, largest):
    h, hp = _i64(hist)
    out, op = _i64(np.zeros(3))
    ok = lib().orc_select_bin(hp, h.size, k, 1 if largest else 0, op)
    return (int(out[0]), int(out[1]), int(out[2])) if ok else None


def count_bin_ex(data, left, right, histlen, withscale, largest):
    d = np.ascontiguousarray(data, dtype=np.float32)
    h, hp = _i64(np.zeros(histlen))
    lib().orc_count_bin_ex(d.ctypes.data, d.size, left, right, histlen, int(withscale), int(largest), hp)
    return h


def rank_histogram(data, flags, shift, width):
    d = np.ascontiguousarray(data, dtype=np.float32)
    h, hp = _i64(np.zeros(1 << width))
    lib().orc_rank_histogram(d.ctypes.data, d.size, flags, shift, width, hp)
    return h


def filter_task(data, k, kth_rank, boundary, flags):
    d = np.ascontiguousarray(data, dtype=np.float32)
    vals = np.zeros(max(k, 1), dtype=np.float32)
    pos, pp = _i64(np.zeros(max(k, 1)))
    c = lib().orc_filter(d.ctypes.data, d.size, k, kth_rank, boundary, flags, vals.ctypes.data, pp)
    return vals[:c], pos[:c]


def _topk(fn, val_in, task_lens, k, flags, idx_in=None, idx_dtype=np.int32, threads=0, stats=False):
    v = np.ascontiguousarray(val_in, dtype=np.float32)
    tl, tlp = _i64(task_lens)
    T = tl.size
    idx_bytes = np.dtype(idx_dtype).itemsize
    if idx_in is not None:
        ii = np.ascontiguousarray(idx_in, dtype=idx_dtype)
        flags |= WITHIDXIN
        iptr = ii.ctypes.data
    else:
        flags &= ~WITHIDXIN
        iptr = None
    vo = np.zeros((T, k), dtype=np.float32)          # row t = column t of the [k, T] output
    io = np.zeros((T, k), dtype=idx_dtype)
    if fn == "exact":
        rc = lib().orc_topk_exact(v.ctypes.data, iptr, idx_bytes, tlp, T, k, flags,
                                  vo.ctypes.data, io.ctypes.data, threads)
        st = None
    else:
        st, sp = _i64(np.zeros((max(T, 1), lib().orc_nstat())))
        rc = lib().orc_topk_schedule(v.ctypes.data, iptr, idx_bytes, tlp, T, k, flags,
                                     vo.ctypes.data, io.ctypes.data, sp, threads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return (vo, io, st) if stats else (vo, io)


def topk_exact(val_in, task_lens, k, flags, idx_in=None, idx_dtype=np.int32, threads=0):
    """[T, k] values and indices (row t = task t) by the contract's direct order."""
    return _topk("exact", val_in, task_lens, k, flags, idx_in, idx_dtype, threads)


def topk_schedule(val_in, task_lens, k, flags, idx_in=None, idx_dtype=np.int32, threads=0, stats=False):
    """Same outputs by the reference's 3-pass schedule (plus per-task pass stats)."""
    return _topk("schedule", val_in, task_lens, k, flags, idx_in, idx_dtype, threads, stats)


def max_threads():
    return lib().orc_max_threads()


def reorder(v, descending):
    """Permutation that stably re-sorts a list of values by key (ascending or
    descending): equal keys keep their relative order.  Turns an oracle
    output in one order (ties by position) into the other order."""
    kk = keys(v).astype(np.int64)
    return np.lexsort((np.arange(v.size), -kk if descending else kk))


# Key types beyond Float32: numpy dtype of the element bit patterns per type.
KEY_F32, KEY_F16, KEY_BF16, KEY_I32 = 0, 1, 2, 3
KEY_DTYPES = {KEY_F32: np.float32, KEY_F16: np.float16, KEY_BF16: np.uint16, KEY_I32: np.int32}


def key_typed(bits, kt):
    return lib().orc_key_typed(int(bits), int(kt))


def topk_exact_typed(kt, val_in, task_lens, k, flags, idx_in=None, idx_dtype=np.int32, threads=0):
    """[T, k] values (the key type's bit patterns in KEY_DTYPES[kt]) and
    indices, by the contract's direct order over the type's key.  BFloat16
    data is passed as uint16 bit patterns."""
    v = np.ascontiguousarray(val_in, dtype=KEY_DTYPES[kt])
    tl, tlp = _i64(task_lens)
    T = tl.size
    idx_bytes = np.dtype(idx_dtype).itemsize
    if idx_in is not None:
        ii = np.ascontiguousarray(idx_in, dtype=idx_dtype)
        flags |= WITHIDXIN
        iptr = ii.ctypes.data
    else:
        flags &= ~WITHIDXIN
        iptr = None
    vo = np.zeros((T, k), dtype=KEY_DTYPES[kt])
    io = np.zeros((T, k), dtype=idx_dtype)
    rc = lib().orc_topk_exact_typed(kt, v.ctypes.data, iptr, idx_bytes, tlp, T, k, flags, vo.ctypes.data,
                                    io.ctypes.data, threads)
    if rc != 0:
        raise ValueError("oracle rejected the arguments")
    return vo, io


def bf16_to_f32(bits):
    """BFloat16 bit patterns (uint16) as the float32 values they denote."""
    return (np.asarray(bits, np.uint16).astype(np.uint32) << np.uint32(16)).view(np.float32)


def f32_to_bf16(x):
    """Float32 values rounded to BFloat16 (nearest, ties to even) bit patterns."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    out = r.astype(np.uint16)
    nan = np.isnan(np.asarray(x, np.float32))
    out[nan] = ((u[nan] >> np.uint64(16)) | np.uint64(0x40)).astype(np.uint16)
    return out


def sample_topp(vals_desc, m, temperature, top_p, u):
    """Temperature / top-p sampling restated in float64 (the oracle of
    radik_plan_sample): per row, over the first m descending values v,
    w_i = exp((v_i - v_1) / T) (1 where v_i == v_1), c = cumsum(w), the
    nucleus n = first i with c_i >= top_p * c_m (m if none), the sample r =
    first i <= n with c_i > u * c_n.  Returns (rank r - 1, slack) per row;
    slack is the float64 distance of the two decisions from their nearest
    boundary, relative to c_m (a float32 evaluation may legitimately land on
    the other side of a boundary closer than its rounding error)."""
    vals_desc = np.asarray(vals_desc, np.float64)
    ranks = np.full(len(m), -1, np.int64)
    slack = np.full(len(m), np.inf)
    for t, mt in enumerate(m):
        if mt <= 0:
            continue
        v = vals_desc[t, :mt]
        if not temperature > 0:
            ranks[t] = 0
            continue
        w = np.where(v == v[0], 1.0, np.exp((v - v[0]) / np.float64(temperature)))
        c = np.cumsum(w)
        z = c[-1]
        want = top_p * z
        hit = np.nonzero(c >= want)[0]
        n = hit[0] + 1 if hit.size else mt
        cn = c[n - 1]
        th = u[t] * cn
        r = int(np.nonzero(c[:n] > th)[0][0]) + 1 if np.any(c[:n] > th) else n
        ranks[t] = r - 1
        d1 = np.min(np.abs(c - want)) / z
        d2 = np.min(np.abs(c[:n] - th)) / z
        slack[t] = min(d1, d2)
    return ranks, slack


# Float64 keys: restated with numpy (the C restatement works on 32-bit keys).
KEY_F64 = 4
KEY_DTYPES[KEY_F64] = np.float64


def keys64(x):
    """Order-preserving 64-bit key of float64 values: the key transform of
    RadiK.jl src/utils.jl:29-33 widened to 64 bits (bits ^ (sign ? ~0 :
    1 << 63)); unsigned order of the keys is IEEE totalOrder."""
    b = np.ascontiguousarray(x, np.float64).view(np.uint64)
    return np.where(b >> np.uint64(63) != 0, ~b, b ^ np.uint64(1 << 63))


def topk_exact_f64(val_in, task_lens, k, flags, idx_in=None, idx_dtype=np.int32):
    """[T, k] float64 values and indices by the contract's direct order:
    per task the m = min(k, n) smallest (rank, position), rank = ~key for
    LARGEST else key; written sorted by (ASCEND ? key : ~key, position)
    (the final sort, src/algorithm.jl:340-352); indices idx_in[pos] or
    pos + 1; slots past m hold 0.0 / -1 (algorithm.jl:355-365)."""
    x = np.ascontiguousarray(val_in, np.float64)
    lens = [int(n) for n in task_lens]
    T = len(lens)
    vo = np.zeros((T, k), np.float64)
    io = np.full((T, k), -1, idx_dtype)
    ones = np.uint64(0xFFFFFFFFFFFFFFFF)
    off = 0
    for t, n in enumerate(lens):
        seg = x[off:off + n]
        key = keys64(seg)
        rank = key ^ ones if flags & LARGEST else key
        pos = np.arange(n, dtype=np.int64)
        m = min(k, n)
        sel = np.lexsort((pos, rank))[:m]
        okey = key[sel] if flags & ASCEND else key[sel] ^ ones
        sel = sel[np.lexsort((sel, okey))]
        vo[t, :m] = seg[sel]
        if idx_in is not None:
            io[t, :m] = np.asarray(idx_in)[off + sel]
        else:
            io[t, :m] = sel + 1
        off += n
    return vo, io
