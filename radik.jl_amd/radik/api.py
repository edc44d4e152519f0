"""Python mirror of RadiK.jl's exported API over libradik_hip.so.

Names, argument meaning and error behaviour follow the reference
(src/RadiK.jl:20-23, src/algorithm.jl:58-559); Julia's `!` becomes a trailing
underscore and Val switches become booleans:

  RadiKWorkspace(n, num_tasks, I=np.int32)            algorithm.jl:58-83
  topk_radix_select_(val_out, idx_out, ws, val_in, idx_in, task_lens, k,
      LARGEST=True, ASCEND=True, WITHSCALE=False, WITHIDXIN=False,
      WITHPACKING=True, VLOAD_SIZE_BYTES=32)           algorithm.jl:162-368
  topk_(val_out, idx_out, data, [task_lens,] k; indices, largest, rev)
                                                      algorithm.jl:477-526
  topk(data, [task_lens,] k; indices, largest, rev)   algorithm.jl:528-559

Arrays are DeviceArray (device.py); outputs are [k, num_tasks] column-major.
Indices are 1-based positions within each task unless `indices` are given.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import RadiKError
from .device import DeviceArray


def _stream_handle(stream):
    return None if stream is None else stream.handle


class RadiKWorkspace:
    """Device workspace for topk_radix_select_ (algorithm.jl:58-83).

    n: longest task served; num_tasks: most tasks per call; I: np.int32 or
    np.int64 counters (the reference's workspace integer type)."""

    def __init__(self, n, num_tasks, I=np.int32, stream=None):
        self.n = int(n)
        self.num_tasks = int(num_tasks)
        self.I = np.dtype(I)
        if self.I.itemsize not in (4, 8):
            raise ValueError("I must be a 32- or 64-bit integer type")
        h = ctypes.c_void_p()
        _lib.check(_lib.load().radik_ws_create(ctypes.byref(h), self.n, self.num_tasks, self.I.itemsize,
                                               _stream_handle(stream)))
        self.handle = h.value

    def __del__(self):
        try:
            if self.handle:
                _lib.load().radik_ws_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _flags(LARGEST, ASCEND, WITHSCALE, WITHIDXIN, WITHPACKING):
    f = 0
    f |= _lib.RADIK_LARGEST if LARGEST else 0
    f |= _lib.RADIK_ASCEND if ASCEND else 0
    f |= _lib.RADIK_WITHSCALE if WITHSCALE else 0
    f |= _lib.RADIK_WITHIDXIN if WITHIDXIN else 0
    f |= _lib.RADIK_WITHPACKING if WITHPACKING else 0
    return f


KEYS = {"f32": 0, "f16": 1, "bf16": 2, "i32": 3, "f64": 4}
_KEY_OF_DTYPE = {np.dtype(np.float32): "f32", np.dtype(np.float16): "f16", np.dtype(np.int32): "i32",
                 np.dtype(np.float64): "f64"}


def key_type(data, key=None):
    """Key type code of `data`: float32, float16, int32 and float64 arrays
    by their dtype; bfloat16 data are uint16 bit patterns with key="bf16"
    (float64 keys: topk / topk_ / topk_radix_select_ only, k <= 8192).  The
    reference takes Float32 only (algorithm.jl:492-494) and raises for
    anything else; here the other types run the same path on their own
    order-preserving keys."""
    if key is not None:
        if key not in KEYS:
            raise RadiKError(_lib.RADIK_EINVAL, f"unknown key type {key!r}; one of {sorted(KEYS)}")
        if key == "bf16" and np.dtype(data.dtype) != np.uint16:
            raise RadiKError(_lib.RADIK_EINVAL, "bfloat16 data are passed as uint16 bit patterns")
        if key != "bf16" and _KEY_OF_DTYPE.get(np.dtype(data.dtype)) != key:
            raise RadiKError(_lib.RADIK_EINVAL, f"data of dtype {data.dtype} cannot be read as {key}")
        return KEYS[key]
    k = _KEY_OF_DTYPE.get(np.dtype(data.dtype))
    if k is None:
        # algorithm.jl:492-494 (the reference raises MethodError earlier, Q10)
        raise RadiKError(_lib.RADIK_EINVAL,
                         f"RadiK supports Float32, Float16, BFloat16 (uint16 bits, key='bf16'), Int32 and Float64 "
                         f"input data. Got {data.dtype}.")
    return KEYS[k]


def _check_f32(data):
    if np.dtype(data.dtype) != np.float32:
        # algorithm.jl:492-494 (the reference raises MethodError earlier, Q10)
        raise RadiKError(_lib.RADIK_EINVAL,
                         f"RadiK.jl only supports Float32 input data. Got {data.dtype}.")


def topk_radix_select_(val_out, idx_out, ws, val_in, idx_in, task_lens, k, LARGEST=True, ASCEND=True,
                       WITHSCALE=False, WITHIDXIN=False, WITHPACKING=True, VLOAD_SIZE_BYTES=32, stream=None,
                       key=None):
    """Writes the top-k of every task into val_out/idx_out ([k, T]); returns them.
    val_out has the element type of val_in (key: see key_type)."""
    kt = key_type(val_in, key)
    if np.dtype(val_out.dtype) != np.dtype(val_in.dtype):
        raise RadiKError(_lib.RADIK_EINVAL, "val_out must have the element type of val_in")
    lens = [int(x) for x in task_lens]
    flags = _flags(LARGEST, ASCEND, WITHSCALE, WITHIDXIN, WITHPACKING) | (kt << 8)
    idx_bytes = np.dtype(idx_out.dtype).itemsize
    # the library sees pointers only: sizes are checked here
    n_in, n_out = sum(lens), int(k) * len(lens)
    if val_in.size < n_in:
        raise RadiKError(_lib.RADIK_ESIZE, f"val_in holds {val_in.size} elements, task_lens sum to {n_in}")
    if WITHIDXIN and idx_in is not None and idx_in.size < n_in:
        raise RadiKError(_lib.RADIK_ESIZE, f"idx_in holds {idx_in.size} indices, task_lens sum to {n_in}")
    if val_out.size < n_out or idx_out.size < n_out:
        raise RadiKError(_lib.RADIK_ESIZE, f"val_out/idx_out must hold k x num_tasks = {n_out} results")
    if WITHIDXIN and idx_in is not None and np.dtype(idx_in.dtype).itemsize != idx_bytes:
        raise TypeError("idx_in and idx_out must share their element width")
    _lib.check(_lib.load().radik_topk_radix_select(
        ws.handle, val_out.ptr, idx_out.ptr, idx_bytes, val_in.ptr,
        idx_in.ptr if (idx_in is not None and WITHIDXIN) else None,
        _lib.i64_array(lens), len(lens), int(k), flags, int(VLOAD_SIZE_BYTES), _stream_handle(stream)))
    return val_out, idx_out


def topk_(val_out, idx_out, data, *args, indices=None, largest=True, rev=False, stream=None, key=None):
    """topk_(val_out, idx_out, data, [task_lens,] k; indices, largest, rev).

    Validates sizes (ErrorException messages of algorithm.jl:498-509) and
    allocates a workspace for the call (algorithm.jl:503)."""
    if len(args) == 1:
        task_lens, k = [data.size], args[0]
    elif len(args) == 2:
        task_lens, k = args
    else:
        raise TypeError("topk_(val_out, idx_out, data, [task_lens,] k)")
    kt = key_type(data, key)
    if np.dtype(val_out.dtype) != np.dtype(data.dtype):
        raise RadiKError(_lib.RADIK_EINVAL, "val_out must have the element type of data")
    lens = [int(x) for x in task_lens]
    T = len(lens)
    shape = val_out.shape if len(val_out.shape) == 2 else (val_out.shape[0], 1)
    ishape = idx_out.shape if len(idx_out.shape) == 2 else (idx_out.shape[0], 1)
    if tuple(shape) != (int(k), T):
        raise RadiKError(_lib.RADIK_ESIZE, f"val_out must have size ({int(k)}, {T}), got {tuple(val_out.shape)}")
    if tuple(ishape) != (int(k), T):
        raise RadiKError(_lib.RADIK_ESIZE, f"idx_out must have size ({int(k)}, {T}), got {tuple(idx_out.shape)}")
    if data.size < sum(lens):
        raise RadiKError(_lib.RADIK_ESIZE, f"data holds {data.size} elements, task_lens sum to {sum(lens)}")
    n_ind = 0 if indices is None else indices.size
    if indices is not None and n_ind > 0:
        if np.dtype(indices.dtype) != np.dtype(idx_out.dtype):
            raise TypeError("indices and idx_out must share their element type")
    idx_bytes = np.dtype(idx_out.dtype).itemsize
    _lib.check(_lib.load().radik_topk_typed(
        kt, val_out.ptr, idx_out.ptr, idx_bytes, shape[0], shape[1], data.ptr,
        indices.ptr if n_ind else None, n_ind, _lib.i64_array(lens), T, int(k),
        1 if largest else 0, 1 if rev else 0, _stream_handle(stream)))
    return val_out, idx_out


def topk(data, *args, indices=None, largest=True, rev=False, stream=None, key=None):
    """topk(data, k) -> (values[k], indices[k]);
    topk(data, task_lens, k) -> (values[k, T], indices[k, T]).

    Index element type: Int32, or that of `indices` (algorithm.jl:548)."""
    if len(args) == 1:
        k = int(args[0])
        v, i = topk(data, [data.size], k, indices=indices, largest=largest, rev=rev, stream=stream, key=key)
        return v.column(0), i.column(0)
    if len(args) != 2:
        raise TypeError("topk(data, [task_lens,] k)")
    task_lens, k = args
    key_type(data, key)
    T = len(task_lens)
    k = int(k)
    idx_type = np.int32 if indices is None else indices.dtype
    val_out = DeviceArray.zeros((k, T), data.dtype)
    idx_out = DeviceArray.zeros((k, T), idx_type)
    topk_(val_out, idx_out, data, task_lens, k, indices=indices, largest=largest, rev=rev, stream=stream, key=key)
    return val_out, idx_out

