"""Persistent batched form and serving engine over libradik_hip.so.

Plan: one pass sequence for many tasks of uneven length, each with its own k,
staged once and launched eagerly on every execute, or replayed from a
captured hipGraph with use_graph=True (no reference counterpart: the
reference launches one grid per pass, RadiK.jl src/algorithm.jl:221-334).
Server: continuous batching of top-k requests against a device budget.
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray


class Plan:
    def __init__(self, task_lens, k, largest=True, rev=False, withscale=False, indices_dtype=None,
                 idx_dtype=np.int32, use_graph=False, stream=None, tuned=True, key="f32"):
        lens = [int(x) for x in task_lens]
        ks = [int(k)] * len(lens) if np.isscalar(k) else [int(x) for x in k]
        if len(ks) != len(lens):
            raise ValueError("k must be a scalar or one value per task")
        self.task_lens = lens
        self.ks = ks
        self.idx_dtype = np.dtype(idx_dtype)
        flags = _lib.RADIK_WITHPACKING
        flags |= _lib.RADIK_LARGEST if largest else 0
        flags |= 0 if rev else _lib.RADIK_ASCEND
        flags |= _lib.RADIK_WITHSCALE if withscale else 0
        flags |= _lib.RADIK_WITHIDXIN if indices_dtype is not None else 0
        from .api import KEYS
        if key not in KEYS:
            raise _lib.RadiKError(_lib.RADIK_EINVAL, f"unknown key type {key!r}")
        self.key = key
        self.val_dtype = np.dtype({"f32": np.float32, "f16": np.float16, "bf16": np.uint16, "i32": np.int32,
                                   "f64": np.float64}[key])
        flags |= KEYS[key] << 8
        self.flags = flags
        h = ctypes.c_void_p()
        _lib.check(_lib.load().radik_plan_create(ctypes.byref(h), _lib.i64_array(lens), _lib.i64_array(ks),
                                                 len(lens), flags, self.idx_dtype.itemsize,
                                                 None if stream is None else stream.handle))
        self.handle = h.value
        _lib.check(_lib.load().radik_plan_set_graph(self.handle, 1 if use_graph else 0))
        self.options = None
        if tuned:
            from . import autotune
            self.options = autotune.apply(self)

    @property
    def out_size(self):
        return int(sum(self.ks))

    def execute(self, val_out, idx_out, val_in, idx_in=None, stream=None):
        # buffer sizes are checked here: the library only sees pointers
        n_in = sum(self.task_lens)
        if val_in.size < n_in:
            raise _lib.RadiKError(_lib.RADIK_ESIZE, f"val_in holds {val_in.size} elements, the plan reads {n_in}")
        if (self.flags & _lib.RADIK_WITHIDXIN) and (idx_in is None or idx_in.size < n_in):
            raise _lib.RadiKError(_lib.RADIK_ESIZE, f"idx_in must hold {n_in} indices")
        if val_out.size < self.out_size or idx_out.size < self.out_size:
            raise _lib.RadiKError(_lib.RADIK_ESIZE, f"val_out/idx_out must hold {self.out_size} results")
        if np.dtype(val_in.dtype) != self.val_dtype or np.dtype(val_out.dtype) != self.val_dtype:
            raise _lib.RadiKError(_lib.RADIK_EINVAL, f"val_in/val_out must be {self.val_dtype} for key {self.key}")
        if np.dtype(idx_out.dtype).itemsize != self.idx_dtype.itemsize:
            raise _lib.RadiKError(_lib.RADIK_EINVAL, f"idx_out must be {self.idx_dtype}")
        if idx_in is not None and np.dtype(idx_in.dtype).itemsize != self.idx_dtype.itemsize:
            raise _lib.RadiKError(_lib.RADIK_EINVAL, f"idx_in must be {self.idx_dtype}")
        _lib.check(_lib.load().radik_plan_execute(self.handle, val_out.ptr, idx_out.ptr, val_in.ptr,
                                                  None if idx_in is None else idx_in.ptr,
                                                  None if stream is None else stream.handle))

    def set_graph(self, use_graph):
        """Replay a captured hipGraph (True) or launch the kernels eagerly."""
        _lib.check(_lib.load().radik_plan_set_graph(self.handle, 1 if use_graph else 0))

    def set_options(self, vload_bytes=0, blocks_per_cu=0, small_cap=0, refine_cap=0, predict=True):
        """Tuning knobs of the launch sequence; results do not depend on them."""
        _lib.check(_lib.load().radik_plan_set_options(self.handle, int(vload_bytes), int(blocks_per_cu),
                                                      int(small_cap), int(refine_cap), 1 if predict else 0))

    def sample(self, val_out, idx_out, temperature, top_p, uniforms=None, tokens=None, ranks=None, stream=None):
        """Temperature / top-p sampling from the results of the last execute
        (radik_plan_sample; the plan must be largest=True, rev=True).
        Returns (tokens int64[T], ranks int32[T]) device arrays."""
        T = len(self.task_lens)
        if tokens is None:
            tokens = DeviceArray.zeros(T, np.int64)
        if ranks is None:
            ranks = DeviceArray.zeros(T, np.int32)
        if uniforms is not None and uniforms.size < T:
            raise _lib.RadiKError(_lib.RADIK_ESIZE, f"uniforms must hold {T} values")
        if tokens.size < T or ranks.size < T:
            raise _lib.RadiKError(_lib.RADIK_ESIZE, f"tokens/ranks must hold {T} entries")
        _lib.check(_lib.load().radik_plan_sample(self.handle, val_out.ptr, idx_out.ptr, float(temperature),
                                                 float(top_p), None if uniforms is None else uniforms.ptr,
                                                 tokens.ptr, ranks.ptr, None if stream is None else stream.handle))
        return tokens, ranks

    def info(self):
        """How the last execute staged the tasks (radik_plan_info)."""
        v = (ctypes.c_int64 * 9)()
        _lib.check(_lib.load().radik_plan_info(self.handle, v, 9))
        names = ("n_long", "long_elements", "n_medium", "medium_elements", "n_short", "short_elements",
                 "hist_chunks", "sample_groups", "has_big")
        return {nm: int(v[i]) for i, nm in enumerate(names)}

    def set_option(self, name, value):
        """One launch knob by name (radik_plan_set_option); results do not depend on it."""
        _lib.check(_lib.load().radik_plan_set_option(self.handle, name.encode(), int(value)))

    def set_timing(self, on):
        _lib.check(_lib.load().radik_plan_set_timing(self.handle, 1 if on else 0))

    def timing(self):
        ms = (ctypes.c_double * 8)()
        n = ctypes.c_int64()
        _lib.check(_lib.load().radik_plan_timing(self.handle, ms, ctypes.byref(n)))
        names = ("sample", "predict", "hist", "select", "compact", "finish", "sort", "small")
        return {nm: ms[i] for i, nm in enumerate(names)}, n.value

    def __del__(self):
        try:
            if self.handle:
                _lib.load().radik_plan_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Server:
    def __init__(self, budget_bytes=16 << 30, max_inflight=4, largest=True, rev=True):
        flags = (_lib.RADIK_LARGEST if largest else 0) | (0 if rev else _lib.RADIK_ASCEND) | _lib.RADIK_WITHPACKING
        h = ctypes.c_void_p()
        _lib.check(_lib.load().radik_server_create(ctypes.byref(h), int(budget_bytes), int(max_inflight), flags))
        self.handle = h.value

    def submit(self, x, rows, cols, k, val_out, idx_out):
        t = ctypes.c_int64()
        _lib.check(_lib.load().radik_server_submit(self.handle, x.ptr, int(rows), int(cols), int(k), val_out.ptr,
                                                   idx_out.ptr, ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        lat = ctypes.c_double()
        rc = _lib.load().radik_server_poll(self.handle, int(ticket), ctypes.byref(lat))
        if rc == _lib.RADIK_ENOTREADY:
            return None
        _lib.check(rc)
        return lat.value

    def wait(self, ticket):
        lat = ctypes.c_double()
        _lib.check(_lib.load().radik_server_wait(self.handle, int(ticket), ctypes.byref(lat)))
        return lat.value

    def drain(self):
        _lib.check(_lib.load().radik_server_drain(self.handle))

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.load().radik_server_stats(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"launches": a.value, "rows": b.value, "max_coalesced": c.value}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().radik_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
