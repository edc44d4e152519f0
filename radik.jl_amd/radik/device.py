"""Minimal device arrays, streams and events over the library's helpers.

A DeviceArray is a HIP allocation with a numpy dtype and a shape; 2-D arrays
are column-major like the Julia arrays the reference takes ([k, num_tasks]
outputs, RadiK.jl src/algorithm.jl:545-549).
"""
import ctypes

import numpy as np

from . import _lib


class DeviceArray:
    def __init__(self, shape, dtype, _ptr=None):
        if isinstance(shape, int):
            shape = (shape,)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        if _ptr is None:
            p = ctypes.c_void_p()
            _lib.check(_lib.load().radik_malloc(ctypes.byref(p), self.nbytes))
            self.ptr = p.value
            self._owner = True
        else:
            self.ptr = _ptr
            self._owner = False

    def __del__(self):
        try:
            if getattr(self, "_owner", False) and self.ptr:
                _lib.load().radik_free(self.ptr)
                self.ptr = None
        except Exception:
            pass

    def __len__(self):
        return self.size

    @property
    def eltype(self):
        return self.dtype

    @classmethod
    def from_host(cls, arr, dtype=None):
        a = np.asarray(arr, dtype=dtype)
        if a.ndim == 2:
            flat = np.asfortranarray(a).ravel(order="F")
        else:
            flat = np.ascontiguousarray(a).ravel()
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            _lib.check(_lib.load().radik_memcpy_h2d(d.ptr, flat.ctypes.data, d.nbytes))
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        if d.nbytes:
            _lib.check(_lib.load().radik_memset(d.ptr, 0, d.nbytes))
        return d

    def to_host(self):
        flat = np.empty(self.size, dtype=self.dtype)
        if self.nbytes:
            _lib.check(_lib.load().radik_memcpy_d2h(flat.ctypes.data, self.ptr, self.nbytes))
        if len(self.shape) == 2:
            return flat.reshape(self.shape[::-1]).T
        return flat.reshape(self.shape)

    def fill(self, value):
        h = np.full(self.size, value, dtype=self.dtype)
        if self.nbytes:
            _lib.check(_lib.load().radik_memcpy_h2d(self.ptr, h.ctypes.data, self.nbytes))
        return self

    def column(self, j):
        """View of column j of a 2-D array (no copy)."""
        rows = self.shape[0]
        v = DeviceArray((rows,), self.dtype, _ptr=self.ptr + j * rows * self.dtype.itemsize)
        v._parent = self  # keeps the allocation alive
        return v

    def view(self, offset_elems, count):
        v = DeviceArray((count,), self.dtype, _ptr=self.ptr + offset_elems * self.dtype.itemsize)
        v._parent = self
        return v


class Stream:
    def __init__(self):
        p = ctypes.c_void_p()
        _lib.check(_lib.load().radik_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self):
        _lib.check(_lib.load().radik_stream_sync(self.handle))

    def __del__(self):
        try:
            if self.handle:
                _lib.load().radik_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        p = ctypes.c_void_p()
        _lib.check(_lib.load().radik_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream=None):
        _lib.check(_lib.load().radik_event_record(self.handle, stream.handle if stream else None))

    def synchronize(self):
        _lib.check(_lib.load().radik_event_sync(self.handle))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _lib.check(_lib.load().radik_event_elapsed(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _lib.load().radik_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def synchronize():
    _lib.check(_lib.load().radik_device_sync())


def device_count():
    n = ctypes.c_int()
    _lib.check(_lib.load().radik_device_count(ctypes.byref(n)))
    return n.value


def device_info():
    cus = ctypes.c_int()
    lds = ctypes.c_int64()
    hbm = ctypes.c_int64()
    name = ctypes.create_string_buffer(64)
    _lib.check(_lib.load().radik_device_info(ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(hbm), name, 64))
    return {"cus": cus.value, "lds_per_block": lds.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}


def lds_caps():
    a, b, c = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    _lib.check(_lib.load().radik_lds_caps(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
    return {"small_cap": a.value, "refine_cap": b.value, "sort_cap": c.value}
