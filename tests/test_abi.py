"""The C-ABI library loads and exports every symbol the headers declare, and
its host-side validation mirrors the reference's errors (no GPU call)."""
import ctypes
import os
import re

import numpy as np
import pytest

import radik
from radik import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("radik.h", "radik_devmem.h"):
        with open(os.path.join(ROOT, "include", h)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(radik_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 30
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert decl == set(_lib.SIGNATURES), "binding table and headers disagree"


def test_version():
    assert b"gfx950" in _lib.load().radik_version()


def test_topk_size_mismatch_is_error():
    # test/test_algorithm.jl:259-264: wrong-size outputs raise ErrorException
    L = _lib.load()
    lens = _lib.i64_array([1000])
    rc = L.radik_topk(None, None, 4, 50, 1, None, None, 0, lens, 1, 100, 1, 0, None)
    assert rc == _lib.RADIK_ESIZE
    assert b"val_out must have size (100, 1), got (50, 1)" in L.radik_last_error()


def test_topk_indices_length_is_error():
    # test/test_algorithm.jl:266-268: indices of the wrong length
    L = _lib.load()
    lens = _lib.i64_array([1000])
    rc = L.radik_topk(None, None, 4, 100, 1, None, ctypes.c_void_p(16), 500, lens, 1, 100, 1, 0, None)
    assert rc == _lib.RADIK_EINDICES
    assert b"indices length must equal data length. Got 500 vs 1000" in L.radik_last_error()


def test_python_mirror_raises_reference_errors():
    class Fake:
        def __init__(self, shape, dtype):
            self.shape, self.dtype, self.ptr = shape, np.dtype(dtype), 0
            self.size = int(np.prod(shape))
    data = Fake((1000,), np.float32)
    with pytest.raises(radik.RadiKError) as e:
        radik.topk_(Fake((50, 1), np.float32), Fake((50, 1), np.int32), data, [1000], 100)
    assert e.value.code == _lib.RADIK_ESIZE
    with pytest.raises(radik.RadiKError):
        radik.topk_(Fake((100, 1), np.float32), Fake((100, 1), np.int32), Fake((10,), np.float64), [10], 100)


def test_bad_arguments_rejected_before_any_gpu_call():
    L = _lib.load()
    assert L.radik_topk_radix_select(None, None, None, 4, None, None, None, 0, 1, 0, 32, None) == _lib.RADIK_EINVAL
    lens = _lib.i64_array([-1])
    rc = L.radik_topk(None, None, 4, 1, 1, None, None, 0, lens, 1, 1, 1, 0, None)
    assert rc == _lib.RADIK_EINVAL
    h = ctypes.c_void_p()
    assert L.radik_plan_create(ctypes.byref(h), lens, lens, 1, 0, 4, None) == _lib.RADIK_EINVAL
    assert L.radik_server_create(ctypes.byref(h), 0, 1, 0) == _lib.RADIK_EINVAL


def test_noalloc_limits_rejected_before_any_gpu_call():
    # the allocation-free entry point validates like radik_topk and rejects
    # inputs beyond its fixed storage with RADIK_ELIMIT (host-side checks)
    L = _lib.load()
    n = (1 << 22) + 1
    rc = L.radik_topk_noalloc(None, None, 4, 8, 1, None, None, 0, _lib.i64_array([n]), 1, 8, 1, 0, None)
    assert rc == _lib.RADIK_ELIMIT and b"at most 4194304 elements" in L.radik_last_error()
    rc = L.radik_topk_noalloc(None, None, 4, 8, 513, None, None, 0, _lib.i64_array([2049] * 513), 513, 8, 1, 0,
                              None)
    assert rc == _lib.RADIK_ELIMIT and b"longer than 2048" in L.radik_last_error()
    rc = L.radik_topk_noalloc(None, None, 4, 8, 8193, None, None, 0, _lib.i64_array([1] * 8193), 8193, 8, 1, 0,
                              None)
    assert rc == _lib.RADIK_ELIMIT
    rc = L.radik_topk_noalloc(None, None, 4, 50, 1, None, None, 0, _lib.i64_array([1000]), 1, 100, 1, 0, None)
    assert rc == _lib.RADIK_ESIZE
