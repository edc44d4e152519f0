"""ctypes binding of libradik_hip.so (include/radik.h, include/radik_devmem.h).

The library is built in-tree (radik.jl_amd/lib/libradik_hip.so).  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libradik_hip.so")

RADIK_LARGEST = 0x01
RADIK_ASCEND = 0x02
RADIK_WITHSCALE = 0x04
RADIK_WITHIDXIN = 0x08
RADIK_WITHPACKING = 0x10
RADIK_KEY_F32, RADIK_KEY_F16, RADIK_KEY_BF16, RADIK_KEY_I32, RADIK_KEY_F64 = 0x000, 0x100, 0x200, 0x300, 0x400
RADIK_F64_MAX_K = 8192

RADIK_OK = 0
RADIK_EINVAL = 1
RADIK_ESIZE = 2
RADIK_EINDICES = 3
RADIK_EWORKSPACE = 4
RADIK_EHIP = 5
RADIK_ENOMEM = 6
RADIK_EADMIT = 7
RADIK_ENOTREADY = 8
RADIK_ELIMIT = 9


class RadiKError(RuntimeError):
    """Error raised by the library; .code holds the RADIK_E* status.

    Size and index-length mismatches carry the reference's ErrorException
    messages (RadiK.jl src/algorithm.jl:498-509)."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I64P = ctypes.POINTER(ctypes.c_int64)
_U32 = ctypes.c_uint32
_INT = ctypes.c_int

# name -> (restype, argtypes); every exported function of the two headers.
SIGNATURES = {
    "radik_version": (ctypes.c_char_p, []),
    "radik_last_error": (ctypes.c_char_p, []),
    "radik_ws_create": (_INT, [ctypes.POINTER(_P), _I64, _I64, _INT, _P]),
    "radik_ws_destroy": (_INT, [_P]),
    "radik_topk_radix_select": (_INT, [_P, _P, _P, _INT, _P, _P, _I64P, _I64, _I64, _U32, _INT, _P]),
    "radik_topk": (_INT, [_P, _P, _INT, _I64, _I64, _P, _P, _I64, _I64P, _I64, _I64, _INT, _INT, _P]),
    "radik_topk_typed": (_INT, [_INT, _P, _P, _INT, _I64, _I64, _P, _P, _I64, _I64P, _I64, _I64, _INT, _INT, _P]),
    "radik_topk_noalloc": (_INT, [_P, _P, _INT, _I64, _I64, _P, _P, _I64, _I64P, _I64, _I64, _INT, _INT, _P]),
    "radik_plan_create": (_INT, [ctypes.POINTER(_P), _I64P, _I64P, _I64, _U32, _INT, _P]),
    "radik_plan_execute": (_INT, [_P, _P, _P, _P, _P, _P]),
    "radik_plan_set_graph": (_INT, [_P, _INT]),
    "radik_plan_set_options": (_INT, [_P, _INT, _INT, _INT, _INT, _INT]),
    "radik_plan_set_option": (_INT, [_P, ctypes.c_char_p, _I64]),
    "radik_plan_set_timing": (_INT, [_P, _INT]),
    "radik_plan_timing": (_INT, [_P, ctypes.POINTER(ctypes.c_double), _I64P]),
    "radik_plan_info": (_INT, [_P, _I64P, _INT]),
    "radik_plan_sample": (_INT, [_P, _P, _P, ctypes.c_float, ctypes.c_float, _P, _P, _P, _P]),
    "radik_plan_destroy": (_INT, [_P]),
    "radik_server_create": (_INT, [ctypes.POINTER(_P), _I64, _INT, _U32]),
    "radik_server_submit": (_INT, [_P, _P, _I64, _I64, _I64, _P, _P, _I64P]),
    "radik_server_poll": (_INT, [_P, _I64, ctypes.POINTER(ctypes.c_double)]),
    "radik_server_wait": (_INT, [_P, _I64, ctypes.POINTER(ctypes.c_double)]),
    "radik_server_drain": (_INT, [_P]),
    "radik_server_stats": (_INT, [_P, _I64P, _I64P, _I64P]),
    "radik_server_destroy": (_INT, [_P]),
    "radik_diag_pass1": (_INT, [_P, _I64P, _I64, _I64, _U32, _P, _P, _P]),
    "radik_diag_scan": (_INT, [_P, _P, _P, _I64, _P]),
    "radik_diag_copy": (_INT, [_P, _P, _I64, _P]),
    "radik_diag_plan_chunks": (_INT, [_I64, _I64, _I64, _INT, _I64, _I64, _I64P, _I64, _I64P]),
    "radik_diag_route": (_INT, [_I64P, _I64P, _I64, _INT, _INT, _INT, ctypes.POINTER(ctypes.c_int32), _I64P]),
    "radik_diag_batch": (_INT, [_I64P, _I64, _I64, _I64, _I64, _I64P]),
    "radik_diag_footprint": (_INT, [_I64, _I64, _I64P]),
    "radik_device_count": (_INT, [ctypes.POINTER(_INT)]),
    "radik_set_device": (_INT, [_INT]),
    "radik_device_info": (_INT, [ctypes.POINTER(_INT), _I64P, _I64P, ctypes.c_char_p, _INT]),
    "radik_lds_caps": (_INT, [ctypes.POINTER(_U32), ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "radik_malloc": (_INT, [ctypes.POINTER(_P), ctypes.c_size_t]),
    "radik_free": (_INT, [_P]),
    "radik_memcpy_h2d": (_INT, [_P, _P, ctypes.c_size_t]),
    "radik_memcpy_d2h": (_INT, [_P, _P, ctypes.c_size_t]),
    "radik_memcpy_d2d": (_INT, [_P, _P, ctypes.c_size_t]),
    "radik_memset": (_INT, [_P, _INT, ctypes.c_size_t]),
    "radik_device_sync": (_INT, []),
    "radik_stream_create": (_INT, [ctypes.POINTER(_P)]),
    "radik_stream_destroy": (_INT, [_P]),
    "radik_stream_sync": (_INT, [_P]),
    "radik_event_create": (_INT, [ctypes.POINTER(_P)]),
    "radik_event_destroy": (_INT, [_P]),
    "radik_event_record": (_INT, [_P, _P]),
    "radik_event_sync": (_INT, [_P]),
    "radik_event_elapsed": (_INT, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
}

_lib = None


def load():
    """Loads the library (no GPU call is made)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RadiKError(RADIK_EHIP, f"{LIB_PATH} is not built; run radik.jl_amd/build.py (or make -C radik.jl_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != RADIK_OK:
        msg = load().radik_last_error().decode(errors="replace")
        raise RadiKError(rc, msg)
    return rc


def i64_array(values):
    arr = (ctypes.c_int64 * len(values))(*[int(v) for v in values])
    return arr
