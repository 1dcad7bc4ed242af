"""ctypes binding of libmdtw.so (the C ABI declared in include/mdtw.h).

The library is built in-tree (cudadtw_amd/libmdtw.so) by
``make -C cudadtw_amd/csrc``.  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdtw.so")

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_QUERY_TOO_SHORT = 2
ERR_QUERY_TOO_LONG = 3
ERR_BAND = 4
ERR_STREAM_TOO_SHORT = 5
ERR_NO_QUERY = 6
ERR_ALLOC = 7
ERR_HIP = 8
ERR_NOT_FINITE = 9
ERR_K = 10
ERR_UNSUPPORTED = 11
ERR_WORKSPACE = 12

MAX_QUERY = 8192
MAX_K = 1024
METRIC_L1 = 0
METRIC_L2 = 1
STATS_FP64 = 0
STATS_REFCOMPAT = 1
STATS_NONE = 2
ALGO_CDTW = 0
ALGO_ED = 1
ALGO_LDTW = 2
ALGO_ED_FFT = 3


class SearchInfo(C.Structure):
    _fields_ = [
        ("windows", C.c_int64),
        ("seeded", C.c_int64),
        ("lb_kim_pruned", C.c_int64),
        ("lb_keogh_pruned", C.c_int64),
        ("dp_evaluated", C.c_int64),
        ("dp_abandoned", C.c_int64),
        ("found", C.c_int32),
        ("threshold", C.c_float),
        ("elapsed_ms", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = C.c_void_p
_I32 = C.c_int32
_I64 = C.c_int64
_F = C.c_float
_FP = C.POINTER(C.c_float)

# name -> (restype, argtypes); every symbol include/mdtw.h declares.
SIGNATURES = {
    "mdtw_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mdtw_destroy": (C.c_int, [_P]),
    "mdtw_set_stream": (C.c_int, [_P, _P]),
    "mdtw_set_workspace_limit": (C.c_int, [_P, C.c_size_t]),
    "mdtw_avg_std_f64": (C.c_int, [_P, _P, _I64, _I32, _P, _P]),
    "mdtw_ed_f64": (C.c_int, [_P, _P, _I64, C.POINTER(C.c_double), _I32, _P]),
    "mdtw_ed_search_f64": (C.c_int, [_P, _P, _I64, C.POINTER(C.c_double), _I32,
                                     C.POINTER(_I64), C.POINTER(C.c_double),
                                     C.POINTER(SearchInfo)]),
    "mdtw_synchronize": (C.c_int, [_P]),
    "mdtw_status_string": (C.c_char_p, [C.c_int]),
    "mdtw_device_info": (C.c_int, [_P, C.c_char_p, _I32, C.POINTER(_I32)]),
    "mdtw_set_query": (C.c_int, [_P, _FP, _FP, _FP, _I32, _I32]),
    "mdtw_znormalize": (C.c_int, [_FP, _I32]),
    "mdtw_znormalize_f64": (C.c_int, [C.POINTER(C.c_double), _I32]),
    "mdtw_envelope": (C.c_int, [_FP, _I32, _FP, _FP, _I32]),
    "mdtw_avg_std": (C.c_int, [_P, _P, _I64, _I32, _P, _P, _I32]),
    "mdtw_lb_kim": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, _I32]),
    "mdtw_lb_keogh": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, _I32]),
    "mdtw_cdtw": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, _I32, _I32]),
    "mdtw_cdtw_wide": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, _I32, _I32]),
    "mdtw_ldtw": (C.c_int, [_P, _P, _P, _P, _P, _I64, _I32, _I32]),
    "mdtw_ed": (C.c_int, [_P, _P, _P, _P, _P, _I64]),
    "mdtw_ed_fft": (C.c_int, [_P, _P, _I64, _I32, _P]),
    "mdtw_cdtw_backward": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, _I32, _I32, _I32, _P, _P,
                                     _P]),
    "mdtw_search": (C.c_int, [_P, _P, _I64, _P, _P, _I32, _I32, _I32, _I32, _I32, _I64,
                              C.POINTER(_I64), _FP, C.POINTER(SearchInfo)]),
    "mdtw_search_noalloc": (C.c_int, [_P, _P, _I64, _P, _P, _I32, _I32, _I32, _I32, _I32, _I64,
                                      C.POINTER(_I64), _FP, C.POINTER(SearchInfo)]),
    "mdtw_batch_create": (C.c_int, [_P, _FP, C.POINTER(_I32), C.POINTER(_I32), _I32, _I32,
                                    C.POINTER(_P)]),
    "mdtw_batch_destroy": (C.c_int, [_P]),
    "mdtw_batch_tune": (C.c_int, [_P, _P, _I64, _I32, _I32, _I64, C.POINTER(_I32)]),
    "mdtw_batch_shape": (C.c_int, [_P, _I32, C.POINTER(_I32)]),
    "mdtw_batch_set_shape": (C.c_int, [_P, _I32, C.POINTER(_I32)]),
    "mdtw_batch_search": (C.c_int, [_P, _P, _I64, _I32, _I32, C.POINTER(_I64), C.POINTER(_I64),
                                    _FP, C.POINTER(SearchInfo), _FP]),
    "mdtw_batch_backward": (C.c_int, [_P, _P, _I64, _I32, _I32, C.POINTER(_I64), _FP, _I32, _FP,
                                      _FP, _FP]),
    "mdtw_stream_create": (C.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I64, _I64,
                                     C.POINTER(_P)]),
    "mdtw_stream_destroy": (C.c_int, [_P]),
    "mdtw_stream_push": (C.c_int, [_P, _FP, _I64]),
    "mdtw_stream_finish": (C.c_int, [_P]),
    "mdtw_stream_result": (C.c_int, [_P, C.POINTER(_I64), _FP, C.POINTER(SearchInfo)]),
    "mdtw_stream_position": (C.c_int, [_P, C.POINTER(_I64), C.POINTER(_I64)]),
    "mdtw_stream_save": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "mdtw_stream_restore": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(_P)]),
    "mdtw_threshold": (C.c_int, [_P, _P, _P, _I64, _F, C.POINTER(_I64)]),
    "mdtw_argmin": (C.c_int, [_P, _P, _I64, C.POINTER(_I64), _FP]),
}

_lib = None


def load() -> C.CDLL:
    """Load libmdtw.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C cudadtw_amd/csrc` "
            "(or __graft_entry__.build()); there is no fallback implementation")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MdtwError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = load().mdtw_status_string(status).decode()
        super().__init__(f"{where}: {msg} (status {status})")
        self.status = status


def check(status: int, where: str) -> None:
    if status != OK:
        raise MdtwError(status, where)
