"""ctypes wrapper of the CPU oracle (liboracle.so, built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker -- never by the product
package.  Each function mirrors the C function it wraps; see oracle.h for
the reference file:line each restates and what pins it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")

_lib = None
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_L = C.POINTER(C.c_int64)
_U8 = C.POINTER(C.c_uint8)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "oracle"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        sig = {
            "or_znormalize_f32": (None, [_F, C.c_int32]),
            "or_znormalize_f64": (None, [_D, C.c_int32]),
            "or_envelope": (None, [_F, C.c_int32, _F, _F, C.c_int32]),
            "or_stats_refcompat": (None, [_F, C.c_int64, C.c_int32, C.c_int64, _F, _F]),
            "or_stats_fp64": (None, [_F, C.c_int64, C.c_int32, _F, _F, _D, _D]),
            "or_lb_kim": (None, [_F, _F, _F, _F, C.c_int32, _I, C.c_int64, C.c_int32, _F]),
            "or_lb_keogh": (None, [_F, _F, _F, _F, _F, C.c_int32, _I, C.c_int64, C.c_int32, _F]),
            "or_cdtw": (None, [_F, _F, _F, _F, C.c_int32, _I, C.c_int64, C.c_int32, C.c_int32,
                               C.c_int32, _F]),
            "or_cdtw_f64": (None, [_D, _F, _D, _D, C.c_int32, _I, C.c_int64, C.c_int32, C.c_int32,
                                   _D]),
            "or_cdtw_grad_f64": (None, [_D, _F, _D, _D, C.c_int32, _I, C.c_int64, C.c_int32,
                                        C.c_int32, C.c_int32, _D, _D, _D, _D]),
            "or_cdtw_grad": (None, [_D, _F, _D, _D, C.c_int32, _I, C.c_int64, C.c_int32,
                                    C.c_int32, C.c_int32, C.c_int32, _D, _D, _D, _D]),
            "or_ldtw": (None, [_F, _F, _F, _F, C.c_int32, _I, C.c_int64, C.c_int32, C.c_int32, _F,
                               _U8]),
            "or_ldtw_raw": (None, [_F, _F, C.c_int32, _I, C.c_int64, C.c_int32, C.c_int32, _F]),
            "or_ed_f64": (None, [_D, _F, C.c_int64, C.c_int32, _D, _D, _D]),
            "or_stats_f64d": (None, [_D, C.c_int64, C.c_int32, _D, _D]),
            "or_set_threads": (None, [C.c_int32]),
            "or_max_threads": (C.c_int32, []),
            "or_ed_f64d": (None, [_D, _D, C.c_int64, C.c_int32, _D, _D, _D]),
            "or_argmin": (C.c_int64, [_F, C.c_int64]),
            "or_topk_greedy": (C.c_int32, [_F, C.c_int64, C.c_int32, C.c_int64, _L, _F]),
            "or_ucr_dtw": (C.c_int64, [_D, C.c_int64, _D, C.c_int32, C.c_int32, _D, _L]),
            "or_ucr_dtw_par": (C.c_int64, [_D, C.c_int64, _D, C.c_int32, C.c_int32, C.c_int32,
                                           _D, _L]),
            "or_ucr_ed": (C.c_int64, [_D, C.c_int64, _D, C.c_int32, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _f(a):
    return a.ctypes.data_as(_F)


def _d(a):
    return a.ctypes.data_as(_D)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _idx(idx):
    if idx is None:
        return None, None
    a = np.ascontiguousarray(idx, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


def znormalize_f32(x):
    a = _c32(x).copy()
    lib().or_znormalize_f32(_f(a), a.size)
    return a


def znormalize_f64(x):
    a = _c64(x).copy()
    lib().or_znormalize_f64(_d(a), a.size)
    return a


def envelope(q, w):
    q = _c32(q)
    lo = np.empty_like(q)
    up = np.empty_like(q)
    lib().or_envelope(_f(q), int(w), _f(lo), _f(up), q.size)
    return lo, up


def stats_refcompat(s, m, epoch=0):
    s = _c32(s)
    L = s.size - m + 1
    avg = np.empty(L, np.float32)
    sd = np.empty(L, np.float32)
    lib().or_stats_refcompat(_f(s), s.size, int(m), int(epoch), _f(avg), _f(sd))
    return avg, sd


def stats_fp64(s, m, want64=False):
    s = _c32(s)
    L = s.size - m + 1
    avg = np.empty(L, np.float32)
    sd = np.empty(L, np.float32)
    a64 = np.empty(L, np.float64) if want64 else None
    s64 = np.empty(L, np.float64) if want64 else None
    lib().or_stats_fp64(_f(s), s.size, int(m), _f(avg), _f(sd),
                        _d(a64) if want64 else None, _d(s64) if want64 else None)
    return (avg, sd, a64, s64) if want64 else (avg, sd)


def lb_kim(s, avg, sd, zq, idx=None, count=None, metric=1):
    s, avg, sd, zq = _c32(s), _c32(avg), _c32(sd), _c32(zq)
    ia, ip = _idx(idx)
    n = ia.size if ia is not None else (count if count is not None else s.size - zq.size + 1)
    out = np.empty(n, np.float32)
    lib().or_lb_kim(_f(s), _f(avg), _f(sd), _f(zq), zq.size, ip, n, int(metric), _f(out))
    return out


def lb_keogh(s, avg, sd, lo, up, idx=None, count=None, metric=1):
    s, avg, sd, lo, up = _c32(s), _c32(avg), _c32(sd), _c32(lo), _c32(up)
    ia, ip = _idx(idx)
    n = ia.size if ia is not None else (count if count is not None else s.size - lo.size + 1)
    out = np.empty(n, np.float32)
    lib().or_lb_keogh(_f(s), _f(avg), _f(sd), _f(lo), _f(up), lo.size, ip, n, int(metric),
                      _f(out))
    return out


def cdtw(zq, s, avg, sd, w, idx=None, count=None, metric=1, fused=1):
    s, avg, sd, zq = _c32(s), _c32(avg), _c32(sd), _c32(zq)
    ia, ip = _idx(idx)
    n = ia.size if ia is not None else (count if count is not None else s.size - zq.size + 1)
    out = np.empty(n, np.float32)
    lib().or_cdtw(_f(zq), _f(s), _f(avg), _f(sd), zq.size, ip, n, int(w), int(metric),
                  int(fused), _f(out))
    return out


def cdtw_f64(zq, s, avg64, sd64, w, idx=None, count=None, metric=1):
    s, avg64, sd64, zq = _c32(s), _c64(avg64), _c64(sd64), _c64(zq)
    ia, ip = _idx(idx)
    n = ia.size if ia is not None else (count if count is not None else s.size - zq.size + 1)
    out = np.empty(n, np.float64)
    lib().or_cdtw_f64(_d(zq), _f(s), _d(avg64), _d(sd64), zq.size, ip, n, int(w), int(metric),
                      _d(out))
    return out


def cdtw_grad_f64(zq64, s, avg64, sd64, w, idx, g, metric=1, sqrt_dist=1, q_raw=None,
                  path32=False):
    """FP64 query gradient of sum_p g[p] * dist_p (or_cdtw_grad; path32: along
    the optimal path of the FP32 DP).  Returns (grad (M,), dist (count,))."""
    zq64, s, avg64, sd64 = _c64(zq64), _c32(s), _c64(avg64), _c64(sd64)
    idx, ip = _idx(idx)
    g = _c64(g)
    m = zq64.size
    grad = np.empty(m, np.float64)
    dist = np.empty(idx.size, np.float64)
    qr = None if q_raw is None else _c64(q_raw)
    lib().or_cdtw_grad(_d(zq64), _f(s), _d(avg64), _d(sd64), m, ip, idx.size, int(w),
                       int(metric), int(sqrt_dist), int(bool(path32)), _d(g),
                       None if qr is None else _d(qr), _d(dist), _d(grad))
    return grad, dist


def ldtw(zq, s, avg, sd, w, idx=None, count=None, metric=1, want_moves=False):
    s, avg, sd, zq = _c32(s), _c32(avg), _c32(sd), _c32(zq)
    ia, ip = _idx(idx)
    n = ia.size if ia is not None else (count if count is not None else s.size - zq.size + 1)
    out = np.empty(n, np.float32)
    moves = np.empty((n, max(2 * zq.size - 2, 1)), np.uint8) if want_moves else None
    lib().or_ldtw(_f(zq), _f(s), _f(avg), _f(sd), zq.size, ip, n, int(w), int(metric), _f(out),
                  moves.ctypes.data_as(_U8) if want_moves else None)
    return (out, moves) if want_moves else out


def ldtw_raw(q, s, w, idx=None, count=None, metric=1):
    """ldtw.cuh:18-71 serial_ldtw over windows of s (no z-normalisation)."""
    s, q = _c32(s), _c32(q)
    ia, ip = _idx(idx)
    n = ia.size if ia is not None else (count if count is not None else s.size - q.size + 1)
    out = np.empty(n, np.float32)
    lib().or_ldtw_raw(_f(q), _f(s), q.size, ip, n, int(w), int(metric), _f(out))
    return out


def ed_f64(zq64, s, avg64, sd64):
    zq64, s, avg64, sd64 = _c64(zq64), _c32(s), _c64(avg64), _c64(sd64)
    L = s.size - zq64.size + 1
    out = np.empty(L, np.float64)
    lib().or_ed_f64(_d(zq64), _f(s), s.size, zq64.size, _d(avg64), _d(sd64), _d(out))
    return out


def stats_f64d(s, m):
    """FP64 window statistics of an FP64 stream (two-pass)."""
    s = _c64(s)
    L = s.size - m + 1
    avg = np.empty(L, np.float64)
    sd = np.empty(L, np.float64)
    lib().or_stats_f64d(_d(s), s.size, int(m), _d(avg), _d(sd))
    return avg, sd


def ed_f64d(zq64, s, avg64, sd64):
    """Squared z-normalised ED of every window of an FP64 stream."""
    zq64, s, avg64, sd64 = _c64(zq64), _c64(s), _c64(avg64), _c64(sd64)
    L = s.size - zq64.size + 1
    out = np.empty(L, np.float64)
    lib().or_ed_f64d(_d(zq64), _d(s), s.size, zq64.size, _d(avg64), _d(sd64), _d(out))
    return out


def set_threads(n):
    """OpenMP threads of the oracle's loops (1 = the reference's one-core form)."""
    lib().or_set_threads(int(n))


def max_threads():
    return int(lib().or_max_threads())


def ucr_reference(binary, data, query, r=None):
    """Runs UCR_DTW_BIN / UCR_ED_BIN as compiled from the reference sources
    into oracle/_ref (never a shipped binary): FP64 files, argv
    ``data query m [R] n``.  Returns (location, distance, seconds as the
    program reports them: CPU time of its search loop), or None when the
    program is not built here."""
    import tempfile

    exe = os.path.join(REF_DIR, binary)
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        dp, qp = os.path.join(td, "d.bin"), os.path.join(td, "q.bin")
        _c64(data).tofile(dp)
        _c64(query).tofile(qp)
        args = [exe, dp, qp, str(query.size)] + ([str(int(r))] if r is not None else []) + \
               [str(data.size)]
        out = subprocess.run(args, capture_output=True, text=True, check=True).stdout
    loc = dist = secs = None
    for line in out.splitlines():
        key, _, val = line.partition(":")
        key = key.strip()
        if key == "Location":
            loc = int(val)
        elif key == "Distance":
            dist = float(val)
        elif key == "Total Execution Time":
            secs = float(val.split()[0])
    return loc, dist, secs


def argmin(x):
    x = _c32(x)
    return int(lib().or_argmin(_f(x), x.size))


def topk_greedy(d, k, excl):
    d = _c32(d)
    idx = np.full(k, -1, np.int64)
    dist = np.full(k, np.inf, np.float32)
    got = lib().or_topk_greedy(_f(d), d.size, int(k), int(excl), idx.ctypes.data_as(_L), _f(dist))
    return idx[:got], dist[:got]


def ucr_dtw(data, query, r):
    data, query = _c64(data), _c64(query)
    dist = C.c_double()
    counts = np.zeros(4, np.int64)
    loc = lib().or_ucr_dtw(_d(data), data.size, _d(query), query.size, int(r), C.byref(dist),
                           counts.ctypes.data_as(_L))
    return int(loc), float(dist.value), counts


def ucr_dtw_par(data, query, r, threads):
    data, query = _c64(data), _c64(query)
    dist = C.c_double()
    counts = np.zeros(4, np.int64)
    loc = lib().or_ucr_dtw_par(_d(data), data.size, _d(query), query.size, int(r), int(threads),
                               C.byref(dist), counts.ctypes.data_as(_L))
    return int(loc), float(dist.value), counts


def ucr_ed(data, query):
    data, query = _c64(data), _c64(query)
    dist = C.c_double()
    loc = lib().or_ucr_ed(_d(data), data.size, _d(query), query.size, C.byref(dist))
    return int(loc), float(dist.value)
