"""Python mirror of the reference's host interface over the mdtw C ABI.

The reference (gravitino/cudadtw) is a set of header templates driven like
this (GPU-dtw.cu:67-96)::

    znormalize(zquery, M); envelope(zquery, W, zlower, zupper, M);
    cudaMemcpyToSymbol(Czquery/Czlower/Czupper, ...);
    avg_std<double>(Subject, AvgS, StdS, M, N);
    prune_cdtw<EUCLIDEAN>(Subject, AvgS, StdS, M, N, W);   // prints result

Here the same steps are ``znormalize``, ``Engine.set_query`` (query and
envelope into the engine), ``Engine.avg_std`` and ``Engine.prune_cdtw`` /
``Engine.search``, which *return* the match instead of printing it.  Device
arrays are torch tensors on the engine's GPU; every call runs on torch's
current stream of that device, so it orders with surrounding torch work.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (ALGO_CDTW, ALGO_ED, ALGO_ED_FFT, ALGO_LDTW, METRIC_L1, METRIC_L2, STATS_FP64,
                   STATS_NONE, STATS_REFCOMPAT, check)

__all__ = ["Engine", "SearchResult", "Batch", "BatchResult", "znormalize", "envelope",
           "band_width",
           "ALGO_CDTW", "ALGO_ED", "ALGO_ED_FFT", "ALGO_LDTW", "METRIC_L1", "METRIC_L2",
           "STATS_FP64", "STATS_NONE", "STATS_REFCOMPAT"]


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def znormalize(series, double: bool = False, keep_f64: bool = False) -> np.ndarray:
    """stats.cuh:8-28 znormalize: returns the z-normalised copy,
    bit-identical to the reference's host routine.  double=False is
    value_t = float (the cDTW / LDTW drivers, GPU-dtw.cu:67); double=True is
    value_t = double (the ED driver, GPU-ed.cu:62), rounded to FP32 after
    unless keep_f64 (the FP64 ED entry points take it as is)."""
    if double:
        x = np.ascontiguousarray(series, dtype=np.float64).copy()
        check(_lib.load().mdtw_znormalize_f64(x.ctypes.data_as(C.POINTER(C.c_double)), x.size),
              "znormalize")
        return x if keep_f64 else x.astype(np.float32)
    x = np.ascontiguousarray(series, dtype=np.float32).copy()
    check(_lib.load().mdtw_znormalize(_fp(x), x.size), "znormalize")
    return x


def envelope(series, w: int):
    """bounds.cuh:13-64 envelope: (lower, upper) of a query for band w."""
    q = np.ascontiguousarray(series, dtype=np.float32)
    lo = np.empty_like(q)
    up = np.empty_like(q)
    check(_lib.load().mdtw_envelope(_fp(q), int(w), _fp(lo), _fp(up), q.size), "envelope")
    return lo, up


def band_width(m: int, percent: float, mode: str = "round") -> int:
    """Absolute Sakoe-Chiba band for a query of length m.  ``mode="truncate"``
    is the reference CLI's ``W = M*(P*0.01)`` (GPU-dtw.cu:27)."""
    if mode == "truncate":
        return int(m * (int(percent) * 0.01))
    return int(np.floor(m * percent / 100.0 + 0.5))


@dataclass
class SearchResult:
    positions: np.ndarray   # int64[k], -1 where fewer than k matches exist
    distances: np.ndarray   # float32[k]; sqrt of squared cDTW/ED for L2
    info: dict

    @property
    def location(self) -> int:
        return int(self.positions[0])

    @property
    def distance(self) -> float:
        return float(self.distances[0])


class Engine:
    """One device's search engine: owns the query state (the reference's
    __constant__ Czquery/Czlower/Czupper) and reusable scratch."""

    def __init__(self, device: int = 0):
        import torch

        self._torch = torch
        self.lib = _lib.load()
        self.device = int(device)
        self._dev = torch.device("cuda", self.device)
        h = C.c_void_p()
        check(self.lib.mdtw_create(self.device, C.byref(h)), "mdtw_create")
        self.h = h
        self.m = 0
        self.w = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.mdtw_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- plumbing ----------------------------------------------------------
    def _stream(self):
        s = self._torch.cuda.current_stream(self._dev).cuda_stream
        check(self.lib.mdtw_set_stream(self.h, C.c_void_p(s)), "mdtw_set_stream")

    def _f32(self, t, name):
        torch = self._torch
        if not (isinstance(t, torch.Tensor) and t.device == self._dev and t.dtype == torch.float32
                and t.is_contiguous()):
            raise TypeError(f"{name}: expected a contiguous float32 tensor on {self._dev}")
        return C.c_void_p(t.data_ptr())

    def _i32(self, t, name):
        torch = self._torch
        if t is None:
            return None
        if not (isinstance(t, torch.Tensor) and t.device == self._dev and t.dtype == torch.int32
                and t.is_contiguous()):
            raise TypeError(f"{name}: expected a contiguous int32 tensor on {self._dev}")
        return C.c_void_p(t.data_ptr())

    def empty(self, n, dtype=None):
        return self._torch.empty(int(n), dtype=dtype or self._torch.float32, device=self._dev)

    def info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int32()
        check(self.lib.mdtw_device_info(self.h, name, 256, C.byref(cus)), "mdtw_device_info")
        return name.value.decode(), cus.value

    # -- query -------------------------------------------------------------
    def set_query(self, zquery, w: int, lower=None, upper=None):
        """cudaMemcpyToSymbol(Czquery/Czlower/Czupper) (GPU-dtw.cu:82-84).
        zquery must already be z-normalised (``znormalize``)."""
        q = np.ascontiguousarray(zquery, dtype=np.float32)
        lo = up = None
        if lower is not None:
            lo = np.ascontiguousarray(lower, dtype=np.float32)
            up = np.ascontiguousarray(upper, dtype=np.float32)
        self._stream()
        check(self.lib.mdtw_set_query(self.h, _fp(q), _fp(lo) if lo is not None else None,
                                      _fp(up) if up is not None else None, q.size, int(w)),
              "mdtw_set_query")
        self.m = q.size
        self.w = int(w)

    # -- statistics and bounds ---------------------------------------------
    def avg_std(self, series, m: int, mode: int = STATS_FP64, out=None):
        """stats.cuh:115-142 avg_std: per-window mean and sigma (FP32)."""
        n = series.numel()
        L = n - m + 1
        avg, sd = out if out is not None else (self.empty(max(L, 1)), self.empty(max(L, 1)))
        self._stream()
        check(self.lib.mdtw_avg_std(self.h, self._f32(series, "series"), n, int(m),
                                    self._f32(avg, "avg"), self._f32(sd, "std"), int(mode)),
              "mdtw_avg_std")
        return avg, sd

    def _count(self, indices, count, series):
        if indices is not None:
            return indices.numel()
        if count is not None:
            return int(count)
        return series.numel() - self.m + 1

    def lb_kim(self, series, avg, sd, indices=None, count=None, metric=METRIC_L2):
        """bounds.cuh:413-441 lb_kim (all windows, or the gathered indices)."""
        n = self._count(indices, count, series)
        out = self.empty(max(n, 1))
        self._stream()
        check(self.lib.mdtw_lb_kim(self.h, self._f32(series, "series"), self._f32(avg, "avg"),
                                   self._f32(sd, "std"), self._f32(out, "out"),
                                   self._i32(indices, "indices"), n, int(metric)),
              "mdtw_lb_kim")
        return out[:n]

    def lb_keogh(self, series, avg, sd, indices=None, count=None, metric=METRIC_L2):
        """bounds.cuh:443-465 lb_keogh against the query envelope."""
        n = self._count(indices, count, series)
        out = self.empty(max(n, 1))
        self._stream()
        check(self.lib.mdtw_lb_keogh(self.h, self._f32(series, "series"), self._f32(avg, "avg"),
                                     self._f32(sd, "std"), self._f32(out, "out"),
                                     self._i32(indices, "indices"), n, int(metric)),
              "mdtw_lb_keogh")
        return out[:n]

    # -- distance profiles -------------------------------------------------
    def gpu_cdtw(self, series, avg, sd, w: int, indices=None, count=None, metric=METRIC_L2,
                 wide: bool = False, out=None):
        """cdtw.cuh:450-487 gpu_cdtw: squared banded cDTW per window."""
        n = self._count(indices, count, series)
        if out is None:
            out = self.empty(max(n, 1))
        fn = self.lib.mdtw_cdtw_wide if wide else self.lib.mdtw_cdtw
        self._stream()
        check(fn(self.h, self._f32(series, "series"), self._f32(avg, "avg"), self._f32(sd, "std"),
                 self._f32(out, "out"), self._i32(indices, "indices"), n, int(w), int(metric)),
              "mdtw_cdtw")
        return out[:n]

    def ldtw(self, series, avg, sd, w: int, count=None, metric=METRIC_L2):
        """ldtw.cuh:294-391 ztsh2_ldtw: greedy LDTW of every window; with
        avg = sd = None the raw windows (ldtw.cuh:198-291 tsh2_ldtw)."""
        n = self._count(None, count, series)
        out = self.empty(max(n, 1))
        raw = avg is None and sd is None
        self._stream()
        check(self.lib.mdtw_ldtw(self.h, self._f32(series, "series"),
                                 None if raw else self._f32(avg, "avg"),
                                 None if raw else self._f32(sd, "std"), self._f32(out, "out"), n,
                                 int(w), int(metric)), "mdtw_ldtw")
        return out[:n]

    def ed(self, series, avg, sd, count=None):
        """Squared z-normalised Euclidean distance per window (ed.cuh:61-69)."""
        n = self._count(None, count, series)
        out = self.empty(max(n, 1))
        self._stream()
        check(self.lib.mdtw_ed(self.h, self._f32(series, "series"), self._f32(avg, "avg"),
                               self._f32(sd, "std"), self._f32(out, "out"), n), "mdtw_ed")
        return out[:n]

    def ed_fft(self, series, stats_mode: int = STATS_FP64):
        """ed.cuh:71-153 (FFT cross-correlation, FP64): squared z-normalised
        ED of every window, for long queries (M <= 8192)."""
        n = series.numel() - self.m + 1
        out = self.empty(max(n, 1))
        self._stream()
        check(self.lib.mdtw_ed_fft(self.h, self._f32(series, "series"), series.numel(),
                                   int(stats_mode), self._f32(out, "out")), "mdtw_ed_fft")
        return out[:n]

    # -- gradient ----------------------------------------------------------
    def cdtw_backward(self, series, avg, sd, indices, grad, w: int, metric: int = METRIC_L2,
                      sqrt_dist: bool = True, query_raw=None, want_dist: bool = False):
        """Query gradient of sum_p grad[p] * dist_p over the windows
        ``indices`` (mdtw_cdtw_backward; no reference counterpart).  dist_p is
        the banded cDTW of the engine's query (sqrt when ``sqrt_dist``).  With
        ``query_raw`` (the raw query whose z-normalisation is the engine's
        query) the gradient is with respect to the raw query, otherwise with
        respect to the z-normalised one.  Returns (grad[M], dist or None)."""
        n = indices.numel()
        out = self.empty(self.m)
        dist = self.empty(max(n, 1)) if want_dist else None
        self._stream()
        check(self.lib.mdtw_cdtw_backward(
            self.h, self._f32(series, "series"), self._f32(avg, "avg"), self._f32(sd, "std"),
            self._i32(indices, "indices"), self._f32(grad, "grad"), n, int(w), int(metric),
            1 if sqrt_dist else 0,
            self._f32(query_raw, "query_raw") if query_raw is not None else None,
            self._f32(dist, "dist") if dist is not None else None, self._f32(out, "grad_query")),
            "mdtw_cdtw_backward")
        return out, (dist[:n] if dist is not None else None)

    # -- search ------------------------------------------------------------
    def search(self, series, w: int, algo: int = ALGO_CDTW, metric: int = METRIC_L2, k: int = 1,
               excl: int = 0, avg=None, sd=None, stats_mode: int = STATS_FP64) -> SearchResult:
        """Best (k = 1) or top-k matches (cdtw.cuh:510-592, ed.cuh:71-153)."""
        pos = np.full(k, -1, dtype=np.int64)
        dist = np.full(k, np.inf, dtype=np.float32)
        info = _lib.SearchInfo()
        self._stream()
        check(self.lib.mdtw_search(
            self.h, self._f32(series, "series"), series.numel(),
            self._f32(avg, "avg") if avg is not None else None,
            self._f32(sd, "std") if sd is not None else None, int(w), int(algo), int(metric),
            int(stats_mode), int(k), int(excl), pos.ctypes.data_as(C.POINTER(C.c_int64)),
            _fp(dist), C.byref(info)), "mdtw_search")
        return SearchResult(pos, dist, info.as_dict())

    def prune_cdtw(self, subject, avg, sd, w: int, metric: int = METRIC_L2):
        """cdtw.cuh:510-592 prune_cdtw: returns (location, distance) instead
        of printing "Location:" / "Distance:" (cdtw.cuh:506-507)."""
        r = self.search(subject, w, ALGO_CDTW, metric, 1, 0, avg, sd)
        return r.location, r.distance

    def calculate_ed(self, subject, avg=None, sd=None, zquery64=None):
        """ed.cuh:71-153 calculate_ed: (location, distance) of the best
        z-normalised Euclidean match.  A float64 ``subject`` takes the FP64
        path of the reference's ED driver (GPU-ed.cu): ``zquery64`` is then
        the FP64 z-normalised query (``znormalize(q, double=True,
        keep_f64=True)``) and the window statistics are computed in FP64."""
        if subject.dtype == self._torch.float64:
            return self.ed_search_f64(subject, zquery64)
        r = self.search(subject, 0, ALGO_ED, METRIC_L2, 1, 0, avg, sd)
        return r.location, r.distance

    # -- FP64 streams (the reference's ED driver) ----------------------------
    def _f64(self, t, name):
        torch = self._torch
        if not (isinstance(t, torch.Tensor) and t.device == self._dev and t.dtype == torch.float64
                and t.is_contiguous()):
            raise TypeError(f"{name}: expected a contiguous float64 tensor on {self._dev}")
        return C.c_void_p(t.data_ptr())

    def avg_std_f64(self, series, m: int):
        """stats.cuh:115-142 avg_std<double> on an FP64 stream: FP64 window
        means and population sigmas."""
        L = series.numel() - m + 1
        avg = self._torch.empty(max(L, 1), dtype=self._torch.float64, device=self._dev)
        sd = self._torch.empty(max(L, 1), dtype=self._torch.float64, device=self._dev)
        self._stream()
        check(self.lib.mdtw_avg_std_f64(self.h, self._f64(series, "series"), series.numel(),
                                        int(m), self._f64(avg, "avg"), self._f64(sd, "std")),
              "mdtw_avg_std_f64")
        return avg[:L], sd[:L]

    def ed_f64(self, series, zquery64):
        """FP64 squared z-normalised ED of every window of an FP64 stream."""
        q = np.ascontiguousarray(zquery64, dtype=np.float64)
        L = series.numel() - q.size + 1
        out = self._torch.empty(max(L, 1), dtype=self._torch.float64, device=self._dev)
        self._stream()
        check(self.lib.mdtw_ed_f64(self.h, self._f64(series, "series"), series.numel(),
                                   q.ctypes.data_as(C.POINTER(C.c_double)), q.size,
                                   self._f64(out, "out")), "mdtw_ed_f64")
        return out[:L]

    def ed_search_f64(self, series, zquery64):
        """ed.cuh:71-153 calculate_ed on an FP64 stream: (location,
        distance), FP64 throughout."""
        q = np.ascontiguousarray(zquery64, dtype=np.float64)
        pos = C.c_int64(-1)
        dist = C.c_double()
        info = _lib.SearchInfo()
        self._stream()
        check(self.lib.mdtw_ed_search_f64(self.h, self._f64(series, "series"), series.numel(),
                                          q.ctypes.data_as(C.POINTER(C.c_double)), q.size,
                                          C.byref(pos), C.byref(dist), C.byref(info)),
              "mdtw_ed_search_f64")
        return pos.value, dist.value

    def set_workspace_limit(self, nbytes: int):
        """Cap the engine's device workspace (0: none); calls needing more
        raise MdtwError(ERR_WORKSPACE) before any launch."""
        check(self.lib.mdtw_set_workspace_limit(self.h, int(nbytes)),
              "mdtw_set_workspace_limit")

    # -- batched search ----------------------------------------------------
    def batch(self, zqueries, bands, pipelines: int = 0) -> "Batch":
        """Queries of mixed length and band searched together against one
        stream (mdtw_batch_*; no reference counterpart)."""
        return Batch(self, zqueries, bands, pipelines)

    # -- streaming search --------------------------------------------------
    def stream(self, w: int, algo: int = ALGO_CDTW, metric: int = METRIC_L2, k: int = 1,
               excl: int = 0, stats_mode: int = STATS_FP64,
               epoch_windows: int = 0) -> "Stream":
        """A search over host chunks of a stream (mdtw_stream_*), answering
        as mdtw_search would over everything pushed."""
        return Stream(self, w, algo, metric, k, excl, stats_mode, epoch_windows)

    def restore_stream(self, state: bytes) -> "Stream":
        """A stream continuing from Stream.save() (same query required)."""
        return Stream(self, state=state)

    # -- primitives --------------------------------------------------------
    def threshold(self, values, indices, thr: float) -> int:
        """cub_util.cuh:154-172 threshold: stable in-place compaction of the
        indices whose value is <= thr; returns the new length."""
        cnt = C.c_int64()
        self._stream()
        check(self.lib.mdtw_threshold(self.h, self._f32(values, "values"),
                                      self._i32(indices, "indices"), values.numel(),
                                      C.c_float(thr), C.byref(cnt)), "mdtw_threshold")
        return cnt.value

    def argmin(self, values):
        """Minimum value and its lowest index (cub_util.cuh:7-27 reduce +
        cdtw.cuh:493-508 report_result)."""
        idx = C.c_int64()
        val = C.c_float()
        self._stream()
        check(self.lib.mdtw_argmin(self.h, self._f32(values, "values"), values.numel(),
                                   C.byref(idx), C.byref(val)), "mdtw_argmin")
        return idx.value, val.value

    def synchronize(self):
        check(self.lib.mdtw_synchronize(self.h), "mdtw_synchronize")


@dataclass
class BatchResult:
    positions: np.ndarray   # int64[nq, k], -1 where fewer than k matches exist
    distances: np.ndarray   # float32[nq, k]
    infos: list             # per-query search info (dict)
    elapsed_ms: float       # device time of the whole batch (one graph launch)


class Stream:
    """Streaming search: host chunks of any size are pushed, every full
    device epoch is searched while the next one is copied, and the answer
    equals the one-shot search over the concatenated samples (the reference's
    epoch loops, stats.cuh:115-142 and UCR_DTW_BIN.cpp:579-707)."""

    def __init__(self, engine: Engine, w: int = 0, algo: int = ALGO_CDTW,
                 metric: int = METRIC_L2, k: int = 1, excl: int = 0,
                 stats_mode: int = STATS_FP64, epoch_windows: int = 0, state: bytes = None):
        self.engine = engine
        h = C.c_void_p()
        engine._stream()
        if state is None:
            check(engine.lib.mdtw_stream_create(engine.h, int(w), int(algo), int(metric),
                                                int(stats_mode), int(k), int(excl),
                                                int(epoch_windows), C.byref(h)),
                  "mdtw_stream_create")
            self.k = int(k)
        else:
            buf = C.create_string_buffer(bytes(state), len(state))
            check(engine.lib.mdtw_stream_restore(engine.h, buf, len(state), C.byref(h)),
                  "mdtw_stream_restore")
            # k sits after the magic and five int32 fields of the saved header
            self.k = int(np.frombuffer(bytes(state), dtype="<i4", count=1, offset=8 + 20)[0])
        self.h = h

    def push(self, samples) -> None:
        x = np.ascontiguousarray(samples, dtype=np.float32)
        self.engine._stream()
        check(self.engine.lib.mdtw_stream_push(self.h, _fp(x), x.size), "mdtw_stream_push")

    def finish(self) -> None:
        self.engine._stream()
        check(self.engine.lib.mdtw_stream_finish(self.h), "mdtw_stream_finish")

    def result(self) -> SearchResult:
        pos = np.full(self.k, -1, dtype=np.int64)
        dist = np.full(self.k, np.inf, dtype=np.float32)
        info = _lib.SearchInfo()
        self.engine._stream()
        check(self.engine.lib.mdtw_stream_result(self.h, pos.ctypes.data_as(C.POINTER(C.c_int64)),
                                                 _fp(dist), C.byref(info)), "mdtw_stream_result")
        return SearchResult(pos, dist, info.as_dict())

    def position(self):
        """(samples pushed, windows searched or in flight)."""
        a, b = C.c_int64(), C.c_int64()
        check(self.engine.lib.mdtw_stream_position(self.h, C.byref(a), C.byref(b)),
              "mdtw_stream_position")
        return a.value, b.value

    def save(self) -> bytes:
        size = C.c_size_t()
        self.engine._stream()
        check(self.engine.lib.mdtw_stream_save(self.h, None, 0, C.byref(size)), "mdtw_stream_save")
        buf = C.create_string_buffer(size.value)
        check(self.engine.lib.mdtw_stream_save(self.h, buf, size.value, C.byref(size)),
              "mdtw_stream_save")
        return buf.raw[:size.value]

    def close(self):
        if getattr(self, "h", None):
            self.engine.lib.mdtw_stream_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A set of z-normalised queries of mixed length and band, searched
    together: every query gets mdtw_search's answer, the searches run on
    concurrent pipelines replayed from one captured hipGraph."""

    def __init__(self, engine: Engine, zqueries, bands, pipelines: int = 0):
        self.engine = engine
        qs = [np.ascontiguousarray(q, dtype=np.float32) for q in zqueries]
        self.lengths = np.array([q.size for q in qs], dtype=np.int32)
        self.bands = np.ascontiguousarray(bands, dtype=np.int32)
        if self.bands.size != len(qs):
            raise ValueError("one band per query")
        packed = np.concatenate(qs) if qs else np.zeros(0, np.float32)
        h = C.c_void_p()
        engine._stream()
        check(engine.lib.mdtw_batch_create(
            engine.h, _fp(packed), self.lengths.ctypes.data_as(C.POINTER(C.c_int32)),
            self.bands.ctypes.data_as(C.POINTER(C.c_int32)), len(qs), int(pipelines),
            C.byref(h)), "mdtw_batch_create")
        self.h = h
        self.nq = len(qs)

    def close(self):
        if getattr(self, "h", None):
            self.engine.lib.mdtw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, series, k: int = 1, metric: int = METRIC_L2, cell_budget: int = 0) -> int:
        """Pick launch shapes per (length class, band kind) bucket by timing."""
        eng = self.engine
        nb = C.c_int32()
        eng._stream()
        check(eng.lib.mdtw_batch_tune(self.h, eng._f32(series, "series"), series.numel(),
                                      int(metric), int(k), int(cell_budget), C.byref(nb)),
              "mdtw_batch_tune")
        return nb.value

    def shape(self, i: int) -> dict:
        """Launch shape query i's searches use (mdtw_batch_shape): DP grid,
        anti-diagonal grid, cascade grid and DP formulation."""
        out = np.zeros(4, dtype=np.int32)
        check(self.engine.lib.mdtw_batch_shape(self.h, int(i), out.ctypes.data_as(C.POINTER(C.c_int32))),
              "mdtw_batch_shape")
        return {"dp_blocks_per_cu": int(out[0]), "wide_waves_per_cu": int(out[1]),
                "cascade_blocks_per_cu": int(out[2]),
                "dp_form": ("auto", "lb", "wide")[int(out[3])]}

    def set_shape(self, i: int, dp_blocks_per_cu: int = 8, wide_waves_per_cu: int = 32,
                  cascade_blocks_per_cu: int = 64, dp_form: str = "auto") -> None:
        """Fix the launch shape of query i's bucket (mdtw_batch_set_shape)."""
        v = np.array([dp_blocks_per_cu, wide_waves_per_cu, cascade_blocks_per_cu,
                      ("auto", "lb", "wide").index(dp_form)], dtype=np.int32)
        check(self.engine.lib.mdtw_batch_set_shape(self.h, int(i),
                                                   v.ctypes.data_as(C.POINTER(C.c_int32))),
              "mdtw_batch_set_shape")

    def search(self, series, k: int = 1, excl=None, metric: int = METRIC_L2) -> BatchResult:
        eng = self.engine
        pos = np.full((self.nq, k), -1, dtype=np.int64)
        dist = np.full((self.nq, k), np.inf, dtype=np.float32)
        infos = (_lib.SearchInfo * self.nq)()
        ms = C.c_float()
        ex = None
        if excl is not None:
            ex = np.ascontiguousarray(np.broadcast_to(np.asarray(excl, np.int64), (self.nq,)))
        eng._stream()
        check(eng.lib.mdtw_batch_search(
            self.h, eng._f32(series, "series"), series.numel(), int(metric), int(k),
            ex.ctypes.data_as(C.POINTER(C.c_int64)) if ex is not None else None,
            pos.ctypes.data_as(C.POINTER(C.c_int64)), _fp(dist), infos, C.byref(ms)),
            "mdtw_batch_search")
        return BatchResult(pos, dist, [i.as_dict() for i in infos], ms.value)

    def backward(self, series, positions, grads, metric: int = METRIC_L2,
                 sqrt_dist: bool = True, raw_queries=None):
        """Query gradients of sum_j grads[i, j] * dist(i, positions[i, j]) for
        every query (mdtw_batch_backward).  positions/grads: (nq, k) arrays
        (position < 0 skipped).  raw_queries: the raw queries (list), giving
        gradients with respect to them; otherwise with respect to the
        z-normalised queries.  Returns (list of nq gradient arrays, dist
        (nq, k))."""
        eng = self.engine
        pos = np.ascontiguousarray(positions, dtype=np.int64)
        k = pos.shape[1]
        g = np.ascontiguousarray(grads, dtype=np.float32).reshape(pos.shape)
        out = np.empty(int(self.lengths.sum()), np.float32)
        dist = np.empty(pos.shape, np.float32)
        raw = None
        if raw_queries is not None:
            raw = np.ascontiguousarray(np.concatenate(
                [np.asarray(q, np.float32) for q in raw_queries]), dtype=np.float32)
        eng._stream()
        check(eng.lib.mdtw_batch_backward(
            self.h, eng._f32(series, "series"), series.numel(), int(metric), int(k),
            pos.ctypes.data_as(C.POINTER(C.c_int64)), _fp(g), 1 if sqrt_dist else 0,
            _fp(raw) if raw is not None else None, _fp(out), _fp(dist)),
            "mdtw_batch_backward")
        splits = np.cumsum(self.lengths)[:-1]
        return np.split(out, splits), dist
