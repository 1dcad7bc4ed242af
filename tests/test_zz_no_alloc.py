"""mdtw_search_noalloc (the allocation-free search) against mdtw_search, both
called through the C ABI.

No reference counterpart.  The bar: identical positions, distance bits and
timing-independent search counters to mdtw_search's on the same inputs, for every algorithm it
serves, k = 1 and top-k, statistics computed or passed in, on two calls in a
row of which the second has the smaller input (the fixed scratch array keeps
the first call's contents); sizes beyond the array and the forms it does not
serve are refused before any launch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from cudadtw_amd import (ALGO_CDTW, ALGO_ED, ALGO_ED_FFT, ALGO_LDTW, METRIC_L1, METRIC_L2,
                         STATS_FP64, STATS_NONE, STATS_REFCOMPAT, _lib, synth)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# Counters that do not depend on timing.  dp_abandoned does: for k = 1 a DP
# abandons against the best distance other waves have published so far,
# which differs from one run of the same search to the next.
FIELDS = ("windows", "seeded", "lb_kim_pruned", "lb_keogh_pruned", "dp_evaluated", "found",
          "threshold")


def call(fn, engine, ds, n, w, algo, metric, stats, k, excl, avg=None, sd=None):
    pos = np.full(k, -7, np.int64)
    dist = np.full(k, -7.0, np.float32)
    info = _lib.SearchInfo()
    engine._stream()
    st = fn(engine.h, C.c_void_p(ds.data_ptr()), n,
            C.c_void_p(avg.data_ptr()) if avg is not None else None,
            C.c_void_p(sd.data_ptr()) if sd is not None else None, w, algo, metric, stats, k,
            excl, pos.ctypes.data_as(C.POINTER(C.c_int64)),
            dist.ctypes.data_as(C.POINTER(C.c_float)), C.byref(info))
    return st, pos, dist, info.as_dict()


def both(engine, ds, n, w, algo, metric, stats, k, excl, avg=None, sd=None):
    lib = _lib.load()
    a = call(lib.mdtw_search, engine, ds, n, w, algo, metric, stats, k, excl, avg, sd)
    b = call(lib.mdtw_search_noalloc, engine, ds, n, w, algo, metric, stats, k, excl, avg, sd)
    assert a[0] == _lib.OK and b[0] == _lib.OK, (a[0], b[0])
    np.testing.assert_array_equal(b[1], a[1])
    np.testing.assert_array_equal(b[2].view(np.uint32), a[2].view(np.uint32))
    for f in FIELDS:
        assert b[3][f] == a[3][f], (f, b[3][f], a[3][f])
    return b


CASES = [(ALGO_CDTW, METRIC_L2, STATS_FP64, 1, 0), (ALGO_CDTW, METRIC_L2, STATS_FP64, 8, 64),
         (ALGO_CDTW, METRIC_L1, STATS_FP64, 4, 2), (ALGO_LDTW, METRIC_L2, STATS_FP64, 1, 0),
         (ALGO_LDTW, METRIC_L1, STATS_FP64, 5, 48), (ALGO_ED, METRIC_L2, STATS_FP64, 1, 0),
         (ALGO_ED, METRIC_L2, STATS_FP64, 6, 64), (ALGO_CDTW, METRIC_L2, STATS_NONE, 3, 40)]


@pytest.mark.parametrize("algo,metric,stats,k,excl", CASES)
def test_noalloc_equals_search_two_calls(engine, algo, metric, stats, k, excl):
    m = 128
    s = synth.ecg_like(400_000, 11 + k)
    q, _ = synth.cut_query(s, m, 12)
    zq = q if stats == STATS_NONE else O.znormalize_f32(q)
    w = 0 if algo == ALGO_ED else 12
    engine.set_query(zq, w)
    ds = torch.from_numpy(s).cuda()
    big = both(engine, ds, s.size, w, algo, metric, stats, k, excl)
    small = both(engine, ds, 60_000, w, algo, metric, stats, k, excl)  # a prefix
    assert big[3]["windows"] == s.size - m + 1 and small[3]["windows"] == 60_000 - m + 1
    assert big[3]["found"] >= 1 and small[3]["found"] >= 1


def test_noalloc_given_statistics_and_band_changes(engine):
    m = 200
    s = synth.random_walk(300_000, 3)
    zq = O.znormalize_f32(synth.random_walk(m, 4))
    ds = torch.from_numpy(s).cuda()
    for w, k in ((20, 1), (120, 4), (7, 2), (20, 1)):  # envelope and padded query per band
        engine.set_query(zq, 3)
        avg, sd = engine.avg_std(ds, m)
        both(engine, ds, s.size, w, ALGO_CDTW, METRIC_L2, STATS_FP64, k, 100, avg, sd)
        both(engine, ds, 50_000, w, ALGO_CDTW, METRIC_L2, STATS_FP64, k, 100)


def test_noalloc_refusals(engine):
    lib = _lib.load()
    m = 64
    engine.set_query(O.znormalize_f32(synth.random_walk(m, 5)), 6)
    ds = torch.zeros(1 << 24, device="cuda")
    # 2^24 windows need ~256 MiB of scratch: refused before any launch
    st = call(lib.mdtw_search_noalloc, engine, ds, 1 << 24, 6, ALGO_CDTW, METRIC_L2, STATS_FP64,
              1, 0)
    assert st[0] == _lib.ERR_WORKSPACE
    assert np.all(st[1] == -1) and np.all(np.isinf(st[2]))
    s = torch.from_numpy(synth.random_walk(10_000, 6)).cuda()
    for algo, stats in ((ALGO_ED_FFT, STATS_FP64), (ALGO_CDTW, STATS_REFCOMPAT)):
        st = call(lib.mdtw_search_noalloc, engine, s, 10_000, 6, algo, METRIC_L2, stats, 1, 0)
        assert st[0] == _lib.ERR_UNSUPPORTED
    # argument errors as mdtw_search's
    st = call(lib.mdtw_search_noalloc, engine, s, 10_000, 6, ALGO_CDTW, METRIC_L2, STATS_FP64, 0,
              0)
    assert st[0] == _lib.ERR_K
    # the ordinary search still serves the large input
    st = call(lib.mdtw_search, engine, ds, 1 << 24, 6, ALGO_CDTW, METRIC_L2, STATS_FP64, 1, 0)
    assert st[0] == _lib.OK
