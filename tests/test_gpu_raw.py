"""Searches on raw values (MDTW_STATS_NONE) against the oracle.

The reference's serial_ldtw / tsh2_ldtw (ldtw.cuh:18-71,198-291) compare the
query with the stream's raw values.  MDTW_STATS_NONE gives every window mean
0 and sigma 1, which every kernel's (x - mean) / sigma maps to x exactly, so
the bars are the z-normalised ones: LDTW bit-exact with the oracle's
serial_ldtw restatement, cDTW bit-exact with the oracle DP over the same
(identity) statistics, searches = argmin / greedy top-k of those profiles,
ED within 1e-5 of FP64.  The stream is metal-like with a level offset of
about 7000, where z-normalised and raw answers differ.
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import ALGO_ED, ALGO_LDTW, METRIC_L1, METRIC_L2, STATS_NONE, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def setup(n, m, seed, noise=0.02):
    s = synth.metal_like(n, seed, offset=7000.0)
    q, pos = synth.cut_query(s, m, seed + 1, noise=noise)
    return s, q, pos


def test_identity_statistics(engine):
    s, q, _ = setup(5000, 64, 1)
    avg, sd = engine.avg_std(dev(s), 64, mode=STATS_NONE)
    assert np.all(host(avg) == 0.0) and np.all(host(sd) == 1.0)
    assert host(avg).size == 5000 - 64 + 1


@pytest.mark.parametrize("metric", [METRIC_L2, METRIC_L1])
@pytest.mark.parametrize("w", [0, 5, 19])
def test_raw_ldtw_profile_and_search(engine, metric, w):
    s, q, _ = setup(1 << 13, 96, 2)
    ds = dev(s)
    engine.set_query(q, w)
    avg, sd = engine.avg_std(ds, q.size, mode=STATS_NONE)
    got = host(engine.ldtw(ds, avg, sd, w, metric=metric))
    want = O.ldtw_raw(q, s, w, metric=metric)
    np.testing.assert_array_equal(got, want)
    # no statistics at all: the tsh2_ldtw form of the entry point
    np.testing.assert_array_equal(host(engine.ldtw(ds, None, None, w, metric=metric)), want)
    r = engine.search(ds, w, algo=ALGO_LDTW, metric=metric, stats_mode=STATS_NONE)
    assert r.positions[0] == O.argmin(want)


@pytest.mark.parametrize("metric", [METRIC_L2, METRIC_L1])
@pytest.mark.parametrize("m,w", [(64, 6), (128, 13), (256, 51), (200, 150)])
def test_raw_cdtw_profile_and_search(engine, metric, m, w):
    if metric == METRIC_L1 and w > 112:
        w = 112
    s, q, pos = setup(1 << 14, m, 3 + m)
    ds = dev(s)
    L = s.size - m + 1
    zeros, ones = np.zeros(L, np.float32), np.ones(L, np.float32)
    prof = O.cdtw(q, s, zeros, ones, w, metric=metric)
    engine.set_query(q, w)
    avg, sd = engine.avg_std(ds, m, mode=STATS_NONE)
    np.testing.assert_array_equal(host(engine.gpu_cdtw(ds, avg, sd, w, metric=metric)), prof)
    r = engine.search(ds, w, metric=metric, stats_mode=STATS_NONE)
    best = O.argmin(prof)
    assert r.positions[0] == best
    want = np.sqrt(prof[best]) if metric == METRIC_L2 else prof[best]
    assert r.distances[0] == np.float32(want)
    assert abs(best - pos) <= max(w, 2)  # the cut window, up to a warp of W
    # top-k with an exclusion zone: greedy over the oracle profile
    k, excl = 5, m // 2
    rk = engine.search(ds, w, metric=metric, k=k, excl=excl, stats_mode=STATS_NONE)
    idx, dist = O.topk_greedy(prof, k, excl)
    np.testing.assert_array_equal(rk.positions, idx)
    want_d = np.sqrt(dist) if metric == METRIC_L2 else dist
    np.testing.assert_array_equal(rk.distances, want_d.astype(np.float32))


def test_raw_ed_search(engine):
    s, q, pos = setup(1 << 15, 128, 5)
    ds = dev(s)
    L = s.size - 128 + 1
    ed = O.ed_f64(q.astype(np.float64), s, np.zeros(L), np.ones(L))
    engine.set_query(q, 0)
    r = engine.search(ds, 0, algo=ALGO_ED, k=3, excl=64, stats_mode=STATS_NONE)
    idx, dist = O.topk_greedy(ed.astype(np.float32), 3, 64)
    assert r.positions[0] == idx[0] == int(np.argmin(ed))
    np.testing.assert_allclose(r.distances[0] ** 2, ed[idx[0]], rtol=1e-5)


def test_raw_and_znormalised_answers_differ(engine):
    # On offset data the raw search finds the cut window; this guards against
    # the mode being ignored (a z-normalised query against raw windows would
    # give distances of order M * 7000^2).
    s, q, pos = setup(1 << 14, 128, 7)
    ds = dev(s)
    engine.set_query(q, 12)
    r = engine.search(ds, 12, stats_mode=STATS_NONE)
    assert abs(int(r.positions[0]) - pos) <= 12
    assert r.distances[0] < 1.0
