"""End-to-end search parity on the GPU.

The pruned search must return exactly what the unpruned profile gives
(pruning never changes the answer; ties go to the lowest index), its
locations must match what the reference's UCR_DTW_BIN / UCR_ED_BIN print for
the same inputs (tests/golden/ucr_golden.json), and top-k must equal the
greedy exclusion-zone selection over the full profile.
"""
import json
import os

import numpy as np
import pytest
import torch

from cudadtw_amd import (ALGO_CDTW, ALGO_ED, ALGO_LDTW, METRIC_L1, METRIC_L2, STATS_FP64,
                         STATS_REFCOMPAT, synth)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ucr_golden.json")))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def golden_inputs(name):
    if name == "config1_random_walk":
        s, q, w = synth.config1()
        return s, q, w
    if name == "ecg_2p18_m128":
        return synth.ecg_like(1 << 18, 5), synth.ecg_query(128, 2), 13
    if name == "ecg_2p18_m256_cut":
        e = synth.ecg_like(1 << 18, 5)
        return e, synth.cut_query(e, 256, 9)[0], 26
    mt = synth.metal_like(1 << 18, 3)
    return mt, synth.cut_query(mt, 512, 4)[0], 51


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_search_matches_profile_and_ucr(engine, name):
    s, q, w = golden_inputs(name)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    ds = dev(s)
    r = engine.search(ds, w)
    # unpruned profile over every window, same kernel arithmetic
    avg, sd = engine.avg_std(ds, q.size)
    prof = host(engine.gpu_cdtw(ds, avg, sd, w))
    i = int(np.argmin(prof))
    assert r.location == i
    assert r.distances[0] == np.float32(np.sqrt(prof[i]))
    assert r.location == GOLDEN[name]["ucr_dtw"]["location"]
    assert abs(r.distance / GOLDEN[name]["ucr_dtw"]["distance"] - 1) < 1e-3
    info = r.info
    assert info["windows"] == s.size - q.size + 1
    assert info["lb_kim_pruned"] + info["lb_keogh_pruned"] > 0
    # ED best match; the reference's ED driver z-normalises the query in FP64
    # (GPU-ed.cu reads doubles; stats.cuh:8-28 with value_t = double)
    engine.set_query(O.znormalize_f64(q.astype(np.float64)).astype(np.float32), 0)
    re_ = engine.search(ds, 0, algo=ALGO_ED)
    assert re_.location == GOLDEN[name]["ucr_ed"]["location"]
    assert abs(re_.distance / GOLDEN[name]["ucr_ed"]["distance"] - 1) < 1e-4


def test_refcompat_search_matches_restated_reference(engine):
    s, q, w = synth.config1()
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    r = engine.search(dev(s), w, stats_mode=STATS_REFCOMPAT)
    ra, rs = O.stats_refcompat(s, q.size)
    prof = O.cdtw(zq, s, ra, rs, w, fused=1)
    assert r.location == O.argmin(prof)
    assert abs(r.distance / np.sqrt(prof.min()) - 1) < 1e-5


@pytest.mark.parametrize("w", [13, 51, 120, 122, 136, 150, 200, 230, 240])
def test_search_with_given_stats_and_wide_band(engine, w):
    s = synth.ecg_like(1 << 17, 21)
    q, pos = synth.cut_query(s, 512, 22)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, 512)
    r = engine.search(ds, w, avg=avg, sd=sd)
    prof = host(engine.gpu_cdtw(ds, avg, sd, w))
    assert r.location == int(np.argmin(prof))
    assert r.distances[0] == np.float32(np.sqrt(prof.min()))
    loc, dist = engine.prune_cdtw(ds, avg, sd, w)
    assert (loc, dist) == (r.location, r.distance)
    # top-k through the same DP (abandoning against the top-k bound)
    rk = engine.search(ds, w, k=4, excl=256, avg=avg, sd=sd)
    idx, d2 = O.topk_greedy(prof, 4, 256)
    assert np.array_equal(rk.positions[:len(idx)], idx)
    assert np.array_equal(rk.distances[:len(idx)], np.sqrt(d2).astype(np.float32))


# (1024, 16) / (300, 3) / (1024, 0): k up to MDTW_MAX_K, so the k-th smallest
# seed (a radix selection over up to 16384 seed distances) sets the first
# bound; excl 0 and 3 give more than 2^15 buckets, the multi-block tighten.
@pytest.mark.parametrize("algo,k,excl", [(ALGO_CDTW, 16, 64), (ALGO_CDTW, 4, 0),
                                         (ALGO_LDTW, 8, 64), (ALGO_ED, 8, 64), (ALGO_CDTW, 1, 0),
                                         (ALGO_LDTW, 1, 0), (ALGO_CDTW, 1024, 16),
                                         (ALGO_CDTW, 300, 3), (ALGO_ED, 1024, 0)])
def test_topk_equals_greedy_over_profile(engine, algo, k, excl):
    m, w = 128, 13
    s = synth.ecg_like(1 << 17, 31)
    zq = O.znormalize_f32(synth.ecg_query(m, 2))
    engine.set_query(zq, w)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, m)
    if algo == ALGO_CDTW:
        prof = host(engine.gpu_cdtw(ds, avg, sd, w))
    elif algo == ALGO_LDTW:
        prof = host(engine.ldtw(ds, avg, sd, w))
    else:
        prof = host(engine.ed(ds, avg, sd))
    r = engine.search(ds, w, algo=algo, k=k, excl=excl)
    idx, d2 = O.topk_greedy(prof, k, excl)
    assert np.array_equal(r.positions[:len(idx)], idx)
    assert np.array_equal(r.distances[:len(idx)], np.sqrt(d2).astype(np.float32))
    assert r.info["found"] == len(idx)


def test_topk_more_picks_than_seeds(engine):
    # A short stream where fewer seed windows (>= 2*excl apart) exist than k:
    # no seed bound (every window is a candidate), and the greedy selection
    # may find fewer than k matches.
    m, w, k, excl = 64, 6, 64, 32
    s = synth.ecg_like(4000, 33)
    zq = O.znormalize_f32(synth.ecg_query(m, 2))
    engine.set_query(zq, w)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, m)
    prof = host(engine.gpu_cdtw(ds, avg, sd, w))
    r = engine.search(ds, w, k=k, excl=excl)
    idx, d2 = O.topk_greedy(prof, k, excl)
    assert r.info["seeded"] < k
    assert r.info["found"] == len(idx)
    assert np.array_equal(r.positions[:len(idx)], idx)
    assert np.array_equal(r.distances[:len(idx)], np.sqrt(d2).astype(np.float32))


def test_metal_topk16_ldtw(engine):
    # Metal-like stream, |Q| = 512 cut and perturbed, W = 10%, top-16 with an
    # exclusion zone of |Q|/2, cDTW and LDTW.
    m, w = 512, 51
    s = synth.metal_like(1 << 17, 40)
    q, pos = synth.cut_query(s, m, 41)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, m)
    for algo, prof in ((ALGO_CDTW, host(engine.gpu_cdtw(ds, avg, sd, w))),
                       (ALGO_LDTW, host(engine.ldtw(ds, avg, sd, w)))):
        r = engine.search(ds, w, algo=algo, k=16, excl=m // 2)
        idx, d2 = O.topk_greedy(prof, 16, m // 2)
        assert np.array_equal(r.positions, idx)
        assert np.array_equal(r.distances, np.sqrt(d2).astype(np.float32))


def test_ties_go_to_lowest_index(engine):
    # A block repeated with the statistics tile period (2048 windows) makes
    # every copy's window statistics and distance bit-identical: exact ties.
    block = synth.random_walk(2048, 3)
    s = np.tile(block, 4).astype(np.float32)
    q = block[300:428].copy()
    zq = O.znormalize_f32(q)
    engine.set_query(zq, 13)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, 128)
    prof = host(engine.gpu_cdtw(ds, avg, sd, 13))
    assert prof[300] == prof[2348] == prof[4396] == prof.min()
    r = engine.search(ds, 13)
    assert r.location == 300
    r2 = engine.search(ds, 13, k=2, excl=64)
    assert list(r2.positions) == [300, 2348]
    assert r2.distances[0] == r2.distances[1]


def test_short_query_and_single_window(engine):
    s = synth.random_walk(5000, 12)
    for m, w in ((8, 2), (3, 1)):
        zq = O.znormalize_f32(synth.random_walk(m, 13))
        engine.set_query(zq, w)
        ds = dev(s)
        r = engine.search(ds, w)
        avg, sd = engine.avg_std(ds, m)
        prof = host(engine.gpu_cdtw(ds, avg, sd, w))
        assert r.location == int(np.argmin(prof))
    zq = O.znormalize_f32(s[:64])
    engine.set_query(zq, 6)
    r = engine.search(dev(s[:64].copy()), 6)
    assert r.location == 0 and r.distance < 1e-3


def test_l1_search(engine):
    s = synth.ecg_like(1 << 16, 50)
    zq = O.znormalize_f32(synth.ecg_query(128, 2))
    engine.set_query(zq, 13)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, 128)
    prof = host(engine.gpu_cdtw(ds, avg, sd, 13, metric=METRIC_L1))
    r = engine.search(ds, 13, metric=METRIC_L1)
    assert r.location == int(np.argmin(prof)) and r.distances[0] == prof.min()


def test_search_deterministic(engine):
    s = synth.ecg_like(1 << 18, 60)
    zq = O.znormalize_f32(synth.ecg_query(256, 2))
    engine.set_query(zq, 26)
    ds = dev(s)
    a = engine.search(ds, 26, k=8, excl=128)
    b = engine.search(ds, 26, k=8, excl=128)
    assert np.array_equal(a.positions, b.positions)
    assert np.array_equal(a.distances, b.distances)
