"""Streaming search (mdtw_stream_*) against the oracle and the one-shot search.

No reference API counterpart (the reference's epoch loops are internal:
stats.cuh:115-142, UCR_DTW_BIN.cpp:579-707).  The bar: after finish, the
streamed answer -- positions and distances -- equals the greedy selection
over the oracle's profile of the whole stream (cDTW and LDTW bit for bit,
ED within 1e-5 of the FP64 oracle) across many epoch boundaries, and is
bit-identical to mdtw_search over the whole stream resident on the device
for every algorithm, k = 1 and top-k, any chunking of the pushes, matches
that straddle epoch boundaries, and a save / restore in the middle.
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import (ALGO_CDTW, ALGO_ED, ALGO_LDTW, METRIC_L1, METRIC_L2, STATS_FP64,
                         STATS_NONE, STATS_REFCOMPAT, synth)
from cudadtw_amd._lib import MdtwError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

EPOCH = 1 << 16


def push_chunks(st, s, seed):
    rng = np.random.default_rng(seed)
    i = 0
    while i < s.size:
        n = int(rng.choice([1, 17, 999, 4096, 30000, 70001, 150000]))
        st.push(s[i:i + n])
        i += n


def one_shot(engine, s, w, **kw):
    return engine.search(torch.from_numpy(s).cuda(), w, **kw)


def same(a, b):
    np.testing.assert_array_equal(a.positions, b.positions)
    np.testing.assert_array_equal(a.distances, b.distances)


@pytest.mark.parametrize("algo,metric,k", [(ALGO_CDTW, METRIC_L2, 1), (ALGO_CDTW, METRIC_L2, 6),
                                           (ALGO_CDTW, METRIC_L1, 4), (ALGO_ED, METRIC_L2, 1),
                                           (ALGO_ED, METRIC_L2, 5), (ALGO_LDTW, METRIC_L2, 1),
                                           (ALGO_LDTW, METRIC_L1, 3)])
def test_stream_equals_one_shot(engine, algo, metric, k):
    s = synth.ecg_like(300_000, 21)
    q, _ = synth.cut_query(s, 128, 22)
    w = 0 if algo == ALGO_ED else 12
    engine.set_query(O.znormalize_f32(q), w)
    excl = 64 if k > 1 else 0
    want = one_shot(engine, s, w, algo=algo, metric=metric, k=k, excl=excl)
    st = engine.stream(w, algo=algo, metric=metric, k=k, excl=excl, epoch_windows=EPOCH)
    push_chunks(st, s, 5 + k)
    st.finish()
    got = st.result()
    same(got, want)
    assert got.info["windows"] == s.size - 128 + 1
    assert got.info["found"] == want.info["found"]
    st.close()


def oracle_profile(engine, s, zq, w, algo, metric):
    """The oracle's profile of every window: sequential_cdtw (cdtw.cuh:23-88)
    or z_serial_ldtw (ldtw.cuh:74-127) over the GPU's window statistics (a
    window's statistics depend only on its own samples and its 1024-window
    tile, so every epoch sees the same values), or FP64 ED over FP64
    two-pass statistics."""
    m = zq.size
    if algo == ALGO_ED:
        _, _, a64, s64 = O.stats_fp64(s, m, want64=True)
        return O.ed_f64(zq.astype(np.float64), s, a64, s64)
    avg, sd = engine.avg_std(torch.from_numpy(s).cuda(), m)
    torch.cuda.synchronize()
    ha, hs = avg.cpu().numpy(), sd.cpu().numpy()
    if algo == ALGO_CDTW:
        return O.cdtw(zq, s, ha, hs, w, metric=metric, fused=1)
    return O.ldtw(zq, s, ha, hs, w, metric=metric)


@pytest.mark.parametrize("algo,metric,k,excl", [(ALGO_CDTW, METRIC_L2, 1, 0),
                                                (ALGO_CDTW, METRIC_L2, 7, 48),
                                                (ALGO_CDTW, METRIC_L1, 5, 2),
                                                (ALGO_LDTW, METRIC_L2, 6, 64),
                                                (ALGO_LDTW, METRIC_L1, 1, 0),
                                                (ALGO_ED, METRIC_L2, 5, 64)])
def test_stream_against_oracle_profile(engine, algo, metric, k, excl):
    # Nine epochs of 8192 windows; pushes of random sizes cross every boundary.
    m = 96
    epoch = 8192
    s = synth.ecg_like(9 * epoch + m - 1 - 333, 80 + k)
    q, _ = synth.cut_query(s, m, 81)
    zq = O.znormalize_f32(q)
    w = 0 if algo == ALGO_ED else 9
    engine.set_query(zq, w)
    st = engine.stream(w, algo=algo, metric=metric, k=k, excl=excl, epoch_windows=epoch)
    push_chunks(st, s, 90 + k)
    st.finish()
    got = st.result()
    st.close()
    assert got.info["windows"] == s.size - m + 1
    prof = oracle_profile(engine, s, zq, w, algo, metric)
    if algo == ALGO_ED:
        # FP32 search against the FP64 oracle: the same picks, distances
        # within 1e-5 relative (no near-ties at these ranks on this stream).
        order = np.lexsort((np.arange(prof.size), prof))
        idx, dist = [], []
        for i in order:
            if len(idx) == k:
                break
            if all(abs(int(i) - j) >= excl for j in idx):
                idx.append(int(i))
                dist.append(np.sqrt(prof[i]))
        np.testing.assert_array_equal(got.positions, idx)
        np.testing.assert_allclose(got.distances, dist, rtol=1e-5)
        return
    idx, d = O.topk_greedy(prof, k, excl)
    np.testing.assert_array_equal(got.positions[:idx.size], idx)
    want = np.sqrt(d).astype(np.float32) if metric == METRIC_L2 else d
    np.testing.assert_array_equal(got.distances[:idx.size], want)
    assert got.info["found"] == idx.size


def test_matches_across_epoch_boundaries(engine):
    # exact copies of the query planted so that the best windows start just
    # before, at and just after epoch boundaries (windows >= EPOCH - 127 use
    # samples of two device buffers' worth of pushes)
    rng = np.random.default_rng(3)
    s = synth.random_walk(5 * EPOCH, 30)
    q = synth.random_walk(128, 31)
    for p in (EPOCH - 64, 2 * EPOCH, 3 * EPOCH + 1, 4 * EPOCH - 127):
        s[p:p + 128] = q * (1.0 + 0.01 * rng.standard_normal()) + rng.standard_normal()
    zq = O.znormalize_f32(q)
    engine.set_query(zq, 10)
    for k in (1, 4):
        want = one_shot(engine, s, 10, k=k, excl=64)
        st = engine.stream(10, k=k, excl=64, epoch_windows=EPOCH)
        push_chunks(st, s, k)
        st.finish()
        same(st.result(), want)
        if k == 4:
            assert sorted(want.positions.tolist()) == [EPOCH - 64, 2 * EPOCH, 3 * EPOCH + 1,
                                                       4 * EPOCH - 127]
        st.close()


@pytest.mark.parametrize("k", [1, 5])
def test_save_and_restore_mid_stream(engine, k):
    s = synth.ecg_like(400_000, 40)
    q, _ = synth.cut_query(s, 200, 41)
    engine.set_query(O.znormalize_f32(q), 20)
    want = one_shot(engine, s, 20, k=k, excl=100)
    st = engine.stream(20, k=k, excl=100, epoch_windows=EPOCH)
    cut = 170_123
    push_chunks(st, s[:cut], 1)
    state = st.save()
    st.close()
    st2 = engine.restore_stream(state)
    assert st2.position()[0] == cut
    push_chunks(st2, s[cut:], 2)
    st2.finish()
    same(st2.result(), want)
    st2.close()


def test_raw_values_stream(engine):
    s = synth.metal_like(250_000, 50, offset=7000.0)
    q, _ = synth.cut_query(s, 96, 51, noise=0.02)
    engine.set_query(q, 9)
    for k in (1, 3):
        want = one_shot(engine, s, 9, k=k, excl=48, stats_mode=STATS_NONE)
        st = engine.stream(9, k=k, excl=48, stats_mode=STATS_NONE, epoch_windows=EPOCH)
        push_chunks(st, s, 9)
        st.finish()
        same(st.result(), want)
        st.close()


def test_partial_results_and_edges(engine):
    s = synth.ecg_like(200_000, 60)
    q, _ = synth.cut_query(s, 128, 61)
    engine.set_query(O.znormalize_f32(q), 12)
    st = engine.stream(12, epoch_windows=EPOCH)
    st.push(s[:100])
    r = st.result()
    assert r.positions[0] == -1 and r.info["windows"] == 0
    st.push(s[100:EPOCH + 127 + 5000])  # one full epoch
    r = st.result()
    assert r.info["windows"] == EPOCH
    prefix = one_shot(engine, s[:EPOCH + 127], 12)
    same(r, prefix)
    st.finish()
    with pytest.raises(MdtwError):
        st.push(s[:10])
    st.close()
    # too short for one window
    st = engine.stream(12, epoch_windows=EPOCH)
    st.push(s[:127])
    st.finish()
    assert st.result().positions[0] == -1
    st.close()


def test_stream_argument_errors(engine):
    q = O.znormalize_f32(synth.random_walk(64, 1))
    engine.set_query(q, 6)
    with pytest.raises(MdtwError):
        engine.stream(6, stats_mode=STATS_REFCOMPAT)
    with pytest.raises(MdtwError):
        engine.stream(6, epoch_windows=1000)  # not a multiple of 1024
    with pytest.raises(MdtwError):
        engine.stream(6, algo=ALGO_ED, metric=METRIC_L1)
    st = engine.stream(6, epoch_windows=EPOCH)
    st.push(synth.random_walk(5000, 2))
    state = st.save()
    st.close()
    engine.set_query(O.znormalize_f32(synth.random_walk(64, 3)), 6)
    with pytest.raises(MdtwError):
        engine.restore_stream(state)  # another query
    with pytest.raises(MdtwError):
        engine.restore_stream(b"garbage" * 20)
    # the stream keeps its own query copy: a later set_query changes nothing
    engine.set_query(q, 6)
    s = synth.ecg_like(100_000, 4)
    want = one_shot(engine, s, 6, stats_mode=STATS_FP64)
    st = engine.stream(6, epoch_windows=EPOCH)
    engine.set_query(O.znormalize_f32(synth.random_walk(64, 9)), 6)
    st.push(s)
    st.finish()
    same(st.result(), want)
    st.close()


def test_restore_after_longer_queries_in_between(engine):
    # The state checks the query by hash over its own samples only: setting
    # longer queries before and after must not change it.
    q = O.znormalize_f32(synth.random_walk(64, 11))
    engine.set_query(O.znormalize_f32(synth.random_walk(900, 12)), 30)
    engine.set_query(q, 6)
    s = synth.ecg_like(120_000, 13)
    want = one_shot(engine, s, 6, k=3, excl=32)
    st = engine.stream(6, k=3, excl=32, epoch_windows=EPOCH)
    st.push(s[:70_000])
    state = st.save()
    st.close()
    engine.set_query(O.znormalize_f32(synth.random_walk(2000, 14)), 100)
    engine.set_query(q, 6)
    st = engine.restore_stream(state)
    st.push(s[70_000:])
    st.finish()
    same(st.result(), want)
    st.close()
