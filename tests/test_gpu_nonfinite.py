"""Non-finite samples in the stream (a robustness case the reference leaves
them undefined).  The rule, in the FP64 and raw statistics modes: a window
holding a NaN or an infinity has NaN mean and sigma, hence NaN distances and
bounds -- never a match, never kept by a top-k bound -- and every other
window is computed exactly as if the stream were clean.  The oracle states
the same rule (or_stats_fp64, ldtw_one) and its DPs carry the NaN through, so
profiles are compared bit for bit with NaN in the same places, and searches
against the greedy selection over the oracle's profile with NaN excluded.
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import (ALGO_CDTW, ALGO_ED, ALGO_ED_FFT, ALGO_LDTW, METRIC_L1, METRIC_L2,
                         STATS_NONE, synth)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# non-finite samples at the stream's start, at and next to 1024-window tile
# starts, and in the middle of tiles
PLANTS = {0: np.nan, 1024: np.inf, 1023: -np.inf, 5000: np.nan, 5003: np.nan, 20480 + 7: np.inf,
          33333: -np.nan}


def planted(n, seed, m):
    s = synth.ecg_like(n, seed)
    for p, v in PLANTS.items():
        s[p] = v
    bad = np.zeros(n - m + 1, bool)
    for p in PLANTS:
        bad[max(0, p - m + 1):p + 1] = True
    return s, bad


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("m", [64, 300])
def test_statistics_nan_exactly_on_windows_with_nonfinite_samples(engine, m):
    s, bad = planted(40_000, 3, m)
    avg, sd = engine.avg_std(torch.from_numpy(s).cuda(), m)
    a, d = host(avg), host(sd)
    assert np.array_equal(np.isnan(a), bad) and np.array_equal(np.isnan(d), bad)
    oa, od, a64, s64 = O.stats_fp64(s, m, want64=True)
    assert np.array_equal(np.isnan(oa), bad)
    ok = ~bad
    assert np.max(np.abs(d[ok] - s64[ok]) / s64[ok]) < 2e-7
    assert np.max(np.abs(a[ok] - a64[ok])) < 1e-6 * np.nanmax(np.abs(s[np.isfinite(s)]))
    # raw mode: mean 0 / sigma 1, NaN on the same windows
    avg, sd = engine.avg_std(torch.from_numpy(s).cuda(), m, STATS_NONE)
    a, d = host(avg), host(sd)
    assert np.array_equal(np.isnan(a), bad) and np.all(a[ok] == 0) and np.all(d[ok] == 1)


@pytest.mark.parametrize("w,metric", [(12, METRIC_L2), (12, METRIC_L1), (150, METRIC_L2),
                                      (260, METRIC_L2)])
def test_profiles_bit_exact_with_nan_windows(engine, w, metric):
    m = 300
    s, bad = planted(40_000, 4, m)
    q, _ = synth.cut_query(np.nan_to_num(s, nan=0.0, posinf=0.0, neginf=0.0), m, 5)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    ds = torch.from_numpy(s).cuda()
    avg, sd = engine.avg_std(ds, m)
    ha, hs = host(avg), host(sd)
    got = host(engine.gpu_cdtw(ds, avg, sd, w, metric=metric))
    ref = O.cdtw(zq, s, ha, hs, w, metric=metric)
    assert np.array_equal(got, ref, equal_nan=True)
    assert np.array_equal(np.isnan(got), bad)
    if w <= 150:
        lg = host(engine.ldtw(ds, avg, sd, w, metric=metric))
        lr = O.ldtw(zq, s, ha, hs, w, metric=metric)
        assert np.array_equal(lg, lr, equal_nan=True) and np.array_equal(np.isnan(lg), bad)
    if metric == METRIC_L2:
        for kind in ("ed", "kim", "keogh"):
            v = host(engine.ed(ds, avg, sd) if kind == "ed" else
                     engine.lb_kim(ds, avg, sd) if kind == "kim" else engine.lb_keogh(ds, avg, sd))
            assert np.array_equal(np.isnan(v), bad), kind
        fft = host(engine.ed_fft(ds))
        assert np.array_equal(np.isnan(fft), bad)
        _, _, a64, s64 = O.stats_fp64(s, m, want64=True)
        e64 = O.ed_f64(zq.astype(np.float64), s, a64, s64)
        ok = ~bad
        assert np.max(np.abs(fft[ok] - e64[ok]) / np.maximum(e64[ok], 1.0)) < 1e-5


@pytest.mark.parametrize("algo,k", [(ALGO_CDTW, 1), (ALGO_CDTW, 5), (ALGO_LDTW, 3), (ALGO_ED, 1),
                                    (ALGO_ED, 4), (ALGO_ED_FFT, 4)])
def test_searches_never_return_nan_windows(engine, algo, k):
    m, w = 128, 12
    s, bad = planted(60_000, 6, m)
    # the best matches sit right next to the non-finite samples
    q = s[5010:5010 + m].copy()
    zq = O.znormalize_f32(q)
    ww = 0 if algo in (ALGO_ED, ALGO_ED_FFT) else w
    engine.set_query(zq, ww)
    ds = torch.from_numpy(s).cuda()
    avg, sd = engine.avg_std(ds, m)
    ha, hs = host(avg), host(sd)
    if algo == ALGO_CDTW:
        prof = O.cdtw(zq, s, ha, hs, w)
    elif algo == ALGO_LDTW:
        prof = O.ldtw(zq, s, ha, hs, w)
    else:  # FP64 ED over FP64 statistics (NaN on the same windows)
        _, _, a64, s64 = O.stats_fp64(s, m, want64=True)
        prof = O.ed_f64(zq.astype(np.float64), s, a64, s64).astype(np.float32)
    r = engine.search(ds, ww, algo=algo, k=k, excl=m // 2)
    assert not np.any(bad[r.positions[r.positions >= 0]])
    idx, d2 = O.topk_greedy(prof, k, m // 2)
    np.testing.assert_array_equal(r.positions[:idx.size], idx)
    if algo in (ALGO_CDTW, ALGO_LDTW):
        np.testing.assert_array_equal(r.distances[:idx.size], np.sqrt(d2).astype(np.float32))
    else:
        # ED^2 within 1e-5 of the FP64 value (relative, absolute below 1)
        got2 = r.distances[:idx.size].astype(np.float64) ** 2
        assert np.all(np.abs(got2 - d2) <= 1e-5 * np.maximum(d2, 1.0))
    assert r.positions[0] == 5010
    # streamed in epochs: the same answer
    if algo != ALGO_ED_FFT:
        st = engine.stream(ww, algo=algo, k=k, excl=m // 2, epoch_windows=8192)
        st.push(s)
        st.finish()
        rs = st.result()
        st.close()
        np.testing.assert_array_equal(rs.positions, r.positions)
        np.testing.assert_array_equal(rs.distances, r.distances)


def test_workspace_cap(engine):
    from cudadtw_amd import _lib
    from cudadtw_amd._lib import MdtwError

    m, w = 128, 12
    s = synth.ecg_like(1 << 20, 9)
    zq = O.znormalize_f32(synth.cut_query(s, m, 10)[0])
    engine.set_query(zq, w)
    ds = torch.from_numpy(s).cuda()
    want = engine.search(ds, w, k=4, excl=64)
    engine.set_workspace_limit(1 << 20)  # far below the ~20 MiB this search needs
    try:
        with pytest.raises(MdtwError) as ei:
            engine.search(ds, w, k=4, excl=64)
        assert ei.value.status == _lib.ERR_WORKSPACE
        small = engine.search(ds[:20_000].contiguous(), w)  # fits: still served
        assert small.location >= 0
    finally:
        engine.set_workspace_limit(0)
    again = engine.search(ds, w, k=4, excl=64)
    np.testing.assert_array_equal(again.positions, want.positions)
    np.testing.assert_array_equal(again.distances, want.distances)
