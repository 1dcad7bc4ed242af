"""Per-kernel parity on the GPU against the CPU oracle, on seeded inputs.

Mirrors the reference's own tests: tests/test_dtw.cu (GPU DP vs the CPU DP
over all windows in random order, with the CPU given the GPU's statistics,
:84-86, :129-130), tests/test_dtwbounds.cu (LB_Kim, LB_Keogh <= cDTW on
every window) and tests/test_threshold.cu (compaction of 100 shuffled
values at threshold 22).  Integer/index results must be identical; the DP,
bounds and LDTW are bit-exact with the oracle (same operation sequence);
window statistics and ED are compared to FP64 within stated tolerances.
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import (ALGO_ED, METRIC_L1, METRIC_L2, STATS_FP64, STATS_NONE, STATS_REFCOMPAT,
                         _lib, synth)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def prepare(engine, s, q, w, mode=STATS_FP64):
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, zq.size, mode)
    return zq, ds, avg, sd, host(avg), host(sd)


# ---- window statistics -----------------------------------------------------------------
@pytest.mark.parametrize("kind,m", [("walk", 11), ("walk", 128), ("walk", 1000), ("walk", 2048),
                                    ("walk", 4099), ("ecg", 256), ("metal7000", 512)])
def test_stats_fp64_vs_two_pass(engine, kind, m):
    n = 1 << 16
    if kind == "walk":
        s = synth.random_walk(n, 11)
    elif kind == "ecg":
        s = synth.ecg_like(n, 11)
    else:
        s = synth.metal_like(n, 11, offset=7000.0)
    avg, sd = engine.avg_std(dev(s), m, STATS_FP64)
    a, d = host(avg), host(sd)
    oa, od, a64, s64 = O.stats_fp64(s, m, want64=True)
    assert a.shape == oa.shape
    # FP32 outputs of an FP64 computation: within 1 ulp of the FP64 value
    assert np.max(np.abs(a - a64) / np.maximum(np.abs(a64), 1e-30)) < 1.2e-7 or \
        np.max(np.abs(a - a64)) < 1e-6 * np.abs(s).max()
    rel = np.abs(d - s64) / s64
    assert rel.max() < 2e-7, rel.max()


def _two_pass(s, wins, m):
    """FP64 two-pass mean and population sigma of the listed windows."""
    x = np.lib.stride_tricks.sliding_window_view(s.astype(np.float64), m)[wins]
    mu = x.mean(axis=1)
    return mu, np.sqrt(((x - mu[:, None]) ** 2).mean(axis=1))


@pytest.mark.parametrize("m", [256, 1500])
def test_stats_sweep_over_many_tiles(engine, m):
    # More 1024-window tiles than the statistics kernel has resident waves
    # (each wave then takes several tiles, the next one's loads in flight),
    # a partial last tile, a first window longer than a tile (m = 1500),
    # and two tiles holding a non-finite sample: a sample of windows from
    # every part against FP64 two-pass statistics, and NaN exactly for the
    # windows holding the planted samples.
    n = (1 << 23) + 777
    s = synth.random_walk(n, 21)
    bad = [3_000_123, 6_500_000]
    s[bad[0]] = np.nan
    s[bad[1]] = np.inf
    L = n - m + 1
    avg, sd = engine.avg_std(dev(s), m, STATS_FP64)
    a, d = host(avg), host(sd)
    hit = np.zeros(L, bool)
    for p in bad:
        hit[max(0, p - m + 1):p + 1] = True
    assert np.all(np.isnan(a[hit])) and np.all(np.isnan(d[hit]))
    assert np.all(np.isfinite(a[~hit])) and np.all(np.isfinite(d[~hit]))
    wins = np.unique(np.concatenate([np.arange(0, L, 997), np.arange(L - 3000, L),
                                     np.arange(bad[0] - m - 40, bad[0] - m + 1),
                                     np.arange(bad[1] + 1, bad[1] + 41)]))
    wins = wins[~hit[wins]]
    mu, sg = _two_pass(s, wins, m)
    scale = np.abs(s[np.isfinite(s)]).max()
    assert np.max(np.abs(a[wins] - mu)) < 1e-6 * scale
    rel = np.abs(d[wins] - sg) / sg
    assert rel.max() < 2e-7, rel.max()


def test_stats_sweep_fp64_stream(engine):
    # The FP64-stream kernel holds two workgroups per CU, so 2^22 windows
    # are already several tiles per wave; tolerances as for the FP64 ED path.
    m = 700
    n = (1 << 22) + 9
    s = synth.metal_like(n, 5, offset=7000.0).astype(np.float64)
    s += 1e-3 * np.random.default_rng(5).standard_normal(n)
    avg, sd = engine.avg_std_f64(dev(s, torch.float64), m)
    a, d = host(avg), host(sd)
    L = n - m + 1
    wins = np.unique(np.concatenate([np.arange(0, L, 1009), np.arange(L - 2000, L)]))
    mu, sg = _two_pass(s, wins, m)
    np.testing.assert_allclose(a[wins], mu, rtol=1e-12)
    np.testing.assert_allclose(d[wins], sg, rtol=1e-9)


def test_stats_short_streams_and_constant_segments(engine):
    for n, m in ((11, 11), (12, 11), (300, 100)):
        s = synth.random_walk(n, 2)
        _, sd = engine.avg_std(dev(s), m)
        _, od = O.stats_fp64(s, m)
        assert np.max(np.abs(host(sd) - od) / od) < 2e-7
    s = np.concatenate([np.full(5000, 3.25, np.float32), synth.random_walk(5000, 4) + 1000,
                        np.zeros(3000, np.float32)])
    avg, sd = engine.avg_std(dev(s), 256)
    a, d = host(avg), host(sd)
    assert np.all(d[:5000 - 255] == 1.0) and np.all(a[:5000 - 255] == 3.25)
    assert np.all(d[-2745:] == 1.0) and np.all(a[-2745:] == 0.0)


def test_stats_refcompat_exact_data_every_window(engine):
    # Samples on a 1/16 grid in [-256, 256): every FP64 prefix of x and of x^2
    # is exact (at most 36 and 48 significant bits), so the summation order --
    # the reference's CUB scan, the restatement's sequential loop, this
    # kernel's blocked scan -- cannot matter, and the whole restated arithmetic
    # must agree bit for bit: the epoch plan with a tail merged into the last
    # epoch (stats.cuh:120-141), the first sample each epoch's in-place scan
    # drops (cub_util.cuh:35-36), the FP64 difference / M cast to FP32 and the
    # fmaf finalize (stats.cuh:30-52).
    m = 128
    width = 10 ** 7 - m + 1
    n = 2 * width + m - 3  # third chunk of M - 3 samples merges into the second
    s = (np.random.default_rng(5).integers(-4096, 4096, size=n) / 16.0).astype(np.float32)
    avg, sd = engine.avg_std(dev(s), m, STATS_REFCOMPAT)
    a, d = host(avg), host(sd)
    oa, od = O.stats_refcompat(s, m)
    assert np.array_equal(a, oa) and np.array_equal(d, od)
    # the two epoch starts leave their first sample out
    for lo in (0, width):
        w = s[lo:lo + m].astype(np.float64)
        assert a[lo] == np.float32((w.sum() - w[0]) / m) != np.float32(w.sum() / m)


def test_stats_refcompat_matches_restatement(engine):
    # General data: the FP64 prefixes are summed in a different order here
    # (blocked) than in the restatement (sequential); the reference's CUB order
    # is unknown, so no order is "the" reference's.  A window's mean or mean
    # square may then round to a different FP32 value only when its FP64
    # quotient lies within the two orders' difference of an FP32 rounding
    # boundary.  That difference is bounded here by 64 roundings at the
    # magnitude of the window's two prefixes, divided by M (the two prefixes
    # of a window share all but M terms of their error in the sequential order,
    # and the blocked scan adds a handful of roundings); windows within that
    # margin are named and excused (0.9% here, nearly all windows whose mean
    # is close to 0, a small difference of two large prefixes), and every other
    # window must be bit-identical, sigma included.
    s, q, _ = synth.config1()
    m = 128
    avg, sd = engine.avg_std(dev(s), m, STATS_REFCOMPAT)
    a, d = host(avg), host(sd)
    oa, od = O.stats_refcompat(s, m)
    x = s.astype(np.float64)
    x[0] = 0.0  # the epoch's first sample (one epoch: 2^20 < 10^7 samples)
    p1 = np.concatenate(([0.0], np.cumsum(x)))
    p2 = np.concatenate(([0.0], np.cumsum(x * x)))
    excused = np.zeros(a.size, bool)
    for p, o in ((p1, oa), (p2, None)):
        quot = (p[m:] - p[:-m]) / m
        f = quot.astype(np.float32)
        if o is not None:
            assert np.array_equal(f, o)  # the restatement's own quotients
        half = np.spacing(np.abs(f)).astype(np.float64) / 2
        tau = 64 * np.spacing(np.maximum(np.abs(p[m:]), np.abs(p[:-m]))) / m
        near = np.minimum(np.abs(quot - (f.astype(np.float64) - half)),
                          np.abs(quot - (f.astype(np.float64) + half))) < tau
        excused |= near
    assert excused.mean() < 0.01, excused.mean()
    keep = ~excused
    assert np.array_equal(a[keep], oa[keep]), np.flatnonzero((a != oa) & keep)[:10]
    assert np.array_equal(d[keep], od[keep]), np.flatnonzero((d != od) & keep)[:10]
    # excused windows: one FP32 step at most in the mean
    assert np.all(np.abs(a[excused] - oa[excused]) <= np.spacing(np.abs(oa[excused])))


# ---- cDTW profile: bit-exact with the restated sequential DP ------------------------
@pytest.mark.parametrize("m,w", [(128, 13), (128, 0), (128, 1), (128, 26), (256, 51), (100, 29),
                                 (100, 30), (256, 59), (256, 60), (256, 89), (256, 90),
                                 (256, 112), (37, 5), (11, 10), (5, 2), (3, 1), (2, 1),
                                 (300, 113), (300, 114), (300, 115), (300, 120), (400, 136),
                                 (400, 137), (300, 144), (400, 145), (400, 177), (500, 208),
                                 (500, 209), (600, 240), (700, 241), (401, 137), (333, 150),
                                 (501, 201), (601, 240), (151, 160), (3, 150), (333, 215),
                                 (334, 216), (335, 232), (337, 233), (3, 233), (2, 200),
                                 (300, 121), (300, 122), (4, 130), (500, 225), (500, 226)])
def test_cdtw_bit_exact_random_order(engine, m, w):
    n = 6000
    s = synth.random_walk(n, m + w)
    q = synth.random_walk(m, 99)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, q, w)
    L = n - m + 1
    idx = np.random.default_rng(0).permutation(L).astype(np.int32)
    got = host(engine.gpu_cdtw(ds, avg, sd, w, indices=dev(idx, torch.int32)))
    ref = O.cdtw(zq, s, ha, hs, w, idx=idx, fused=1)
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))


@pytest.mark.parametrize("metric,w,m", [(METRIC_L2, 160, 300), (METRIC_L1, 205, 250),
                                        (METRIC_L2, 240, 260)])
def test_cdtw_split_band_many_passes(engine, metric, w, m):
    # every window of a 2^20-sample stream (several grid passes of the split
    # kernel), checked against the restated DP on 600 sampled windows
    n = 1 << 20
    s = synth.ecg_like(n, w + 5)
    q, _ = synth.cut_query(s, m, 3)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, q, w)
    got = host(engine.gpu_cdtw(ds, avg, sd, w, metric=metric))
    L = n - m + 1
    pick = np.sort(np.random.default_rng(5).choice(L, 600, replace=False)).astype(np.int32)
    ref = O.cdtw(zq, s, ha, hs, w, idx=pick, metric=int(metric), fused=1)
    assert np.array_equal(got[pick], ref), np.max(np.abs(got[pick] - ref))


@pytest.mark.parametrize("w", [0, 13, 62, 63, 64, 100, 113, 127, 128, 200, 300])
def test_cdtw_wide_kernel_bit_exact(engine, w):
    m, n = 512, 3000
    s = synth.ecg_like(n, w)
    q, _ = synth.cut_query(s, m, 1)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, q, w)
    got = host(engine.gpu_cdtw(ds, avg, sd, w, wide=True))
    ref = O.cdtw(zq, s, ha, hs, w, fused=1)
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    if w > 112:  # the regular entry point: register-band (W <= 240) or this kernel
        got2 = host(engine.gpu_cdtw(ds, avg, sd, w))
        assert np.array_equal(got2, ref)


@pytest.mark.parametrize("m,w", [(300, 112), (300, 113), (300, 115), (400, 136), (400, 137),
                                 (500, 200), (600, 240), (700, 241), (401, 151), (601, 239), (300, 129),
                                 (300, 130), (401, 131), (402, 140), (403, 200), (300, 121),
                                 (300, 122), (500, 226)])
def test_cdtw_l1_wide_bands_bit_exact(engine, m, w):
    # L1 bands above 112: one lane per window to W = 121 (LDS tail from 115),
    # the band split over two waves from 122 to 240, the anti-diagonal kernel
    # beyond; both against the restated DP.
    n = 4000
    s = synth.random_walk(n, m + w + 7)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, synth.random_walk(m, 98), w)
    L = n - m + 1
    idx = np.random.default_rng(1).permutation(L).astype(np.int32)
    got = host(engine.gpu_cdtw(ds, avg, sd, w, indices=dev(idx, torch.int32), metric=METRIC_L1))
    ref = O.cdtw(zq, s, ha, hs, w, idx=idx, metric=0, fused=1)
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    wide = host(engine.gpu_cdtw(ds, avg, sd, w, metric=METRIC_L1, wide=True))
    assert np.array_equal(wide[idx], ref)


@pytest.mark.parametrize("wide", [False, True])
def test_cdtw_l1_bit_exact(engine, wide):
    m, w, n = 64, 6, 4000
    s = synth.random_walk(n, 5)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, synth.random_walk(m, 6), w)
    got = host(engine.gpu_cdtw(ds, avg, sd, w, metric=METRIC_L1, wide=wide))
    ref = O.cdtw(zq, s, ha, hs, w, metric=0, fused=1)
    assert np.array_equal(got, ref)


def test_cdtw_vs_fp64_error_stated(engine):
    # Error of the FP32 DP against an FP64 DP (stats and DP in double).
    m, w, n = 128, 13, 20000
    s = synth.random_walk(n, 8)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, synth.random_walk(m, 9), w)
    got = host(engine.gpu_cdtw(ds, avg, sd, w)).astype(np.float64)
    _, _, a64, s64 = O.stats_fp64(s, m, want64=True)
    q64 = O.znormalize_f64(synth.random_walk(m, 9).astype(np.float64))
    ref = O.cdtw_f64(q64, s, a64, s64, w)
    rel = np.abs(got - ref) / ref
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4


# ---- bounds (tests/test_dtwbounds.cu) -----------------------------------------------------
@pytest.mark.parametrize("metric", [METRIC_L2, METRIC_L1])
def test_bounds_bit_exact_and_below_cdtw(engine, metric):
    m, w, n = 128, 13, 30000
    s = synth.ecg_like(n, 3)
    q = synth.ecg_query(m, 2)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, q, w)
    L = n - m + 1
    lo, up = O.envelope(zq, w)
    kim = host(engine.lb_kim(ds, avg, sd, metric=metric))
    keo = host(engine.lb_keogh(ds, avg, sd, metric=metric))
    assert np.array_equal(kim, O.lb_kim(s, ha, hs, zq, metric=metric))
    assert np.array_equal(keo, O.lb_keogh(s, ha, hs, lo, up, metric=metric))
    idx = np.random.default_rng(1).permutation(L).astype(np.int32)
    kim_g = host(engine.lb_kim(ds, avg, sd, indices=dev(idx, torch.int32), metric=metric))
    assert np.array_equal(kim_g, kim[idx])
    d = host(engine.gpu_cdtw(ds, avg, sd, w, metric=metric))
    assert np.all(kim <= d + 1e-4)
    assert np.all(keo <= d + 1e-4)


# ---- LDTW -------------------------------------------------------------------------------------
@pytest.mark.parametrize("m,w,metric", [(64, 6, 1), (128, 13, 1), (100, 10, 0), (11, 3, 1)])
def test_ldtw_bit_exact(engine, m, w, metric):
    n = 8000
    s = synth.random_walk(n, m)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, synth.random_walk(m, 1), w)
    got = host(engine.ldtw(ds, avg, sd, w, metric=metric))
    ref = O.ldtw(zq, s, ha, hs, w, metric=metric)
    assert np.array_equal(got, ref)
    d = host(engine.gpu_cdtw(ds, avg, sd, w, metric=metric))
    assert np.all(got >= d - 1e-4 * d)


# ---- ED -----------------------------------------------------------------------------------------
@pytest.mark.parametrize("m", [16, 128, 256, 1000])
def test_ed_vs_fp64(engine, m):
    n = 1 << 15
    s = synth.ecg_like(n, m)
    q = synth.ecg_query(m, 2) if m >= 200 else synth.random_walk(m, 4)
    zq, ds, avg, sd, ha, hs = prepare(engine, s, q, 0)
    got = host(engine.ed(ds, avg, sd)).astype(np.float64)
    _, _, a64, s64 = O.stats_fp64(s, m, want64=True)
    ref = O.ed_f64(zq.astype(np.float64), s, a64, s64)
    assert np.max(np.abs(got - ref) / np.maximum(ref, 1.0)) < 1e-5
    assert int(np.argmin(got)) == int(np.argmin(ref))


@pytest.mark.parametrize("n,m,kind", [(300, 256, "ecg"), (1000, 256, "ecg"), (1 << 15, 16, "ecg"),
                                        (40_000, 1000, "ecg"), (7777, 333, "ecg"),
                                        (70_000, 2048, "ecg"), (50_000, 300, "metal7000"),
                                        (20_000, 128, "walk"), (5_003, 12, "walk"),
                                        (4_099, 3, "walk")])
def test_ed_exact_over_its_statistics(engine, n, m, kind):
    # The ED kernels sum in FP64 (centred query, group-shifted samples), so a
    # window's distance is the exact direct form over the FP32 statistics it
    # is given, rounded once to FP32: within one FP32 rounding of the oracle's
    # FP64 direct sum over those same statistics (and within 1e-5 of the sum
    # over FP64 statistics, test_ed_vs_fp64).  Windows are grouped by their
    # index relative to the stream pointer, so a stream one float off 16-byte
    # alignment gives the same bits.
    s = (synth.ecg_like(n, m) if kind == "ecg" else synth.random_walk(n, m) if kind == "walk"
         else synth.metal_like(n, m, offset=7000.0))
    q = synth.ecg_query(m, 2) if kind == "ecg" and m >= 200 else synth.cut_query(s, m, 4)[0]
    # (centred first: the FP32 query z-normalisation cancels at a level of 7000)
    zq = O.znormalize_f32((q - q.astype(np.float64).mean()).astype(np.float32))
    engine.set_query(zq, 0)
    ds = dev(s)
    padded = dev(np.concatenate([[0.0], s]).astype(np.float32))
    ds_off = padded[1:]
    assert ds.data_ptr() % 16 == 0 and ds_off.data_ptr() % 16 == 4
    avg, sd = engine.avg_std(ds, m)
    a = host(engine.ed(ds, avg, sd))
    b = host(engine.ed(ds_off, avg, sd))
    assert np.array_equal(a, b)
    ref = O.ed_f64(zq.astype(np.float64), s, host(avg).astype(np.float64),
                   host(sd).astype(np.float64))
    assert np.all(np.abs(a - ref) <= 1.2e-7 * ref + 1e-12 * m)


@pytest.mark.parametrize("m,stats", [(256, "own"), (600, "own"), (1000, "own"), (2048, "own"),
                                     (600, "given"), (777, "raw")])
def test_ed_search_kernels_agree_with_profile(engine, m, stats):
    # Searches evaluate windows with other kernels than the profile: strided
    # seeds one window per lane, and for queries of 512 or more over the
    # search's own statistics an FP32 screen followed by the exact evaluation
    # of the windows it keeps.  The 1024 best windows (no exclusion zone) and
    # the best match must carry the profile's bits, also with non-finite
    # samples planted inside the groups of the best windows (their
    # neighbours then sum their own samples).
    n = 60_000
    s = synth.ecg_like(n, m)
    q, pos = synth.cut_query(s, m, 3, noise=0.02)
    s = s.copy()
    for off in (-9, 5 + m, 3 * m + 7):  # near the cut, outside its window
        if 0 <= pos + off < n:
            s[pos + off] = np.nan if off < 0 else np.inf
    ds = dev(s)
    if stats == "raw":
        zq = q.astype(np.float32)
        engine.set_query(zq, 0)
        avg, sd = engine.avg_std(ds, m, STATS_NONE)
        kw = dict(stats_mode=STATS_NONE)
    else:
        zq = O.znormalize_f32(q)
        engine.set_query(zq, 0)
        avg, sd = engine.avg_std(ds, m)
        kw = dict(avg=avg, sd=sd) if stats == "given" else {}
    prof = host(engine.ed(ds, avg, sd))
    assert np.isnan(prof[max(0, pos - 9 - m + 1):pos - 9 + 1]).all()
    for k, excl in ((1024, 0), (1, 0), (8, m // 2)):
        r = engine.search(ds, 0, algo=ALGO_ED, k=k, excl=excl, **kw)
        idx, d2 = O.topk_greedy(prof, k, excl)
        np.testing.assert_array_equal(r.positions[:idx.size], idx)
        np.testing.assert_array_equal(r.distances[:idx.size], np.sqrt(d2).astype(np.float32))


def test_ed_topk_rounds_split_inside_a_group(engine):
    # Top-k searches run a first round over L/16 windows and a second over the
    # rest; with L = 1,048,600 the second starts at window 65,537, inside a
    # 16-window group, so both rounds evaluate that group's windows (each
    # reporting its own).  The kept windows must still carry the profile's bits.
    m = 256
    L = 1_048_600
    s = synth.ecg_like(L + m - 1, 12)
    q, _ = synth.cut_query(s, m, 13, noise=0.05)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, 0)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, m)
    prof = host(engine.ed(ds, avg, sd))
    for k, excl in ((8, m // 2), (300, 0)):
        r = engine.search(ds, 0, algo=ALGO_ED, k=k, excl=excl)
        idx, d2 = O.topk_greedy(prof, k, excl)
        np.testing.assert_array_equal(r.positions[:idx.size], idx)
        np.testing.assert_array_equal(r.distances[:idx.size], np.sqrt(d2).astype(np.float32))


# ---- primitives (tests/test_threshold.cu) -------------------------------------------------------
def test_threshold_compaction(engine):
    vals = np.random.default_rng(2).permutation(100).astype(np.float32)
    idx = np.arange(100, dtype=np.int32)
    dv, di = dev(vals), dev(idx, torch.int32)
    n = engine.threshold(dv, di, 22.0)
    kept = host(di)[:n]
    assert n == 23
    assert np.array_equal(kept, np.nonzero(vals <= 22)[0].astype(np.int32))
    big = np.random.default_rng(3).standard_normal(1 << 20).astype(np.float32)
    bi = np.arange(big.size, dtype=np.int32)
    dbi = dev(bi, torch.int32)
    n2 = engine.threshold(dev(big), dbi, 0.5)
    assert np.array_equal(host(dbi)[:n2], np.nonzero(big <= 0.5)[0].astype(np.int32))


def test_argmin_lowest_index(engine):
    v = np.full(1 << 20, 5.0, np.float32)
    v[[7, 100, 500000]] = 1.0
    i, x = engine.argmin(dev(v))
    assert (i, x) == (7, 1.0)
    v2 = np.random.default_rng(4).standard_normal(12345).astype(np.float32)
    i2, x2 = engine.argmin(dev(v2))
    assert i2 == int(np.argmin(v2)) and x2 == v2.min()


# ---- error paths -------------------------------------------------------------------------------
def test_error_codes(engine):
    from cudadtw_amd import Engine
    from cudadtw_amd._lib import MdtwError

    s = dev(synth.random_walk(1000, 1))
    engine.set_query(O.znormalize_f32(synth.random_walk(8, 2)), 2)
    avg, sd = engine.avg_std(s, 8)
    with pytest.raises(MdtwError) as ei:
        engine.lb_kim(s, avg, sd)
    assert ei.value.status == _lib.ERR_QUERY_TOO_SHORT
    with pytest.raises(MdtwError) as ei:
        engine.set_query(np.array([0.0, np.nan, 1.0], np.float32), 1)
    assert ei.value.status == _lib.ERR_NOT_FINITE
    engine.set_query(O.znormalize_f32(synth.random_walk(64, 2)), 6)
    with pytest.raises(MdtwError) as ei:
        engine.search(dev(synth.random_walk(32, 1)), 6)
    assert ei.value.status == _lib.ERR_STREAM_TOO_SHORT
    with pytest.raises(MdtwError) as ei:
        engine.search(s, 6, k=0)
    assert ei.value.status == _lib.ERR_K
    with pytest.raises(MdtwError) as ei:
        engine.search(s, -1)
    assert ei.value.status == _lib.ERR_BAND
    fresh = Engine(0)
    with pytest.raises(MdtwError) as ei:
        fresh.search(s, 6)
    assert ei.value.status == _lib.ERR_NO_QUERY
    fresh.close()
