"""Batched search (mdtw_batch_*) against per-query plain searches and the
oracle.

No reference counterpart (the reference runs one query per process).  The
bar: every query's batched result -- positions and distances -- is
identical to its own mdtw_search (the same kernels and bounds, so the same
bits), which is itself checked against the oracle's full profile
(test_gpu_search.py); graph replays and tuned launch shapes change nothing.
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import METRIC_L1, METRIC_L2, band_width, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def make_queries(stream, specs, seed):
    """specs: (M, band percent, kind); kind 'cut' = a perturbed stream window,
    'walk' = an independent random walk."""
    rng = np.random.default_rng(seed)
    qs, bands = [], []
    for i, (m, p, kind) in enumerate(specs):
        if kind == "cut":
            q, _ = synth.cut_query(stream, m, seed + i)
        else:
            q = synth.random_walk(m, seed + 100 + i)
        qs.append(O.znormalize_f32(q.astype(np.float32)))
        bands.append(band_width(m, p))
    return qs, np.array(bands, np.int32)


SPECS = [(8, 20, "walk"), (16, 5, "cut"), (64, 10, "walk"), (100, 20, "cut"), (128, 10, "cut"),
         (200, 5, "walk"), (256, 10, "cut"), (300, 40, "walk"), (333, 35, "cut"), (97, 0, "cut"),
         (512, 10, "walk"), (40, 100, "cut")]


@pytest.mark.parametrize("k", [1, 4])
@pytest.mark.parametrize("metric", [METRIC_L2, METRIC_L1])
def test_batch_equals_plain_search(engine, k, metric):
    s = synth.ecg_like(1 << 16, 5)
    ds = dev(s)
    qs, bands = make_queries(s, SPECS, 17)
    assert any(b > 112 for b in bands)  # the anti-diagonal kernel is covered too
    b = engine.batch(qs, bands, pipelines=3)
    r = b.search(ds, k=k, metric=metric)
    for i, (q, w) in enumerate(zip(qs, bands)):
        engine.set_query(q, int(w))
        p = engine.search(ds, int(w), metric=metric, k=k, excl=q.size // 2)
        np.testing.assert_array_equal(r.positions[i], p.positions, err_msg=f"query {i}")
        np.testing.assert_array_equal(r.distances[i], p.distances, err_msg=f"query {i}")
        assert r.infos[i]["found"] == int(np.sum(p.positions >= 0))
    r2 = b.search(ds, k=k, metric=metric)  # graph replay
    np.testing.assert_array_equal(r2.positions, r.positions)
    np.testing.assert_array_equal(r2.distances, r.distances)
    b.close()


def test_batch_against_oracle_profile(engine):
    s = synth.random_walk(1 << 13, 3)
    ds = dev(s)
    qs, bands = make_queries(s, [(24, 10, "cut"), (50, 20, "walk"), (77, 5, "cut")], 9)
    b = engine.batch(qs, bands)
    r = b.search(ds, k=3, excl=[12, 25, 38])
    for i, (q, w) in enumerate(zip(qs, bands)):
        # the oracle's DP over the GPU's own window statistics: bit-exact
        avg, sd = engine.avg_std(ds, q.size)
        torch.cuda.synchronize()
        prof = O.cdtw(q, s, avg.cpu().numpy(), sd.cpu().numpy(), int(w))
        idx, dist = O.topk_greedy(prof, 3, [12, 25, 38][i])
        np.testing.assert_array_equal(r.positions[i, :idx.size], idx)
        np.testing.assert_array_equal(r.distances[i, :idx.size], np.sqrt(dist))
    b.close()


def test_batch_tuned_shapes_keep_results(engine):
    s = synth.ecg_like(1 << 15, 8)
    ds = dev(s)
    qs, bands = make_queries(s, SPECS[1:9], 23)
    b = engine.batch(qs, bands, pipelines=2)
    r0 = b.search(ds, k=2)
    nb = b.tune(ds, k=2, cell_budget=1 << 30)
    assert nb >= 3
    r1 = b.search(ds, k=2)
    np.testing.assert_array_equal(r1.positions, r0.positions)
    np.testing.assert_array_equal(r1.distances, r0.distances)
    b.close()


def test_batch_every_dp_form_gives_the_same_results(engine):
    # The tuner may run a bucket's DP as the register band with or without
    # the cumulative LB_Keogh bound, or as the anti-diagonal kernel; every
    # form must return the same positions and distance bits.  Queries cover
    # the register-band, LDS-tail / split and anti-diagonal band classes.
    s = synth.ecg_like(1 << 15, 9)
    ds = dev(s)
    specs = [(96, 10), (128, 25), (300, 60), (320, 118), (400, 130), (600, 200), (700, 260)]
    qs, bands = [], []
    for j, (m, w) in enumerate(specs):
        q, _ = synth.cut_query(s, m, 40 + j, noise=0.1)
        qs.append(O.znormalize_f32(q))
        bands.append(w)
    b = engine.batch(qs, bands, pipelines=2)
    r0 = b.search(ds, k=3)
    for form in ("lb", "wide"):
        for i in range(len(specs)):
            b.set_shape(i, dp_blocks_per_cu=4, wide_waves_per_cu=16, cascade_blocks_per_cu=16,
                        dp_form=form)
            assert b.shape(i)["dp_form"] == form
        r = b.search(ds, k=3)
        np.testing.assert_array_equal(r.positions, r0.positions)
        np.testing.assert_array_equal(r.distances, r0.distances)
    with pytest.raises(Exception):
        b.set_shape(0, dp_blocks_per_cu=0)
    b.close()


def test_batch_argument_errors(engine):
    from cudadtw_amd._lib import MdtwError
    q = O.znormalize_f32(synth.random_walk(32, 1))
    q2 = O.znormalize_f32(synth.random_walk(2048, 1))
    with pytest.raises(MdtwError):
        engine.batch([q2], [1500])         # band beyond the supported range
    with pytest.raises(MdtwError):
        engine.batch([q, q * np.nan], [3, 3])
    b = engine.batch([q], [3])
    with pytest.raises(MdtwError):
        b.search(dev(np.zeros(16, np.float32)))  # stream shorter than the query
    with pytest.raises(MdtwError):
        b.search(dev(synth.random_walk(1000, 2)), k=0)
    b.close()


@pytest.mark.parametrize("raw", [False, True])
def test_batch_backward_equals_per_query(engine, raw):
    s = synth.ecg_like(1 << 15, 12)
    ds = dev(s)
    # (1024, 10): its direction words exceed LDS, so it runs query by query
    # beside the batch kernels; (150, 90) is the 4-slots-per-lane class.
    specs = [(64, 10, "cut"), (100, 20, "walk"), (128, 5, "cut"), (150, 90, "walk"),
             (40, 10, "cut"), (1024, 10, "walk")]
    rng = np.random.default_rng(4)
    raws = []
    qs, bands = [], []
    for i, (m, p, kind) in enumerate(specs):
        q = (synth.cut_query(s, m, 60 + i)[0] if kind == "cut"
             else synth.random_walk(m, 70 + i)).astype(np.float32)
        raws.append(q)
        qs.append(O.znormalize_f32(q))
        bands.append(band_width(m, p))
    bands = np.array(bands, np.int32)
    b = engine.batch(qs, bands, pipelines=3)
    k = 4
    r = b.search(ds, k=k)
    grads = rng.uniform(0.5, 2.0, size=(len(qs), k)).astype(np.float32)
    pos = r.positions.copy()
    pos[1, 2] = -1  # a skipped entry
    pos[4, :] = -1  # a query with no windows
    gb, db = b.backward(ds, pos, grads, raw_queries=raws if raw else None)
    for i, (q, w) in enumerate(zip(qs, bands)):
        engine.set_query(q, int(w))
        avg, sd = engine.avg_std(ds, q.size)
        keep = pos[i] >= 0
        idx = torch.from_numpy(pos[i][keep].astype(np.int32)).cuda()
        g1, d1 = engine.cdtw_backward(ds, avg, sd, idx, dev(grads[i][keep]), int(w),
                                      query_raw=dev(raws[i]) if raw else None, want_dist=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(gb[i], g1.cpu().numpy(), err_msg=f"query {i}")
        np.testing.assert_array_equal(db[i][keep], d1.cpu().numpy())
        assert np.all(np.isinf(db[i][~keep]))
        np.testing.assert_array_equal(db[i][keep], r.distances[i][keep])
    b.close()


# Mixed-batch shapes: lengths 64..1024, bands 5..20 % (three of them W >= 113,
# the register-band kernel's wide range), half cut from the stream and
# perturbed, half independent walks; top-8 with exclusion |Q|/2.
C4_SPECS = [(64, 5, "cut"), (64, 20, "walk"), (96, 10, "cut"), (128, 15, "walk"),
            (160, 20, "cut"), (200, 5, "walk"), (256, 10, "cut"), (300, 12, "walk"),
            (384, 8, "cut"), (448, 6, "walk"), (512, 10, "cut"), (600, 20, "walk"),
            (760, 15, "cut"), (1024, 15, "cut"), (900, 5, "walk"), (1000, 6, "cut")]


def test_batch_mixed_shapes_against_oracle(engine):
    """Every query's top-8 equals the greedy selection over the oracle's DP
    (sequential_cdtw, cdtw.cuh:23-88, over the GPU's window statistics).  The
    oracle evaluates every window whose LB_Keogh (bounds.cuh:111-144, a lower
    bound of the same DP: test_oracle.py) does not exceed the batch's 8th
    distance plus a rounding margin; every other window is farther than all
    eight picks, so the greedy picks over the evaluated windows are the
    picks over the whole profile (checked: all eight fall below the margin)."""
    k = 8
    s = synth.ecg_like(1 << 18, 44)
    ds = dev(s)
    qs, bands = make_queries(s, C4_SPECS, 50)
    assert sum(int(w) >= 113 for w in bands) >= 3
    b = engine.batch(qs, bands)
    r = b.search(ds, k=k)
    b.close()
    for i, (q, w) in enumerate(zip(qs, bands)):
        m, w = q.size, int(w)
        avg, sd = engine.avg_std(ds, m)
        torch.cuda.synchronize()
        ha, hs = avg.cpu().numpy(), sd.cpu().numpy()
        lo, up = O.envelope(q, w)
        lb = O.lb_keogh(s, ha, hs, lo, up)
        t = np.float64(r.distances[i, -1]) ** 2
        limit = t * (1 + 1e-4) + 1e-3 if np.isfinite(t) else np.inf
        cand = np.flatnonzero(lb <= limit).astype(np.int32)
        prof = np.full(s.size - m + 1, np.inf, np.float32)
        prof[cand] = O.cdtw(q, s, ha, hs, w, idx=cand)
        idx, d2 = O.topk_greedy(prof, k, m // 2)
        assert idx.size == k and np.all(d2 <= limit), f"query {i}"
        np.testing.assert_array_equal(r.positions[i], idx, err_msg=f"query {i} (M={m}, W={w})")
        np.testing.assert_array_equal(r.distances[i], np.sqrt(d2).astype(np.float32),
                                      err_msg=f"query {i}")
