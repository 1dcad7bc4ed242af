"""FFT ED (mdtw_ed_fft, MDTW_ALGO_ED_FFT) against the FP64 oracle.

The reference computes ED through FP64 cuFFT cross-correlation
(ed.cuh:71-153); cuFFT's rounding is not reproducible here (parity
unpinned), so the bar is the quantity it computes: the FP64 direct sum
(or_ed_f64, pinned against numpy and the UCR_ED golden location) over FP64
two-pass statistics, within 2e-6 relative plus 1e-9 * M absolute (FP64
transform error, then the FP32 store); the searches return the oracle
profile's argmin / greedy top-k.  Streams span several overlap-save blocks
(including a partial last block), carry a level offset, and hold an exactly
constant stretch (sigma := 1 windows).
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import ALGO_ED, ALGO_ED_FFT, STATS_REFCOMPAT, synth
from cudadtw_amd._lib import MdtwError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def stream_for(n, seed, offset=0.0):
    s = synth.random_walk(n, seed).astype(np.float64) * 0.5 + offset
    s[1000:1000 + 9000] = s[1000]  # exactly constant stretch
    return s.astype(np.float32)


def oracle_profile(s, zq):
    _, _, a64, s64 = O.stats_fp64(s, zq.size, want64=True)
    return O.ed_f64(zq.astype(np.float64), s, a64, s64)


@pytest.mark.parametrize("m", [16, 100, 256, 1000, 2048, 3000, 4097, 8192])
@pytest.mark.parametrize("offset", [0.0, 7000.0])
def test_fft_profile_matches_fp64(engine, m, offset):
    n = 3 * (1 << 15) + 777
    s = stream_for(n, 10 + m, offset)
    q, _ = synth.cut_query(s, m, 20 + m, noise=0.3)
    # FP64 z-normalisation: the reference's FP32 form (stats.cuh:21) fails
    # at a 7000 offset (E[x^2] - mean^2 cancels to a negative)
    zq = O.znormalize_f64(q.astype(np.float64)).astype(np.float32)
    engine.set_query(zq, 0)
    got = host(engine.ed_fft(dev(s))).astype(np.float64)
    want = oracle_profile(s, zq)
    err = np.abs(got - want)
    bound = 2e-6 * want + 1e-9 * m
    worst = float(np.max(err / np.maximum(bound, 1e-300)))
    print(f"m={m} offset={offset}: max abs err {err.max():.3e}, worst err/bound {worst:.3f}")
    assert got.size == n - m + 1
    assert np.all(err <= bound)
    # constant windows: sigma := 1, ED^2 = sum q^2
    np.testing.assert_allclose(got[1000:1000 + 9000 - m + 1][:50],
                               np.sum(zq.astype(np.float64) ** 2), rtol=2e-6)


def test_fft_matches_direct_ed_profile(engine):
    s = synth.ecg_like(1 << 16, 3)
    q, _ = synth.cut_query(s, 256, 4)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, 0)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, 256)
    direct = host(engine.ed(ds, avg, sd)).astype(np.float64)
    fft = host(engine.ed_fft(ds)).astype(np.float64)
    want = oracle_profile(s, zq)
    # both within 1e-5 relative of FP64 (direct: FP32 arithmetic)
    assert np.all(np.abs(direct - want) <= 1e-5 * want + 1e-6 * 256)
    assert np.all(np.abs(fft - want) <= 2e-6 * want + 1e-9 * 256)


@pytest.mark.parametrize("m,k", [(512, 1), (2000, 1), (512, 5), (6000, 3)])
def test_fft_search_equals_oracle_selection(engine, m, k):
    n = 1 << 17
    s = synth.ecg_like(n, 30 + m)
    q, pos = synth.cut_query(s, m, 31 + m, noise=0.1)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, 0)
    excl = m // 2
    want = oracle_profile(s, zq).astype(np.float32)
    r = engine.search(dev(s), 0, algo=ALGO_ED_FFT, k=k, excl=excl)
    idx, dist = O.topk_greedy(want, k, excl)
    np.testing.assert_array_equal(r.positions[:idx.size], idx)
    np.testing.assert_allclose(r.distances[:idx.size] ** 2, dist, rtol=1e-5)
    rd = engine.search(dev(s), 0, algo=ALGO_ED, k=k, excl=excl)
    np.testing.assert_array_equal(rd.positions, r.positions)


def test_fft_argument_errors(engine):
    s = dev(synth.random_walk(20000, 1))
    engine.set_query(O.znormalize_f32(synth.random_walk(300, 2)), 0)
    with pytest.raises(MdtwError):
        engine.search(s, 0, algo=ALGO_ED_FFT, stats_mode=STATS_REFCOMPAT)
    avg, sd = engine.avg_std(s, 300)
    with pytest.raises(MdtwError):
        engine.search(s, 0, algo=ALGO_ED_FFT, avg=avg, sd=sd)
    with pytest.raises(MdtwError):
        engine.ed_fft(dev(np.zeros(100, np.float32)))  # shorter than the query
