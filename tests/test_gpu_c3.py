"""The metal-sensor workload's shapes against the oracle.

Metal-like stream (piecewise-constant levels, drift, N(0, 0.02) noise),
queries of |Q| = 512 and 1024 cut from it and perturbed, W = 10% (51 / 102),
top-16 with an exclusion zone of |Q| / 2, cDTW and greedy LDTW.

- LDTW profiles are bit-exact with the oracle's restatement of
  z_serial_ldtw (ldtw.cuh:74-127) at these lengths, for both metrics, with
  the query staged in LDS (M <= 2048) and read from global memory beyond
  (M = 2100); the oracle receives the GPU's window statistics, as the
  reference's own tests give the CPU the GPU's statistics
  (tests/test_dtw.cu:84-86).
- The top-16 searches equal the greedy exclusion-zone selection over the
  oracle's full profiles (O.cdtw = sequential_cdtw, cdtw.cuh:23-88; O.ldtw),
  positions and distance bits, on the level-0 stream and on the
  offset-7000 stream of the real recordings.
"""
import numpy as np
import pytest
import torch

from cudadtw_amd import ALGO_CDTW, ALGO_LDTW, METRIC_L1, METRIC_L2, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def metal_case(n, m, seed, offset=0.0):
    s = synth.metal_like(n, seed, offset=offset)
    q, _ = synth.cut_query(s, m, seed + 1)
    # The reference's FP32 znormalize (stats.cuh:8-28, sqrt(E[x^2] - mean^2)
    # in single precision) cancels to NaN on a query near 7000; such a caller
    # z-normalises in FP64 (the same function with value_t = double).
    zq = O.znormalize_f32(q) if offset == 0.0 else \
        O.znormalize_f64(q.astype(np.float64)).astype(np.float32)
    return s, zq


@pytest.mark.parametrize("m,w,metric", [(512, 51, METRIC_L2), (512, 51, METRIC_L1),
                                        (1024, 102, METRIC_L2), (1024, 102, METRIC_L1),
                                        (2048, 204, METRIC_L2), (2100, 210, METRIC_L2),
                                        (2100, 210, METRIC_L1)])
def test_ldtw_profile_bit_exact_c3_lengths(engine, m, w, metric):
    n = m + 12000
    s, zq = metal_case(n, m, 100 + m + metric)
    engine.set_query(zq, w)
    ds = dev(s)
    avg, sd = engine.avg_std(ds, m)
    got = host(engine.ldtw(ds, avg, sd, w, metric=metric))
    ref = O.ldtw(zq, s, host(avg), host(sd), w, metric=metric)
    assert np.array_equal(got, ref), (np.flatnonzero(got != ref)[:5], np.max(np.abs(got - ref)))


@pytest.mark.parametrize("m,w", [(512, 51), (1024, 102)])
@pytest.mark.parametrize("algo", [ALGO_CDTW, ALGO_LDTW])
@pytest.mark.parametrize("offset", [0.0, 7000.0])
def test_top16_equals_greedy_over_oracle_profile(engine, m, w, algo, offset):
    n = 1 << 16 if m <= 512 else 1 << 15  # the oracle's DP: ~1e10 cells at most
    s, zq = metal_case(n, m, 7 + m, offset)
    k, excl = 16, m // 2
    engine.set_query(zq, w)
    ds = dev(s)
    r = engine.search(ds, w, algo=algo, k=k, excl=excl)
    avg, sd = engine.avg_std(ds, m)
    ha, hs = host(avg), host(sd)
    prof = O.cdtw(zq, s, ha, hs, w, fused=1) if algo == ALGO_CDTW else O.ldtw(zq, s, ha, hs, w)
    idx, d2 = O.topk_greedy(prof, k, excl)
    assert idx.size == k
    np.testing.assert_array_equal(r.positions, idx)
    np.testing.assert_array_equal(r.distances, np.sqrt(d2).astype(np.float32))
    assert r.info["found"] == k
    assert r.info["windows"] == n - m + 1
