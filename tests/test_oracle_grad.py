"""Pins the FP64 cDTW query gradient of the oracle (or_cdtw_grad_f64).

The reference has no gradient; what fixes this one is mathematics: FP64
central differences (eps = 1e-3) of the oracle's own FP64 cDTW
(or_cdtw_f64, restating cdtw.cuh:23-88) through the FP64 z-normalisation,
the W = 0 closed form d/dq_i sum_p g_p (q_i - z_i)^2 = 2 sum_p g_p (q_i - z_i),
and the path cost: the returned distance is or_cdtw_f64's.
"""
import numpy as np
import pytest

from cudadtw_amd import synth
from oracle import oracle as O


def _case(m, seed, n=6000):
    rng = np.random.default_rng(seed)
    s = synth.random_walk(n, seed)
    q = np.cumsum(rng.standard_normal(m))
    _, _, a64, s64 = O.stats_fp64(s, m, want64=True)
    idx = rng.choice(n - m + 1, size=5, replace=False).astype(np.int32)
    g = rng.uniform(-1.0, 2.0, size=idx.size)
    return s, q, a64, s64, idx, g


def _loss(q, s, a64, s64, w, idx, g, metric, sqrt_dist):
    z = O.znormalize_f64(q.copy())
    d = O.cdtw_f64(z, s, a64, s64, w, idx=idx, metric=metric)
    return float(np.sum(g * (np.sqrt(d) if sqrt_dist else d)))


@pytest.mark.parametrize("m,w,sqrt_dist", [(32, 3, 1), (48, 0, 1), (64, 7, 0), (40, 39, 1),
                                           (50, 100, 0)])
def test_grad_matches_central_differences(m, w, sqrt_dist):
    s, q, a64, s64, idx, g = _case(m, m + w)
    zq = O.znormalize_f64(q.copy())
    grad, dist = O.cdtw_grad_f64(zq, s, a64, s64, w, idx, g, sqrt_dist=sqrt_dist, q_raw=q)
    d64 = O.cdtw_f64(zq, s, a64, s64, w, idx=idx)
    np.testing.assert_array_equal(dist, np.sqrt(d64) if sqrt_dist else d64)
    eps = 1e-3
    fd = np.empty(m)
    for i in range(m):
        e = np.zeros(m)
        e[i] = eps
        fd[i] = (_loss(q + e, s, a64, s64, w, idx, g, 1, sqrt_dist)
                 - _loss(q - e, s, a64, s64, w, idx, g, 1, sqrt_dist)) / (2 * eps)
    assert np.linalg.norm(fd - grad) / np.linalg.norm(grad) < 1e-6


def test_grad_w0_closed_form_and_zero_sum():
    m = 40
    s, q, a64, s64, idx, g = _case(m, 3)
    zq = O.znormalize_f64(q.copy())
    grad, _ = O.cdtw_grad_f64(zq, s, a64, s64, 0, idx, g, sqrt_dist=0)
    want = np.zeros(m)
    for gp, p in zip(g, idx):
        z = (s[p:p + m].astype(np.float64) - a64[p]) / s64[p]
        want += gp * 2.0 * (zq - z)
    np.testing.assert_allclose(grad, want, rtol=1e-12, atol=1e-12)
    # through the z-normalisation the gradient is orthogonal to 1 and to zq
    graw, _ = O.cdtw_grad_f64(zq, s, a64, s64, 5, idx, g, sqrt_dist=1, q_raw=q)
    assert abs(graw.sum()) < 1e-9 * np.abs(graw).sum()
    assert abs(graw @ zq) < 1e-9 * np.abs(graw).sum() * np.abs(zq).max()


def test_grad_l1_sign_form():
    m, w = 36, 4
    s, q, a64, s64, idx, g = _case(m, 9)
    zq = O.znormalize_f64(q.copy())
    grad, dist = O.cdtw_grad_f64(zq, s, a64, s64, w, idx, g, metric=0, sqrt_dist=0)
    d64 = O.cdtw_f64(zq, s, a64, s64, w, idx=idx, metric=0)
    np.testing.assert_array_equal(dist, d64)
    eps = 1e-6
    fd = np.empty(m)
    for i in range(m):
        e = np.zeros(m)
        e[i] = eps
        fp = np.sum(g * O.cdtw_f64(zq + e, s, a64, s64, w, idx=idx, metric=0))
        fm = np.sum(g * O.cdtw_f64(zq - e, s, a64, s64, w, idx=idx, metric=0))
        fd[i] = (fp - fm) / (2 * eps)
    np.testing.assert_allclose(grad, fd, atol=1e-6)


def test_grad_path32_distance_is_or_cdtw():
    # The FP32-path mode follows or_cdtw's FP32 DP: its distance is that DP's
    # value bit for bit, and for L2 its gradient is within FP32 rounding of
    # the FP64 one on a generic case.
    m, w = 96, 9
    s, q, a64, s64, idx, g = _case(m, 21)
    avg, sd = O.stats_fp64(s, m)
    zq = O.znormalize_f32(q.astype(np.float32))
    for metric in (0, 1):
        prof = O.cdtw(zq, s, avg, sd, w, idx=idx, metric=metric)
        g32, d32 = O.cdtw_grad_f64(zq.astype(np.float64), s, avg.astype(np.float64),
                                   sd.astype(np.float64), w, idx, g, metric=metric,
                                   sqrt_dist=0, path32=True)
        np.testing.assert_array_equal(d32.astype(np.float32), prof)
    g64, _ = O.cdtw_grad_f64(zq.astype(np.float64), s, avg.astype(np.float64),
                             sd.astype(np.float64), w, idx, g, metric=1)
    g32, _ = O.cdtw_grad_f64(zq.astype(np.float64), s, avg.astype(np.float64),
                             sd.astype(np.float64), w, idx, g, metric=1, path32=True)
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-5
