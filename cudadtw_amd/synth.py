"""Deterministic synthetic streams and queries for the benchmark workloads.

There is no dataset access; these stand in for the reference's ECG and metal
recordings (data/ecg, data/metal) with the shapes the workloads name.  Every
generator is seeded (numpy PCG64 ``default_rng``), computed in float64 and
cast to float32 once, so a (seed, length) pair always gives the same array.

  random_walk        cumulative sum of N(0, 1) increments
  ecg_like           Gaussian-bump beats (P, Q, R, S, T) on a period of
                     250 +- U{0..15} samples, amplitude 1 +- U(0.1),
                     N(0, 0.05) noise and a 25,000-sample baseline wander
  metal_like         piecewise-constant levels (geometric segment lengths,
                     mean 2000; N(0, 1) level values) with a linear drift and
                     N(0, 0.02) noise; ``offset`` lifts the whole stream (the
                     real metal data sits near 7000)
  ecg_query          one noisy beat window cut from an independent stream
  cut_query          a window cut from a stream and perturbed with noise
"""
from __future__ import annotations

import numpy as np

# Beat template: (amplitude, centre, width) of the P, Q, R, S, T bumps in
# samples from the beat onset.
_BUMPS = ((0.12, 50.0, 8.0), (-0.10, 88.0, 3.0), (1.00, 95.0, 3.0), (-0.20, 102.0, 3.0),
          (0.30, 160.0, 14.0))
_PERIOD = 250
_JITTER = 15
_MAXP = _PERIOD + _JITTER + 1


def _template() -> np.ndarray:
    t = np.arange(_MAXP, dtype=np.float64)
    out = np.zeros_like(t)
    for a, c, w in _BUMPS:
        out += a * np.exp(-0.5 * ((t - c) / w) ** 2)
    return out


def random_walk(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.standard_normal(n)).astype(np.float32)


def ecg_like(n: int, seed: int, chunk: int = 1 << 24) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nbeats = n // (_PERIOD - _JITTER) + 2
    periods = _PERIOD + rng.integers(-_JITTER, _JITTER + 1, size=nbeats)
    amps = 1.0 + rng.uniform(-0.1, 0.1, size=nbeats)
    onsets = np.concatenate(([0], np.cumsum(periods)[:-1]))
    tmpl = _template()
    out = np.empty(n, dtype=np.float32)
    phase = rng.uniform(0, 2 * np.pi)
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        idx = np.arange(lo, hi, dtype=np.int64)
        b = np.searchsorted(onsets, idx, side="right") - 1
        off = idx - onsets[b]
        x = amps[b] * tmpl[np.minimum(off, _MAXP - 1)]
        x += 0.05 * rng.standard_normal(hi - lo)
        x += 0.2 * np.sin(2 * np.pi * idx / 25000.0 + phase)
        out[lo:hi] = x.astype(np.float32)
    return out


def metal_like(n: int, seed: int, offset: float = 0.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nseg = max(4, int(n / 2000 * 2) + 4)
    lens = rng.geometric(1.0 / 2000.0, size=nseg)
    while lens.sum() < n:
        lens = np.concatenate((lens, rng.geometric(1.0 / 2000.0, size=nseg)))
    levels = rng.standard_normal(lens.size)
    x = np.repeat(levels, lens)[:n].astype(np.float64)
    x += np.linspace(0.0, 2.0, n)
    x += 0.02 * rng.standard_normal(n)
    x += offset
    return x.astype(np.float32)


def ecg_query(m: int = 256, seed: int = 2) -> np.ndarray:
    """A window of m samples starting at a beat onset of an independent
    ECG-like stream (one noisy beat)."""
    s = ecg_like(max(4 * m, 4 * _MAXP), seed)
    rng = np.random.default_rng(seed + 1000)
    start = int(rng.integers(0, _PERIOD))
    # align to the nearest beat onset: the R peak sits 95 samples in
    r = int(np.argmax(s[start:start + _MAXP])) + start
    begin = max(0, r - 95)
    return s[begin:begin + m].copy()


def cut_query(stream: np.ndarray, m: int, seed: int, noise: float = 0.05):
    """(query, position): a window of the stream perturbed with N(0, noise)."""
    rng = np.random.default_rng(seed)
    pos = int(rng.integers(0, stream.size - m + 1))
    q = stream[pos:pos + m].astype(np.float64) + noise * rng.standard_normal(m)
    return q.astype(np.float32), pos


def config1():
    """Random walk stream of 2^20 samples (seed 0), query a 128-step
    independent walk (seed 1), band 10% -> W = 13."""
    return random_walk(1 << 20, 0), random_walk(128, 1), 13


def config2_stream(log2n: int = 27, seed: int = 0):
    return ecg_like(1 << log2n, seed)


def batch_queries(stream: np.ndarray, nq: int = 1024, seed: int = 4, mmin: int = 64,
                  mmax: int = 1024, pmin: float = 5.0, pmax: float = 20.0):
    """Mixed-shape query set: lengths uniform in [mmin, mmax], band percent
    uniform in [pmin, pmax] (W = round(p * M / 100)); even-numbered queries
    are stream windows perturbed with N(0, 0.05), odd ones independent
    random walks.  Returns (raw queries, bands, kinds)."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(mmin, mmax + 1, size=nq)
    pct = rng.uniform(pmin, pmax, size=nq)
    qs, bands, kinds = [], [], []
    for i in range(nq):
        m = int(lens[i])
        if i % 2 == 0:
            q, _ = cut_query(stream, m, seed * 100003 + i)
            kinds.append("cut")
        else:
            q = random_walk(m, seed * 100003 + i)
            kinds.append("walk")
        qs.append(q)
        bands.append(int(np.floor(m * pct[i] / 100.0 + 0.5)))
    return qs, np.array(bands, np.int32), kinds


def config4(log2n: int = 26, seed: int = 4):
    """1024 queries, |Q| uniform in [64, 1024], bands 5-20 %, half cut from
    and perturbed against the stream, half independent walks; one ECG-like
    stream of 2^26 samples."""
    s = ecg_like(1 << log2n, seed)
    qs, bands, kinds = batch_queries(s, 1024, seed)
    return s, qs, bands, kinds


def config5(log2n: int = 24, seed: int = 5, nq: int = 256):
    """256 queries, |Q| uniform in [128, 512], band 10 %; ECG-like stream of
    2^24 samples (half cut, half walks)."""
    s = ecg_like(1 << log2n, seed)
    qs, bands, kinds = batch_queries(s, nq, seed, 128, 512, 10.0, 10.0)
    return s, qs, bands, kinds
