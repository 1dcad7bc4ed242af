"""The three-wave split of a wide band (cudadtw_amd/csrc/cdtw_pair.hpp,
trio_dp), restated step by step in float64 on the CPU: the band's 2W + 1
cells split over three waves by rows, wave k on column s - k at step s, two
phases per step.  Checks that the schedule reproduces the banded DP
(cdtw.cuh:90-157 in the reference computes the same recurrence) and that
the three-part cut the kernel abandons on is a lower bound of the distance
at every step, so abandoning on it never drops a window that could match."""
import numpy as np
import pytest


def banded_dp(q, z, w):
    m = len(q)
    d = np.full((m + 1, m + 1), np.inf)
    d[0, 0] = 0.0
    for j in range(m):
        for i in range(max(0, j - w), min(m, j + w + 1)):
            d[i + 1, j + 1] = (q[i] - z[j]) ** 2 + min(d[i, j], d[i, j + 1], d[i + 1, j])
    return d[m, m]


def three_wave_schedule(q, z, w, stale_upper=False):
    """`stale_upper` plants a mistake: phase 2 takes the upper wave's first
    cell of the step before (a missing barrier's effect).  Returns the distance and, for every step where all three waves hold a
    column, the minimum over the cut (bottom's column s, middle's s - 1,
    top's s - 2) after phase 2."""
    m = len(q)
    p = (2 * w + 1) // 3          # cells per wave in phase 1
    tp = 2 * w - 1 - 2 * p        # the top wave's band cells
    c = w - (p + 1)               # band centre in the middle wave
    assert 1 <= tp <= p and 0 <= c < p
    qx = np.concatenate([np.full(w, np.inf), q, np.full(w + 40, np.inf)])
    e = [np.full(p + 1, np.inf) for _ in range(3)]
    e[1][c] = 0.0
    lo_last = [np.inf] * 3
    up_first = [np.inf] * 3
    cuts = []
    for s in range(m + 1):
        act = [0 <= s - k < m for k in range(3)]
        for k in range(3):                      # phase 1
            if not act[k]:
                continue
            j, off = s - k, k * (p + 1)
            prev = np.inf if k == 0 else lo_last[k - 1]
            for t in range(p):
                v = (qx[j + off + t] - z[j]) ** 2 + min(prev, e[k][t], e[k][t + 1])
                e[k][t] = prev = v
            if k == 2:
                e[k][tp:p] = np.inf             # past the band's end
            up_first[k] = e[k][0]
        new_lo = list(lo_last)
        seen = old_up if stale_upper and s > 0 else up_first
        for k in range(2):                      # phase 2, after the barrier
            if not act[k]:
                continue
            j, off = s - k, k * (p + 1)
            r = seen[k + 1] if j > 0 else np.inf
            e[k][p] = (qx[j + off + p] - z[j]) ** 2 + min(e[k][p - 1], e[k][p], r)
            new_lo[k] = e[k][p]
        lo_last = new_lo
        old_up = list(up_first)
        if all(act):
            cuts.append(min(e[0].min(), e[1].min(), e[2].min()))
    return e[1][c], cuts


@pytest.mark.parametrize("m,w", [(300, 159), (3, 200), (2, 180), (161, 160), (250, 205),
                                 (400, 225), (240, 224), (10, 160), (1, 159), (97, 171),
                                 (120, 10), (60, 4)])
def test_three_wave_schedule_is_the_banded_dp(m, w):
    rng = np.random.default_rng(m * 1000 + w)
    q, z = rng.standard_normal(m), rng.standard_normal(m)
    want = banded_dp(q, z, w)
    got, cuts = three_wave_schedule(q, z, w)
    assert got == pytest.approx(want, rel=1e-12)
    # the abandon cut never exceeds the final distance
    assert all(cm <= want * (1 + 1e-12) for cm in cuts)


def test_a_stale_boundary_value_is_caught():
    # the same schedule with phase 2 reading the upper wave's previous step
    # gives a different distance, so the equality above pins the ordering (a
    # narrow band, whose part boundaries lie next to the diagonal the path
    # follows; the schedule is the same for any W)
    m, w = 120, 10
    rng = np.random.default_rng(7)
    q, z = rng.standard_normal(m), rng.standard_normal(m)
    want = banded_dp(q, z, w)
    assert three_wave_schedule(q, z, w)[0] == pytest.approx(want, rel=1e-12)
    assert three_wave_schedule(q, z, w, stale_upper=True)[0] != pytest.approx(want, rel=1e-6)
