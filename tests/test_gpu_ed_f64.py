"""The ED driver's FP64 path (GPU-ed.cu:20-82): an FP64 stream, avg_std<double>
(stats.cuh:115-142), calculate_ed on doubles (ed.cuh:71-153), searched in FP64
end to end -- no narrowing to FP32 anywhere.

Against the oracle: FP64 window statistics within 1e-9 relative of two-pass
FP64 (on a stream near 7000, where the reference's FP32 finalize fails); the
FP64 profile within 1e-9 relative (+ 1e-9 M absolute) of the direct FP64
z-normalised ED; the best window equal to the oracle profile's argmin and to
UCR_ED_BIN's printed location on the golden inputs
(tests/golden/ucr_golden.json).  cuFFT's own rounding is not reproducible
here; the FP64 closed form pins the value instead.
"""
import json
import os

import numpy as np
import pytest
import torch

from cudadtw_amd import _lib, synth, znormalize
from cudadtw_amd._lib import MdtwError
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ucr_golden.json")))


def dev64(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


@pytest.mark.parametrize("m", [16, 128, 512, 2048, 8192])
def test_ed_f64_profile_and_best_on_offset_stream(engine, m):
    s = synth.metal_like(1 << 16, 70 + m, offset=7000.0).astype(np.float64)
    s += 1e-3 * np.random.default_rng(m).standard_normal(s.size)  # below FP32 resolution
    q = s[1234:1234 + m] + 0.05 * np.random.default_rng(1).standard_normal(m)
    zq = znormalize(q, double=True, keep_f64=True)
    ds = dev64(s)
    avg, sd = engine.avg_std_f64(ds, m)
    torch.cuda.synchronize()
    oa, osd = O.stats_f64d(s, m)
    np.testing.assert_allclose(avg.cpu().numpy(), oa, rtol=1e-12)
    np.testing.assert_allclose(sd.cpu().numpy(), osd, rtol=1e-9)
    prof = engine.ed_f64(ds, zq).cpu().numpy()
    ref = O.ed_f64d(zq, s, oa, osd)
    assert np.all(np.abs(prof - ref) <= 1e-9 * ref + 1e-9 * m), np.max(np.abs(prof - ref))
    loc, dist = engine.ed_search_f64(ds, zq)
    assert loc == int(np.argmin(ref))
    assert abs(dist - np.sqrt(ref.min())) <= 1e-9 * np.sqrt(ref.min()) + 1e-9
    assert loc == int(np.argmin(prof)) and dist == np.sqrt(prof.min())


def golden_inputs(name):
    if name == "config1_random_walk":
        s, q, _ = synth.config1()
        return s, q
    if name == "ecg_2p18_m128":
        return synth.ecg_like(1 << 18, 5), synth.ecg_query(128, 2)
    if name == "ecg_2p18_m256_cut":
        e = synth.ecg_like(1 << 18, 5)
        return e, synth.cut_query(e, 256, 9)[0]
    mt = synth.metal_like(1 << 18, 3)
    return mt, synth.cut_query(mt, 512, 4)[0]


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_ed_f64_ucr_golden_location(engine, name):
    s, q = golden_inputs(name)
    zq = znormalize(q, double=True, keep_f64=True)
    loc, dist = engine.calculate_ed(dev64(s), zquery64=zq)
    assert loc == GOLDEN[name]["ucr_ed"]["location"]
    assert abs(dist / GOLDEN[name]["ucr_ed"]["distance"] - 1) < 1e-5  # UCR prints 6 digits


def test_ed_f64_argument_errors(engine):
    ds = dev64(synth.random_walk(5000, 1))
    with pytest.raises(MdtwError) as ei:
        engine.ed_search_f64(ds, np.array([0.0, np.nan, 1.0]))
    assert ei.value.status == _lib.ERR_NOT_FINITE
    with pytest.raises(MdtwError) as ei:
        engine.ed_search_f64(ds, np.zeros(8193))
    assert ei.value.status == _lib.ERR_QUERY_TOO_LONG
    with pytest.raises(MdtwError) as ei:
        engine.ed_search_f64(dev64(np.zeros(10)), np.ones(11))
    assert ei.value.status == _lib.ERR_STREAM_TOO_SHORT
