"""Argument limits of the search entry points, checked on the host before any
launch (the reference's index_t is int, so its streams hold < 2^31 windows;
here candidate lists are int32 as well, cdtw.cuh:450-453).

- More than 2^31 - 1 windows: MDTW_ERR_UNSUPPORTED from mdtw_search,
  mdtw_batch_search and mdtw_stream_create (epoch size), with nothing read
  from the (short) device buffer passed in.
- Exclusion zones beyond the stream (2^62, a "no overlap" sentinel) give the
  same answer as a zone of the stream's length, and negative zones mean none.
- A truncated or corrupt saved stream state is an argument error, never a
  crash or an allocation failure.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from cudadtw_amd import _lib, synth
from cudadtw_amd._lib import MdtwError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HUGE = 1 << 62


def test_more_than_2p31_windows_rejected(engine):
    lib = _lib.load()
    m = 128
    engine.set_query(O.znormalize_f32(synth.random_walk(m, 1)), 12)
    buf = torch.zeros(1024, device="cuda")
    pos = np.full(1, -1, np.int64)
    dist = np.full(1, np.inf, np.float32)
    for n in ((1 << 31) + m - 1, (1 << 32) + m - 1, (1 << 33)):
        st = lib.mdtw_search(engine.h, C.c_void_p(buf.data_ptr()), n, None, None, 12, 0, 1, 0, 1,
                             0, pos.ctypes.data_as(C.POINTER(C.c_int64)),
                             dist.ctypes.data_as(C.POINTER(C.c_float)), None)
        assert st == _lib.ERR_UNSUPPORTED, (n, st)
    # the largest accepted count is 2^31 - 1 windows; a batch checks every query
    b = engine.batch([O.znormalize_f32(synth.random_walk(64, 2))], np.array([6], np.int32))
    with pytest.raises(MdtwError) as ei:
        bpos = np.full(1, -1, np.int64)
        bdist = np.full(1, np.inf, np.float32)
        _lib.check(lib.mdtw_batch_search(
            b.h, C.c_void_p(buf.data_ptr()), (1 << 31) + 63, 1, 1, None,
            bpos.ctypes.data_as(C.POINTER(C.c_int64)), bdist.ctypes.data_as(C.POINTER(C.c_float)),
            None, None), "batch")
    assert ei.value.status == _lib.ERR_UNSUPPORTED
    b.close()
    for ew in (1 << 31, 1 << 40):
        with pytest.raises(MdtwError) as ei:
            engine.stream(12, epoch_windows=ew)
        assert ei.value.status == _lib.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("k", [1, 4])
def test_huge_and_negative_exclusion_zones(engine, k):
    m, w = 96, 9
    s = synth.ecg_like(40_000, 5)
    q, _ = synth.cut_query(s, m, 6)
    zq = O.znormalize_f32(q)
    engine.set_query(zq, w)
    ds = torch.from_numpy(s).cuda()
    L = s.size - m + 1
    avg, sd = engine.avg_std(ds, m)
    torch.cuda.synchronize()
    prof = O.cdtw(zq, s, avg.cpu().numpy(), sd.cpu().numpy(), w)
    for excl, ref_excl in ((HUGE, L), (L + 5, L), (-7, 0)):
        r = engine.search(ds, w, k=k, excl=excl)
        idx, d2 = O.topk_greedy(prof, k, ref_excl)
        np.testing.assert_array_equal(r.positions[:idx.size], idx)
        np.testing.assert_array_equal(r.distances[:idx.size], np.sqrt(d2).astype(np.float32))
        assert np.all(r.positions[idx.size:] == -1)
        b = engine.batch([zq], np.array([w], np.int32))
        rb = b.search(ds, k=k, excl=[excl])
        b.close()
        np.testing.assert_array_equal(rb.positions[0], r.positions)
        np.testing.assert_array_equal(rb.distances[0], r.distances)
    # a stream with a huge zone: every epoch sees at most its own windows
    st = engine.stream(w, k=k, excl=HUGE, epoch_windows=8192)
    st.push(s)
    st.finish()
    rs = st.result()
    st.close()
    idx, d2 = O.topk_greedy(prof, k, L)
    np.testing.assert_array_equal(rs.positions[:idx.size], idx)
    np.testing.assert_array_equal(rs.distances[:idx.size], np.sqrt(d2).astype(np.float32))


def test_corrupt_saved_state_rejected(engine):
    m, w = 64, 6
    engine.set_query(O.znormalize_f32(synth.random_walk(m, 3)), w)
    st = engine.stream(w, k=3, excl=32, epoch_windows=4096)
    s = synth.ecg_like(20_000, 4)
    st.push(s)
    state = bytearray(st.save())
    st.close()
    # a valid state restores
    engine.restore_stream(bytes(state)).close()
    # truncated
    for cut in (1, 8, len(state) // 2, len(state) - 1):
        with pytest.raises(MdtwError) as ei:
            engine.restore_stream(bytes(state[:cut]))
        assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    # the header is: magic (8), six int32, excl, ew, hash, base, fill, windows_done (int64),
    # best_d, thr (float), best_pos, seeded, kim, keogh, dp, abandoned, ncand (int64)
    off_ncand = 8 + 6 * 4 + 8 * 6 + 4 * 2 + 8 * 6
    for bad in (1 << 60, (1 << 64) - 1, 10 ** 9):
        corrupt = bytearray(state)
        corrupt[off_ncand:off_ncand + 8] = (bad % (1 << 64)).to_bytes(8, "little")
        with pytest.raises(MdtwError) as ei:
            engine.restore_stream(bytes(corrupt))
        assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
    off_best = 8 + 6 * 4 + 8 * 6 + 4 * 2
    corrupt = bytearray(state)
    corrupt[off_best:off_best + 8] = (10 ** 12).to_bytes(8, "little")
    with pytest.raises(MdtwError) as ei:
        engine.restore_stream(bytes(corrupt))
    assert ei.value.status == _lib.ERR_INVALID_ARGUMENT
