"""Data formats of the reference (cudadtw_amd.data), on the CPU.

Raw .bin layout = the bytes of struct.pack('<f' / '<d') (what array('f') /
array('d') writes on a little-endian host, data/metal/extract.py:45-55, and
what GPU-dtw.cu:59-64 / GPU-ed.cu:51-56 fread); metal text concatenation =
extract.py:5-29 on a synthetic folder (order, NaN lines dropped); and, where
the reference tree is present, the real metal subject's length (11,348,848
values) and the ECG query file.
"""
import os
import struct

import numpy as np
import pytest

from cudadtw_amd import data

REF = "/root/reference/data"


def test_bin_round_trip_and_byte_layout(tmp_path):
    x = np.array([1.5, -2.25, 7031.60720648, np.pi], np.float64)
    p32, p64 = tmp_path / "s.bin", tmp_path / "d.bin"
    data.write_bin(p32, x, np.float32)
    data.write_bin(p64, x, np.float64)
    assert p32.read_bytes() == b"".join(struct.pack("<f", v) for v in x)
    assert p64.read_bytes() == b"".join(struct.pack("<d", v) for v in x)
    np.testing.assert_array_equal(data.read_bin(p32), x.astype(np.float32))
    np.testing.assert_array_equal(data.read_bin(p64, np.float64), x)
    np.testing.assert_array_equal(data.read_bin(p32, count=2), x[:2].astype(np.float32))


def test_metal_text_concatenation(tmp_path):
    os.makedirs(tmp_path / "parent")
    parts = [["1.0", "nan", "2.5"], ["NaN", "3.25", "", "4"], ["-5e-1"]]
    for i, lines in enumerate(parts):
        (tmp_path / "parent" / f"parent_{i:04d}.txt").write_text("\n".join(lines) + "\n")
    got = data.load_metal(tmp_path, "parent", files=3)
    np.testing.assert_array_equal(got, [1.0, 2.5, 3.25, 4.0, -0.5])
    with pytest.raises(FileNotFoundError):
        data.load_metal(tmp_path, "parent", files=4)


def test_extract_queries_seeded_and_in_range(tmp_path):
    x = np.arange(1000, dtype=np.float64)
    a = data.extract_queries(x, 100, 7, seed=3)
    b = data.extract_queries(x, 100, 7, seed=3)
    for qa, qb in zip(a, b):
        np.testing.assert_array_equal(qa, qb)
        assert qa.size == 100 and qa[0] + 99 == qa[-1] and 0 <= qa[0] < 900
    paths = data.write_extracted(tmp_path / "out", x, a)
    assert len(paths) == 2 * (1 + 7)
    np.testing.assert_array_equal(data.read_bin(tmp_path / "out" / "double_query2.bin",
                                                np.float64), a[2])
    q, s = data.split_ecg(np.arange(100005.0))
    assert q.size == 100000 and s.tolist() == [100000.0, 100001, 100002, 100003, 100004]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "metal", "parent")),
                    reason="reference data folder absent")
def test_real_metal_subject_length():
    x = data.load_metal(os.path.join(REF, "metal"), "parent")
    assert x.size == 11348848
    assert np.all(np.isfinite(x))
    # the first value of parent_0000.txt
    assert x[0] == 7031.60720648


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF, "ecg", "ECG_query.mat")),
                    reason="reference data folder absent")
def test_real_ecg_query():
    q = data.load_ecg_query(os.path.join(REF, "ecg", "ECG_query.mat"))
    assert q.shape == (421,) and np.all(np.isfinite(q))
