"""The C ABI library: loads without a GPU, exports every entry point that
include/mdtw.h declares, and its host-only helpers (no device call) are
bit-exact with the oracle's restatements of stats.cuh:8-28 and
bounds.cuh:13-64."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cudadtw_amd import _lib, envelope, znormalize
from cudadtw_amd import synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mdtw.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mdtw_\w+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding covers the same set
    assert set(names) == set(_lib.SIGNATURES)


def test_status_strings():
    lib = _lib.load()
    for code in range(13):
        assert lib.mdtw_status_string(code)
    assert lib.mdtw_status_string(0) == b"ok"


@pytest.mark.parametrize("seed", range(5))
def test_host_znormalize_bit_exact(seed):
    x = synth.random_walk(int(64 + 100 * seed), seed)
    assert np.array_equal(znormalize(x), O.znormalize_f32(x))


@pytest.mark.parametrize("seed", range(3))
def test_host_znormalize_f64_bit_exact(seed):
    x = synth.metal_like(300 + 50 * seed, seed, offset=7000.0)
    got = znormalize(x, double=True)
    assert np.array_equal(got, O.znormalize_f64(x.astype(np.float64)).astype(np.float32))


@pytest.mark.parametrize("w", [0, 1, 5, 13, 40, 127, 500])
def test_host_envelope_bit_exact(w):
    q = O.znormalize_f32(synth.random_walk(128, 3))
    lo, up = envelope(q, w)
    olo, oup = O.envelope(q, w)
    assert np.array_equal(lo, olo) and np.array_equal(up, oup)


def test_invalid_arguments_rejected_on_host():
    lib = _lib.load()
    assert lib.mdtw_znormalize(None, 10) == _lib.ERR_INVALID_ARGUMENT
    x = np.zeros(4, np.float32)
    p = x.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.mdtw_envelope(p, -1, p, p, 4) == _lib.ERR_INVALID_ARGUMENT
    assert lib.mdtw_create(0, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.mdtw_destroy(None) == _lib.ERR_INVALID_ARGUMENT
