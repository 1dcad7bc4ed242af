"""The GPU-DTW / GPU-ED / GPU-LDTW compatible drivers (cudadtw_amd/bin, C++
over include/mdtw_compat.hpp) on raw .bin files, against the Python API on
the same inputs: same argv and band truncation (GPU-dtw.cu:27), same result
lines (cdtw.cuh:506-507, ed.cuh:149-150, ldtw.cuh:189-191)."""
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from cudadtw_amd import ALGO_ED, band_width, synth, znormalize

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cudadtw_amd", "bin")


def run(prog, *args):
    path = os.path.join(BIN, prog)
    if not os.path.exists(path):
        pytest.fail(f"{path} missing: build with make -C cudadtw_amd/cli")
    r = subprocess.run([path] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_gpu_dtw_cli(engine, tmp_path):
    s = synth.ecg_like(1 << 18, 3)
    q, _ = synth.cut_query(s, 200, 4)
    (tmp_path / "q.bin").write_bytes(q.astype("<f4").tobytes())
    (tmp_path / "s.bin").write_bytes(s.astype("<f4").tobytes())
    out = run("gpu_dtw", tmp_path / "q.bin", tmp_path / "s.bin", 200, s.size, 10)
    loc = int(re.search(r"Location: (\d+)", out).group(1))
    dist = float(re.search(r"Distance: ([0-9.eE+-]+)", out).group(1))
    w = band_width(200, 10, mode="truncate")
    assert w == 20 and f"window = {w}" in out
    engine.set_query(znormalize(q), w)
    r = engine.search(torch.from_numpy(s).cuda(), w)
    assert loc == r.location
    assert abs(dist - r.distance) <= 1e-5 * r.distance


def test_gpu_ed_cli(engine, tmp_path):
    s = synth.ecg_like(1 << 17, 5)
    q, _ = synth.cut_query(s, 128, 6)
    (tmp_path / "q.bin").write_bytes(q.astype("<f8").tobytes())
    (tmp_path / "s.bin").write_bytes(s.astype("<f8").tobytes())
    out = run("gpu_ed", tmp_path / "q.bin", tmp_path / "s.bin", 128, s.size)
    loc = int(re.search(r"location: (\d+)", out).group(1))
    dist = float(re.search(r"distance: ([0-9.eE+-]+)", out).group(1))
    engine.set_query(znormalize(q, double=True), 0)
    r = engine.search(torch.from_numpy(s).cuda(), 0, algo=ALGO_ED)
    assert loc == r.location
    assert abs(dist - r.distance) <= 1e-5 * r.distance


def test_gpu_ldtw_cli(engine, tmp_path):
    s = synth.random_walk(1 << 15, 7)
    q = synth.random_walk(64, 8)
    (tmp_path / "q.bin").write_bytes(q.astype("<f4").tobytes())
    (tmp_path / "s.bin").write_bytes(s.astype("<f4").tobytes())
    out = run("gpu_ldtw", tmp_path / "q.bin", tmp_path / "s.bin", 64, s.size, 10)
    m = re.search(r"best alignment found at index (\d+) with distance ([0-9.eE+-]+)", out)
    w = band_width(64, 10, mode="truncate")
    engine.set_query(znormalize(q), w)
    ds = torch.from_numpy(s).cuda()
    avg, sd = engine.avg_std(ds, 64)
    prof = engine.ldtw(ds, avg, sd, w).cpu().numpy()
    i = int(np.argmin(prof))
    assert int(m.group(1)) == i
    assert abs(float(m.group(2)) - np.sqrt(prof[i])) <= 1e-5 * np.sqrt(prof[i])
