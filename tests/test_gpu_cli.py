"""The GPU-DTW / GPU-ED / GPU-LDTW compatible drivers (cudadtw_amd/bin, C++
over include/mdtw_compat.hpp) on raw .bin files: same argv and band
truncation (GPU-dtw.cu:27), same result lines (cdtw.cuh:506-507,
ed.cuh:149-150, ldtw.cuh:189-191).

gpu_dtw and gpu_ed must print the locations that the reference's UCR_DTW_BIN
/ UCR_ED_BIN (built from the reference sources) print for the same inputs
(tests/golden/ucr_golden.json, "ucr_dtw_p10" at the band P = 10 percent
gives); gpu_ldtw's three lines (GPU, all cores, one core; GPU-ldtw.cu:93-141)
must name the window the oracle's z_serial_ldtw restatement selects."""
import json
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from cudadtw_amd import ALGO_ED, band_width, synth, znormalize
from oracle import oracle as O

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


GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "ucr_golden.json")))


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
def test_cli_drivers_print_ucr_golden_locations(name, tmp_path):
    s, q = golden_inputs(name)
    g = GOLDEN[name]
    m = q.size
    (tmp_path / "q.bin").write_bytes(q.astype("<f4").tobytes())
    (tmp_path / "s.bin").write_bytes(s.astype("<f4").tobytes())
    out = run("gpu_dtw", tmp_path / "q.bin", tmp_path / "s.bin", m, s.size, 10)
    w = g["ucr_dtw_p10"]["w"]
    assert f"window = {w}" in out
    assert int(re.search(r"Location: (\d+)", out).group(1)) == g["ucr_dtw_p10"]["location"]
    dist = float(re.search(r"Distance: ([0-9.eE+-]+)", out).group(1))
    assert abs(dist / g["ucr_dtw_p10"]["distance"] - 1) < 1e-3  # UCR prints 6 digits
    # GPU-ED reads FP64 files (GPU-ed.cu:51-56)
    (tmp_path / "q64.bin").write_bytes(q.astype("<f8").tobytes())
    (tmp_path / "s64.bin").write_bytes(s.astype("<f8").tobytes())
    out = run("gpu_ed", tmp_path / "q64.bin", tmp_path / "s64.bin", m, s.size)
    assert int(re.search(r"location: (\d+)", out).group(1)) == g["ucr_ed"]["location"]
    dist = float(re.search(r"distance: ([0-9.eE+-]+)", out).group(1))
    assert abs(dist / g["ucr_ed"]["distance"] - 1) < 1e-4


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
    lines = re.findall(r"best alignment found at index (\d+) with distance ([0-9.eE+-]+)", out)
    assert len(lines) == 3, out  # GPU, all cores, one core
    for label in ("on the GPU", "on all cores", "on a single core"):
        assert f"Miliseconds to find best alignment {label}" in out
    w = band_width(64, 10, mode="truncate")
    zq = znormalize(q)
    engine.set_query(zq, w)
    ds = torch.from_numpy(s).cuda()
    avg, sd = engine.avg_std(ds, 64)
    torch.cuda.synchronize()
    prof = O.ldtw(zq, s, avg.cpu().numpy(), sd.cpu().numpy(), w)
    i = O.argmin(prof)
    for idx, dist in lines:
        assert int(idx) == i
        assert abs(float(dist) - np.sqrt(prof[i])) <= 1e-5 * np.sqrt(prof[i])
