"""Writes tests/golden/ucr_golden.json: best-match locations and distances
that the reference's own CPU comparators (UCR Suite/UCR_DTW_BIN.cpp and
UCR_ED_BIN.cpp, compiled from the reference sources into oracle/_ref by
oracle/Makefile) print for seeded synthetic inputs.  Run in the build
container, where the reference tree is mounted:

    make -C oracle ref && python tests/golden/make_golden.py

The binaries take FP64 files and ``data query m R n`` (R > 1 is an absolute
band, UCR_DTW_BIN.cpp:419-423).  Only their printed (location, distance) is
kept, with the SHA-256 of each input so a generator change is caught.

Each case is run twice through UCR_DTW_BIN: at its own band W ("ucr_dtw") and
at the band the reference's GPU-DTW driver derives from P = 10 percent,
W = M*(P*0.01) truncated (GPU-dtw.cu:27) ("ucr_dtw_p10"), which is what a
driver invoked as ``gpu_dtw q.bin s.bin M N 10`` searches with.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from cudadtw_amd import synth  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def run(binary, args):
    out = subprocess.run([os.path.join(REF, binary)] + [str(a) for a in args],
                         capture_output=True, text=True, check=True).stdout
    loc = dist = None
    for line in out.splitlines():
        if line.startswith("Location"):
            loc = int(line.split(":")[1])
        if line.startswith("Distance"):
            dist = float(line.split(":")[1])
    return loc, dist


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def cases():
    s, q, w = synth.config1()
    yield "config1_random_walk", s, q, w
    e = synth.ecg_like(1 << 18, 5)
    qe = synth.ecg_query(128, 2)
    yield "ecg_2p18_m128", e, qe, 13
    qe2, _ = synth.cut_query(e, 256, 9)
    yield "ecg_2p18_m256_cut", e, qe2, 26
    mt = synth.metal_like(1 << 18, 3)
    qm, _ = synth.cut_query(mt, 512, 4)
    yield "metal_2p18_m512_cut", mt, qm, 51


def main():
    golden = {}
    with tempfile.TemporaryDirectory() as td:
        for name, s, q, w in cases():
            dp, qp = os.path.join(td, "d.bin"), os.path.join(td, "q.bin")
            s.astype(np.float64).tofile(dp)
            q.astype(np.float64).tofile(qp)
            dtw = run("UCR_DTW_BIN", [dp, qp, q.size, w, s.size])
            w10 = int(q.size * (10 * 0.01))
            dtw10 = run("UCR_DTW_BIN", [dp, qp, q.size, w10, s.size])
            ed = run("UCR_ED_BIN", [dp, qp, q.size, s.size])
            golden[name] = {"n": int(s.size), "m": int(q.size), "w": int(w),
                            "stream_sha256": sha(s), "query_sha256": sha(q),
                            "ucr_dtw": {"location": dtw[0], "distance": dtw[1]},
                            "ucr_dtw_p10": {"w": w10, "location": dtw10[0],
                                            "distance": dtw10[1]},
                            "ucr_ed": {"location": ed[0], "distance": ed[1]}}
            print(name, golden[name]["ucr_dtw"], golden[name]["ucr_ed"])
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ucr_golden.json"),
              "w") as f:
        json.dump(golden, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
