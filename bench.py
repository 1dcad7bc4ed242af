"""Benchmark: cDTW GCUPS and windows/s on one MI355X.

Workload (one step): an ECG-like FP32 stream of 2^27 samples (Gaussian-bump
beats, period 250 +- 15, N(0, 0.05) noise, baseline wander), a 256-sample
noisy-beat query, Sakoe-Chiba band W = 26 (10%): window statistics (FP64
prefix sums), the banded cDTW of every one of the 134,217,473 windows, and
the best-match argmin.  ``value`` is that step's throughput in GCUPS (band
cells M(2W+1) - W(W+1) per window, all windows, divided by the step time),
inputs resident in HBM.

Also reported on the same line: the same step at W = 13 and 51, the pruned
best-match search (windows/s and effective GCUPS, as the reference's
GPU-DTW), z-normalised ED search windows/s, the dominant kernel's roofline
(FP32 VALU) from HIP events on the stream it runs on, and a CPU baseline
timed on this host: the restated reference CPU path (brute-force banded DTW
over all windows, OpenMP, cdtw.cuh:90-157) on a bounded sample, plus the
single-threaded UCR-cascade port.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 vector (packed)
PEAK_HBM_GBS = 8000.0


def cells(m, w):
    w = min(w, m - 1)
    return m * (2 * w + 1) - w * (w + 1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(stream, zq, q_raw, m, w):
    """Restated reference CPU path on a bounded sample (all host threads),
    and the UCR cascade port (one thread)."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nwin = 1 << 19
    sample = stream[:nwin + m - 1]
    avg, sd = O.stats_fp64(sample, m)
    t0 = time.perf_counter()
    O.cdtw(zq, sample, avg, sd, w, fused=0)
    t_bf = time.perf_counter() - t0
    gcups = nwin * cells(m, w) / t_bf / 1e9
    nu = 1 << 21
    us = stream[:nu].astype(np.float64)
    t0 = time.perf_counter()
    loc, d2, counts = O.ucr_dtw(us, q_raw.astype(np.float64), w)
    t_ucr = time.perf_counter() - t0
    L_ucr = nu - m + 1
    return ({"value": round(gcups, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
             "sample": f"first {nwin} windows of the stream, M={m}, W={w}; brute-force "
                       f"banded DTW over every window (cdtw.cuh:90-157 restated, OpenMP), "
                       f"{t_bf:.2f} s"},
            {"windows_per_s": round(L_ucr / t_ucr, 1),
             "effective_gcups": round(L_ucr * cells(m, w) / t_ucr / 1e9, 4), "cores": 1,
             "kind": "port", "sample": f"first {nu} samples, M={m}, W={w}, UCR_DTW cascade "
                                       f"port, {t_ucr:.2f} s, location {loc}"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU, and the "
              "search path has no multi-GPU code (see DESIGN.md for how it would shard).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from cudadtw_amd import ALGO_ED, Engine, synth
    from cudadtw_amd import znormalize

    m, w = 256, 26
    log(f"generating ECG-like stream 2^{args.log2n} ...")
    t0 = time.time()
    stream = synth.config2_stream(args.log2n, 0)
    q_raw = synth.ecg_query(m, 2)
    zq = znormalize(q_raw)
    log(f"  {time.time() - t0:.1f} s")
    n = stream.size
    L = n - m + 1

    eng = Engine(0)
    name, cus = eng.info()
    ds = torch.from_numpy(stream).to("cuda:0")
    avg = eng.empty(L)
    sd = eng.empty(L)
    prof = eng.empty(L)
    st = torch.cuda.current_stream()

    def step(wb, evs=None):
        eng.avg_std(ds, m, out=(avg, sd))
        if evs is not None:
            evs[0].record(st)
        eng.gpu_cdtw(ds, avg, sd, wb, out=prof)
        if evs is not None:
            evs[1].record(st)
        return eng.argmin(prof)

    if args.calib_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    results = {}
    for wb in (w, 13, 51):
        eng.set_query(zq, wb)
        steps = args.steps if wb == w else max(2, args.steps // 4)
        for _ in range(args.warmup):
            step(wb)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t0 = time.perf_counter()
        for i in range(steps):
            best = step(wb, evs[i])
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        dp_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
        results[wb] = {"ms_per_step": dt * 1e3, "dp_ms": dp_ms,
                       "gcups": L * cells(m, wb) / dt / 1e9,
                       "dp_gcups": L * cells(m, wb) / (dp_ms / 1e3) / 1e9,
                       "best": [int(best[0]), float(np.sqrt(best[1]))]}
        log(f"W={wb}: {dt*1e3:.2f} ms/step, {results[wb]['gcups']:.1f} GCUPS, DP kernel "
            f"{dp_ms:.2f} ms ({results[wb]['dp_gcups']:.1f} GCUPS), best {results[wb]['best']}")

    # pruned best-match search (the reference's GPU-DTW path), stats included
    eng.set_query(zq, w)
    r = eng.search(ds, w)
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        r = eng.search(ds, w)
    t_search = (time.perf_counter() - t0) / reps
    log(f"search: {t_search*1e3:.2f} ms, location {r.location} distance {r.distance:.5f}, "
        f"info {r.info}")
    assert r.location == results[w]["best"][0], "pruned search disagrees with the full profile"

    # ED search
    re_ = eng.search(ds, 0, algo=ALGO_ED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        re_ = eng.search(ds, 0, algo=ALGO_ED)
    t_ed = (time.perf_counter() - t0) / reps
    log(f"ED search: {t_ed*1e3:.2f} ms, location {re_.location}")

    main_res = results[w]
    flops_per_window = 5 * cells(m, w) + 2 * m
    achieved = L * flops_per_window / (main_res["dp_ms"] / 1e3) / 1e12
    out = {
        "metric": "cDTW GCUPS + windows/s at 1 GPU; batched & bwd vs plain, same GPU",
        "value": round(main_res["gcups"], 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_res["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ecg_like stream 2^{args.log2n} samples, query M={m} noisy "
                               f"beat, band W={w} (10%): stats + cDTW of all {L} windows "
                               f"+ argmin",
                   "n": n, "m": m, "w": w, "device": name, "cus": cus},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "kernel": "band_kernel<26,true> (register-band cDTW)",
                     "flops_per_window": flops_per_window},
        "gcups_by_band": {str(k): round(v["gcups"], 2) for k, v in results.items()},
        "dp_kernel_gcups_by_band": {str(k): round(v["dp_gcups"], 2) for k, v in results.items()},
        "pruned_search": {"ms": round(t_search * 1e3, 3),
                          "windows_per_s": round(L / t_search, 1),
                          "effective_gcups": round(L * cells(m, w) / t_search / 1e9, 2),
                          "location": r.location, "distance": round(r.distance, 6),
                          "dp_evaluated": r.info["dp_evaluated"],
                          "lb_kim_pruned": r.info["lb_kim_pruned"],
                          "lb_keogh_pruned": r.info["lb_keogh_pruned"]},
        "ed_search": {"ms": round(t_ed * 1e3, 3), "windows_per_s": round(L / t_ed, 1),
                      "location": re_.location},
    }
    if not args.no_cpu:
        log("CPU baseline ...")
        cb, ucr = cpu_baseline(stream, zq, q_raw, m, w)
        out["cpu_baseline"] = cb
        out["cpu_ucr_cascade"] = ucr
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
