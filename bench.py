"""Benchmark: cDTW GCUPS and windows/s on one MI355X; batched and backward vs
plain on the same GPU.

Main workload (one step, ``value``): an ECG-like FP32 stream of 2^27 samples
(Gaussian-bump beats, period 250 +- 15, N(0, 0.05) noise, baseline wander),
a 256-sample noisy-beat query, Sakoe-Chiba band W = 26 (10%): window
statistics (FP64 prefix sums), the banded cDTW of every one of the
134,217,473 windows, and the best-match argmin.  ``value`` is that step's
throughput in GCUPS (band cells M(2W+1) - W(W+1) per window, all windows,
divided by the step time), inputs resident in HBM.

Also on the same line, one object per workload:
  c2      the same step at W = 13 and 51; the pruned best-match search
          (windows/s, effective GCUPS); z-normalised ED search,
  c1      random-walk stream 2^20, walk query M = 128, W = 13 (best match),
  c3      metal-like stream 2^26, cut queries M = 512 / 1024, W = 10 %,
          top-16 with exclusion M/2, cDTW and LDTW,
  c4      mixed batch: queries |Q| in [64, 1024], bands 5-20 %, top-8, one
          captured-graph batch launch vs the same queries as one plain search
          each (default: every 32nd of the 1024 queries; --c4 full: all),
          plus 1024 short queries on a 2^18 stream, where per-query fixed
          costs dominate (--c4 short: only that case),
  c5      256 queries |Q| in [128, 512], W = 10 %, 2^24 stream: top-8 forward,
          query-gradient backward (recompute GCUPS), 50-step template
          refinement,
plus the dominant kernel's roofline (FP32 VALU) from HIP events on the
stream it runs on, and a CPU baseline timed on this host (the restated
reference CPU path, brute-force banded DTW over all windows with OpenMP,
cdtw.cuh:90-157, on a bounded sample; and the single-threaded UCR-cascade
port).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--c4 sample|full|short|off]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 vector (packed)
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 vector
PEAK_HBM_GBS = 8000.0


def cells(m, w):
    w = min(w, m - 1)
    return m * (2 * w + 1) - w * (w + 1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def cpu_baseline(stream, zq, q_raw, m, w):
    """CPU comparators on this host, each on a bounded sample of the same
    workload: the reference's CPU paths restated (brute-force banded DTW over
    every window, cdtw.cuh:90-157 with OpenMP and :23-88 on one core; greedy
    LDTW, ldtw.cuh:151-174, all cores and one core), the UCR-suite cascade
    (project code from its paper: one thread, and one thread per core), and
    UCR_DTW_BIN / UCR_ED_BIN compiled from the reference sources where they
    are built (oracle/_ref).  Returns (cpu_baseline, extra comparators)."""
    from oracle import oracle as O

    threads = O.max_threads()
    host = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "omp_threads": threads}
    c = cells(m, w)

    def brute(nwin, nthreads):
        O.set_threads(nthreads)
        sample = stream[:nwin + m - 1]
        avg, sd = O.stats_fp64(sample, m)
        t0 = time.perf_counter()
        O.cdtw(zq, sample, avg, sd, w, fused=0)
        dt = time.perf_counter() - t0
        O.set_threads(threads)
        return nwin * c / dt / 1e9, dt

    gcups, t_bf = brute(1 << 19, threads)
    g1, t_1 = brute(1 << 15, 1)
    cb = {"value": round(gcups, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
          "sample": f"first {1 << 19} windows of the stream, M={m}, W={w}; brute-force banded "
                    f"DTW over every window (cdtw.cuh:90-157 restated, OpenMP), {t_bf:.2f} s",
          **host}
    extra = {"host": host,
             "cdtw_one_core": {"gcups": round(g1, 4), "cores": 1, "kind": "port",
                               "sample": f"first {1 << 15} windows, sequential_cdtw "
                                         f"(cdtw.cuh:23-88 restated), {t_1:.2f} s"}}
    # greedy LDTW, all cores and one core (GPU-ldtw.cu:108-141)
    for name, nwin, nthreads in (("ldtw_all_cores", 1 << 21, threads), ("ldtw_one_core", 1 << 17, 1)):
        O.set_threads(nthreads)
        sample = stream[:nwin + m - 1]
        avg, sd = O.stats_fp64(sample, m)
        t0 = time.perf_counter()
        O.ldtw(zq, sample, avg, sd, w)
        dt = time.perf_counter() - t0
        extra[name] = {"windows_per_s": round(nwin / dt, 1),
                       "g_steps_per_s": round(nwin * (2 * m - 2) / dt / 1e9, 4),
                       "cores": nthreads, "kind": "port",
                       "sample": f"first {nwin} windows, M={m}, W={w}, z_stream_ldtw "
                                 f"(ldtw.cuh:151-174 restated), {dt:.2f} s"}
    O.set_threads(threads)
    # UCR-suite cascade: one thread (as UCR runs), one per core
    q64 = q_raw.astype(np.float64)
    for name, nu, nthreads in (("ucr_dtw_one_thread", 1 << 20, 1),
                               ("ucr_dtw_all_threads", 1 << 22, threads)):
        us = stream[:nu].astype(np.float64)
        t0 = time.perf_counter()
        if nthreads == 1:
            loc, d2, _ = O.ucr_dtw(us, q64, w)
        else:
            loc, d2, _ = O.ucr_dtw_par(us, q64, w, nthreads)
        dt = time.perf_counter() - t0
        L_u = nu - m + 1
        extra[name] = {"windows_per_s": round(L_u / dt, 1),
                       "effective_gcups": round(L_u * c / dt / 1e9, 4), "cores": nthreads,
                       "kind": "port", "sample": f"first {nu} samples, M={m}, W={w}, UCR cascade "
                                                 f"(KDD 2012), {dt:.2f} s, location {loc}"}
    ue = stream[:1 << 22].astype(np.float64)
    t0 = time.perf_counter()
    loc, _ = O.ucr_ed(ue, q64)
    dt = time.perf_counter() - t0
    extra["ucr_ed_one_thread"] = {"windows_per_s": round((ue.size - m + 1) / dt, 1), "cores": 1,
                                  "kind": "port",
                                  "sample": f"first {ue.size} samples, M={m}, early-abandoning "
                                            f"z-normalised ED, {dt:.2f} s, location {loc}"}
    # the reference's own UCR programs, built from its sources
    for name, binary, nu, r in (("ucr_dtw_bin", "UCR_DTW_BIN", 1 << 20, w),
                                ("ucr_ed_bin", "UCR_ED_BIN", 1 << 22, None)):
        us = stream[:nu].astype(np.float64)
        t0 = time.perf_counter()
        res = O.ucr_reference(binary, us, q64, r)
        dt = time.perf_counter() - t0
        if res is None:
            extra[name] = {"kind": "reference", "skipped": "oracle/_ref not built on this host"}
            continue
        loc, dist, secs = res
        secs = secs if secs else dt
        extra[name] = {"windows_per_s": round((nu - m + 1) / secs, 1), "cores": 1,
                       "kind": "reference",
                       "sample": f"first {nu} samples, M={m}" + (f", R={r}" if r else "") +
                                 f", {binary} from the reference sources, {secs:.2f} s "
                                 f"(its own clock), location {loc}, distance {dist}"}
    return cb, extra


def timed(torch, fn, reps=1):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = fn()
    torch.cuda.synchronize()
    return r, (time.perf_counter() - t0) / reps


def ed_vendor_fft(torch, ds, zq):
    """ED of every window in the reference's form (ed.cuh:71-153) with
    torch.fft (rocFFT on this platform), FP64: a comparison for the
    hand-written transform, not part of the library.  Returns (argmin, ED^2)."""
    x = ds.double()
    n, m = x.numel(), len(zq)
    q = torch.zeros(n, dtype=torch.float64, device=ds.device)
    q[:m] = torch.from_numpy(np.asarray(zq, dtype=np.float64)).to(ds.device)
    corr = torch.fft.irfft(torch.fft.rfft(x) * torch.conj(torch.fft.rfft(q)), n=n)[:n - m + 1]
    c = x[0]
    p1 = torch.cumsum(torch.cat([x.new_zeros(1), x - c]), 0)
    p2 = torch.cumsum(torch.cat([x.new_zeros(1), (x - c) ** 2]), 0)
    s1 = (p1[m:] - p1[:-m]) / m
    var = (p2[m:] - p2[:-m]) / m - s1 * s1
    sig = torch.where(var > 0, var.clamp(min=0).sqrt(), torch.ones_like(var))
    mu = s1 + c
    ed2 = 2 * m - 2 * (corr - mu * float(np.sum(np.asarray(zq, np.float64)))) / sig
    i = int(torch.argmin(ed2))
    return i, float(ed2[i])


def section_c2(eng, torch, args):
    from cudadtw_amd import ALGO_ED, ALGO_ED_FFT, synth, znormalize

    m, w = 256, 26
    log(f"[c2] generating ECG-like stream 2^{args.log2n} ...")
    stream = synth.config2_stream(args.log2n, 0)
    q_raw = synth.ecg_query(m, 2)
    zq = znormalize(q_raw)
    n = stream.size
    L = n - m + 1
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
        log(f"[c2] W={wb}: {dt*1e3:.2f} ms/step, {results[wb]['gcups']:.1f} GCUPS, DP kernel "
            f"{dp_ms:.2f} ms ({results[wb]['dp_gcups']:.1f} GCUPS), best {results[wb]['best']}")

    eng.set_query(zq, w)
    r, t_search = timed(torch, lambda: eng.search(ds, w), reps=5)
    log(f"[c2] search: {t_search*1e3:.2f} ms, location {r.location} distance {r.distance:.5f}, "
        f"info {r.info}")
    assert r.location == results[w]["best"][0], "pruned search disagrees with the full profile"
    re_, t_ed = timed(torch, lambda: eng.search(ds, 0, algo=ALGO_ED), reps=5)
    log(f"[c2] ED search: {t_ed*1e3:.2f} ms, location {re_.location}")

    # The memory-side kernels over every window of the 2^27 stream, timed with
    # HIP events on the stream they run on: window statistics (reads 4 B,
    # writes 8 B per window), LB_Kim and LB_Keogh and the ED profile (read the
    # sample plus mean and sigma, 12 B, write 4 B); GB/s of these algorithmic
    # bytes against the 8 TB/s HBM peak.
    eng.set_query(zq, w)
    mem = {}
    for name, nbytes, fn in (
            ("stats", 12, lambda: eng.avg_std(ds, m, out=(avg, sd))),
            ("lb_kim", 16, lambda: eng.lb_kim(ds, avg, sd)),
            ("lb_keogh", 16, lambda: eng.lb_keogh(ds, avg, sd)),
            ("ed_profile", 16, lambda: eng.ed(ds, avg, sd))):
        fn()
        st.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(3):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 3
        gbps = L * nbytes / (ms / 1e3) / 1e9
        mem[name] = {"ms": round(ms, 3), "gb_per_s": round(gbps, 1),
                     "frac_of_8tbs": round(gbps / 8000.0, 3), "bytes_per_window": nbytes}
    # the ED profile against its own roof: it sums in FP64, one fused
    # multiply-add per window-element (2M flop per window) against the FP64
    # vector peak, and the issue its loop needs (19 full-rate FP64 operations
    # per 16 window-elements, 4 SIMD cycles per wave64 operation) against
    # the chip's 256 CU x 4 SIMD at 2.4 GHz
    ed_ms = mem["ed_profile"]["ms"]
    ed_tflops = L * 2 * m / (ed_ms / 1e3) / 1e12
    issue = L * m / 64 * (19.0 / 16.0) * 4.0 / (ed_ms / 1e3) / (256 * 4 * 2.4e9)
    mem["ed_profile"]["roofline"] = {"bound": "valu", "achieved": round(ed_tflops, 2),
                                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s (FP64)",
                                     "frac": round(ed_tflops / PEAK_FP64_TFLOPS, 4),
                                     "flops_per_window": 2 * m,
                                     "valu_issue_frac": round(issue, 3)}
    # its HBM bytes per launch, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # at this shape (profiles/ed_traffic.json), scaled to the windows
    epath = os.path.join(ROOT, "profiles", "ed_traffic.json")
    if os.path.exists(epath):
        with open(epath) as f:
            et = json.load(f)
        if et.get("m") == m:
            mem["ed_profile"]["roofline"]["traffic"] = round(et["bytes_per_window"] * L)
            mem["ed_profile"]["roofline"]["traffic_unit"] = "bytes per launch (HBM, rocprofv3 PMC)"
    # a plain device copy of the stream (torch's copy kernel): the HBM rate a
    # 1 read : 1 write pattern reaches here, beside the kernels above
    dst = torch.empty_like(ds)
    dst.copy_(ds)
    st.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(3):
        dst.copy_(ds)
    e1.record(st)
    torch.cuda.synchronize()
    cms = e0.elapsed_time(e1) / 3
    mem["device_copy"] = {"ms": round(cms, 3), "bytes": 2 * 4 * n,
                          "gb_per_s": round(2 * 4 * n / (cms / 1e3) / 1e9, 1)}
    del dst
    log(f"[c2] memory-side kernels over all windows: {mem}")

    # DP rate on wide bands (|Q| = 1024 at 10 / 15.6 / 20 %: W = 102 in the
    # four-column register kernel, 160 and 205 in the one-wave-per-SIMD
    # register kernel that takes four skewed columns per step), profile mode
    # over a 2^24-sample prefix of the stream.
    wide = {}
    n24 = min(n, 1 << 24)
    q1k = znormalize(synth.cut_query(stream[:n24], 1024, 11)[0])
    for wb in (102, 160, 205):
        eng.set_query(q1k, wb)
        a1, s1 = eng.avg_std(ds[:n24], 1024)
        p1 = eng.gpu_cdtw(ds[:n24], a1, s1, wb)
        st.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        eng.gpu_cdtw(ds[:n24], a1, s1, wb, out=p1)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        wide[str(wb)] = round((n24 - 1023) * cells(1024, wb) / (ms / 1e3) / 1e9, 1)
    log(f"[c2] wide-band DP GCUPS (|Q| = 1024): {wide}")

    # ED through FP64 FFT cross-correlation (the reference's form, ed.cuh:71-153)
    # against the direct ED search, for the C2 query and two long queries
    # cut from the stream (|Q| = 2048, 8192; the FFT path's limit is 8192).
    ed_forms = {}
    for mq in (m, 2048, 8192):
        zz = zq if mq == m else znormalize(synth.cut_query(stream, mq, 7 + mq)[0])
        eng.set_query(zz, 0)
        rd, td = timed(torch, lambda: eng.search(ds, 0, algo=ALGO_ED), reps=3)
        rf, tf = timed(torch, lambda: eng.search(ds, 0, algo=ALGO_ED_FFT), reps=3)
        # comparison only: the reference's own form through the vendor FFT
        # (torch.fft is rocFFT here), ed.cuh:71-153 -- one FP64 transform of
        # the whole stream and of the zero-padded query, the product, the
        # inverse, FP64 window statistics from FP64 prefix sums, then
        # ED^2 = 2M - 2 (corr - mu sum q) / sigma and the argmin
        (loc_v, _), tv = timed(torch, lambda: ed_vendor_fft(torch, ds, zz), reps=2)
        Lq = n - mq + 1
        ed_forms[str(mq)] = {"direct_ms": round(td * 1e3, 3), "fft_ms": round(tf * 1e3, 3),
                             "fft_windows_per_s": round(Lq / tf, 1),
                             "direct_location": rd.location, "fft_location": rf.location,
                             "same_location": rd.location == rf.location,
                             "rocfft_reference_form_ms": round(tv * 1e3, 3),
                             "rocfft_reference_form_location": loc_v}
        log(f"[c2] ED |Q|={mq}: direct {td*1e3:.2f} ms (location {rd.location}), FFT "
            f"{tf*1e3:.2f} ms (location {rf.location}), the reference's whole-stream form "
            f"through rocFFT {tv*1e3:.2f} ms (location {loc_v})")
    eng.set_query(zq, w)

    # The same search fed from host memory: one-shot after a host-to-device
    # copy, and the streaming search (2^24-window epochs, pushes of 2^24
    # samples, copies overlapped with the search).  PCIe included in both.
    chunk = 1 << 24

    def copied():
        return eng.search(torch.from_numpy(stream).to("cuda:0"), w)

    def streamed():
        sst = eng.stream(w, epoch_windows=chunk)
        for i in range(0, n, chunk):
            sst.push(stream[i:i + chunk])
        sst.finish()
        res = sst.result()
        sst.close()
        return res

    rcp, t_copy = timed(torch, copied, reps=2)
    rst, t_stream = timed(torch, streamed, reps=2)
    assert rst.location == r.location and rst.distance == r.distance, "streamed answer differs"
    assert rcp.location == r.location
    log(f"[c2] from host memory: copy + search {t_copy*1e3:.1f} ms, streamed "
        f"{t_stream*1e3:.1f} ms (same answer)")
    out = {
        "main": results[w], "L": L, "n": n, "m": m, "w": w,
        "gcups_by_band": {str(k): round(v["gcups"], 2) for k, v in results.items()},
        "dp_kernel_gcups_by_band": {str(k): round(v["dp_gcups"], 2) for k, v in results.items()},
        "dp_kernel_gcups_wide_m1024": wide,
        "memory_kernels_all_windows": mem,
        "pruned_search": {"ms": round(t_search * 1e3, 3), "windows_per_s": round(L / t_search, 1),
                          "effective_gcups": round(L * cells(m, w) / t_search / 1e9, 2),
                          "location": r.location, "distance": round(r.distance, 6),
                          "dp_evaluated": r.info["dp_evaluated"],
                          "lb_kim_pruned": r.info["lb_kim_pruned"],
                          "lb_keogh_pruned": r.info["lb_keogh_pruned"]},
        "ed_search": {"ms": round(t_ed * 1e3, 3), "windows_per_s": round(L / t_ed, 1),
                      "location": re_.location},
        "ed_direct_vs_fft": ed_forms,
        "from_host": {"copy_then_search_ms": round(t_copy * 1e3, 2),
                      "streamed_ms": round(t_stream * 1e3, 2),
                      "streamed_windows_per_s": round(L / t_stream, 1),
                      "epoch_windows": chunk, "push_samples": chunk},
    }
    return out, stream, q_raw, zq


def section_c1(eng, torch):
    from cudadtw_amd import synth, znormalize

    s, q, w = synth.config1()
    ds = torch.from_numpy(s).to("cuda:0")
    eng.set_query(znormalize(q), w)
    r, t = timed(torch, lambda: eng.search(ds, w), reps=5)
    L = s.size - q.size + 1
    log(f"[c1] search {t*1e3:.2f} ms, location {r.location} distance {r.distance:.5f}")
    return {"workload": "random walk 2^20 (seed 0), walk query M=128 (seed 1), W=13, best match",
            "ms": round(t * 1e3, 3), "windows_per_s": round(L / t, 1),
            "effective_gcups": round(L * cells(128, w) / t / 1e9, 2),
            "location": r.location, "distance": round(r.distance, 6)}


def section_c3(eng, torch):
    from cudadtw_amd import ALGO_CDTW, ALGO_LDTW, synth, znormalize

    log("[c3] generating metal-like stream 2^26 ...")
    s = synth.metal_like(1 << 26, 3)
    ds = torch.from_numpy(s).to("cuda:0")
    out = {"workload": "metal-like 2^26, cut queries (N(0,0.05)), W=10%, top-16, excl M/2"}
    for m in (512, 1024):
        q, pos0 = synth.cut_query(s, m, 30 + m)
        w = int(np.floor(m * 0.1 + 0.5))
        eng.set_query(znormalize(q), w)
        avg, sd = eng.avg_std(ds, m)
        L = s.size - m + 1
        for name, algo in (("cdtw", ALGO_CDTW), ("ldtw", ALGO_LDTW)):
            r, t = timed(torch, lambda: eng.search(ds, w, algo=algo, k=16, excl=m // 2,
                                                   avg=avg, sd=sd), reps=2)
            out[f"{name}_m{m}"] = {"w": w, "ms": round(t * 1e3, 3),
                                   "windows_per_s": round(L / t, 1),
                                   "first": int(r.positions[0]), "cut_at": pos0,
                                   "found": int(r.info["found"])}
            log(f"[c3] {name} M={m} W={w}: {t*1e3:.2f} ms, top match {r.positions[0]} "
                f"(cut at {pos0})")
    return out


def c4_short(eng, torch, k=8):
    """The batch-vs-plain comparison where per-query fixed costs dominate:
    1024 short queries on a short stream (launch- and synchronisation-bound),
    to show what one graph launch over concurrent pipelines removes."""
    from cudadtw_amd import synth, znormalize

    s2 = synth.ecg_like(1 << 18, 9)
    q2, b2, _ = synth.batch_queries(s2, 1024, 9, 64, 128)
    z2 = [znormalize(q) for q in q2]
    ds2 = torch.from_numpy(s2).to("cuda:0")
    bat2 = eng.batch(z2, b2, pipelines=4)
    bat2.search(ds2, k=k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r2 = bat2.search(ds2, k=k)
    t_b2 = time.perf_counter() - t0
    p2 = np.full((len(z2), k), -1, np.int64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (z, w) in enumerate(zip(z2, b2)):
        eng.set_query(z, int(w))
        p2[i] = eng.search(ds2, int(w), k=k, excl=z.size // 2).positions
    torch.cuda.synchronize()
    t_p2 = time.perf_counter() - t0
    same2 = bool(np.array_equal(p2, r2.positions))
    bat2.close()
    log(f"[c4] short queries: batch {t_b2*1e3:.1f} ms vs plain {t_p2*1e3:.1f} ms "
        f"({t_p2 / t_b2:.2f}x), identical: {same2}")
    return {"workload": "ECG-like 2^18, 1024 queries |Q| 64-128, bands 5-20%, top-8 "
                        "(not the C4 shape)",
            "batch_ms": round(t_b2 * 1e3, 2), "plain_ms": round(t_p2 * 1e3, 2),
            "speedup": round(t_p2 / t_b2, 3), "identical": same2}


def section_c4(eng, torch, args):
    if args.c4 == "short":
        return {"launch_bound_case": c4_short(eng, torch)}
    from cudadtw_amd import synth, znormalize

    log("[c4] generating ECG-like stream 2^26 and 1024 queries ...")
    s, qs, bands, kinds = synth.config4()
    sel = list(range(0, len(qs), 1 if args.c4 == "full" else args.c4_stride))
    zqs = [znormalize(qs[i]) for i in sel]
    bs = bands[sel]
    ds = torch.from_numpy(s).to("cuda:0")
    k = 8
    Ls = np.array([s.size - z.size + 1 for z in zqs], np.float64)
    cellsum = float(sum(L * cells(z.size, int(b)) for L, z, b in zip(Ls, zqs, bs)))
    # One batch over the selected queries, or (--c4-split G) consecutive
    # groups of G queries, one batch each (progress per group for the full
    # set, whose single launch runs for minutes).  Per group: tuning and the
    # capturing first search untimed, the second search timed; then one
    # plain search per query of the group, timed.
    G = args.c4_split if args.c4_split > 0 else len(zqs)
    t_batch = t_plain = t_tune = 0.0
    nb = 0
    same = True
    forms = {}  # DP formulation the tuner picked, per query
    for g0 in range(0, len(zqs), G):
        gz, gb = zqs[g0:g0 + G], bs[g0:g0 + G]
        batch = eng.batch(gz, gb, pipelines=4)
        t0 = time.perf_counter()
        nb += batch.tune(ds, k=k, cell_budget=1 << 38)
        t_tune += time.perf_counter() - t0
        for i, (z, w) in enumerate(zip(gz, gb)):
            sh = batch.shape(i)
            key = f"{'W<=112' if w <= 112 else 'W113-240' if w <= 240 else 'W>240'}:{sh['dp_form']}"
            forms[key] = forms.get(key, 0) + 1
        rb = batch.search(ds, k=k)  # captures the graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rb = batch.search(ds, k=k)
        tb = time.perf_counter() - t0
        pos = np.full((len(gz), k), -1, np.int64)
        dist = np.full((len(gz), k), np.inf, np.float32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, (z, w) in enumerate(zip(gz, gb)):
            eng.set_query(z, int(w))
            r = eng.search(ds, int(w), k=k, excl=z.size // 2)
            pos[i], dist[i] = r.positions, r.distances
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        gsame = bool(np.array_equal(pos, rb.positions) and np.array_equal(dist, rb.distances))
        batch.close()
        t_batch += tb
        t_plain += tp
        same = same and gsame
        log(f"[c4] queries {g0}..{g0 + len(gz) - 1}: batch {tb:.3f} s (graph device time "
            f"{rb.elapsed_ms:.1f} ms), plain {tp:.3f} s, identical: {gsame}")
    log(f"[c4] {len(zqs)} queries in batches of {G}: batch {t_batch:.3f} s, plain "
        f"{t_plain:.3f} s; speedup {t_plain / t_batch:.2f}x; identical results: {same}")

    short = c4_short(eng, torch, k)
    return {"workload": f"ECG-like 2^26, {len(zqs)} of 1024 queries "
                        f"({'all' if args.c4 == 'full' else 'every %d-th' % args.c4_stride}), "
                        "|Q| 64-1024, bands 5-20%, half cut+N(0,0.05), half walks, top-8, "
                        "excl |Q|/2",
            "queries": len(zqs), "queries_per_batch": G,
            "wide_band_queries": int(np.sum(bs > 112)),
            "batch_s": round(t_batch, 4), "plain_s": round(t_plain, 4),
            "speedup": round(t_plain / t_batch, 3), "identical": same,
            "tune_s": round(t_tune, 2), "buckets_tuned": nb, "tuned_dp_forms": forms,
            "launch_bound_case": short,
            "windows_per_s_batch": round(float(Ls.sum()) / t_batch, 1),
            "effective_gcups_batch": round(cellsum / t_batch / 1e9, 2)}


def section_c5(eng, torch, args):
    from cudadtw_amd import synth, znormalize
    from cudadtw_amd.autograd import topk_cdtw

    log("[c5] generating ECG-like stream 2^24 and 256 queries ...")
    s, qs, bands, _ = synth.config5()
    ds = torch.from_numpy(s).to("cuda:0")
    k = 8
    zqs = [znormalize(q) for q in qs]
    # plain: one search and one backward call per query
    t_stats = t_fwd = t_bwd = 0.0
    rec_cells = 0
    fwd_cells = 0.0
    plain_pos = np.full((len(qs), k), -1, np.int64)
    for i, (q, zq, w) in enumerate(zip(qs, zqs, bands)):
        m, w = q.size, int(w)
        eng.set_query(zq, w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        avg, sd = eng.avg_std(ds, m)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        r = eng.search(ds, w, k=k, excl=m // 2, avg=avg, sd=sd)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        found = int(np.sum(r.positions >= 0))
        plain_pos[i] = r.positions
        idx = torch.from_numpy(r.positions[:found].astype(np.int32)).to("cuda:0")
        g = torch.ones(found, dtype=torch.float32, device="cuda:0")
        qd = torch.from_numpy(q).to("cuda:0")
        eng.cdtw_backward(ds, avg, sd, idx, g, w, query_raw=qd)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        t_stats += t1 - t0
        t_fwd += t2 - t1
        t_bwd += t3 - t2
        rec_cells += found * cells(m, w)
        fwd_cells += (s.size - m + 1) * cells(m, w)
    log(f"[c5] plain: stats {t_stats:.3f} s, forward top-{k} {t_fwd:.3f} s, backward "
        f"{t_bwd:.3f} s ({rec_cells / t_bwd / 1e9:.2f} GCUPS recompute)")
    # batched: one graph launch for all forwards, one call for all backwards
    batch = eng.batch(zqs, bands, pipelines=4)
    rb = batch.search(ds, k=k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rb = batch.search(ds, k=k)
    t_bfwd = time.perf_counter() - t0
    ones = np.ones((len(qs), k), np.float32)
    batch.backward(ds, rb.positions, ones, raw_queries=qs)
    t0 = time.perf_counter()
    grads, _ = batch.backward(ds, rb.positions, ones, raw_queries=qs)
    t_bbwd = time.perf_counter() - t0
    same = bool(np.array_equal(rb.positions, plain_pos))
    batch.close()
    log(f"[c5] batched: forward {t_bfwd:.3f} s, backward {t_bbwd*1e3:.2f} ms "
        f"({rec_cells / t_bbwd / 1e9:.2f} GCUPS recompute), same matches: {same}")
    # template refinement (DBA-style): gradient steps on query 0's template
    q0 = qs[0]
    w0 = int(bands[0])
    avg, sd = eng.avg_std(ds, q0.size)
    qt = torch.tensor(q0, dtype=torch.float32, requires_grad=True)
    opt = torch.optim.SGD([qt], lr=0.05)
    losses = []
    t0 = time.perf_counter()
    for _ in range(50):
        opt.zero_grad()
        d, _ = topk_cdtw(qt, eng, ds, avg, sd, w0, k, q0.size // 2)
        loss = d.sum()
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    t_dba = time.perf_counter() - t0
    log(f"[c5] refinement: loss {losses[0]:.4f} -> {losses[-1]:.4f} in 50 steps, {t_dba:.2f} s")
    return {"workload": "ECG-like 2^24, 256 queries |Q| 128-512, W=10%, loss = sum of top-8 "
                        "distances, backward dLoss/dQuery (raw query, through z-normalisation)",
            "plain": {"stats_s": round(t_stats, 4), "forward_s": round(t_fwd, 4),
                      "backward_s": round(t_bwd, 4),
                      "forward_effective_gcups": round(fwd_cells / t_fwd / 1e9, 2),
                      "backward_recompute_gcups": round(rec_cells / t_bwd / 1e9, 3)},
            "batched": {"forward_s": round(t_bfwd, 4), "backward_s": round(t_bbwd, 5),
                        "forward_effective_gcups": round(fwd_cells / t_bfwd / 1e9, 2),
                        "backward_recompute_gcups": round(rec_cells / t_bbwd / 1e9, 3),
                        "backward_over_forward": round(t_bbwd / t_bfwd, 5),
                        "same_matches_as_plain": same},
            "refinement": {"steps": 50, "query": 0, "m": int(q0.size), "w": w0,
                           "loss_first": round(losses[0], 5), "loss_last": round(losses[-1], 5),
                           "seconds": round(t_dba, 3)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=27)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sections", default="c2,c1,c3,c4,c5")
    ap.add_argument("--c4", choices=["sample", "full", "short", "off"], default="sample")
    ap.add_argument("--c4-stride", type=int, default=32)
    ap.add_argument("--c4-split", type=int, default=0,
                    help="c4: batches of this many queries (0: one batch)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one 1 GiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the search path has no multi-GPU code "
              "(it would shard by window start, one stream slice plus an M-1 halo per GPU).",
              file=sys.stderr)
        sys.exit(2)
    sections = set(args.sections.split(","))

    import torch

    from cudadtw_amd import Engine

    eng = Engine(0)
    name, cus = eng.info()
    m, w = 256, 26
    extra = {}
    c2, stream, q_raw, zq = section_c2(eng, torch, args)
    main_res = c2.pop("main")
    L = c2["L"]
    if "c1" in sections:
        extra["c1"] = section_c1(eng, torch)
    if "c3" in sections:
        extra["c3"] = section_c3(eng, torch)
    if "c4" in sections and args.c4 != "off":
        extra["c4"] = section_c4(eng, torch, args)
    if "c5" in sections:
        extra["c5"] = section_c5(eng, torch, args)

    flops_per_window = 5 * cells(m, w) + 2 * m
    achieved = L * flops_per_window / (main_res["dp_ms"] / 1e3) / 1e12
    # HBM bytes per window of this kernel at this shape, measured by rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (calibrated against a device copy).
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "dp_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("m") == m and t.get("w") == w:
            traffic = round(t["bytes_per_window"] * L)
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
                   "n": c2["n"], "m": m, "w": w, "device": name, "cus": cus},
        "roofline": {"bound": "valu", "achieved": round(achieved, 2),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                     "traffic_unit": "bytes per launch (HBM, rocprofv3 PMC)",
                     "kernel": "band_kernel<26,true> (register-band cDTW)",
                     "flops_per_window": flops_per_window},
        "c2": {k: v for k, v in c2.items() if k not in ("L", "n", "m", "w")},
    }
    out.update(extra)
    if not args.no_cpu:
        log("CPU baseline ...")
        cb, comparators = cpu_baseline(stream, zq, q_raw, m, w)
        out["cpu_baseline"] = cb
        out["cpu_comparators"] = comparators
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
