"""AddressSanitizer and UndefinedBehaviorSanitizer over host code, on the CPU.

- The oracle (oracle/oracle.c, oracle/cascade.c) built with
  -fsanitize=address,undefined and driven through every function on small
  inputs, including the awkward ones: bands at and beyond M - 1, epochs
  shorter than the query, a window list, non-finite samples, gradients,
  top-k with more picks than windows.
- The host-only entry points of the C ABI (include/mdtw.h) -- the
  z-normalisation and envelope helpers, status strings, and every entry
  point's argument checks that return before touching a device -- from
  libmdtw's own sources with the host side instrumented
  (hipcc -Xarch_host -fsanitize=...), device code unchanged.
Any sanitizer report fails the run (halt_on_error, no recovery).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")

ORACLE_DRIVER = r'''
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include "oracle.h"

static float* walk(int n, unsigned s) {
    float* x = malloc(sizeof(float) * n); float v = 0;
    for (int i = 0; i < n; ++i) { s = s * 1664525u + 1013904223u; v += (float)((s >> 8) % 2001) / 1000.0f - 1.0f; x[i] = v; }
    return x;
}
int main(void) {
    const int n = 1500, m = 40;
    const int L = n - m + 1;
    float* s = walk(n, 1);
    float* q = walk(m, 2);
    or_znormalize_f32(q, m);
    double* q64 = malloc(sizeof(double) * m);
    for (int i = 0; i < m; ++i) q64[i] = q[i];
    or_znormalize_f64(q64, m);
    float *lo = malloc(4 * m), *up = malloc(4 * m);
    for (int w = 0; w < m + 5; w += 7) or_envelope(q, w, lo, up, m);
    or_envelope(q, 5, lo, up, m);
    float *avg = malloc(4 * L), *sd = malloc(4 * L), *out = malloc(4 * L);
    double *a64 = malloc(8 * L), *s64 = malloc(8 * L), *o64 = malloc(8 * L);
    long long epochs[] = {0, m + 2, 3 * m, 1000, 10000000};
    for (int e = 0; e < 5; ++e) or_stats_refcompat(s, n, m, epochs[e], avg, sd);
    or_stats_fp64(s, n, m, avg, sd, a64, s64);
    int* idx = malloc(4 * L);
    for (int i = 0; i < L; ++i) idx[i] = (i * 7919) % L;
    for (int metric = 0; metric < 2; ++metric) {
        or_lb_kim(s, avg, sd, q, m, NULL, L, metric, out);
        or_lb_kim(s, avg, sd, q, m, idx, L, metric, out);
        or_lb_keogh(s, avg, sd, lo, up, m, idx, L, metric, out);
        for (int w = 0; w <= m + 3; w += 13) {
            or_cdtw(q, s, avg, sd, m, idx, L, w, metric, 0, out);
            or_cdtw(q, s, avg, sd, m, NULL, L, w, metric, 1, out);
            or_cdtw_f64(q64, s, a64, s64, m, idx, 50, w, metric, o64);
            uint8_t* mv = malloc((size_t)L * (2 * m - 2));
            or_ldtw(q, s, avg, sd, m, NULL, L, w, metric, out, mv);
            free(mv);
            or_ldtw_raw(q, s, m, idx, L, w, metric, out);
        }
        double* g = malloc(8 * 30); double* grad = malloc(8 * m); double* dist = malloc(8 * 30);
        for (int i = 0; i < 30; ++i) g[i] = 1.0;
        or_cdtw_grad_f64(q64, s, a64, s64, m, idx, 30, 5, metric, 1, g, q64, dist, grad);
        or_cdtw_grad(q64, s, a64, s64, m, idx, 30, 5, metric, 1, 1, g, NULL, dist, grad);
        free(g); free(grad); free(dist);
    }
    or_ed_f64(q64, s, n, m, a64, s64, o64);
    (void)or_argmin(out, L);
    long long pos[64]; float dd[64];
    (void)or_topk_greedy(out, L, 64, 30, (int64_t*)pos, dd);
    (void)or_topk_greedy(out, 5, 64, 0, (int64_t*)pos, dd);
    double* sd64 = malloc(8 * n);
    for (int i = 0; i < n; ++i) sd64[i] = s[i];
    double dist; int64_t counts[4];
    (void)or_ucr_dtw(sd64, n, q64, m, 5, &dist, counts);
    (void)or_ucr_dtw(sd64, n, q64, m, m + 10, &dist, counts);
    (void)or_ucr_dtw_par(sd64, n, q64, m, 5, 3, &dist, counts);
    (void)or_ucr_ed(sd64, n, q64, m, &dist);
    or_stats_f64d(sd64, n, m, a64, s64);
    or_ed_f64d(q64, sd64, n, m, a64, s64, o64);
    /* non-finite samples */
    s[0] = NAN; s[700] = INFINITY; s[1499] = -INFINITY;
    or_stats_fp64(s, n, m, avg, sd, a64, s64);
    or_cdtw(q, s, avg, sd, m, NULL, L, 5, 1, 1, out);
    or_ldtw(q, s, avg, sd, m, NULL, L, 5, 1, out, NULL);
    or_ldtw_raw(q, s, m, NULL, L, 5, 0, out);
    (void)or_topk_greedy(out, L, 8, 10, (int64_t*)pos, dd);
    puts("oracle driver ok");
    return 0;
}
'''

ABI_DRIVER = r'''
#include <math.h>
#include <stdio.h>
#include "mdtw.h"
int main(void) {
    float q[100], lo[100], up[100];
    double q64[100];
    for (int i = 0; i < 100; ++i) { q[i] = (float)((i * 37) % 11) - 5.0f; q64[i] = q[i]; }
    if (mdtw_znormalize(q, 100) != MDTW_OK || mdtw_znormalize_f64(q64, 100) != MDTW_OK) return 1;
    for (int w = 0; w < 110; w += 9)
        if (mdtw_envelope(q, w, lo, up, 100) != MDTW_OK) return 2;
    if (mdtw_znormalize(NULL, 3) == MDTW_OK || mdtw_envelope(q, -1, lo, up, 100) == MDTW_OK) return 3;
    for (int c = -1; c < 20; ++c) if (!mdtw_status_string(c)) return 4;
    int64_t pos[4]; float dist[4]; double d64;
    /* argument checks that return before any device work */
    if (mdtw_create(0, NULL) == MDTW_OK) return 5;
    if (mdtw_destroy(NULL) == MDTW_OK) return 6;
    if (mdtw_search(NULL, NULL, 10, NULL, NULL, 1, 0, 1, 0, 1, 0, pos, dist, NULL) == MDTW_OK) return 7;
    if (mdtw_search_noalloc(NULL, NULL, 10, NULL, NULL, 1, 0, 1, 0, 1, 0, pos, dist, NULL) == MDTW_OK) return 8;
    if (mdtw_ed_search_f64(NULL, NULL, 10, q64, 100, pos, &d64, NULL) == MDTW_OK) return 9;
    if (mdtw_stream_restore(NULL, q, sizeof(q), NULL) == MDTW_OK) return 10;
    if (mdtw_batch_destroy(NULL) == MDTW_OK) return 11;
    if (mdtw_set_workspace_limit(NULL, 0) == MDTW_OK) return 12;
    puts("abi host driver ok");
    return 0;
}
'''


def run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, **kw)


def test_oracle_under_asan_ubsan(tmp_path):
    src = tmp_path / "driver.c"
    src.write_text(ORACLE_DRIVER)
    exe = tmp_path / "driver"
    od = os.path.join(ROOT, "oracle")
    r = run(["gcc", "-O1", "-g", "-std=c11", "-fopenmp", "-ffp-contract=off", "-mfma", "-mavx2",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", "-I", od, str(src), os.path.join(od, "oracle.c"),
             os.path.join(od, "cascade.c"), "-lm", "-o", str(exe)])
    assert r.returncode == 0, r.stderr
    r = run([str(exe)], env=SAN_ENV, timeout=600)
    assert r.returncode == 0 and "oracle driver ok" in r.stdout, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr


def _library_units():
    """The translation units of libmdtw.so, as the library's Makefile lists them."""
    text = open(os.path.join(ROOT, "cudadtw_amd", "csrc", "Makefile")).read()
    m = re.search(r"^SRCS\s*:=(.*?)(?:\n(?!\s)|\Z)", text, re.S | re.M)
    return [f[:-4] for f in m.group(1).replace("\\", " ").split() if f.endswith(".hip")]


OBJS = _library_units()
SANITIZED = ["engine", "search", "stream", "batch", "noalloc", "ed_fft"]


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="no hipcc")
def test_abi_host_code_under_asan_ubsan(tmp_path):
    """The host side of the library (where argument checks and state
    parsing live) instrumented; the kernels' translation units are linked
    as built (their host side is launch code only)."""
    objdir = os.path.join(ROOT, "build", "obj")
    missing = [o for o in OBJS if not os.path.exists(os.path.join(objdir, o + ".o"))]
    if missing:
        pytest.skip(f"library objects not built ({missing[:3]}...): make -C cudadtw_amd/csrc")
    csrc = os.path.join(ROOT, "cudadtw_amd", "csrc")
    flags = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=all",
             "-I", os.path.join(ROOT, "include"), "-I", csrc]
    objs = []
    for o in OBJS:
        if o in SANITIZED:
            out = tmp_path / (o + ".o")
            r = run(["/opt/rocm/bin/hipcc"] + flags + ["-c", os.path.join(csrc, o + ".hip"),
                                                        "-o", str(out)])
            assert r.returncode == 0, r.stderr[-3000:]
            objs.append(str(out))
        else:
            objs.append(os.path.join(objdir, o + ".o"))
    drv = tmp_path / "abi.c"
    drv.write_text(ABI_DRIVER)
    exe = tmp_path / "abi"
    r = run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-Xarch_host", "-fsanitize=address",
             "-Xarch_host", "-fsanitize=undefined", "-x", "c++", str(drv), "-x", "none"] + objs +
            ["-I", os.path.join(ROOT, "include"), "-o", str(exe)])
    assert r.returncode == 0, r.stderr[-3000:]
    r = run([str(exe)], env=SAN_ENV, timeout=300)
    assert r.returncode == 0 and "abi host driver ok" in r.stdout, r.stdout + r.stderr[-3000:]
