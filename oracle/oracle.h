/*
 * oracle.h -- CPU restatement of the reference's z-normalised subsequence
 * search path (gravitino/cudadtw, include/ *.cuh), plus a port of the UCR
 * Suite cascade the reference benchmarks against.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (cudadtw_amd, libmdtw.so) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Floating-point
 * policy: compiled with -ffp-contract=off; every fused multiply-add is an
 * explicit fmaf()/fma(), so the rounding sequence is fixed and identical to
 * the HIP kernels that claim bit-exact parity with it.
 *
 * Parity-pinning status of each function (what checks it, in tests/):
 *   or_znormalize_f32/f64   closed form: mean 0, population sigma 1 (FP64).
 *   or_envelope             brute-force clamped max/min, exact equality.
 *   or_stats_fp64           numpy two-pass per-window mean/std (independent).
 *   or_stats_refcompat      exact small-integer data (FP32 finalize exact),
 *                           epoch-boundary invariance, constant series -> 1.
 *   or_lb_kim               generic FP64 corner-block DTW on tiny inputs;
 *                           LB_Kim <= cDTW on every window.
 *   or_lb_keogh             numpy closed form; LB_Keogh <= cDTW.
 *   or_cdtw                 FP64 full-matrix brute force; W=0 -> ED^2;
 *                           W >= M-1 -> textbook unconstrained DTW.
 *   or_ldtw                 path property check (every move is the greedy
 *                           choice under the reference's tie rules, the value
 *                           is the path's cost) and LDTW >= cDTW.
 *   or_ldtw_raw             W = 0 closed form (FP32 running sum of squared
 *                           differences); equal to or_ldtw with mean 0 and
 *                           sigma 1; LDTW >= FP64 cDTW of the raw values.
 *   or_ed_f64               numpy direct sum; UCR_ED_BIN location (golden).
 *   or_topk_greedy          brute-force enumeration on tiny inputs.
 *   or_cdtw_grad_f64        FP64 central differences of or_cdtw_f64 (through
 *                           or_znormalize_f64 for the raw-query form); W = 0
 *                           closed form 2(q - z); path cost equals or_cdtw_f64.
 *   or_ucr_dtw / or_ucr_ed  UCR_DTW_BIN / UCR_ED_BIN compiled from the
 *   (cascade.c)             reference sources (oracle/_ref), location golden.
 */
#ifndef MDTW_ORACLE_H
#define MDTW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* metric: 1 = squared Euclidean local cost, 0 = Manhattan (|d|). */

/* stats.cuh:8-28, value_t = float (GPU-dtw.cu:238). */
void or_znormalize_f32(float* x, int32_t n);
/* stats.cuh:8-28, value_t = double (GPU-ed.cu:62). */
void or_znormalize_f64(double* x, int32_t n);

/* bounds.cuh:13-64 (Lemire two-deque envelope). */
void or_envelope(const float* q, int32_t w, float* lo, float* up, int32_t m);

/* stats.cuh:85-142: FP64 prefix per 10^7-sample epoch with M-1 overlap,
 * FP64 window difference / M cast to FP32, FP32 finalize fmaf(-X, X, X2),
 * sigma := 1 when the FP32 variance is not positive.  epoch <= 0 means the
 * reference's 10,000,000. */
void or_stats_refcompat(const float* s, int64_t n, int32_t m, int64_t epoch,
                        float* avg, float* sd);

/* Accurate window statistics: two-pass FP64 mean and population sigma per
 * window; sigma := 1 where the variance is exactly 0 (stats.cuh:50 guard).
 * avg64/sd64 may be NULL. */
void or_stats_fp64(const float* s, int64_t n, int32_t m, float* avg, float* sd,
                   double* avg64, double* sd64);

/* bounds.cuh:152-281 (idx != NULL) and :283-407 (idx == NULL: all windows). */
void or_lb_kim(const float* s, const float* avg, const float* sd,
               const float* zq, int32_t m, const int32_t* idx, int64_t count,
               int32_t metric, float* out);

/* bounds.cuh:111-144 (gathered) / :70-109 (all windows). */
void or_lb_keogh(const float* s, const float* avg, const float* sd,
                 const float* lo, const float* up, int32_t m,
                 const int32_t* idx, int64_t count, int32_t metric, float* out);

/* cdtw.cuh:23-88 sequential_cdtw.  fused = 1 rounds each cell once
 * (fmaf(d, d, min), the reference GPU kernels' arithmetic, cdtw.cuh:339-350);
 * fused = 0 rounds the square then the add (the reference's host build).
 * Output is the squared distance D[M][M]. */
void or_cdtw(const float* zq, const float* s, const float* avg, const float* sd,
             int32_t m, const int32_t* idx, int64_t count, int32_t w,
             int32_t metric, int32_t fused, float* out);

/* FP64 companion of or_cdtw (z-normalisation and DP in double). */
void or_cdtw_f64(const double* zq, const float* s, const double* avg,
                 const double* sd, int32_t m, const int32_t* idx, int64_t count,
                 int32_t w, int32_t metric, double* out);

/* Gradient of  sum_p g[p] * dist_p  with respect to the query, FP64
 * throughout.  dist_p is the banded cDTW of the z-normalised query zq against
 * window idx[p] (or p) -- the D[M][M] of or_cdtw_f64 (cdtw.cuh:23-88), or its
 * sqrt when sqrt_dist (the reported distance, cdtw.cuh:507).  The DP keeps,
 * per cell, which predecessor gave the minimum (ties: diagonal if it is not
 * above either other, then (i-1, j) if not above (i, j-1), else (i, j-1));
 * the path is walked back from (M-1, M-1) and each cell (i, j) on it adds
 * d cost / d zq_i = 2 (zq_i - z_j) (metric 1) or sign(zq_i - z_j) (metric 0).
 * sqrt_dist with D = 0 contributes 0 (the sqrt is not differentiable there).
 * q_raw NULL: grad = d/d zq.  Otherwise zq must be or_znormalize_f64(q_raw)
 * and grad is carried through the z-normalisation (population sigma):
 * dL/dq = (G - mean(G) - zq * mean(G * zq)) / sigma.  dist may be NULL. */
void or_cdtw_grad_f64(const double* zq, const float* s, const double* avg,
                      const double* sd, int32_t m, const int32_t* idx,
                      int64_t count, int32_t w, int32_t metric,
                      int32_t sqrt_dist, const double* g, const double* q_raw,
                      double* dist, double* grad);

/* or_cdtw_grad_f64 (path32 = 0), or with path32 = 1 the path of the FP32 DP
 * of or_cdtw (fused = 1; z = (s - avg) / sd and d = q - z in FP32, zq and the
 * statistics taken as FP32), i.e. the path an FP32 implementation follows;
 * the gradient terms are still summed in FP64.  dist then gets the FP32
 * distance. */
void or_cdtw_grad(const double* zq, const float* s, const double* avg,
                  const double* sd, int32_t m, const int32_t* idx,
                  int64_t count, int32_t w, int32_t metric,
                  int32_t sqrt_dist, int32_t path32, const double* g,
                  const double* q_raw, double* dist, double* grad);

/* ldtw.cuh:74-127 z_serial_ldtw, for windows idx[0..count) (idx NULL: 0..).
 * moves (optional, count*(2m-2) bytes): 0 = diagonal, 1 = down, 2 = right,
 * 3 = no step (path already at the end). */
void or_ldtw(const float* zq, const float* s, const float* avg, const float* sd,
             int32_t m, const int32_t* idx, int64_t count, int32_t w,
             int32_t metric, float* out, uint8_t* moves);

/* ldtw.cuh:18-71 serial_ldtw (no z-normalisation), windows idx[0..count). */
void or_ldtw_raw(const float* q, const float* s, int32_t m, const int32_t* idx,
                 int64_t count, int32_t w, int32_t metric, float* out);

/* z-normalised Euclidean distance squared, all in FP64 (the quantity
 * ed.cuh:61-69,149 computes through cuFFT: sqrt(2M + bsf) = ED). */
void or_ed_f64(const double* zq, const float* s, int64_t n, int32_t m,
               const double* avg64, const double* sd64, double* out);

/* ldtw.cuh:177-192 determine_min: strict '<', lowest index wins ties. */
int64_t or_argmin(const float* x, int64_t n);

/* Greedy top-k with exclusion zone: repeatedly take the smallest (distance,
 * index) not within |i - j| < excl of an earlier pick.  Returns the number of
 * picks (< k only when every window is excluded). */
int32_t or_topk_greedy(const float* d, int64_t n, int32_t k, int64_t excl,
                       int64_t* out_idx, float* out_dist);

/* FP64 streams: two-pass FP64 window statistics and the z-normalised squared
 * ED of every window (ed.cuh:61-69's quantity, computed directly). */
void or_stats_f64d(const double* s, int64_t n, int32_t m, double* avg, double* sd);
void or_ed_f64d(const double* zq, const double* s, int64_t n, int32_t m, const double* avg,
                const double* sd, double* out);

/* OpenMP threads of the functions above (1: the reference's one-core loops). */
void or_set_threads(int32_t n);
int32_t or_max_threads(void);

/* cascade.c: the UCR-suite search restated from its paper (KDD 2012), pinned
 * to UCR_DTW_BIN / UCR_ED_BIN compiled from the reference sources.  Best
 * window of query (z-normalised here) in data with band r; *dist gets the
 * squared distance; counts (optional) = {LB_Kim, LB_Keogh EQ, LB_Keogh EC
 * prunes, windows}. */
int64_t or_ucr_dtw(const double* data, int64_t n, const double* query,
                   int32_t m, int32_t r, double* dist, int64_t* counts);
/* The same on `threads` host threads sharing one best-so-far. */
int64_t or_ucr_dtw_par(const double* data, int64_t n, const double* query, int32_t m,
                       int32_t r, int32_t threads, double* dist, int64_t* counts);
/* Early-abandoning z-normalised ED with the query reordered by |q|. */
int64_t or_ucr_ed(const double* data, int64_t n, const double* query,
                  int32_t m, double* dist);

#ifdef __cplusplus
}
#endif

#endif
