/*
 * mdtw.h -- C ABI of the MI355X (gfx950) subsequence similarity-search
 * engine.  Plain pointers and sizes only; every entry point returns an int
 * status (MDTW_OK = 0) instead of exiting, and never prints.
 *
 * Each declaration names the reference interface it replaces
 * (gravitino/cudadtw, file:line).  The reference is header-only C++ that
 * keeps the query in __constant__ memory (cdtw.cuh:9-11) and prints its
 * result (cdtw.cuh:493-508); here the query lives in an engine object and
 * results are returned.
 *
 * Conventions
 *   - d_* arguments are device pointers (HBM), h_* are host pointers.
 *   - Stream data is FP32 (the *_f64 ED entry points: FP64, as the
 *     reference's ED driver); window statistics are FP32 (mean, population
 *     sigma), computed from FP64 sums.
 *   - Non-finite samples (NaN, +-Inf) in the stream: a window holding one has
 *     NaN statistics (FP64 and raw modes) and so NaN distances and bounds --
 *     never a match, never kept by top-k; other windows are unaffected.
 *     REFCOMPAT restates the reference, whose running prefix carries such a
 *     sample into every later window of its epoch.  Queries must be finite
 *     (MDTW_ERR_NOT_FINITE).
 *   - Window i is series[i .. i+M); there are L = N - M + 1 windows.
 *   - DP outputs are squared cDTW distances, as the reference's Cdtw array
 *     (cdtw.cuh:256); reported search distances are sqrt of that
 *     (cdtw.cuh:507).  Ties go to the lowest window index.
 *   - Index arrays on the device are int32 (the reference's index_t = int);
 *     host-facing positions are int64.  One search covers at most 2^31 - 1
 *     windows (MDTW_ERR_UNSUPPORTED beyond; streams: per epoch).
 *   - One engine per host thread.  Calls on one engine are ordered on its
 *     stream; distinct engines may run concurrently.
 */
#ifndef MDTW_H
#define MDTW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define MDTW_OK                    0
#define MDTW_ERR_INVALID_ARGUMENT  1  /* null pointer, negative size          */
#define MDTW_ERR_QUERY_TOO_SHORT   2  /* M < 11 for LB_Kim (bounds.cuh:418)   */
#define MDTW_ERR_QUERY_TOO_LONG    3  /* M > MDTW_MAX_QUERY                   */
#define MDTW_ERR_BAND              4  /* W < 0 or W > M - 1                   */
#define MDTW_ERR_STREAM_TOO_SHORT  5  /* N < M                                */
#define MDTW_ERR_NO_QUERY          6  /* mdtw_set_query not called           */
#define MDTW_ERR_ALLOC             7  /* device or host allocation failed     */
#define MDTW_ERR_HIP               8  /* HIP runtime error                    */
#define MDTW_ERR_NOT_FINITE        9  /* NaN/Inf in the query                 */
#define MDTW_ERR_K                10  /* k < 1 or k > MDTW_MAX_K              */
#define MDTW_ERR_UNSUPPORTED      11  /* option not available                 */
#define MDTW_ERR_WORKSPACE        12  /* request exceeds the workspace cap    */

#define MDTW_MAX_QUERY 8192   /* ed.cuh:10 MAXQUERYSIZE (cdtw.cuh:8 has 4096) */
#define MDTW_MAX_K     1024

/* metric (the reference's template<bool euclidean>, GPU-dtw.cu:7-8) */
#define MDTW_METRIC_L1 0
#define MDTW_METRIC_L2 1

/* window statistics mode */
#define MDTW_STATS_FP64      0  /* tile-local FP64 prefix sums, FP64 finalize */
#define MDTW_STATS_REFCOMPAT 1  /* stats.cuh:85-142: FP64 epoch prefix, FP32
                                   finalize fmaf(-X, X, X2), sigma := 1 if <= 0 */
#define MDTW_STATS_NONE      2  /* no z-normalisation: mean 0, sigma 1 for every
                                   window, so windows are compared raw (the
                                   reference's serial_ldtw / tsh2_ldtw,
                                   ldtw.cuh:18-71,198-291); the query is used
                                   as given */

/* search algorithm */
#define MDTW_ALGO_CDTW 0  /* cdtw.cuh:510-592 prune_cdtw            */
#define MDTW_ALGO_ED   1  /* ed.cuh:71-153 calculate_ed            */
#define MDTW_ALGO_LDTW 2  /* ldtw.cuh:294-391 ztsh2_ldtw + argmin  */
#define MDTW_ALGO_ED_FFT 3 /* ed.cuh:71-153 calculate_ed's FFT form: FP64
                              cross-correlation, for long queries (M <= 8192;
                              statistics MDTW_STATS_FP64, metric L2) */

typedef struct mdtw_engine mdtw_engine;

/* What a search did, stage by stage (UCR_DTW_BIN.cpp:738-741 prints the same
 * prune fractions; cdtw.cuh:576 has the commented-out survival print). */
typedef struct mdtw_search_info {
    int64_t windows;        /* L = N - M + 1                                 */
    int64_t seeded;         /* windows evaluated exactly to seed the bound   */
    int64_t lb_kim_pruned;  /* windows dropped by LB_Kim                     */
    int64_t lb_keogh_pruned;/* windows dropped by LB_Keogh                   */
    int64_t dp_evaluated;   /* windows whose DP (or LDTW/ED) ran             */
    int64_t dp_abandoned;   /* DP runs stopped early by the row-minimum test */
    int32_t found;          /* matches returned (<= k)                       */
    float   threshold;      /* final pruning threshold (squared distance)    */
    float   elapsed_ms;     /* device time of the whole search (HIP events)  */
} mdtw_search_info;

/* ---- engine ------------------------------------------------------------- */
/* Replaces the reference's process-global state: cudaSetDevice (GPU-dtw.cu:19)
 * and the __constant__ query/envelope and texture (cdtw.cuh:9-14). */
int mdtw_create(int device, mdtw_engine** out);
int mdtw_destroy(mdtw_engine* e);
/* Every launch of the engine goes to this stream (hipStream_t; NULL = the
 * default stream).  The reference takes cudaStream_t per call and ignores it
 * in places (stats.cuh:104-108, ed.cuh:111). */
int mdtw_set_stream(mdtw_engine* e, void* hip_stream);
int mdtw_synchronize(mdtw_engine* e);
/* No reference counterpart (the reference mallocs per call, cdtw.cuh:510-592).
 * Caps the device workspace the engine may grow to (bytes; 0 = no cap): a call
 * whose scratch would exceed it returns MDTW_ERR_WORKSPACE before any launch
 * and leaves the engine usable.  A workspace already larger is kept; the cap
 * applies to growth and to each call's need.  Batches and streams allocate
 * their own buffers at creation and are not affected. */
int mdtw_set_workspace_limit(mdtw_engine* e, size_t bytes);
const char* mdtw_status_string(int status);
/* Device name, compute-unit count; for logs. */
int mdtw_device_info(mdtw_engine* e, char* name, int32_t name_len, int32_t* cus);

/* ---- query -------------------------------------------------------------- */
/* Replaces cudaMemcpyToSymbol(Czquery/Czlower/Czupper) (GPU-dtw.cu:82-84).
 * h_zquery is the z-normalised query (mdtw_znormalize); h_lower/h_upper are
 * its envelope for band w (mdtw_envelope), or NULL to have the engine build
 * it for band w.  The query is copied; the host arrays may be reused. */
int mdtw_set_query(mdtw_engine* e, const float* h_zquery, const float* h_lower,
                   const float* h_upper, int32_t m, int32_t w);

/* ---- host helpers (bit-exact restatements) ------------------------------ */
/* stats.cuh:8-28 znormalize<float>: FP32 accumulation in index order
 * (the cDTW/LDTW drivers, GPU-dtw.cu:67, GPU-ldtw.cu:52). */
int mdtw_znormalize(float* h_series, int32_t len);
/* stats.cuh:8-28 znormalize<double> (the ED driver, GPU-ed.cu:62). */
int mdtw_znormalize_f64(double* h_series, int32_t len);
/* bounds.cuh:13-64 envelope (Lemire). */
int mdtw_envelope(const float* h_series, int32_t w, float* h_lower,
                  float* h_upper, int32_t m);

/* ---- window statistics -------------------------------------------------- */
/* stats.cuh:115-142 avg_std<double>(Series, AvgS, StdS, M, N, stream). */
int mdtw_avg_std(mdtw_engine* e, const float* d_series, int64_t n, int32_t m,
                 float* d_avg, float* d_std, int32_t mode);

/* ---- lower bounds (need mdtw_set_query) ---------------------------------- */
/* bounds.cuh:413-441 lb_kim: d_indices NULL = all count windows 0..count-1
 * (register_lb_kim), else the gathered windows (random_lb_kim).  M >= 11. */
int mdtw_lb_kim(mdtw_engine* e, const float* d_series, const float* d_avg,
                const float* d_std, float* d_out, const int32_t* d_indices,
                int64_t count, int32_t metric);
/* bounds.cuh:443-465 lb_keogh against the query envelope. */
int mdtw_lb_keogh(mdtw_engine* e, const float* d_series, const float* d_avg,
                  const float* d_std, float* d_out, const int32_t* d_indices,
                  int64_t count, int32_t metric);

/* ---- distance profiles (need mdtw_set_query) ----------------------------- */
/* cdtw.cuh:450-487 gpu_cdtw: squared banded cDTW of the query against the
 * windows d_indices[0..count) (NULL = windows 0..count-1), band w. */
int mdtw_cdtw(mdtw_engine* e, const float* d_series, const float* d_avg,
              const float* d_std, float* d_out, const int32_t* d_indices,
              int64_t count, int32_t w, int32_t metric);
/* ldtw.cuh:294-391 ztsh2_ldtw: greedy LDTW of every window 0..count-1.
 * d_avg = d_std = NULL: raw windows, ldtw.cuh:198-291 tsh2_ldtw (the query
 * as given to mdtw_set_query, no z-normalisation). */
int mdtw_ldtw(mdtw_engine* e, const float* d_series, const float* d_avg,
              const float* d_std, float* d_out, int64_t count, int32_t w,
              int32_t metric);
/* ed.cuh:61-69 finalize: squared z-normalised ED of windows 0..count-1
 * (the reference stores 2M-shifted values, ed.cuh:68; this stores ED^2),
 * summed in FP64: the exact direct-form distance for the statistics given,
 * rounded once to FP32. */
int mdtw_ed(mdtw_engine* e, const float* d_series, const float* d_avg,
            const float* d_std, float* d_out, int64_t count);

/* ed.cuh:71-153 (calculate_ed's FFT cross-correlation, FP64): squared
 * z-normalised ED of every window 0..n-M, into d_out[n-M+1], with FP64 window
 * statistics computed here (stats_mode MDTW_STATS_FP64 only).  For M up to
 * 8192; faster than mdtw_ed for long queries. */
int mdtw_ed_fft(mdtw_engine* e, const float* d_series, int64_t n,
                int32_t stats_mode, float* d_out);

/* ---- FP64 streams (the reference's ED driver, GPU-ed.cu:20-82) ----------- */
/* stats.cuh:115-142 avg_std<double> on an FP64 stream: FP64 mean and
 * population sigma of every window (sigma := 1 for a constant window), from
 * FP64 tile-local sums. */
int mdtw_avg_std_f64(mdtw_engine* e, const double* d_series, int64_t n, int32_t m,
                     double* d_avg, double* d_std);
/* ed.cuh:71-153 calculate_ed's profile on an FP64 stream: d_out[n-M+1] FP64
 * squared z-normalised ED of every window against h_zquery (M FP64 values,
 * z-normalised with mdtw_znormalize_f64; M <= 8192), through FP64 FFT
 * cross-correlation with FP64 window statistics computed here. */
int mdtw_ed_f64(mdtw_engine* e, const double* d_series, int64_t n, const double* h_zquery,
                int32_t m, double* d_out);
/* ed.cuh:71-153 calculate_ed on an FP64 stream: the best window (lowest index
 * on ties) and its distance sqrt(ED^2), both FP64-computed (the reference
 * prints sqrt(2M + bsf) and the location, ed.cuh:149-150).  At most 2^31 - 1
 * windows. */
int mdtw_ed_search_f64(mdtw_engine* e, const double* d_series, int64_t n,
                       const double* h_zquery, int32_t m, int64_t* h_pos, double* h_dist,
                       mdtw_search_info* info);

/* No reference counterpart: the same profile as mdtw_cdtw, forced through the
 * anti-diagonal kernel that serves bands wider than the register-band kernel
 * (for cross-checking the two kernels against each other). */
int mdtw_cdtw_wide(mdtw_engine* e, const float* d_series, const float* d_avg,
                   const float* d_std, float* d_out, const int32_t* d_indices,
                   int64_t count, int32_t w, int32_t metric);

/* ---- gradient -------------------------------------------------------------- */
/* No reference counterpart (the reference computes distances only; this
 * differentiates its recurrence, cdtw.cuh:23-88).  Gradient with respect to
 * the query of  sum_p d_grad[p] * dist_p,  where dist_p is the banded cDTW
 * of the engine's query against window d_indices[p] (squared, or its sqrt
 * when sqrt_dist, as search distances are reported).  The DP is recomputed
 * (no state is kept from a forward call) and the optimal warping path walked
 * back; ties take the diagonal if it is not above either other predecessor,
 * then (i-1, j) if not above (i, j-1).  With sqrt_dist a zero distance
 * contributes 0.  d_query_raw NULL: the output d_grad_query (M floats) is the
 * gradient with respect to the z-normalised query; otherwise d_query_raw must
 * be the raw query whose z-normalisation was given to mdtw_set_query, and the
 * gradient is carried through that z-normalisation (population sigma) to the
 * raw query.  d_dist (may be NULL) receives dist_p, bit-identical to
 * mdtw_cdtw (sqrt applied when sqrt_dist).  All arrays are device memory. */
int mdtw_cdtw_backward(mdtw_engine* e, const float* d_series, const float* d_avg,
                       const float* d_std, const int32_t* d_indices,
                       const float* d_grad, int64_t count, int32_t w,
                       int32_t metric, int32_t sqrt_dist,
                       const float* d_query_raw, float* d_dist,
                       float* d_grad_query);

/* ---- search -------------------------------------------------------------- */
/* cdtw.cuh:510-592 prune_cdtw / ed.cuh:71-153 calculate_ed / GPU-ldtw.cu:
 * best (k = 1) or top-k matches of the engine's query in d_series[0..n).
 * d_avg/d_std: window statistics for M (NULL = computed here in stats_mode).
 * excl: exclusion zone for k > 1 (a match excludes windows j with
 * |i - j| < excl); ignored for k = 1.  Outputs are host arrays of k entries:
 * positions (ascending distance) and distances (sqrt for L2, as
 * cdtw.cuh:507).  Unused entries get position -1.  info may be NULL. */
int mdtw_search(mdtw_engine* e, const float* d_series, int64_t n,
                const float* d_avg, const float* d_std, int32_t w,
                int32_t algo, int32_t metric, int32_t stats_mode, int32_t k,
                int64_t excl, int64_t* h_pos, float* h_dist,
                mdtw_search_info* info);

/* No reference counterpart; for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * mdtw_search with the same arguments and results, except that it allocates
 * nothing: its scratch (window statistics, candidate lists, kept keys, the
 * padded query, the selected matches) lives in a fixed device array of
 * 128 MiB that each GPU the library runs on holds until the process ends.
 * Limits: algo CDTW, ED or LDTW (ED_FFT: MDTW_ERR_UNSUPPORTED); stats_mode
 * FP64 or NONE, or statistics passed in (REFCOMPAT: MDTW_ERR_UNSUPPORTED);
 * about 16 bytes per window (8 with d_avg/d_std given; 24 for top-k), so up
 * to ~8.3M windows with statistics computed here, or ~5.5M for top-k --
 * beyond: MDTW_ERR_WORKSPACE, before any launch; more than 8192 windows at or
 * below the final top-k bound: MDTW_ERR_WORKSPACE (mdtw_search selects those
 * on the host).  All calls on one GPU share its array: calls that overlap on
 * one GPU (engines on different host threads) race. */
int mdtw_search_noalloc(mdtw_engine* e, const float* d_series, int64_t n,
                        const float* d_avg, const float* d_std, int32_t w,
                        int32_t algo, int32_t metric, int32_t stats_mode, int32_t k,
                        int64_t excl, int64_t* h_pos, float* h_dist,
                        mdtw_search_info* info);

/* ---- batched search ------------------------------------------------------- */
/* No reference counterpart (the reference holds one query per process in
 * __constant__ memory, cdtw.cuh:9-11).  A batch is a set of queries of mixed
 * length and band searched together against one stream: every query's search
 * is the mdtw_search algorithm (results identical to searching the queries
 * one by one), run on several concurrent pipelines and captured into one
 * hipGraph that each mdtw_batch_search replays.  The batch belongs to the
 * engine it was created on; destroy it before the engine. */
typedef struct mdtw_batch mdtw_batch;
/* h_zqueries: the nq z-normalised queries (mdtw_znormalize) concatenated;
 * h_lengths[i] their lengths (1..MDTW_MAX_QUERY), h_bands[i] their absolute
 * bands (clamped to M - 1; at most 1023).  pipelines: concurrent pipelines
 * (0 = 4, at most 16).  The host arrays are copied. */
int mdtw_batch_create(mdtw_engine* e, const float* h_zqueries,
                      const int32_t* h_lengths, const int32_t* h_bands,
                      int32_t nq, int32_t pipelines, mdtw_batch** out);
int mdtw_batch_destroy(mdtw_batch* b);
/* Picks launch shapes per bucket of queries (power-of-two length class x
 * band class: register band W <= 112, LDS-tail / two-wave split band
 * W <= 240, anti-diagonal above) by timing the bucket's median query on a
 * stream prefix of about cell_budget band cells (<= 0: 2^39): the persistent
 * DP grid, the anti-diagonal DP grid and the cascade grid, then the DP's
 * formulation where more than one applies (MDTW_DP_FORM_*).  Affects speed
 * only. */
int mdtw_batch_tune(mdtw_batch* b, const float* d_series, int64_t n,
                    int32_t metric, int32_t k, int64_t cell_budget,
                    int32_t* h_buckets_tuned);
#define MDTW_DP_FORM_AUTO 0  /* register band, column-minimum abandon (anti-diagonal above W = 240) */
#define MDTW_DP_FORM_LB   1  /* register band, also the cumulative LB_Keogh bound (one lane per window) */
#define MDTW_DP_FORM_WIDE 2  /* anti-diagonal kernel, the band's diagonals in LDS */
/* No reference counterpart: the shape query i's searches use (default or
 * tuned): h_shape[0..4) = persistent DP workgroups per CU, anti-diagonal
 * waves per CU, cascade workgroups per CU, MDTW_DP_FORM_*. */
int mdtw_batch_shape(const mdtw_batch* b, int32_t i, int32_t* h_shape);
/* No reference counterpart: sets the shape of query i's bucket (h_shape as
 * above: 1..64 DP workgroups per CU, 1..256 anti-diagonal waves per CU,
 * 1..256 cascade workgroups per CU, an MDTW_DP_FORM_*), as a tuning result
 * kept from an earlier run would.  Affects speed only. */
int mdtw_batch_set_shape(mdtw_batch* b, int32_t i, const int32_t* h_shape);
/* Best (k = 1) or top-k cDTW matches of every query (mdtw_search semantics,
 * window statistics computed per query length in FP64).  h_excl: nq
 * exclusion zones, or NULL for M/2 each.  Outputs: h_pos/h_dist nq*k
 * (query-major), h_info nq entries (may be NULL; elapsed_ms is the whole
 * batch), *h_elapsed_ms the device time of the graph (may be NULL). */
int mdtw_batch_search(mdtw_batch* b, const float* d_series, int64_t n,
                      int32_t metric, int32_t k, const int64_t* h_excl,
                      int64_t* h_pos, float* h_dist, mdtw_search_info* h_info,
                      float* h_elapsed_ms);

/* No reference counterpart.  Query gradients for the whole batch: for every
 * query i, the gradient of  sum_j h_grad[i*k + j] * dist(i, h_pos[i*k + j])
 * (positions < 0 skipped; dist as mdtw_cdtw_backward, sqrt when sqrt_dist,
 * window statistics computed as the batch search computes them).
 * h_query_raw: the raw queries concatenated (gradients with respect to them,
 * through each z-normalisation), or NULL (with respect to the z-normalised
 * queries).  h_grad_out: the nq gradients concatenated (sum of lengths);
 * h_dist (nq*k, may be NULL): the recomputed distances (+inf where skipped). */
int mdtw_batch_backward(mdtw_batch* b, const float* d_series, int64_t n,
                        int32_t metric, int32_t k, const int64_t* h_pos,
                        const float* h_grad, int32_t sqrt_dist,
                        const float* h_query_raw, float* h_grad_out,
                        float* h_dist);

/* ---- streaming search ---------------------------------------------------- */
/* A search over a stream that arrives in host chunks, for streams longer than
 * device memory or without an end.  The reference's counterparts are its
 * epoch loops: avg_std over 10^7-sample epochs with M-1 samples of overlap
 * (stats.cuh:115-142) and UCR_DTW_BIN.cpp:579-707, which reads fixed chunks
 * with M-1 overlap and keeps its best-so-far across them.  The answer after
 * mdtw_stream_finish is the one-shot mdtw_search answer over everything
 * pushed, bit for bit; positions are int64 stream positions.  The stream
 * keeps the engine's query at creation (later mdtw_set_query calls do not
 * affect it) and runs on the engine's stream; it allocates at creation only
 * (two epoch buffers, its own search workspace). */
typedef struct mdtw_stream mdtw_stream;
/* epoch_windows: windows per device epoch, a multiple of 1024 (0: 2^24).
 * stats_mode: MDTW_STATS_FP64 or MDTW_STATS_NONE (REFCOMPAT statistics need
 * the whole stream's prefix: MDTW_ERR_UNSUPPORTED). */
int mdtw_stream_create(mdtw_engine* e, int32_t w, int32_t algo, int32_t metric,
                       int32_t stats_mode, int32_t k, int64_t excl,
                       int64_t epoch_windows, mdtw_stream** out);
int mdtw_stream_destroy(mdtw_stream* s);
/* Appends count host samples (any count; the buffer may be reused on
 * return).  Every completed epoch is searched while the next is copied. */
int mdtw_stream_push(mdtw_stream* s, const float* h_samples, int64_t count);
/* End of stream: searches the last partial epoch.  No push after it. */
int mdtw_stream_finish(mdtw_stream* s);
/* The answer over the windows searched so far (k entries, -1 / +inf past the
 * matches found); info->windows counts those windows. */
int mdtw_stream_result(mdtw_stream* s, int64_t* h_pos, float* h_dist,
                       mdtw_search_info* info);
/* Samples pushed so far and windows searched or in flight. */
int mdtw_stream_position(mdtw_stream* s, int64_t* samples_pushed,
                         int64_t* windows_done);
/* Resume state: h_buf NULL sets *size only; else writes *size <= cap bytes. */
int mdtw_stream_save(mdtw_stream* s, void* h_buf, size_t cap, size_t* size);
/* A stream continuing from a saved state; the engine must hold the same
 * query (checked by hash: MDTW_ERR_INVALID_ARGUMENT otherwise). */
int mdtw_stream_restore(mdtw_engine* e, const void* h_buf, size_t size,
                        mdtw_stream** out);

/* ---- primitives (cub_util.cuh) ------------------------------------------- */
/* cub_util.cuh:154-172 threshold: keeps d_indices[i] with d_values[i] <= thr,
 * in order, in place; the new length goes to *h_count. */
int mdtw_threshold(mdtw_engine* e, const float* d_values, int32_t* d_indices,
                   int64_t count, float thr, int64_t* h_count);
/* cub_util.cuh:7-27 reduce(Min) + report_result (cdtw.cuh:493-508): lowest
 * value and its lowest index among d_values[0..count). */
int mdtw_argmin(mdtw_engine* e, const float* d_values, int64_t count,
                int64_t* h_index, float* h_value);

#ifdef __cplusplus
}
#endif

#endif /* MDTW_H */
