/*
 * mdtw_compat.hpp -- the reference's C++ host entry points, as inline
 * forwarders over the C ABI of mdtw.h.
 *
 * gravitino/cudadtw is header-only CUDA C++ whose drivers (GPU-dtw.cu,
 * GPU-ed.cu, GPU-ldtw.cu) call these templates.  This header declares the
 * same names with the same parameter order and meaning, so a driver ports by
 * swapping its includes and two set-up lines:
 *
 *     cudaMemcpyToSymbol(::Czlower/Czupper/Czquery, ...)   (GPU-dtw.cu:82-84)
 *  -> mdtw_compat::set_query(zquery, zlower, zupper, M, W);
 *     cudaBindTexture(0, Tsubject, Subject, ...)           (GPU-dtw.cu:46)
 *  -> (nothing: the stream is read through its pointer)
 *
 * Device buffers are plain device pointers (hipMalloc'ed by the caller, as
 * the reference's are cudaMalloc'ed).  Streams are passed as void* (a
 * hipStream_t converts implicitly; 0 = the engine's current stream).
 *
 * Differences, on purpose:
 *   - errors throw mdtw_compat::error (the reference prints and exit(1)s
 *     through CUERR, cuda_def.cuh);
 *   - prune_cdtw / calculate_ed print the reference's result lines and
 *     also return the match (the reference returns void / the raw bsf);
 *   - ties go to the lowest window index (the reference's go to LB_Kim
 *     order, cdtw.cuh:532-533);
 *   - avg_std uses FP64 tile-local prefix sums and an FP64 finalize unless
 *     MDTW_COMPAT_REFSTATS is defined, which selects the restatement of the
 *     reference's epoch prefix + FP32 finalize (stats.cuh:85-142).
 *
 * This header needs only <mdtw.h> and the C++ standard library; link with
 * libmdtw.so.
 */
#ifndef MDTW_COMPAT_HPP
#define MDTW_COMPAT_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mdtw.h"

namespace mdtw_compat {

struct error : std::runtime_error {
    int status;
    error(int s, const std::string& where)
        : std::runtime_error(where + ": " + mdtw_status_string(s)), status(s) {}
};

inline void check(int status, const char* where) {
    if (status != MDTW_OK) throw error(status, where);
}

// The process-wide engine: the reference keeps its query in process-global
// __constant__ memory (cdtw.cuh:9-11), so its API has no handle either.
inline mdtw_engine*& engine_slot() {
    static mdtw_engine* e = nullptr;
    return e;
}
// Device 0 on first use (GPU-dtw.cu:19 cudaSetDevice(0)); use_engine to pick
// another engine (the caller keeps ownership of it).
inline mdtw_engine* engine() {
    mdtw_engine*& e = engine_slot();
    if (!e) check(mdtw_create(0, &e), "mdtw_create");
    return e;
}
inline void use_engine(mdtw_engine* e) { engine_slot() = e; }

inline mdtw_engine* on_stream(void* stream) {
    mdtw_engine* e = engine();
    if (stream) check(mdtw_set_stream(e, stream), "mdtw_set_stream");
    return e;
}

// Replaces cudaMemcpyToSymbol(Czquery/Czlower/Czupper) (GPU-dtw.cu:82-84):
// the z-normalised query and its envelope for band W.
inline void set_query(const float* zquery, const float* zlower, const float* zupper, int M,
                      int W) {
    check(mdtw_set_query(engine(), zquery, zlower, zupper, M, W), "mdtw_set_query");
}

// Replaces cudaMemcpyToSymbol(::Czquery, zquery, sizeof(double) * M)
// (GPU-ed.cu:72): the FP64 z-normalised query of the ED driver, kept here for
// calculate_ed on FP64 streams.
inline std::vector<double>& query_f64() {
    static std::vector<double> q;
    return q;
}
inline void set_query(const double* zquery, int M) {
    if (!zquery || M < 1) throw error(MDTW_ERR_INVALID_ARGUMENT, "set_query");
    query_f64().assign(zquery, zquery + M);
}

struct match {
    int64_t location;
    double distance;
};

template <class T>
using req_float = typename std::enable_if<std::is_same<T, float>::value, int>::type;
template <class T>
using req_double = typename std::enable_if<std::is_same<T, double>::value, int>::type;

}  // namespace mdtw_compat

// ---- stats.cuh ----------------------------------------------------------------

// stats.cuh:8-28 znormalize (population sigma; float or double).
template <class value_t, class index_t>
inline void znormalize(value_t* const series, const index_t L) {
    static_assert(std::is_same<value_t, float>::value || std::is_same<value_t, double>::value,
                  "znormalize: float or double series");
    if (std::is_same<value_t, float>::value)
        mdtw_compat::check(mdtw_znormalize(reinterpret_cast<float*>(series),
                                           static_cast<int32_t>(L)),
                           "znormalize");
    else
        mdtw_compat::check(mdtw_znormalize_f64(reinterpret_cast<double*>(series),
                                               static_cast<int32_t>(L)),
                           "znormalize");
}

// stats.cuh:115-142 avg_std<prefx_t>(Series, AvgS, StdS, M, N, stream):
// window mean and sigma of every window of the device stream.  prefx_t must
// be double (the reference's FP64 prefix sums).
template <class prefx_t, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void avg_std(const value_t* const Series, value_t* const AvgS, value_t* const StdS,
                    const index_t M, const index_t N, void* stream = nullptr) {
    static_assert(std::is_same<prefx_t, double>::value, "avg_std: prefix sums are FP64");
#ifdef MDTW_COMPAT_REFSTATS
    const int32_t mode = MDTW_STATS_REFCOMPAT;
#else
    const int32_t mode = MDTW_STATS_FP64;
#endif
    mdtw_compat::check(mdtw_avg_std(mdtw_compat::on_stream(stream), Series,
                                    static_cast<int64_t>(N), static_cast<int32_t>(M), AvgS, StdS,
                                    mode),
                       "avg_std");
}

// stats.cuh:115-142 avg_std<double> on an FP64 stream (value_t = double, the
// ED driver GPU-ed.cu:76): FP64 window means and population sigmas.
template <class prefx_t, class value_t, class index_t, mdtw_compat::req_double<value_t> = 0>
inline void avg_std(const value_t* const Series, value_t* const AvgS, value_t* const StdS,
                    const index_t M, const index_t N, void* stream = nullptr) {
    static_assert(std::is_same<prefx_t, double>::value, "avg_std: prefix sums are FP64");
    mdtw_compat::check(mdtw_avg_std_f64(mdtw_compat::on_stream(stream), Series,
                                        static_cast<int64_t>(N), static_cast<int32_t>(M), AvgS,
                                        StdS),
                       "avg_std");
}

// ---- bounds.cuh ---------------------------------------------------------------

// bounds.cuh:13-64 envelope(series, W, L, U, M) (host).
template <class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void envelope(value_t* series, index_t W, value_t* L, value_t* U, index_t M) {
    mdtw_compat::check(mdtw_envelope(series, static_cast<int32_t>(W), L, U,
                                     static_cast<int32_t>(M)),
                       "envelope");
}

// bounds.cuh:413-427 lb_kim over the windows Indices[0..indices_l).
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void lb_kim(value_t* Subject, value_t* AvgS, value_t* StdS, value_t* Lb_kim, index_t M,
                   index_t* Indices, index_t indices_l, void* stream = nullptr) {
    static_assert(sizeof(index_t) == 4, "device indices are 32-bit");
    (void)M;
    mdtw_compat::check(mdtw_lb_kim(mdtw_compat::on_stream(stream), Subject, AvgS, StdS, Lb_kim,
                                   reinterpret_cast<const int32_t*>(Indices),
                                   static_cast<int64_t>(indices_l),
                                   euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
                       "lb_kim");
}

// bounds.cuh:429-441 lb_kim over all N - M + 1 windows.
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0,
          typename std::enable_if<std::is_integral<index_t>::value, int>::type = 0>
inline void lb_kim(value_t* Subject, value_t* AvgS, value_t* StdS, value_t* Lb_kim, index_t M,
                   index_t N, void* stream = nullptr) {
    mdtw_compat::check(mdtw_lb_kim(mdtw_compat::on_stream(stream), Subject, AvgS, StdS, Lb_kim,
                                   nullptr, static_cast<int64_t>(N - M + 1),
                                   euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
                       "lb_kim");
}

// bounds.cuh:443-453 lb_keogh over the windows Indices[0..indices_l).
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void lb_keogh(value_t* Subject, value_t* AvgS, value_t* StdS, value_t* Lb_keogh,
                     index_t M, index_t* Indices, index_t indices_l, void* stream = nullptr) {
    static_assert(sizeof(index_t) == 4, "device indices are 32-bit");
    (void)M;
    mdtw_compat::check(mdtw_lb_keogh(mdtw_compat::on_stream(stream), Subject, AvgS, StdS,
                                     Lb_keogh, reinterpret_cast<const int32_t*>(Indices),
                                     static_cast<int64_t>(indices_l),
                                     euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
                       "lb_keogh");
}

// bounds.cuh:455-465 lb_keogh over all N - M + 1 windows.
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0,
          typename std::enable_if<std::is_integral<index_t>::value, int>::type = 0>
inline void lb_keogh(value_t* Subject, value_t* AvgS, value_t* StdS, value_t* Lb_keogh,
                     index_t M, index_t N, void* stream) {
    mdtw_compat::check(mdtw_lb_keogh(mdtw_compat::on_stream(stream), Subject, AvgS, StdS,
                                     Lb_keogh, nullptr, static_cast<int64_t>(N - M + 1),
                                     euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
                       "lb_keogh");
}

// ---- cdtw.cuh -----------------------------------------------------------------

// cdtw.cuh:450-487 gpu_cdtw: squared banded cDTW of the query against the
// windows Indices[0..indices_l).  blockdtw (block vs thread kernel choice in
// the reference) does not change the result and is ignored.
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void gpu_cdtw(value_t* Subject, value_t* AvgS, value_t* StdS, value_t* Cdtw, index_t M,
                     index_t* Indices, index_t indices_l, index_t W, bool blockdtw = true,
                     void* stream = nullptr) {
    static_assert(sizeof(index_t) == 4, "device indices are 32-bit");
    (void)M;
    (void)blockdtw;
    mdtw_compat::check(mdtw_cdtw(mdtw_compat::on_stream(stream), Subject, AvgS, StdS, Cdtw,
                                 reinterpret_cast<const int32_t*>(Indices),
                                 static_cast<int64_t>(indices_l), static_cast<int32_t>(W),
                                 euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
                       "gpu_cdtw");
}

// cdtw.cuh:510-592 prune_cdtw: best match of the query in Subject[0..N);
// prints "Location:" / "Distance:" as report_result does (cdtw.cuh:506-507)
// and returns the match.
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline mdtw_compat::match prune_cdtw(value_t* Subject, value_t* AvgS, value_t* StdS, index_t M,
                                     index_t N, index_t W, void* stream = nullptr) {
    (void)M;
    int64_t pos = -1;
    float dist = INFINITY;
    mdtw_compat::check(
        mdtw_search(mdtw_compat::on_stream(stream), Subject, static_cast<int64_t>(N), AvgS, StdS,
                    static_cast<int32_t>(W), MDTW_ALGO_CDTW,
                    euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1, MDTW_STATS_FP64, 1, 0, &pos,
                    &dist, nullptr),
        "prune_cdtw");
    std::cout << "Location: " << pos << std::endl;
    std::cout << "Distance: " << dist << std::endl;
    return {pos, static_cast<double>(dist)};
}

// ---- ed.cuh -------------------------------------------------------------------

// ed.cuh:71-153 calculate_ed: best z-normalised Euclidean match of the query
// (set_query) in Subject[0..N).  The reference computes in FP64 through
// cuFFT; this searches the FP32 stream (Subject float) directly (FP64 sums)
// for M <= 256 and through the FP64 FFT form above that (MDTW_ALGO_ED_FFT,
// faster from M = 384 on: 2.4 vs 2.9 ms for 2^27 samples, flat in M), with
// the window statistics computed here (StdS
// is accepted for signature parity and not read).  Prints the reference's
// "distance:" / "location:" lines (ed.cuh:149-150) and returns its
// bsf = ED^2 - 2M.
template <class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline double calculate_ed(const value_t* const Subject, const value_t* const StdS,
                           const index_t M, const index_t N, void* stream = nullptr) {
    (void)StdS;
    int64_t pos = -1;
    float dist = INFINITY;
    const int32_t algo = M > 256 ? MDTW_ALGO_ED_FFT : MDTW_ALGO_ED;
    mdtw_compat::check(mdtw_search(mdtw_compat::on_stream(stream), Subject,
                                   static_cast<int64_t>(N), nullptr, nullptr, 0, algo,
                                   MDTW_METRIC_L2, MDTW_STATS_FP64, 1, 0, &pos, &dist, nullptr),
                       "calculate_ed");
    const double ed = static_cast<double>(dist);
    std::cout << "\n= result ===================================" << std::endl;
    std::cout << "distance: " << ed << std::endl;
    std::cout << "location: " << pos << std::endl;
    return ed * ed - 2.0 * static_cast<double>(M);
}

// ed.cuh:71-153 calculate_ed on an FP64 stream, as the reference's ED driver
// calls it (GPU-ed.cu:82): the FP64 query given to set_query(const double*,
// M), FP64 window statistics (computed here; StdS is accepted for signature
// parity and not read), FP64 cross-correlation.  Prints the reference's
// result lines and returns its bsf = ED^2 - 2M.
template <class value_t, class index_t, mdtw_compat::req_double<value_t> = 0>
inline double calculate_ed(const value_t* const Subject, const value_t* const StdS,
                           const index_t M, const index_t N, void* stream = nullptr) {
    (void)StdS;
    const std::vector<double>& q = mdtw_compat::query_f64();
    if (static_cast<index_t>(q.size()) != M)
        throw mdtw_compat::error(MDTW_ERR_NO_QUERY, "calculate_ed");
    int64_t pos = -1;
    double dist = INFINITY;
    mdtw_compat::check(mdtw_ed_search_f64(mdtw_compat::on_stream(stream), Subject,
                                          static_cast<int64_t>(N), q.data(),
                                          static_cast<int32_t>(M), &pos, &dist, nullptr),
                       "calculate_ed");
    std::cout << "\n= result ===================================" << std::endl;
    std::cout << "distance: " << dist << std::endl;
    std::cout << "location: " << pos << std::endl;
    return dist * dist - 2.0 * static_cast<double>(M);
}

// ---- ldtw.cuh -----------------------------------------------------------------

namespace mdtw_compat {

// ldtw.cuh:294-391 ztsh2_ldtw<Euclidean><<<...>>>(Q, M, S, N, AvgS, StdS, R, W)
// (a kernel launch in the reference): greedy LDTW of the query (set_query)
// against every window, R[0..N-M+1) squared (L2) distances.
template <bool Euclidean>
inline void ztsh2_ldtw(const float* S, int M, int N, const float* AvgS, const float* StdS,
                       float* R, int W, void* stream = nullptr) {
    check(mdtw_ldtw(on_stream(stream), S, AvgS, StdS, R, static_cast<int64_t>(N - M + 1), W,
                    Euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
          "ztsh2_ldtw");
}

// ldtw.cuh:198-291 tsh2_ldtw<Euclidean><<<...>>>(Q, M, S, N, R, W): the same
// greedy walk on raw values (the query as given to set_query, no window
// statistics).
template <bool Euclidean>
inline void tsh2_ldtw(const float* S, int M, int N, float* R, int W, void* stream = nullptr) {
    check(mdtw_ldtw(on_stream(stream), S, nullptr, nullptr, R, static_cast<int64_t>(N - M + 1), W,
                    Euclidean ? MDTW_METRIC_L2 : MDTW_METRIC_L1),
          "tsh2_ldtw");
}

}  // namespace mdtw_compat

// ---- the reference's CPU paths (host code, as in the reference) ----------------
//
// The reference ships CPU versions of its DP and greedy LDTW for its drivers'
// and tests' comparisons (GPU-ldtw.cu:108-141 prints them beside the GPU
// answer; tests/test_dtw.cu:129-130 compares against them).  They are host
// loops here too, independent of the device path.  Rounding: every operation
// is a separate IEEE single-precision step in the order of the reference's
// expressions; built without FP contraction across statements (clang's
// default), they give the device kernels' bits for LDTW and the unfused host
// rounding of the reference's sequential_cdtw.

namespace mdtw_compat {
namespace host {

inline float cost(float d, bool l2) {
    if (l2) {
        const float c = d * d;
        return c;
    }
    return d < 0 ? -d : d;
}

// One window of the greedy walk (ldtw.cuh:18-127): from (0,0) to (M-1,M-1),
// each step takes the smallest local cost among the diagonal, the cell below
// (i+1, j) and the cell to the right (i, j+1), band-checked; the diagonal only
// when strictly smaller than both, then below when not above the diagonal and
// strictly below the right cell, else right.  zs(j) is the window's j-th
// (z-normalised) sample.
template <class Z>
inline float greedy_walk(const float* q, Z zs, int M, int W, bool l2) {
    const float kFar = 1e10f;  // INFTY, ldtw.cuh:8
    int i = 0, j = 0;
    float d = cost(q[0] - zs(0), l2);
    while (i + 1 != M || j + 1 != M) {
        float dg = kFar, dn = kFar, rt = kFar;
        if (i + 1 < M && j + 1 < M) dg = cost(q[i + 1] - zs(j + 1), l2);
        if (i + 1 < M && std::abs(i - j + 1) <= W) dn = cost(q[i + 1] - zs(j), l2);
        if (j + 1 < M && std::abs(j - i + 1) <= W) rt = cost(q[i] - zs(j + 1), l2);
        if (dg < dn && dg < rt) {
            d += dg;
            ++i;
            ++j;
        } else if (!(dg < dn) && dn < rt) {
            d += dn;
            ++i;
        } else {
            d += rt;
            ++j;
        }
    }
    return d;
}

template <bool Euclidean>
inline void ldtw_windows(const float* query, int M, const float* subject, int N, const float* avgs,
                         const float* stds, float* result, int W, bool parallel) {
    const long L = static_cast<long>(N) - M + 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (parallel)
#endif
    for (long k = 0; k < L; ++k) {
        const float* x = subject + k;
        if (avgs) {
            const float mu = avgs[k], sg = stds[k];
            result[k] = greedy_walk(query, [&](int j) { return (x[j] - mu) / sg; }, M, W,
                                    Euclidean);
        } else {
            result[k] = greedy_walk(query, [&](int j) { return x[j]; }, M, W, Euclidean);
        }
    }
}

// Banded DP of one window over two rolling rows (cdtw.cuh:23-88): row i of the
// query against the band of the window, +inf outside it, D[0][0] = 0.
template <bool euclidean>
inline float cdtw_window(const float* zq, const float* x, float mu, float sg, int M, int W,
                         std::vector<float>& rows) {
    const int lane = M + 1;
    rows.assign(2 * static_cast<size_t>(lane), INFINITY);
    float* prev = rows.data();
    float* cur = rows.data() + lane;
    prev[0] = 0.0f;
    for (int i = 1; i <= M; ++i) {
        const int lo = std::max(0, i - W - 1), hi = std::min(lane, i + W + 1);
        cur[lo] = INFINITY;
        for (int j = lo + 1; j < hi; ++j) {
            const float v = zq[i - 1] - (x[j - 1] - mu) / sg;
            const float c = cost(v, euclidean);
            const float up = prev[j] < cur[j - 1] ? prev[j] : cur[j - 1];
            const float best = prev[j - 1] < up ? prev[j - 1] : up;
            cur[j] = c + best;
        }
        if (hi < lane) cur[hi] = INFINITY;
        std::swap(prev, cur);
    }
    return prev[M];
}

}  // namespace host
}  // namespace mdtw_compat

// cdtw.cuh:437-448 cpu_cdtw (sequential_cdtw :23-88 / openmp_cdtw :90-157):
// host DP of the z-normalised query against the windows indices[0..indices_l)
// of the host subject, squared for euclidean.
template <bool euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void cpu_cdtw(value_t* zquery, value_t* subject, value_t* avgS, value_t* stdS,
                     value_t* cdtw, index_t M, index_t* indices, index_t indices_l, index_t W,
                     bool openmp = true) {
    const long n = static_cast<long>(indices_l);
#ifdef _OPENMP
#pragma omp parallel if (openmp)
#endif
    {
        std::vector<float> rows;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (long p = 0; p < n; ++p) {
            const long w = static_cast<long>(indices[p]);
            cdtw[p] = mdtw_compat::host::cdtw_window<euclidean>(
                zquery, subject + w, avgS[w], stdS[w], static_cast<int>(M), static_cast<int>(W),
                rows);
        }
    }
}

// ldtw.cuh:151-174 z_stream_ldtw / omp_z_stream_ldtw: greedy LDTW of the
// z-normalised query against every window of the host subject (window
// statistics avgs / stds), on one core / on all cores.
template <bool Euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void z_stream_ldtw(const value_t* query, const index_t M, const value_t* subject,
                          const index_t N, const value_t* avgs, const value_t* stds,
                          value_t* result, const index_t W) {
    mdtw_compat::host::ldtw_windows<Euclidean>(query, M, subject, N, avgs, stds, result, W, false);
}
template <bool Euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void omp_z_stream_ldtw(const value_t* query, const index_t M, const value_t* subject,
                              const index_t N, const value_t* avgs, const value_t* stds,
                              value_t* result, const index_t W) {
    mdtw_compat::host::ldtw_windows<Euclidean>(query, M, subject, N, avgs, stds, result, W, true);
}
// ldtw.cuh:130-149 stream_ldtw / omp_stream_ldtw: the same on raw values.
template <bool Euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void stream_ldtw(const value_t* query, const index_t M, const value_t* subject,
                        const index_t N, value_t* result, const index_t W) {
    mdtw_compat::host::ldtw_windows<Euclidean>(query, M, subject, N, nullptr, nullptr, result, W,
                                               false);
}
template <bool Euclidean, class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline void omp_stream_ldtw(const value_t* query, const index_t M, const value_t* subject,
                            const index_t N, value_t* result, const index_t W) {
    mdtw_compat::host::ldtw_windows<Euclidean>(query, M, subject, N, nullptr, nullptr, result, W,
                                               true);
}

// ldtw.cuh:177-192 determine_min (host): strict '<', lowest index wins;
// prints the reference's line.
template <bool Euclidean, class index_t, class value_t>
inline void determine_min(const value_t* result, const index_t L) {
    value_t bsf = static_cast<value_t>(1e10);  // INFTY, ldtw.cuh:8
    index_t bsf_index = 0;
    for (index_t i = 0; i < L; ++i)
        if (result[i] < bsf) {
            bsf = result[i];
            bsf_index = i;
        }
    std::cout << "best alignment found at index " << bsf_index << " with distance "
              << (Euclidean ? std::sqrt(bsf) : bsf) << std::endl
              << std::endl;
}

#endif  // MDTW_COMPAT_HPP
