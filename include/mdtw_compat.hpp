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

#include <cmath>
#include <cstdint>
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

struct match {
    int64_t location;
    double distance;
};

template <class T>
using req_float = typename std::enable_if<std::is_same<T, float>::value, int>::type;

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
// cuFFT; this searches the FP32 stream (Subject float) directly for
// M <= 2048 and through the FP64 FFT form above that (MDTW_ALGO_ED_FFT,
// faster for long queries), with the window statistics computed here (StdS
// is accepted for signature parity and not read).  Prints the reference's
// "distance:" / "location:" lines (ed.cuh:149-150) and returns its
// bsf = ED^2 - 2M.
template <class value_t, class index_t, mdtw_compat::req_float<value_t> = 0>
inline double calculate_ed(const value_t* const Subject, const value_t* const StdS,
                           const index_t M, const index_t N, void* stream = nullptr) {
    (void)StdS;
    int64_t pos = -1;
    float dist = INFINITY;
    const int32_t algo = M > 2048 ? MDTW_ALGO_ED_FFT : MDTW_ALGO_ED;
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
