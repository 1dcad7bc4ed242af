// internal.hpp -- shared declarations of the mdtw HIP library (gfx950).
//
// Floating-point policy (all translation units): -ffp-contract=off, packed
// FP32 disabled; every fused multiply-add is an explicit fmaf, division and
// sqrt are IEEE (correctly rounded).  That fixes the rounding sequence of each
// kernel so it can be compared bit-for-bit with the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "mdtw.h"

namespace mdtw {

constexpr float kInf = __builtin_huge_valf();
constexpr int kWave = 64;

// Registers-per-lane band kernel covers W <= kRegBandMaxW; wider bands use the
// anti-diagonal kernel, which covers W <= kMaxBandW (the reference's own limit
// is W + 2 <= 1024 threads, cdtw.cuh:459).
constexpr int kRegBandMaxW = 57;
constexpr int kMaxBandW = 1023;

// Largest number of windows one search covers.  Candidate lists hold window
// indices as int32 and the kernels widen them with a sign extension, so every
// window index of a search must be below 2^31.
constexpr int64_t kMaxSearchWindows = 0x7fffffffll;

// Exclusion zone as a search uses it: negative means none (0); a zone of L or
// more windows excludes every other window of an L-window search, exactly as
// a zone of L does, so it is capped there.  With L <= kMaxSearchWindows the
// bucket length 2 * excl stays below 2^32 (the device divides in 32 bits).
inline int64_t clamp_excl(int64_t excl, int64_t L) {
    return excl < 0 ? 0 : (excl > L ? L : excl);
}

// Control block in device memory, one per engine.
struct Ctrl {
    unsigned long long best;     // (float bits << 32) | window index
    unsigned long long count;    // candidates appended
    unsigned long long work;     // persistent-kernel work cursor
    unsigned long long nres;     // results appended (top-k)
    unsigned long long kim;      // pruned by LB_Kim
    unsigned long long keogh;    // pruned by LB_Keogh
    unsigned long long dp;       // DP runs started
    unsigned long long abandoned;// DP runs stopped early
    float thr;                   // top-k threshold (squared distance); only falls
    uint32_t unused;
    // Top-k: per-bucket minimum distance (float bits) of the windows whose
    // distance is known exactly, bucket = window / bucket_len; NULL when not
    // kept.  See launch_tighten.
    uint32_t* bmin;
    long long bucket_len;
    long long pad[2];
};

}  // namespace mdtw

struct mdtw_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    char name[256] = {0};

    // Query state: the reference's Czquery/Czlower/Czupper (cdtw.cuh:9-11).
    int32_t m = 0;
    int32_t w_env = -1;
    bool has_query = false;
    float* d_q = nullptr;
    float* d_lo = nullptr;
    float* d_up = nullptr;
    std::vector<float> h_q, h_lo, h_up;
    // Index deques of the envelope (2 x (MDTW_MAX_QUERY + 1)), sized at
    // creation so that building an envelope allocates nothing.
    std::vector<int> env_deques;
    // Query padded with +inf for the register-band kernel, built per band.
    float* d_qx = nullptr;
    size_t qx_cap = 0;
    int32_t qx_len = 0;
    int32_t w_qx = -1;

    // Growing scratch, reused across calls (no per-call malloc), and the cap
    // its growth may not pass (0: none; mdtw_set_workspace_limit).
    char* d_ws = nullptr;
    size_t ws_cap = 0;
    size_t ws_limit = 0;

    mdtw::Ctrl* d_ctrl = nullptr;
    mdtw::Ctrl* h_ctrl = nullptr;  // pinned
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace mdtw {

void set_last_hip_error(hipError_t e);

// MDTW_OK, or MDTW_ERR_HIP with the HIP error kept for mdtw_status_string.
inline int hip_status(hipError_t err) {
    if (err == hipSuccess) return MDTW_OK;
    set_last_hip_error(err);
    return MDTW_ERR_HIP;
}

// Makes the engine's device current for one entry point, restores the caller's.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};
int ws_reserve(mdtw_engine* e, size_t bytes, void** out);
int upload(mdtw_engine* e, void* dst, const void* src, size_t bytes);
void host_envelope(const float* q, int32_t w, float* lo, float* up, int32_t m);
// The same with caller-provided deques of m + 1 entries each (no allocation).
void host_envelope(const float* q, int32_t w, float* lo, float* up, int32_t m, int* ub, int* lb);
int ensure_envelope(mdtw_engine* e, int32_t w);
int ensure_qx(mdtw_engine* e, int32_t w);

// Sub-allocator over one reserved workspace span (256-byte aligned pieces).
struct Carve {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t n) {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += n * sizeof(T);
        return p;
    }
};
inline size_t carve_size(std::initializer_list<size_t> bytes) {
    size_t t = 0;
    for (size_t b : bytes) t = ((t + 255) & ~size_t(255)) + b;
    return t + 256;
}

// FP64 sums of the query the ED kernels use: its mean, and the sum and sum
// of squares of its centred values (q_k - mean), k ascending.
struct EdQuerySums {
    double mean = 0.0;
    double sum = 0.0;
    double sumsq = 0.0;
};
EdQuerySums ed_query_sums(const float* h_q, int32_t m);

// ---- kernel interface shared by the DP-style kernels ----------------------
enum BandMode : int {
    kModeProfile = 0,  // out[g] = distance of window g
    kModeBest = 1,     // atomicMin (distance, index) into ctrl->best; abandon above it
    kModeCollect = 2,  // append keys with distance <= ctrl->thr to res; abandon above it
    kModeSeed = 3,     // out[g] = distance, and atomicMin into ctrl->best
};

// Windows are idx[g] (idx non-null; count read from ctrl->count on the device
// when idx_count_from_ctrl) or base + g * stride, for g < count.
struct BandArgs {
    const float* s = nullptr;
    const float* avg = nullptr;
    const float* sd = nullptr;
    const float* qx = nullptr;  // band kernel: padded query; others: plain query
    int32_t qx_len = 0;         // length of the padded query buffer
    int32_t m = 0;
    const int32_t* idx = nullptr;
    int32_t idx_count_from_ctrl = 0;
    int64_t count = 0;
    int64_t base = 0;
    int64_t stride = 1;
    int32_t mode = kModeProfile;
    int32_t persistent = 0;
    Ctrl* ctrl = nullptr;
    unsigned long long* res = nullptr;
    float* out = nullptr;
    // Cumulative-bound abandon (register-band kernel, idx lists from the
    // cascade): lbt[g] = the window's LB_Keogh total, lo/up = the query
    // envelope.  NULL lbt: column-minimum abandon only.
    const float* lbt = nullptr;
    const float* lo = nullptr;
    const float* up = nullptr;
    // ED kernels: the query's sums (ed_query_sums of the query in qx), and
    // whether the statistics are the search's own FP64 or raw ones (sum of
    // z^2 over a window = M up to rounding), which the FP32 screen's rounding
    // margin assumes.
    EdQuerySums ed;
    int32_t ed_screen = 0;
};

struct BandScalars {
    int32_t m;
    int32_t idx_count_from_ctrl;
    int64_t count;
    int64_t base;
    int64_t stride;
    int32_t mode;
    int32_t persistent;
    int32_t lq_len;  // band kernel: LDS length of one shifted query copy
    int32_t qx_len;  // band kernel: length of the padded query buffer
};

// Launch shapes of the search kernels; the batch autotuner picks them per
// query-length / band bucket, plain searches use the defaults.
// DP formulation of a search (LaunchShape::dp_form).  Every form computes the
// same bits for every window it completes and abandons only above the bound,
// so the form changes the time, never the result:
//   kDpFormAuto:  the register-band kernel of the band (one lane per window,
//                 its band in registers, an LDS tail or split over two waves
//                 as W requires), abandoning on the column minimum; the
//                 anti-diagonal kernel above W = 240;
//   kDpFormLb:    the same kernels, one lane per window also abandoning on
//                 the column minimum plus the LB_Keogh terms of the columns
//                 still to come (where its registers leave the residency);
//                 measured slower than kDpFormAuto on every plain-search
//                 workload here (1-20%), kept for the tuner;
//   kDpFormWide:  the anti-diagonal kernel (the band across a wave, diagonals
//                 through LDS) for any band.
enum DpForm : int32_t { kDpFormAuto = 0, kDpFormLb = 1, kDpFormWide = 2 };

struct LaunchShape {
    int32_t dp_blocks_per_cu = 8;         // persistent register-band DP (4 waves each)
    int32_t wide_waves_per_cu = 32;       // persistent anti-diagonal DP (1 wave each)
    int32_t cascade_blocks_per_cu = 64;   // cap of the cascade grid (4 waves each)
    int32_t dp_form = kDpFormAuto;        // DpForm
};

// Everything one search needs on the device (enqueue_search).
struct SearchPlan {
    hipStream_t st = nullptr;
    int cus = 0;
    const float* s = nullptr;
    int64_t n = 0;
    const float* d_avg = nullptr;  // given window statistics, or NULL: computed
    const float* d_std = nullptr;
    int32_t stats_mode = MDTW_STATS_FP64;
    mdtw_engine* e = nullptr;      // REFCOMPAT statistics only
    const float* q = nullptr;      // z-normalised query, M floats
    const float* lo = nullptr;     // its envelope for band w
    const float* up = nullptr;
    const float* qx = nullptr;     // padded query of the register-band kernel
    int32_t qx_len = 0;
    EdQuerySums ed;                // ED: ed_query_sums of the query
    int32_t m = 0;
    int32_t w = 0;                 // band, <= m - 1
    int32_t algo = MDTW_ALGO_CDTW;
    int32_t metric = MDTW_METRIC_L2;
    int32_t k = 1;
    int64_t excl = 0;
    Ctrl* ctrl = nullptr;
    char* ws = nullptr;            // search_ws_bytes(...) of scratch
    LaunchShape shape;
    // Bounds carried in from windows outside this stream (a streaming search's
    // earlier epochs): k = 1 starts from this key instead of none; k > 1 caps
    // the seed bound.  Both must be distances of windows of the whole search.
    unsigned long long best_init = 0xffffffffffffffffull;
    float thr_cap = __builtin_huge_valf();
    // false: no seeding (the carried bounds above already bound the search).
    bool seed = true;
};
struct SearchBufs {
    float* avg = nullptr;
    float* sd = nullptr;
    unsigned long long* res = nullptr;  // top-k keys (ctrl->nres of them)
    int64_t seeded = 0;
};
size_t search_ws_bytes(int64_t L, int32_t k, int64_t excl, bool own_stats);
// mdtw_search's argument checks (status MDTW_OK or the error to return);
// on success *excl is clamped to the stream and *wc is the band clamped to M - 1.
int check_search_args(const mdtw_engine* e, const float* d_series, int64_t n, const float* d_avg,
                      const float* d_std, int32_t w, int32_t algo, int32_t metric,
                      int32_t stats_mode, int32_t k, const int64_t* h_pos, const float* h_dist,
                      int64_t* excl, int32_t* wc);
// Allocation-free search (noalloc.hip): the device's fixed scratch array is
// looked up once per device at engine creation, with that device current.
int scratch_setup(int device);
// Host greedy top-k with exclusion over (distance, position) pairs sorted
// ascending: a pair is taken unless |position - earlier pick| < excl.
int32_t greedy_pick(const std::vector<std::pair<float, int64_t>>& sorted, int32_t k, int64_t excl,
                    int32_t metric, int64_t* h_pos, float* h_dist);
int64_t search_seed_count(int64_t L, int32_t k, int64_t excl);
// Enqueues the whole search on p.st (no host synchronisation, no allocation;
// safe to capture in a graph when stats_mode is FP64).
int enqueue_search(const SearchPlan& p, SearchBufs* bufs);
// FFT ED (ed_fft.hip): scratch of a profile (k = 1) or a top-k search, and
// the search itself (statistics MDTW_STATS_FP64 only).
size_t ed_fft_ws_bytes(int64_t L, int32_t m, int32_t k);
int search_ed_fft(mdtw_engine* e, const float* d_series, int64_t n, int32_t stats_mode, int32_t k,
                  int64_t excl, int64_t* h_pos, float* h_dist, mdtw_search_info* info);
// Greedy top-k with exclusion over the kept keys on the device (one
// workgroup): out_pos/out_dist get k entries (position -1 / +inf past the
// matches found); distances are sqrt'ed for L2.  k = 1 decodes ctrl->best.
// Sets *overflow when more than kSelectMaxKeys keys are at or below the
// final bound (the caller then selects on the host).
constexpr int32_t kSelectMaxKeys = 8192;
int launch_topk_select(hipStream_t st, const Ctrl* ctrl, const unsigned long long* res,
                       int32_t k, int64_t excl, int32_t metric, int64_t* out_pos,
                       float* out_dist, int32_t* overflow);

// Query gradient of sum_p grad[p] * dist_p over windows idx[0..count) of the
// query q (m, band w): recompute, path walk, z-normalisation Jacobian when
// q_raw is given (backward.hip).  avg/sd: window statistics indexed by window
// (pair_stats false) or by list entry (true).  ws: bwd_ws_bytes(m, w, count).
size_t bwd_ws_bytes(int32_t m, int32_t w, int64_t count);
int enqueue_backward(hipStream_t st, const float* s, const float* avg, const float* sd,
                     bool pair_stats, const int32_t* idx, const float* grad, int64_t count,
                     const float* q, int32_t m, int32_t w, int32_t metric, int32_t sqrt_dist,
                     const float* q_raw, float* dist, float* out, char* ws);

// One query of a batched backward (device table entry).
struct BwdQuery {
    int64_t pair_begin;  // its first entry in the pair lists
    int64_t row_begin;   // its first gradient row in the row buffer (floats)
    int64_t q_off;       // its z-normalised query in the query arena
    int64_t r_off;       // its raw query, gradient sum and output (floats)
    int32_t m, w;
    int32_t pairs;       // its windows in the pair lists
    int32_t batched;     // 1: run by the batch kernels, 0: by enqueue_backward
};
// Whether a query's windows can run in the batch kernels (their direction
// words fit in LDS), and the class (slots per lane) that runs them.
bool bwd_batchable(int32_t m, int32_t w);
int bwd_class(int32_t w);
// Windows list[0..count) (pair indices, queries of class C, m <= max_m):
// pair pr has window pos[pr], statistics mu/sd[pr], upstream gradient
// gin[pr]; writes dist[pr] and the gradient rows tab[query].row_begin + ...
int launch_bwd_path_batch(hipStream_t st, int C, bool l2, int32_t max_m, const float* s,
                          const float* mu, const float* sd, const int32_t* pos,
                          const int32_t* pair_query, const int32_t* list, int64_t count,
                          const BwdQuery* tab, const float* qarena, const float* gin,
                          float* gpair, float* dist, int32_t sqrt_dist);
// Row sums (FP64, pairs ascending) and the z-normalisation Jacobian of every
// batched query of tab[0..nq); gsum and out indexed by r_off.
int launch_bwd_reduce_batch(hipStream_t st, const float* gpair, const BwdQuery* tab, int32_t nq,
                            int32_t max_m, const float* q_raw, double* gsum, float* out);

// ---- kernel launchers (each in its own translation unit) -----------------
// stats.hip
int launch_stats_fp64(hipStream_t st, const float* s, int64_t n, int32_t m,
                      float* avg, float* sd);
// FP64 mean and variance (0 for a constant window) of every window: the same
// tile sums as launch_stats_fp64, kept in FP64.
int launch_stats_f64out(hipStream_t st, const float* s, int64_t n, int32_t m, double* avg,
                        double* var);
// MDTW_STATS_NONE: avg = 0, sd = 1 for the n - m + 1 windows of s (NaN for a
// window holding a non-finite sample).
int launch_stats_none(hipStream_t st, const float* s, int64_t n, int32_t m, float* avg,
                      float* sd);
// FP64 stream: FP64 mean and population sigma (1 for a constant window, NaN
// for a window holding a non-finite sample) of the n - m + 1 windows; with
// sigma false the variance instead (0 for a constant window).
int launch_stats_d(hipStream_t st, const double* s, int64_t n, int32_t m, double* avg,
                   double* sd, bool sigma = true);
inline int launch_stats_d_var(hipStream_t st, const double* s, int64_t n, int32_t m, double* avg,
                              double* var) {
    return launch_stats_d(st, s, n, m, avg, var, false);
}
int launch_stats_refcompat(mdtw_engine* e, const float* s, int64_t n, int32_t m,
                           float* avg, float* sd);
// MDTW_STATS_FP64 statistics of listed windows (window pos[p], length len[p]),
// bit-identical to launch_stats_fp64's for the same window.
int launch_stats_at(hipStream_t st, const float* s, int64_t n, const int32_t* len,
                    const int64_t* pos, int64_t count, float* avg, float* sd);
// bounds.hip
int launch_lb_kim(hipStream_t st, const float* s, const float* avg,
                  const float* sd, const float* q, int32_t m,
                  const int32_t* idx, int64_t count, int32_t metric, float* out);
int launch_lb_keogh(hipStream_t st, const float* s, const float* avg,
                    const float* sd, const float* lo, const float* up, int32_t m,
                    const int32_t* idx, int64_t count, int32_t metric, float* out);
// Search cascade: LB_Kim then LB_Keogh on windows base..base+count-1;
// survivors (bound not above the limit of ctrl->best, or of ctrl->thr) are
// appended to cand (ctrl->count).
// cand_lb (may be NULL) gets each survivor's LB_Keogh total.
int launch_cascade(hipStream_t st, const float* s, const float* avg,
                   const float* sd, const float* q, const float* lo,
                   const float* up, int32_t m, int64_t base, int64_t count,
                   int32_t metric, bool use_best, Ctrl* ctrl, int32_t* cand,
                   float* cand_lb = nullptr, int64_t max_blocks = 256 * 64,
                   int cus = 256);
// cdtw.hip: band kernel (qx = padded query) or, for wide bands, the
// anti-diagonal kernel (q_plain).
int launch_dp(hipStream_t st, int cus, const BandArgs& a, const float* q_plain,
              int32_t w, int32_t metric, bool force_wide,
              const LaunchShape& shape = LaunchShape());
// ldtw.hip (a.qx = plain query)
int launch_ldtw(hipStream_t st, int cus, const BandArgs& a, int32_t w, int32_t metric);
// ed.hip (a.qx = plain query; no persistent mode)
int launch_ed(hipStream_t st, int cus, const BandArgs& a);
// select.hip
// best_init: kBestNone for searches (key of +inf, so load_best_dist is +inf
// until a distance lands), ~0ull for the signed-order argmin.
constexpr unsigned long long kBestNone = (0x7f800000ull << 32) | 0xffffffffull;
int launch_ctrl_reset(hipStream_t st, Ctrl* ctrl, unsigned long long best_init,
                      uint32_t* bmin = nullptr, int64_t bucket_len = 1);
// A search's start in one launch: the control block (as launch_ctrl_reset),
// the nb bucket minima to +inf bits and the ntw tighten-scratch words to 0.
int launch_search_reset(hipStream_t st, Ctrl* ctrl, unsigned long long best_init,
                        uint32_t* bmin, int64_t bucket_len, int64_t nb, uint32_t* tscratch,
                        int64_t ntw);
// Between search rounds: candidate count and work cursor back to 0.
int launch_round_reset(hipStream_t st, Ctrl* ctrl);
// Top-k bound from the bucket minima (nb buckets, +inf where unknown):
// ctrl->thr = min(ctrl->thr, T) with T >= the k-th smallest bucket minimum
// (split: over even and over odd buckets separately, T = the smaller).  With
// buckets of 2*excl windows, windows in distinct buckets of one parity are
// more than 2*excl apart, so one greedy pick (which excludes |i - j| < excl)
// removes at most one of them; the k-th pick is then no worse than the k-th
// smallest of either parity class.  scratch: tighten_ws_words() words.
// reset_round: also set the candidate count and work cursor back to 0 at the
// end (the next round's start, as launch_round_reset).
size_t tighten_ws_words();
int launch_tighten(hipStream_t st, const uint32_t* bmin, int64_t nb, int32_t k, bool split,
                   uint32_t* scratch, Ctrl* ctrl, bool reset_round = false);
int launch_thr_from_best(hipStream_t st, Ctrl* ctrl);
int launch_argmin(hipStream_t st, const float* v, int64_t count, Ctrl* ctrl);
size_t threshold_ws_ints(int64_t count);
int launch_threshold(hipStream_t st, const float* v, int32_t* idx, int64_t count,
                     float thr, int32_t* tmp, int32_t** d_total);
// ctrl->thr = min(cap, k-th smallest of nseeds values (+inf if fewer than k)).
int launch_seed_threshold(hipStream_t st, const float* seeds, int64_t nseeds,
                          int32_t k, Ctrl* ctrl, float cap);
constexpr int64_t kSeedMaxCount = 16384;

}  // namespace mdtw

// ---- device helpers -------------------------------------------------------
namespace mdtw {

__device__ __forceinline__ unsigned long long pack_key(float d, uint32_t idx) {
    // Non-negative floats (and +inf) order like their bit patterns.
    return (static_cast<unsigned long long>(__float_as_uint(d)) << 32) | idx;
}
__device__ __forceinline__ float key_dist(unsigned long long k) {
    return __uint_as_float(static_cast<uint32_t>(k >> 32));
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ float load_best_dist(const Ctrl* c) {
    unsigned long long k = __hip_atomic_load(&c->best, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    return key_dist(k);
}
__device__ __forceinline__ float load_thr(const Ctrl* c) {
    return __hip_atomic_load(&c->thr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The bound a DP abandons against (mode kModeBest: the best distance so far;
// kModeCollect: the top-k threshold), read once per batch of windows and kept
// in a scalar register.  Both only fall while a launch runs (atomicMin on the
// best; the top-k bound changes between launches), so a value read earlier
// is a valid, slightly looser bound.  Re-reading it at every abandon check
// sent every wave of the chip to the same L2 line every few columns, where
// the reads serialise: the search DP spent most of its cycles waiting.
__device__ __forceinline__ float load_abandon_bound(const Ctrl* c, int mode) {
    const float v = mode == kModeBest ? load_best_dist(c) : load_thr(c);
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// z = (x - mu) / sigma, bit-identical to the IEEE quotient: q = a * y with y
// the correctly rounded 1/sigma, then one Markstein correction
// fma(fma(-sigma, q, a), y, q).  Outside the exponent range where the
// residual a - sigma*q is exact (tiny |a|, extreme sigma) the lane divides.
__device__ __forceinline__ float band_z(float x, float mu, float sg, float y, bool fast) {
    const float a = x - mu;
    const float q = a * y;
    float z = fmaf(fmaf(-sg, q, a), y, q);
    if (__builtin_expect(!fast || (a != 0.0f && fabsf(a) < 0x1p-100f), 0)) z = a / sg;
    return z;
}

// band_z's fast path for every sample of a window, decided once per window:
// sigma in range, and either |mu| >= 2^-75 (then a nonzero x - mu of FP32
// values is at least 2^-99: both at or above 2^-76 in magnitude, the
// difference is a multiple of 2^-99; x below 2^-76, it is above 2^-76) or
// sigma == 1 (q = a * 1 is exact and the residual 0).  Then the quotient is
// band_z_fast, with no per-sample test.
__device__ __forceinline__ bool band_z_window_fast(float mu, float sg) {
    return sg >= 0x1p-100f && sg <= 0x1p100f && (fabsf(mu) >= 0x1p-75f || sg == 1.0f);
}
__device__ __forceinline__ float band_z_fast(float x, float mu, float sg, float y) {
    const float a = x - mu;
    const float q = a * y;
    return fmaf(fmaf(-sg, q, a), y, q);
}
// band_z with the window's test hoisted: wfast = band_z_window_fast(mu, sg)
// (loop-invariant); the rare other windows keep the per-sample test.
__device__ __forceinline__ float band_z_w(float x, float mu, float sg, float y, bool wfast) {
    if (__builtin_expect(wfast, 1)) return band_z_fast(x, mu, sg, y);
    return band_z(x, mu, sg, y, sg >= 0x1p-100f && sg <= 0x1p100f);
}

// Exact distance d of window w is known: lower its bucket's minimum.
__device__ __forceinline__ void note_bucket(uint32_t* bmin, long long bucket_len, int64_t w,
                                            float d) {
    if (bmin)
        atomicMin(bmin + static_cast<uint32_t>(w) / static_cast<uint32_t>(bucket_len),
                  __float_as_uint(d));
}

// Pruning limit that keeps pruning exact.  A computed lower bound can exceed
// the computed distance of the same window by rounding: the bound sums m
// rounded terms, the DP rounds along a path of up to 2m cells, and the bound
// kernels z-normalise with a reciprocal (<= 1.5 ulp per value; summed over a
// window that perturbs an L2 bound by at most 1.8e-7 * sqrt(m * bound) by
// Cauchy-Schwarz with sum z^2 = m, an L1 bound by at most 1e-7 * m).  A
// window is pruned only when its bound exceeds prune_limit(threshold); the
// limit is monotone, so a partial sum above it also prunes (early abandon).
__device__ __forceinline__ float prune_limit(float thr, int32_t m, bool l2) {
    const float rel = thr * (float)(3 * m + 8) * 6.0e-8f;
    const float absl = l2 ? 2.5e-7f * sqrtf((float)m * thr) : 1.2e-7f * (float)m;
    return thr + rel + absl + 1e-30f;
}

// Limit for the DP's cumulative test  colmin_j + max(0, T - P_j) > limit,
// where T is the window's LB_Keogh total from the cascade and P_j the same
// terms summed by the DP over columns <= j: T - P_j bounds the cost of the
// columns still to come (each later column holds a path cell, whose cost is
// at least that column's LB_Keogh term).  On top of prune_limit: the
// cancellation in T - P_j (both sums of <= m rounded terms) and the two
// normalisations (reciprocal in the cascade, quotient in the DP), each bounded
// as in prune_limit.
__device__ __forceinline__ float suffix_limit(float thr, float T, int32_t m, bool l2) {
    const float rel = T * (float)(2 * m + 2) * 6.0e-8f;
    const float absl = l2 ? 5.0e-7f * sqrtf((float)m * T) : 2.4e-7f * (float)m;
    return prune_limit(thr, m, l2) + rel + absl;
}

}  // namespace mdtw
