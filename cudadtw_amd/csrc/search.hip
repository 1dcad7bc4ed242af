// search.hip -- mdtw_search: best or top-k matches of the engine's query.
//
// Replaces prune_cdtw (cdtw.cuh:510-592), calculate_ed (ed.cuh:71-153) and
// the GPU-ldtw.cu driver.  The reference sorts all windows by LB_Kim, then
// runs a host loop over doubling chunks (LB_Keogh, compaction, DP, reduce)
// with several host-device synchronisations and mallocs per chunk.  Here the
// whole search is a fixed sequence of launches on one stream with no host
// round trip until the answer:
//   1. window statistics (unless supplied),
//   2. seed: exact distances of up to 16384 evenly strided windows, which
//      sets the bound (k = 1: best seed; k > 1: k-th smallest seed, valid
//      because the seeds are >= 2*excl apart, so no pick can exclude two),
//   3. rounds over window ranges (one for k = 1; for k > 1 the first L/16
//      windows, then the rest -- the reference refines its bound over
//      doubling chunks in LB_Kim order, cdtw.cuh:551-585; here stream order,
//      so no sort is needed), each:
//      cascade (LB_Kim then LB_Keogh, survivors appended to a candidate
//      list), then a persistent DP over the candidates with early
//      abandoning; k = 1: atomicMin of (distance, index) keys (lowest index
//      wins ties), which the next round's cascade prunes against; k > 1:
//      every candidate at or below the bound is kept as a key, its distance
//      lowers its bucket's minimum, and after the round the bound is
//      tightened from the bucket minima (launch_tighten),
//   4. k > 1: the kept keys are sorted and picked greedily with the
//      exclusion zone on the host (they are few).
// Pruning never changes the answer: windows are dropped only when a bound
// strictly exceeds prune_limit(threshold) (rounding-aware), and a DP run is
// abandoned only when its running column minimum strictly exceeds the
// threshold; the threshold only falls and never below the k-th pick's
// distance, and every window at or below the final threshold is kept.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "cdtw_band.hpp"

using namespace mdtw;

namespace {

constexpr int64_t kSeedTarget = 8192;

struct SeedPlan {
    int64_t count;
    int64_t stride;
};

SeedPlan plan_seeds(int64_t L, int32_t k, int64_t excl) {
    int64_t stride = std::max<int64_t>(1, L / kSeedTarget);
    if (k > 1) stride = std::max<int64_t>(stride, 2 * std::max<int64_t>(excl, 0));
    int64_t count = (L - 1) / stride + 1;
    if (count > kSeedMaxCount) {
        stride = (L + kSeedMaxCount - 1) / kSeedMaxCount;
        if (k > 1) stride = std::max<int64_t>(stride, 2 * std::max<int64_t>(excl, 0));
        count = (L - 1) / stride + 1;
    }
    return {count, stride};
}

// Window ranges of the search rounds.  k = 1: one pass (the DP lowers the
// best distance as it goes, and on repetitive streams an early, tighter
// bound barely changes what LB_Keogh prunes).  k > 1: the first L/16
// windows, then the rest, so the bound from the first range's bucket minima
// (far below the k-th seed) prunes the rest.
struct Round {
    int64_t base;
    int64_t count;
};
constexpr int64_t kRoundMin = int64_t(1) << 16;

std::vector<Round> plan_rounds(int64_t L, bool topk) {
    const int64_t first = std::max<int64_t>(kRoundMin, L / 16);
    if (!topk || first >= L) return {{0, L}};
    return {{0, first}, {first, L - first}};
}

int hip_check(hipError_t err) {
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

struct TopkLayout {
    bool split;
    int64_t bucket_len;
    int64_t nb;
};

TopkLayout topk_layout(int64_t L, int32_t k, int64_t excl) {
    // Top-k bound buckets: 2*excl windows (parity split) or single windows.
    TopkLayout t;
    t.split = k > 1 && excl >= 2;
    t.bucket_len = t.split ? 2 * excl : 1;
    t.nb = k > 1 ? (L + t.bucket_len - 1) / t.bucket_len : 0;
    return t;
}

// Greedy top-k with the exclusion zone over the kept (distance, index)
// keys: ascending keys, a key is taken unless within |i - j| < excl of an
// earlier pick.  Returns the number of picks.
int32_t host_topk_select(std::vector<unsigned long long>& keys, int32_t k, int64_t excl,
                         int32_t metric, int64_t* h_pos, float* h_dist) {
    std::sort(keys.begin(), keys.end());
    int32_t found = 0;
    for (unsigned long long key : keys) {
        if (found == k) break;
        const int64_t pos = static_cast<int64_t>(key & 0xffffffffull);
        bool ok = true;
        for (int32_t p = 0; p < found; ++p) {
            int64_t dd = pos - h_pos[p];
            if (dd < 0) dd = -dd;
            if (dd < excl) {
                ok = false;
                break;
            }
        }
        if (!ok) continue;
        const uint32_t bits = static_cast<uint32_t>(key >> 32);
        float d;
        std::memcpy(&d, &bits, sizeof(d));
        h_pos[found] = pos;
        h_dist[found] = metric == MDTW_METRIC_L2 ? std::sqrt(d) : d;
        found++;
    }
    return found;
}

}  // namespace

namespace mdtw {

int32_t greedy_pick(const std::vector<std::pair<float, int64_t>>& sorted, int32_t k, int64_t excl,
                    int32_t metric, int64_t* h_pos, float* h_dist) {
    int32_t found = 0;
    for (const auto& c : sorted) {
        if (found == k) break;
        bool ok = true;
        for (int32_t p = 0; p < found; ++p) {
            const int64_t dd = c.second > h_pos[p] ? c.second - h_pos[p] : h_pos[p] - c.second;
            if (dd < excl) {
                ok = false;
                break;
            }
        }
        if (!ok) continue;
        h_pos[found] = c.second;
        h_dist[found] = metric == MDTW_METRIC_L2 ? std::sqrt(c.first) : c.first;
        found++;
    }
    return found;
}

}  // namespace mdtw

namespace {

// One DP-style launch for the selected algorithm.
int run_algo(const SearchPlan& p, const BandArgs& a) {
    if (p.algo == MDTW_ALGO_CDTW)
        return launch_dp(p.st, p.cus, a, p.q, p.w, p.metric, p.shape.dp_form == kDpFormWide,
                         p.shape);
    BandArgs b = a;
    b.qx = p.q;
    b.ed = p.ed;
    if (p.algo == MDTW_ALGO_LDTW) return launch_ldtw(p.st, p.cus, b, p.w, p.metric);
    return launch_ed(p.st, p.cus, b);
}

}  // namespace

namespace mdtw {

int64_t search_seed_count(int64_t L, int32_t k, int64_t excl) {
    return plan_seeds(L, k, excl).count;
}

size_t search_ws_bytes(int64_t L, int32_t k, int64_t excl, bool own_stats) {
    const SeedPlan sp = plan_seeds(L, k, excl);
    const bool topk = k > 1;
    const TopkLayout tl = topk_layout(L, k, excl);
    int64_t max_round = 0;
    for (const Round& r : plan_rounds(L, topk)) max_round = std::max(max_round, r.count);
    return carve_size({own_stats ? sizeof(float) * L : 0, own_stats ? sizeof(float) * L : 0,
                       sizeof(int32_t) * max_round, sizeof(float) * max_round,
                       topk ? sizeof(unsigned long long) * L : 0, sizeof(float) * sp.count,
                       topk ? sizeof(uint32_t) * tl.nb : 0,
                       topk ? sizeof(uint32_t) * tighten_ws_words() : 0});
}

int enqueue_search(const SearchPlan& p, SearchBufs* bufs) {
    const int32_t m = p.m;
    const int64_t L = p.n - m + 1;
    const int32_t k = p.k;
    const bool topk = k > 1;
    const bool own_stats = p.d_avg == nullptr;
    const SeedPlan sp = plan_seeds(L, k, p.excl);
    const TopkLayout tl = topk_layout(L, k, p.excl);
    const std::vector<Round> rounds = plan_rounds(L, topk);
    int64_t max_round = 0;
    for (const Round& r : rounds) max_round = std::max(max_round, r.count);
    Carve cv{p.ws};
    float* avg = own_stats ? cv.take<float>(L) : const_cast<float*>(p.d_avg);
    float* sd = own_stats ? cv.take<float>(L) : const_cast<float*>(p.d_std);
    int32_t* cand = cv.take<int32_t>(max_round);
    float* cand_lb = cv.take<float>(max_round);
    unsigned long long* res = topk ? cv.take<unsigned long long>(L) : nullptr;
    float* seeds = cv.take<float>(sp.count);
    uint32_t* bmin = topk ? cv.take<uint32_t>(tl.nb) : nullptr;
    uint32_t* tscratch = topk ? cv.take<uint32_t>(tighten_ws_words()) : nullptr;
    bufs->avg = avg;
    bufs->sd = sd;
    bufs->res = res;
    bufs->seeded = sp.count;

    hipStream_t st = p.st;
    const unsigned long long best0 = p.best_init < kBestNone ? p.best_init : kBestNone;
    int rc = launch_search_reset(st, p.ctrl, best0, bmin, tl.bucket_len, topk ? tl.nb : 0,
                                 tscratch, topk ? static_cast<int64_t>(tighten_ws_words()) : 0);
    if (rc == MDTW_OK && own_stats) {
        rc = p.stats_mode == MDTW_STATS_REFCOMPAT ? launch_stats_refcompat(p.e, p.s, p.n, m, avg, sd)
             : p.stats_mode == MDTW_STATS_NONE  ? launch_stats_none(st, p.s, p.n, m, avg, sd)
                                                : launch_stats_fp64(st, p.s, p.n, m, avg, sd);
    }
    if (rc != MDTW_OK) return rc;

    BandArgs a;
    a.s = p.s;
    a.avg = avg;
    a.sd = sd;
    a.qx = p.qx;
    a.qx_len = p.qx_len;
    a.m = m;
    a.ctrl = p.ctrl;
    a.res = res;
    a.ed_screen = own_stats && p.stats_mode != MDTW_STATS_REFCOMPAT;

    // 2. seed
    if (p.seed) {
        BandArgs sa = a;
        sa.count = sp.count;
        sa.base = 0;
        sa.stride = sp.stride;
        sa.mode = topk ? kModeSeed : kModeBest;
        sa.out = seeds;
        rc = run_algo(p, sa);
    } else {
        bufs->seeded = 0;
    }
    if (rc == MDTW_OK && topk)
        rc = launch_seed_threshold(st, seeds, p.seed ? sp.count : 0, k, p.ctrl, p.thr_cap);
    if (rc != MDTW_OK) return rc;

    // 3. rounds: the bound left by earlier rounds (k = 1: the best distance
    // so far; k > 1: the bucket bound) prunes the later ones.
    const int64_t cascade_cap = static_cast<int64_t>(p.cus) * p.shape.cascade_blocks_per_cu;
    for (size_t ri = 0; ri < rounds.size(); ++ri) {
        const Round& r = rounds[ri];
        if (p.algo == MDTW_ALGO_ED) {
            // ED is cheap per window and its own early abandon is the filter.
            BandArgs fa = a;
            fa.base = r.base;
            fa.count = r.count;
            fa.mode = topk ? kModeCollect : kModeBest;
            rc = run_algo(p, fa);
        } else if (m >= 11) {
            // The candidate count and work cursor are 0 here: the search
            // reset cleared them, and between rounds the tighten does.
            const bool lbf = p.shape.dp_form == kDpFormLb;
            rc = launch_cascade(st, p.s, avg, sd, p.q, p.lo, p.up, m, r.base, r.count, p.metric,
                                !topk, p.ctrl, cand, lbf ? cand_lb : nullptr, cascade_cap, p.cus);
            if (rc == MDTW_OK) {
                BandArgs da = a;
                da.idx = cand;
                da.lbt = lbf ? cand_lb : nullptr;
                da.lo = p.lo;
                da.up = p.up;
                da.idx_count_from_ctrl = 1;
                da.persistent = 1;
                da.mode = topk ? kModeCollect : kModeBest;
                rc = run_algo(p, da);
            }
        } else {
            // LB_Kim needs M >= 11 (bounds.cuh:418): evaluate every window.
            BandArgs da = a;
            da.base = r.base;
            da.count = r.count;
            da.mode = topk ? kModeCollect : kModeBest;
            rc = run_algo(p, da);
        }
        if (rc == MDTW_OK && topk)
            rc = launch_tighten(st, bmin, tl.nb, k, tl.split, tscratch, p.ctrl,
                                ri + 1 < rounds.size());
        if (rc != MDTW_OK) return rc;
    }
    return MDTW_OK;
}

}  // namespace mdtw

namespace mdtw {

int check_search_args(const mdtw_engine* e, const float* d_series, int64_t n, const float* d_avg,
                      const float* d_std, int32_t w, int32_t algo, int32_t metric,
                      int32_t stats_mode, int32_t k, const int64_t* h_pos, const float* h_dist,
                      int64_t* excl, int32_t* wc_out) {
    if (!e || !d_series || !h_pos || !h_dist || w < 0) return w < 0 ? MDTW_ERR_BAND
                                                                    : MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    if ((d_avg == nullptr) != (d_std == nullptr)) return MDTW_ERR_INVALID_ARGUMENT;
    if (algo != MDTW_ALGO_CDTW && algo != MDTW_ALGO_ED && algo != MDTW_ALGO_LDTW &&
        algo != MDTW_ALGO_ED_FFT)
        return MDTW_ERR_INVALID_ARGUMENT;
    if (k < 1 || k > MDTW_MAX_K) return MDTW_ERR_K;
    if (metric != MDTW_METRIC_L1 && metric != MDTW_METRIC_L2) return MDTW_ERR_INVALID_ARGUMENT;
    if ((algo == MDTW_ALGO_ED || algo == MDTW_ALGO_ED_FFT) && metric != MDTW_METRIC_L2)
        return MDTW_ERR_UNSUPPORTED;
    // The FFT form needs FP64 window statistics, which it computes itself.
    if (algo == MDTW_ALGO_ED_FFT && (d_avg != nullptr || stats_mode != MDTW_STATS_FP64))
        return MDTW_ERR_UNSUPPORTED;
    if (stats_mode != MDTW_STATS_FP64 && stats_mode != MDTW_STATS_REFCOMPAT &&
        stats_mode != MDTW_STATS_NONE)
        return MDTW_ERR_INVALID_ARGUMENT;
    const int32_t m = e->m;
    if (n < m) return MDTW_ERR_STREAM_TOO_SHORT;
    const int64_t L = n - m + 1;
    if (L > kMaxSearchWindows) return MDTW_ERR_UNSUPPORTED;
    const int32_t wc = std::min(w, m - 1);
    if (algo == MDTW_ALGO_CDTW && wc > kMaxBandW) return MDTW_ERR_UNSUPPORTED;
    *excl = clamp_excl(*excl, L);
    *wc_out = wc;
    return MDTW_OK;
}

}  // namespace mdtw

extern "C" int mdtw_search(mdtw_engine* e, const float* d_series, int64_t n, const float* d_avg,
                           const float* d_std, int32_t w, int32_t algo, int32_t metric,
                           int32_t stats_mode, int32_t k, int64_t excl, int64_t* h_pos,
                           float* h_dist, mdtw_search_info* info) {
    int32_t wc = 0;
    int rc = check_search_args(e, d_series, n, d_avg, d_std, w, algo, metric, stats_mode, k, h_pos,
                               h_dist, &excl, &wc);
    if (rc != MDTW_OK) return rc;
    const int32_t m = e->m;
    const int64_t L = n - m + 1;
    for (int32_t i = 0; i < k; ++i) {
        h_pos[i] = -1;
        h_dist[i] = INFINITY;
    }
    DeviceGuard guard(e->device);
    if (algo == MDTW_ALGO_ED_FFT)
        return search_ed_fft(e, d_series, n, stats_mode, k, excl, h_pos, h_dist, info);

    // Query-side state first (host work and small uploads, outside timing).
    if (algo != MDTW_ALGO_ED) {
        rc = ensure_envelope(e, wc);
        if (rc == MDTW_OK && algo == MDTW_ALGO_CDTW) rc = ensure_qx(e, wc);
        if (rc != MDTW_OK) return rc;
    }
    void* ws = nullptr;
    rc = ws_reserve(e, search_ws_bytes(L, k, excl, d_avg == nullptr), &ws);
    if (rc != MDTW_OK) return rc;

    SearchPlan p;
    p.st = e->stream;
    p.cus = e->cus;
    p.s = d_series;
    p.n = n;
    p.d_avg = d_avg;
    p.d_std = d_std;
    p.stats_mode = stats_mode;
    p.e = e;
    p.q = e->d_q;
    p.lo = e->d_lo;
    p.up = e->d_up;
    p.qx = e->d_qx;
    p.qx_len = e->qx_len;
    if (algo == MDTW_ALGO_ED) p.ed = ed_query_sums(e->h_q.data(), m);
    p.m = m;
    p.w = wc;
    p.algo = algo;
    p.metric = metric;
    p.k = k;
    p.excl = excl;
    p.ctrl = e->d_ctrl;
    p.ws = static_cast<char*>(ws);
    SearchBufs bufs;
    hipStream_t st = e->stream;
    rc = hip_check(hipEventRecord(e->ev0, st));
    if (rc == MDTW_OK) rc = enqueue_search(p, &bufs);
    if (rc == MDTW_OK) rc = hip_check(hipEventRecord(e->ev1, st));
    if (rc == MDTW_OK)
        rc = hip_check(hipMemcpyAsync(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK) rc = hip_check(hipStreamSynchronize(st));
    if (rc != MDTW_OK) return rc;
    const Ctrl c = *e->h_ctrl;
    const bool topk = k > 1;

    int32_t found = 0;
    float thr_out;
    if (!topk) {
        if ((c.best >> 32) < 0x7f800000ull) {
            const uint32_t bits = static_cast<uint32_t>(c.best >> 32);
            float d;
            std::memcpy(&d, &bits, sizeof(d));
            h_pos[0] = static_cast<int64_t>(c.best & 0xffffffffull);
            h_dist[0] = metric == MDTW_METRIC_L2 ? std::sqrt(d) : d;
            found = 1;
            thr_out = d;
        } else {
            thr_out = INFINITY;
        }
    } else {
        thr_out = c.thr;
        std::vector<unsigned long long> keys(static_cast<size_t>(c.nres));
        if (c.nres) {
            rc = hip_check(hipMemcpyAsync(keys.data(), bufs.res,
                                          sizeof(unsigned long long) * c.nres,
                                          hipMemcpyDeviceToHost, st));
            if (rc == MDTW_OK) rc = hip_check(hipStreamSynchronize(st));
            if (rc != MDTW_OK) return rc;
        }
        found = host_topk_select(keys, k, excl, metric, h_pos, h_dist);
    }
    if (info) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
        info->windows = L;
        info->seeded = bufs.seeded;
        info->lb_kim_pruned = static_cast<int64_t>(c.kim);
        info->lb_keogh_pruned = static_cast<int64_t>(c.keogh);
        info->dp_evaluated = static_cast<int64_t>(c.dp);
        info->dp_abandoned = static_cast<int64_t>(c.abandoned);
        info->found = found;
        info->threshold = thr_out;
        info->elapsed_ms = ms;
    }
    return MDTW_OK;
}
