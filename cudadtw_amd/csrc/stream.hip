// stream.hip -- streaming search over a stream that arrives in host chunks.
//
// No reference API counterpart.  The reference handles long streams in
// epochs: avg_std runs over 10^7-sample epochs with M-1 samples of overlap
// (stats.cuh:115-142), and UCR_DTW_BIN.cpp:579-707 reads fixed-size chunks
// with M-1 samples of overlap, keeping its best-so-far across chunks.  Here
// the caller pushes samples of any size; the stream object
//   - copies them into one of two device epoch buffers on its own copy stream
//     (the next epoch's copy overlaps the current epoch's search),
//   - runs the whole search (enqueue_search) on every full epoch of
//     ew windows (ew + M - 1 samples; the next epoch starts with the last
//     M - 1 samples of this one), on the engine's stream,
//   - carries the answer from epoch to epoch: k = 1 the best (distance,
//     position), which starts the next epoch's search as its best; k > 1
//     every window at or below the carried bound T plus T itself.
// Exactness.  Every window belongs to exactly one epoch and its statistics,
// bounds and distance do not depend on the epoch: epochs start at multiples
// of 1024 windows, the tile of the FP64 statistics kernel.  k = 1: an
// epoch's best replaces the carried one only when strictly smaller, so ties
// keep the lower position.  k > 1: T is always the largest distance of k
// windows that are pairwise at least 2*excl apart (an epoch's own seed or
// bucket bound, or the host's bucket bound over the carried windows); one
// pick excludes at most one of them, so the k-th greedy pick over the whole
// stream is at most T, and every window at or below T is kept.  The answer
// is therefore the one-shot mdtw_search answer over the concatenated stream,
// bit for bit (positions and distances).
//
// Save / restore: the state (position, the unfinished epoch's samples, the
// carried answer, counters) serialises into a byte buffer, so a search can
// stop and resume in another process over the same query.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "cdtw_band.hpp"

using namespace mdtw;

namespace {

constexpr int64_t kEpochAlign = 1024;                 // stats tile, in windows
constexpr int64_t kEpochDefault = int64_t(1) << 24;  // windows per epoch
constexpr char kMagic[8] = {'M', 'D', 'T', 'W', 'S', 'T', 'R', '1'};

int hip_ok(hipError_t err) {
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Fixed header of a saved state (little-endian host layout).
struct SavedHeader {
    char magic[8];
    int32_t m, w, algo, metric, stats_mode, k;
    int64_t excl, ew;
    uint64_t query_hash;
    int64_t base, fill, windows_done;
    float best_d, thr;
    int64_t best_pos;
    int64_t seeded, kim, keogh, dp, abandoned;
    int64_t ncand;
};

struct Cand {
    float d;
    int64_t pos;
};

}  // namespace

struct mdtw_stream {
    mdtw_engine* e = nullptr;
    int32_t m = 0, w = 0, algo = 0, metric = 0, stats_mode = 0, k = 1;
    int64_t excl = 0;        // exclusion zone of the whole stream (host selection)
    int64_t excl_epoch = 0;  // the same zone as one epoch's search sees it (<= ew)
    int64_t ew = 0;  // windows per epoch
    int64_t es = 0;  // samples per epoch buffer
    uint64_t query_hash = 0;

    float* d_buf[2] = {nullptr, nullptr};
    float* d_q = nullptr;
    float* d_lo = nullptr;
    float* d_up = nullptr;
    float* d_qx = nullptr;
    int32_t qx_len = 0;
    EdQuerySums ed;  // ED: sums of the query the stream was created with
    char* d_ws = nullptr;
    Ctrl* d_ctrl = nullptr;
    Ctrl* h_ctrl = nullptr;  // pinned
    hipStream_t copy = nullptr;
    hipEvent_t filled = nullptr, start = nullptr, done = nullptr;

    int cur = 0;
    int64_t fill = 0;  // samples in d_buf[cur]
    int64_t base = 0;  // stream position of d_buf[cur][0]
    bool finished = false;

    bool pending = false;
    int64_t pend_base = 0, pend_windows = 0;
    SearchBufs bufs;

    int64_t windows_done = 0;
    float best_d = INFINITY;
    int64_t best_pos = -1;
    float thr = INFINITY;
    std::vector<Cand> cand;
    int64_t seeded = 0, kim = 0, keogh = 0, dp = 0, abandoned = 0;
    double ms = 0.0;

    ~mdtw_stream() {
        if (copy) (void)hipStreamSynchronize(copy);
        if (e && e->stream) (void)hipStreamSynchronize(e->stream);
        for (float* p : {d_buf[0], d_buf[1], d_q, d_lo, d_up, d_qx})
            if (p) (void)hipFree(p);
        if (d_ws) (void)hipFree(d_ws);
        if (d_ctrl) (void)hipFree(d_ctrl);
        if (h_ctrl) (void)hipHostFree(h_ctrl);
        if (filled) (void)hipEventDestroy(filled);
        if (start) (void)hipEventDestroy(start);
        if (done) (void)hipEventDestroy(done);
        if (copy) (void)hipStreamDestroy(copy);
    }
};

namespace {

uint64_t query_hash_of(const mdtw_engine* e) {
    // the first m entries: the host copy is sized for the longest query
    return fnv1a(e->h_q.data(), sizeof(float) * static_cast<size_t>(e->m),
                 fnv1a(&e->m, sizeof(e->m)));
}

int dev_alloc(void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return MDTW_ERR_ALLOC;
    }
    return MDTW_OK;
}

// Device copies of the query, its envelope for band w and the padded query
// of the register-band kernel: the stream keeps answering for the query it
// was created with, whatever mdtw_set_query does later.
int upload_query(mdtw_stream* s) {
    mdtw_engine* e = s->e;
    const int32_t m = s->m, w = s->w;
    std::vector<float> lo(m), up(m);
    host_envelope(e->h_q.data(), w, lo.data(), up.data(), m);
    s->ed = ed_query_sums(e->h_q.data(), m);
    const size_t len = static_cast<size_t>(m) + 2 * static_cast<size_t>(w) + kBandChunk + 8;
    std::vector<float> qx(len, kInf);
    for (int32_t i = 0; i < m; ++i) qx[static_cast<size_t>(i) + w] = e->h_q[i];
    int rc = dev_alloc(reinterpret_cast<void**>(&s->d_q), sizeof(float) * m);
    if (rc == MDTW_OK) rc = dev_alloc(reinterpret_cast<void**>(&s->d_lo), sizeof(float) * m);
    if (rc == MDTW_OK) rc = dev_alloc(reinterpret_cast<void**>(&s->d_up), sizeof(float) * m);
    if (rc == MDTW_OK) rc = dev_alloc(reinterpret_cast<void**>(&s->d_qx), sizeof(float) * len);
    if (rc != MDTW_OK) return rc;
    s->qx_len = static_cast<int32_t>(len);
    hipStream_t st = e->stream;
    rc = hip_ok(hipMemcpyAsync(s->d_q, e->h_q.data(), sizeof(float) * m, hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(s->d_lo, lo.data(), sizeof(float) * m, hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(s->d_up, up.data(), sizeof(float) * m, hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(s->d_qx, qx.data(), sizeof(float) * len, hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(st));
    return rc;
}

int make_stream(mdtw_engine* e, int32_t w, int32_t algo, int32_t metric, int32_t stats_mode,
                int32_t k, int64_t excl, int64_t epoch_windows, mdtw_stream** out) {
    if (!e || !out) return MDTW_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    if (w < 0) return MDTW_ERR_BAND;
    if (algo != MDTW_ALGO_CDTW && algo != MDTW_ALGO_ED && algo != MDTW_ALGO_LDTW)
        return MDTW_ERR_INVALID_ARGUMENT;
    if (metric != MDTW_METRIC_L1 && metric != MDTW_METRIC_L2) return MDTW_ERR_INVALID_ARGUMENT;
    if (algo == MDTW_ALGO_ED && metric != MDTW_METRIC_L2) return MDTW_ERR_UNSUPPORTED;
    // The REFCOMPAT statistics run a global prefix from the stream's start,
    // which no epoch sees; FP64 and NONE are per window.
    if (stats_mode != MDTW_STATS_FP64 && stats_mode != MDTW_STATS_NONE)
        return stats_mode == MDTW_STATS_REFCOMPAT ? MDTW_ERR_UNSUPPORTED
                                                  : MDTW_ERR_INVALID_ARGUMENT;
    if (k < 1 || k > MDTW_MAX_K) return MDTW_ERR_K;
    if (excl < 0) return MDTW_ERR_INVALID_ARGUMENT;
    const int32_t m = e->m;
    const int32_t wc = std::min(w, m - 1);
    if (algo == MDTW_ALGO_CDTW && wc > kMaxBandW) return MDTW_ERR_UNSUPPORTED;
    if (epoch_windows == 0) epoch_windows = kEpochDefault;
    // An epoch is one search: its window indices must stay below 2^31.
    if (epoch_windows < kEpochAlign || epoch_windows % kEpochAlign != 0 ||
        epoch_windows > kMaxSearchWindows)
        return MDTW_ERR_INVALID_ARGUMENT;
    // Zones beyond 2^61 windows behave as 2^61 on any stream that can exist
    // and keep 2 * excl inside int64; within one epoch a zone of ew or more
    // excludes every other window, as a zone of ew does.
    excl = std::min<int64_t>(excl, int64_t(1) << 61);
    const int64_t excl_epoch = clamp_excl(excl, epoch_windows);

    DeviceGuard guard(e->device);
    mdtw_stream* s = new mdtw_stream();
    s->e = e;
    s->m = m;
    s->w = wc;
    s->algo = algo;
    s->metric = metric;
    s->stats_mode = stats_mode;
    s->k = k;
    s->excl = excl;
    s->excl_epoch = excl_epoch;
    s->ew = epoch_windows;
    s->es = epoch_windows + m - 1;
    s->query_hash = query_hash_of(e);
    int rc = upload_query(s);
    for (int b = 0; b < 2 && rc == MDTW_OK; ++b)
        rc = dev_alloc(reinterpret_cast<void**>(&s->d_buf[b]), sizeof(float) * s->es);
    if (rc == MDTW_OK)
        rc = dev_alloc(reinterpret_cast<void**>(&s->d_ws),
                       search_ws_bytes(s->ew, k, excl_epoch, true));
    if (rc == MDTW_OK) rc = dev_alloc(reinterpret_cast<void**>(&s->d_ctrl), sizeof(Ctrl));
    if (rc == MDTW_OK && hipHostMalloc(reinterpret_cast<void**>(&s->h_ctrl), sizeof(Ctrl),
                                       hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        s->h_ctrl = nullptr;
        rc = MDTW_ERR_ALLOC;
    }
    if (rc == MDTW_OK) rc = hip_ok(hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking));
    if (rc == MDTW_OK) rc = hip_ok(hipEventCreateWithFlags(&s->filled, hipEventDisableTiming));
    if (rc == MDTW_OK) rc = hip_ok(hipEventCreate(&s->start));
    if (rc == MDTW_OK) rc = hip_ok(hipEventCreate(&s->done));
    if (rc != MDTW_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return MDTW_OK;
}

// Bucket bound over the carried windows (buckets of 2*excl positions split by
// parity, as the device search's tighten step; single windows for
// excl < 2), then drop every carried window above the bound.
void tighten_host(mdtw_stream* s) {
    const int32_t k = s->k;
    if (static_cast<int32_t>(s->cand.size()) >= k) {
        const bool split = s->excl >= 2;
        const int64_t blen = split ? 2 * s->excl : 1;
        std::unordered_map<int64_t, float> bmin;
        bmin.reserve(s->cand.size() * 2);
        for (const Cand& c : s->cand) {
            auto it = bmin.find(c.pos / blen);
            if (it == bmin.end())
                bmin.emplace(c.pos / blen, c.d);
            else if (c.d < it->second)
                it->second = c.d;
        }
        for (int par = 0; par < (split ? 2 : 1); ++par) {
            std::vector<float> v;
            for (const auto& kv : bmin)
                if (!split || (kv.first & 1) == par) v.push_back(kv.second);
            if (static_cast<int32_t>(v.size()) < k) continue;
            std::nth_element(v.begin(), v.begin() + (k - 1), v.end());
            s->thr = std::min(s->thr, v[k - 1]);
        }
    }
    const float t = s->thr;
    s->cand.erase(std::remove_if(s->cand.begin(), s->cand.end(),
                                 [t](const Cand& c) { return !(c.d <= t); }),
                  s->cand.end());
}

// Wait for the enqueued epoch and fold its answer into the carried one.
int collect(mdtw_stream* s) {
    if (!s->pending) return MDTW_OK;
    int rc = hip_ok(hipEventSynchronize(s->done));
    if (rc != MDTW_OK) return rc;
    const Ctrl c = *s->h_ctrl;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s->start, s->done) == hipSuccess) s->ms += ms;
    s->seeded += s->bufs.seeded;
    s->kim += static_cast<int64_t>(c.kim);
    s->keogh += static_cast<int64_t>(c.keogh);
    s->dp += static_cast<int64_t>(c.dp);
    s->abandoned += static_cast<int64_t>(c.abandoned);
    if (s->k == 1) {
        if ((c.best >> 32) < 0x7f800000ull) {
            const uint32_t bits = static_cast<uint32_t>(c.best >> 32);
            float d;
            std::memcpy(&d, &bits, sizeof(d));
            // Strictly smaller only: an equal distance here lies at a later
            // position than the carried window (and the carried key itself
            // comes back unchanged when nothing beat it).
            if (d < s->best_d) {
                s->best_d = d;
                s->best_pos = s->pend_base + static_cast<int64_t>(c.best & 0xffffffffull);
            }
        }
    } else {
        s->thr = std::min(s->thr, c.thr);
        if (c.nres) {
            std::vector<unsigned long long> keys(static_cast<size_t>(c.nres));
            rc = hip_ok(hipMemcpyAsync(keys.data(), s->bufs.res,
                                       sizeof(unsigned long long) * c.nres,
                                       hipMemcpyDeviceToHost, s->e->stream));
            if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(s->e->stream));
            if (rc != MDTW_OK) return rc;
            for (unsigned long long key : keys) {
                const uint32_t bits = static_cast<uint32_t>(key >> 32);
                float d;
                std::memcpy(&d, &bits, sizeof(d));
                if (d <= s->thr)
                    s->cand.push_back(
                        {d, s->pend_base + static_cast<int64_t>(key & 0xffffffffull)});
            }
        }
        tighten_host(s);
    }
    s->windows_done += s->pend_windows;
    s->pending = false;
    return MDTW_OK;
}

// Search the nsamples samples of the current buffer (their copies are queued
// on the copy stream), then start the next buffer with their last M - 1.
int run_epoch(mdtw_stream* s, int64_t nsamples) {
    mdtw_engine* e = s->e;
    int rc = collect(s);  // the previous epoch's answer sets this one's bound
    if (rc != MDTW_OK) return rc;
    hipStream_t st = e->stream;
    rc = hip_ok(hipEventRecord(s->filled, s->copy));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamWaitEvent(st, s->filled, 0));
    if (rc != MDTW_OK) return rc;

    SearchPlan p;
    p.st = st;
    p.cus = e->cus;
    p.s = s->d_buf[s->cur];
    p.n = nsamples;
    p.stats_mode = s->stats_mode;
    p.e = e;
    p.q = s->d_q;
    p.lo = s->d_lo;
    p.up = s->d_up;
    p.qx = s->d_qx;
    p.qx_len = s->qx_len;
    p.ed = s->ed;
    p.m = s->m;
    p.w = s->w;
    p.algo = s->algo;
    p.metric = s->metric;
    p.k = s->k;
    p.excl = s->excl_epoch;
    p.ctrl = s->d_ctrl;
    p.ws = s->d_ws;
    if (s->k == 1 && s->best_pos >= 0) {
        uint32_t bits;
        std::memcpy(&bits, &s->best_d, sizeof(bits));
        p.best_init = static_cast<unsigned long long>(bits) << 32;  // position 0
    }
    if (s->k > 1) p.thr_cap = s->thr;
    // A bound carried from earlier epochs replaces the seeds' bound.
    p.seed = s->k == 1 ? s->best_pos < 0 : !(s->thr < INFINITY);
    rc = hip_ok(hipEventRecord(s->start, st));
    if (rc == MDTW_OK) rc = enqueue_search(p, &s->bufs);
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(s->h_ctrl, s->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK) rc = hip_ok(hipEventRecord(s->done, st));
    if (rc != MDTW_OK) return rc;
    s->pending = true;
    s->pend_base = s->base;
    s->pend_windows = nsamples - s->m + 1;

    // The other buffer was last searched by the epoch collected above.
    const int nxt = s->cur ^ 1;
    const int64_t keep = s->m - 1;
    if (keep > 0)
        rc = hip_ok(hipMemcpyAsync(s->d_buf[nxt], s->d_buf[s->cur] + (nsamples - keep),
                                   sizeof(float) * keep, hipMemcpyDeviceToDevice, s->copy));
    s->cur = nxt;
    s->base += nsamples - keep;
    s->fill = keep;
    return rc;
}

}  // namespace

extern "C" {

int mdtw_stream_create(mdtw_engine* e, int32_t w, int32_t algo, int32_t metric,
                       int32_t stats_mode, int32_t k, int64_t excl, int64_t epoch_windows,
                       mdtw_stream** out) {
    return make_stream(e, w, algo, metric, stats_mode, k, excl, epoch_windows, out);
}

int mdtw_stream_destroy(mdtw_stream* s) {
    if (!s) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(s->e->device);
    delete s;
    return MDTW_OK;
}

int mdtw_stream_push(mdtw_stream* s, const float* h_samples, int64_t count) {
    if (!s || count < 0 || (count > 0 && !h_samples)) return MDTW_ERR_INVALID_ARGUMENT;
    if (s->finished) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(s->e->device);
    while (count > 0) {
        const int64_t take = std::min(count, s->es - s->fill);
        int rc = hip_ok(hipMemcpyAsync(s->d_buf[s->cur] + s->fill, h_samples, sizeof(float) * take,
                                       hipMemcpyHostToDevice, s->copy));
        if (rc != MDTW_OK) return rc;
        s->fill += take;
        h_samples += take;
        count -= take;
        if (s->fill == s->es) {
            rc = run_epoch(s, s->es);
            if (rc != MDTW_OK) return rc;
        }
    }
    return MDTW_OK;
}

int mdtw_stream_finish(mdtw_stream* s) {
    if (!s) return MDTW_ERR_INVALID_ARGUMENT;
    if (s->finished) return MDTW_OK;
    DeviceGuard guard(s->e->device);
    int rc = MDTW_OK;
    if (s->fill >= s->m && s->fill > s->m - 1) rc = run_epoch(s, s->fill);
    if (rc == MDTW_OK) rc = collect(s);
    if (rc == MDTW_OK) s->finished = true;
    return rc;
}

int mdtw_stream_result(mdtw_stream* s, int64_t* h_pos, float* h_dist, mdtw_search_info* info) {
    if (!s || !h_pos || !h_dist) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(s->e->device);
    int rc = collect(s);
    if (rc != MDTW_OK) return rc;
    for (int32_t i = 0; i < s->k; ++i) {
        h_pos[i] = -1;
        h_dist[i] = INFINITY;
    }
    int32_t found = 0;
    float thr_out = s->thr;
    if (s->k == 1) {
        if (s->best_pos >= 0) {
            h_pos[0] = s->best_pos;
            h_dist[0] = s->metric == MDTW_METRIC_L2 ? std::sqrt(s->best_d) : s->best_d;
            found = 1;
        }
        thr_out = s->best_d;
    } else {
        std::vector<std::pair<float, int64_t>> v;
        v.reserve(s->cand.size());
        for (const Cand& c : s->cand) v.emplace_back(c.d, c.pos);
        std::sort(v.begin(), v.end());
        found = greedy_pick(v, s->k, s->excl, s->metric, h_pos, h_dist);
    }
    if (info) {
        info->windows = s->windows_done;
        info->seeded = s->seeded;
        info->lb_kim_pruned = s->kim;
        info->lb_keogh_pruned = s->keogh;
        info->dp_evaluated = s->dp;
        info->dp_abandoned = s->abandoned;
        info->found = found;
        info->threshold = thr_out;
        info->elapsed_ms = static_cast<float>(s->ms);
    }
    return MDTW_OK;
}

int mdtw_stream_position(mdtw_stream* s, int64_t* samples_pushed, int64_t* windows_done) {
    if (!s) return MDTW_ERR_INVALID_ARGUMENT;
    if (samples_pushed) *samples_pushed = s->base + s->fill;
    if (windows_done) *windows_done = s->windows_done + (s->pending ? s->pend_windows : 0);
    return MDTW_OK;
}

int mdtw_stream_save(mdtw_stream* s, void* h_buf, size_t cap, size_t* size) {
    if (!s || !size) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(s->e->device);
    int rc = collect(s);
    if (rc != MDTW_OK) return rc;
    const size_t need = sizeof(SavedHeader) + sizeof(float) * s->fill +
                        (sizeof(float) + sizeof(int64_t)) * s->cand.size();
    *size = need;
    if (!h_buf) return MDTW_OK;
    if (cap < need) return MDTW_ERR_INVALID_ARGUMENT;
    SavedHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.m = s->m;
    h.w = s->w;
    h.algo = s->algo;
    h.metric = s->metric;
    h.stats_mode = s->stats_mode;
    h.k = s->k;
    h.excl = s->excl;
    h.ew = s->ew;
    h.query_hash = s->query_hash;
    h.base = s->base;
    h.fill = s->fill;
    h.windows_done = s->windows_done;
    h.best_d = s->best_d;
    h.thr = s->thr;
    h.best_pos = s->best_pos;
    h.seeded = s->seeded;
    h.kim = s->kim;
    h.keogh = s->keogh;
    h.dp = s->dp;
    h.abandoned = s->abandoned;
    h.ncand = static_cast<int64_t>(s->cand.size());
    char* out = static_cast<char*>(h_buf);
    std::memcpy(out, &h, sizeof(h));
    out += sizeof(h);
    if (s->fill) {
        rc = hip_ok(hipStreamSynchronize(s->copy));
        if (rc == MDTW_OK)
            rc = hip_ok(hipMemcpy(out, s->d_buf[s->cur], sizeof(float) * s->fill,
                                  hipMemcpyDeviceToHost));
        if (rc != MDTW_OK) return rc;
        out += sizeof(float) * s->fill;
    }
    for (const Cand& c : s->cand) {
        std::memcpy(out, &c.d, sizeof(float));
        std::memcpy(out + sizeof(float), &c.pos, sizeof(int64_t));
        out += sizeof(float) + sizeof(int64_t);
    }
    return MDTW_OK;
}

int mdtw_stream_restore(mdtw_engine* e, const void* h_buf, size_t size, mdtw_stream** out) {
    if (!e || !h_buf || !out || size < sizeof(SavedHeader)) return MDTW_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    SavedHeader h;
    std::memcpy(&h, h_buf, sizeof(h));
    if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    if (h.m != e->m || h.query_hash != query_hash_of(e)) return MDTW_ERR_INVALID_ARGUMENT;
    // Every field is checked before it sizes anything: a truncated or corrupt
    // buffer is an argument error, never an allocation or an overflow.
    if (h.ew < kEpochAlign || h.ew % kEpochAlign != 0 || h.ew > kMaxSearchWindows)
        return MDTW_ERR_INVALID_ARGUMENT;
    if (h.fill < 0 || h.ncand < 0 || h.fill > h.ew + h.m - 1 || h.base < 0 ||
        h.base % kEpochAlign != 0)
        return MDTW_ERR_INVALID_ARGUMENT;
    const size_t fill_bytes = sizeof(float) * static_cast<size_t>(h.fill);
    if (size - sizeof(SavedHeader) < fill_bytes) return MDTW_ERR_INVALID_ARGUMENT;
    const size_t per_cand = sizeof(float) + sizeof(int64_t);
    if (static_cast<uint64_t>(h.ncand) > (size - sizeof(SavedHeader) - fill_bytes) / per_cand)
        return MDTW_ERR_INVALID_ARGUMENT;
    // Windows searched so far start at 0 and end where the buffered samples
    // begin; positions are windows already searched.
    if (h.windows_done < 0 || h.windows_done > h.base) return MDTW_ERR_INVALID_ARGUMENT;
    if (h.best_pos < -1 || h.best_pos >= h.base) return MDTW_ERR_INVALID_ARGUMENT;
    {
        const char* cp = static_cast<const char*>(h_buf) + sizeof(h) + fill_bytes;
        for (int64_t i = 0; i < h.ncand; ++i, cp += per_cand) {
            int64_t pos;
            std::memcpy(&pos, cp + sizeof(float), sizeof(pos));
            if (pos < 0 || pos >= h.base) return MDTW_ERR_INVALID_ARGUMENT;
        }
    }
    mdtw_stream* s = nullptr;
    int rc = make_stream(e, h.w, h.algo, h.metric, h.stats_mode, h.k, h.excl, h.ew, &s);
    if (rc != MDTW_OK) return rc;
    DeviceGuard guard(e->device);
    const char* in = static_cast<const char*>(h_buf) + sizeof(h);
    if (h.fill) {
        rc = hip_ok(hipMemcpy(s->d_buf[0], in, sizeof(float) * h.fill, hipMemcpyHostToDevice));
        in += sizeof(float) * h.fill;
    }
    if (rc != MDTW_OK) {
        delete s;
        return rc;
    }
    s->cur = 0;
    s->fill = h.fill;
    s->base = h.base;
    s->windows_done = h.windows_done;
    s->best_d = h.best_d;
    s->best_pos = h.best_pos;
    s->thr = h.thr;
    s->seeded = h.seeded;
    s->kim = h.kim;
    s->keogh = h.keogh;
    s->dp = h.dp;
    s->abandoned = h.abandoned;
    try {
        s->cand.resize(static_cast<size_t>(h.ncand));
    } catch (...) {
        delete s;
        return MDTW_ERR_ALLOC;
    }
    for (Cand& c : s->cand) {
        std::memcpy(&c.d, in, sizeof(float));
        std::memcpy(&c.pos, in + sizeof(float), sizeof(int64_t));
        in += sizeof(float) + sizeof(int64_t);
    }
    *out = s;
    return MDTW_OK;
}

}  // extern "C"
