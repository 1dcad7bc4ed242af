// batch.hip -- many queries of mixed length and band against one stream.
//
// No reference counterpart: the reference keeps one query per process in
// __constant__ memory (cdtw.cuh:9-11) and its driver runs one search per
// invocation.  A batch uploads all queries once (z-normalised query,
// envelope and padded copy per query, one device arena), then runs every
// query's search (enqueue_search: the plain search's kernels, no host round
// trip) on P concurrent pipelines (streams), each query followed by an
// on-device greedy top-k selection, and finally copies all results to pinned
// host memory.  The whole sequence is captured once into a hipGraph and
// replayed: one graph launch per batch search.  Queries heavy enough to fill
// the device on their own run one after another on the first pipeline; the
// others are assigned longest-first (largest L * cells(M, W) to the
// least-loaded pipeline), so their kernels overlap on the device and fill
// the tails of one query's kernels with another's work.
//
// Launch shapes are picked per (|Q|, band) bucket by mdtw_batch_tune: the
// persistent DP grid (windows per workgroup pass times workgroups), the
// anti-diagonal DP grid and the cascade grid, and then the DP's formulation
// where more than one applies (DpForm: register band with or without the
// cumulative LB_Keogh bound, or the anti-diagonal kernel with the band's
// diagonals in LDS for W = 113..240).  Each candidate runs the search of the
// bucket's median-cost query on a prefix of the stream and the fastest is
// kept.  Shapes change only how work is spread over the device, never the
// result.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "cdtw_band.hpp"

using namespace mdtw;

namespace {

constexpr int kDefaultPipelines = 4;
constexpr int kMaxPipelines = 16;
// Band cells (times the per-cell cost class) of a query from which it runs
// serialized with the other such queries (capture): 2^35, about 2 ms of DP.
constexpr double kHeavyCells = 34359738368.0;

int64_t band_cells(int32_t m, int32_t w) {
    return static_cast<int64_t>(m) * (2 * w + 1) - static_cast<int64_t>(w) * (w + 1);
}

// Bucket of a query for launch-shape tuning: power-of-two length class and
// band class (one lane per window with the band in registers; the LDS-tail
// and two-wave split bands; the anti-diagonal kernel only).
int shape_bucket(int32_t m, int32_t w) {
    int lg = 0;
    while ((1 << (lg + 1)) <= m) ++lg;
    return lg * 3 + (w <= kBandMaxW ? 0 : w <= kBandWideMaxW ? 1 : 2);
}

int hip_ok(hipError_t err) { return hip_status(err); }

}  // namespace

struct mdtw_batch {
    mdtw_engine* e = nullptr;
    int32_t nq = 0;
    std::vector<int32_t> m, w;
    std::vector<int64_t> off, qx_off;  // float offsets into the arena
    std::vector<int32_t> qx_len;
    float* d_arena = nullptr;

    int P = 0;
    std::vector<hipStream_t> streams;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipEvent_t t0 = nullptr, t1 = nullptr;

    std::vector<char*> slot_ws;
    size_t slot_ws_bytes = 0;
    Ctrl* d_ctrls = nullptr;
    Ctrl* h_ctrls = nullptr;  // pinned
    int64_t* d_pos = nullptr;
    int64_t* h_pos = nullptr;  // pinned
    float* d_dist = nullptr;
    float* h_dist = nullptr;   // pinned
    int32_t* d_ovf = nullptr;
    int32_t* h_ovf = nullptr;  // pinned
    int32_t out_k = 0;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool have_graph = false;
    const float* g_s = nullptr;
    int64_t g_n = 0;
    int32_t g_k = 0, g_metric = -1;
    std::vector<int64_t> g_excl;
    uint64_t g_shapes = ~0ull;

    std::map<int, LaunchShape> shapes;
    uint64_t shapes_version = 0;

    // Backward: one device arena for the pair lists, statistics, raw
    // queries and outputs, and one scratch per pipeline (grown on demand).
    char* bwd_buf = nullptr;
    size_t bwd_cap = 0;
    std::vector<char*> bwd_ws;
    size_t bwd_ws_cap = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        have_graph = false;
    }
    void free_outputs() {
        if (d_ctrls) (void)hipFree(d_ctrls);
        if (d_pos) (void)hipFree(d_pos);
        if (d_dist) (void)hipFree(d_dist);
        if (d_ovf) (void)hipFree(d_ovf);
        if (h_ctrls) (void)hipHostFree(h_ctrls);
        if (h_pos) (void)hipHostFree(h_pos);
        if (h_dist) (void)hipHostFree(h_dist);
        if (h_ovf) (void)hipHostFree(h_ovf);
        d_ctrls = h_ctrls = nullptr;
        d_pos = h_pos = nullptr;
        d_dist = h_dist = nullptr;
        d_ovf = h_ovf = nullptr;
        out_k = 0;
    }
    void free_ws() {
        for (char* p : slot_ws)
            if (p) (void)hipFree(p);
        slot_ws.clear();
        slot_ws_bytes = 0;
    }
    ~mdtw_batch() {
        drop_graph();
        free_outputs();
        free_ws();
        if (bwd_buf) (void)hipFree(bwd_buf);
        for (char* p : bwd_ws)
            if (p) (void)hipFree(p);
        if (d_arena) (void)hipFree(d_arena);
        for (hipStream_t s : streams)
            if (s) (void)hipStreamDestroy(s);
        if (origin) (void)hipStreamDestroy(origin);
        if (fork) (void)hipEventDestroy(fork);
        for (hipEvent_t ev : join)
            if (ev) (void)hipEventDestroy(ev);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
    }

    const LaunchShape& shape_of(int32_t i) const {
        static const LaunchShape kDefault;
        auto it = shapes.find(shape_bucket(m[i], w[i]));
        return it == shapes.end() ? kDefault : it->second;
    }

    SearchPlan plan(int32_t i, hipStream_t st, const float* s, int64_t n, int32_t metric,
                    int32_t k, int64_t excl, Ctrl* ctrl, char* ws,
                    const LaunchShape& shape) const {
        SearchPlan p;
        p.st = st;
        p.cus = e->cus;
        p.s = s;
        p.n = n;
        p.stats_mode = MDTW_STATS_FP64;
        p.q = d_arena + off[i];
        p.lo = p.q + m[i];
        p.up = p.lo + m[i];
        p.qx = d_arena + qx_off[i];  // launch_dp picks the kernel for (w, metric)
        p.qx_len = qx_len[i];
        p.m = m[i];
        p.w = w[i];
        p.algo = MDTW_ALGO_CDTW;
        p.metric = metric;
        p.k = k;
        p.excl = excl;
        p.ctrl = ctrl;
        p.ws = ws;
        p.shape = shape;
        return p;
    }
};

namespace {

int ensure_outputs(mdtw_batch* b, int32_t k) {
    if (b->out_k >= k) return MDTW_OK;
    b->drop_graph();
    b->free_outputs();
    const size_t nk = static_cast<size_t>(b->nq) * k;
    if (hipMalloc(&b->d_ctrls, sizeof(Ctrl) * b->nq) != hipSuccess ||
        hipMalloc(&b->d_pos, sizeof(int64_t) * nk) != hipSuccess ||
        hipMalloc(&b->d_dist, sizeof(float) * nk) != hipSuccess ||
        hipMalloc(&b->d_ovf, sizeof(int32_t) * b->nq) != hipSuccess ||
        hipHostMalloc(&b->h_ctrls, sizeof(Ctrl) * b->nq, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&b->h_pos, sizeof(int64_t) * nk, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&b->h_dist, sizeof(float) * nk, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&b->h_ovf, sizeof(int32_t) * b->nq, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        b->free_outputs();
        return MDTW_ERR_ALLOC;
    }
    b->out_k = k;
    return MDTW_OK;
}

int ensure_ws(mdtw_batch* b, size_t bytes) {
    if (b->slot_ws_bytes >= bytes && static_cast<int>(b->slot_ws.size()) == b->P) return MDTW_OK;
    b->drop_graph();
    b->free_ws();
    bytes = (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    for (int p = 0; p < b->P; ++p) {
        char* ptr = nullptr;
        if (hipMalloc(&ptr, bytes) != hipSuccess) {
            (void)hipGetLastError();
            b->free_ws();
            return MDTW_ERR_ALLOC;
        }
        b->slot_ws.push_back(ptr);
    }
    b->slot_ws_bytes = bytes;
    return MDTW_OK;
}

// Records the batch's device work into a graph.
int capture(mdtw_batch* b, const float* s, int64_t n, int32_t metric, int32_t k,
            const std::vector<int64_t>& excl) {
    b->drop_graph();
    const int P = b->P;
    // Longest first onto the least-loaded pipeline.
    std::vector<int32_t> order(b->nq);
    std::iota(order.begin(), order.end(), 0);
    std::vector<double> cost(b->nq);
    // Relative time per cell of the kernels a band lands on (register band
    // at two or more waves per SIMD, ~18 T cells/s; wider register bands,
    // LDS tail or skewed columns at one wave per SIMD, ~9-16; anti-diagonal).
    for (int32_t i = 0; i < b->nq; ++i) {
        const int32_t w = b->w[i];
        const double rate = w <= kBandMaxW ? 1.0 : w <= kBandWideMaxW ? 1.7 : 6.5;
        cost[i] = static_cast<double>(n - b->m[i] + 1) * band_cells(b->m[i], w) * rate;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t c) { return cost[a] > cost[c]; });
    std::vector<double> load(P, 0.0);

    int rc = hip_ok(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    if (rc != MDTW_OK) return rc;
    rc = hip_ok(hipEventRecord(b->fork, b->origin));
    for (int p = 0; p < P && rc == MDTW_OK; ++p)
        rc = hip_ok(hipStreamWaitEvent(b->streams[p], b->fork, 0));
    for (int32_t i : order) {
        if (rc != MDTW_OK) break;
        // A query whose work alone keeps the device busy for milliseconds
        // runs on pipeline 0, one after another: side by side such queries'
        // grid-filling persistent kernels only compete (C4's sample of 32
        // queries on a 2^26 stream: 13.09 s on four pipelines, 12.57 s one
        // after another).  Lighter ones go to the least-loaded pipeline,
        // where their short kernels fill each other's gaps (and the heavy
        // queries' tails).
        const int p = cost[i] >= kHeavyCells
                          ? 0
                          : static_cast<int>(std::min_element(load.begin(), load.end()) -
                                             load.begin());
        load[p] += cost[i];
        SearchPlan pl = b->plan(i, b->streams[p], s, n, metric, k, excl[i], b->d_ctrls + i,
                                b->slot_ws[p], b->shape_of(i));
        SearchBufs bufs;
        rc = enqueue_search(pl, &bufs);
        if (rc == MDTW_OK)
            rc = launch_topk_select(b->streams[p], b->d_ctrls + i, bufs.res, k, excl[i], metric,
                                    b->d_pos + static_cast<size_t>(i) * k,
                                    b->d_dist + static_cast<size_t>(i) * k, b->d_ovf + i);
    }
    for (int p = 0; p < P && rc == MDTW_OK; ++p) {
        rc = hip_ok(hipEventRecord(b->join[p], b->streams[p]));
        if (rc == MDTW_OK) rc = hip_ok(hipStreamWaitEvent(b->origin, b->join[p], 0));
    }
    const size_t nk = static_cast<size_t>(b->nq) * k;
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(b->h_pos, b->d_pos, sizeof(int64_t) * nk,
                                   hipMemcpyDeviceToHost, b->origin));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(b->h_dist, b->d_dist, sizeof(float) * nk,
                                   hipMemcpyDeviceToHost, b->origin));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(b->h_ctrls, b->d_ctrls, sizeof(Ctrl) * b->nq,
                                   hipMemcpyDeviceToHost, b->origin));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(b->h_ovf, b->d_ovf, sizeof(int32_t) * b->nq,
                                   hipMemcpyDeviceToHost, b->origin));
    hipGraph_t g = nullptr;
    const int rc_end = hip_ok(hipStreamEndCapture(b->origin, &g));
    if (rc == MDTW_OK) rc = rc_end;
    if (rc != MDTW_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    b->graph = g;
    rc = hip_ok(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    if (rc != MDTW_OK) {
        b->drop_graph();
        return rc;
    }
    b->have_graph = true;
    b->g_s = s;
    b->g_n = n;
    b->g_k = k;
    b->g_metric = metric;
    b->g_excl = excl;
    b->g_shapes = b->shapes_version;
    return MDTW_OK;
}

// One query searched eagerly on the engine's stream, selected on the host
// (tuning runs, and the rare query whose kept keys overflow the device
// selection).
int run_one(mdtw_batch* b, int32_t i, const float* s, int64_t n, int32_t metric, int32_t k,
            int64_t excl, const LaunchShape& shape, int64_t* pos, float* dist, Ctrl* c_out,
            float* ms) {
    mdtw_engine* e = b->e;
    SearchPlan pl = b->plan(i, e->stream, s, n, metric, k, excl, e->d_ctrl, b->slot_ws[0], shape);
    SearchBufs bufs;
    int rc = hip_ok(hipEventRecord(e->ev0, e->stream));
    if (rc == MDTW_OK) rc = enqueue_search(pl, &bufs);
    if (rc == MDTW_OK) rc = hip_ok(hipEventRecord(e->ev1, e->stream));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost,
                                   e->stream));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(e->stream));
    if (rc != MDTW_OK) return rc;
    const Ctrl c = *e->h_ctrl;
    if (c_out) *c_out = c;
    if (ms) (void)hipEventElapsedTime(ms, e->ev0, e->ev1);
    if (!pos) return MDTW_OK;
    for (int32_t j = 0; j < k; ++j) {
        pos[j] = -1;
        dist[j] = INFINITY;
    }
    std::vector<unsigned long long> keys;
    if (k == 1) {
        if ((c.best >> 32) < 0x7f800000ull) keys.push_back(c.best);
    } else if (c.nres) {
        keys.resize(static_cast<size_t>(c.nres));
        rc = hip_ok(hipMemcpy(keys.data(), bufs.res, sizeof(unsigned long long) * c.nres,
                              hipMemcpyDeviceToHost));
        if (rc != MDTW_OK) return rc;
    }
    std::sort(keys.begin(), keys.end());
    int32_t found = 0;
    for (unsigned long long key : keys) {
        if (found == k) break;
        const int64_t p = static_cast<int64_t>(key & 0xffffffffull);
        bool ok = true;
        for (int32_t j = 0; j < found; ++j)
            if ((p > pos[j] ? p - pos[j] : pos[j] - p) < excl) ok = false;
        if (!ok) continue;
        const uint32_t bits = static_cast<uint32_t>(key >> 32);
        float d;
        std::memcpy(&d, &bits, sizeof(d));
        pos[found] = p;
        dist[found] = metric == MDTW_METRIC_L2 ? std::sqrt(d) : d;
        ++found;
    }
    return MDTW_OK;
}

int check_common(mdtw_batch* b, const float* s, int64_t n, int32_t metric, int32_t k) {
    if (!b || !s) return MDTW_ERR_INVALID_ARGUMENT;
    if (k < 1 || k > MDTW_MAX_K) return MDTW_ERR_K;
    if (metric != MDTW_METRIC_L1 && metric != MDTW_METRIC_L2) return MDTW_ERR_INVALID_ARGUMENT;
    for (int32_t i = 0; i < b->nq; ++i) {
        if (n < b->m[i]) return MDTW_ERR_STREAM_TOO_SHORT;
        if (n - b->m[i] + 1 > kMaxSearchWindows) return MDTW_ERR_UNSUPPORTED;
    }
    return MDTW_OK;
}

std::vector<int64_t> excl_of(const mdtw_batch* b, const int64_t* h_excl, int64_t n) {
    std::vector<int64_t> ex(b->nq);
    for (int32_t i = 0; i < b->nq; ++i)
        ex[i] = clamp_excl(h_excl ? h_excl[i] : b->m[i] / 2, n - b->m[i] + 1);
    return ex;
}

size_t slot_bytes(const mdtw_batch* b, int64_t n, int32_t k, const std::vector<int64_t>& ex) {
    size_t bytes = 0;
    for (int32_t i = 0; i < b->nq; ++i)
        bytes = std::max(bytes, search_ws_bytes(n - b->m[i] + 1, k, ex[i], true));
    return bytes;
}

}  // namespace

extern "C" {

int mdtw_batch_create(mdtw_engine* e, const float* h_zqueries, const int32_t* h_lengths,
                      const int32_t* h_bands, int32_t nq, int32_t pipelines, mdtw_batch** out) {
    if (!e || !h_zqueries || !h_lengths || !h_bands || !out || nq < 1)
        return MDTW_ERR_INVALID_ARGUMENT;
    if (pipelines < 0 || pipelines > kMaxPipelines) return MDTW_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    DeviceGuard guard(e->device);
    auto b = new (std::nothrow) mdtw_batch();
    if (!b) return MDTW_ERR_ALLOC;
    b->e = e;
    b->nq = nq;
    b->P = pipelines ? pipelines : kDefaultPipelines;
    b->m.resize(nq);
    b->w.resize(nq);
    b->off.resize(nq);
    b->qx_off.resize(nq);
    b->qx_len.resize(nq);
    int64_t total = 0, src = 0;
    for (int32_t i = 0; i < nq; ++i) {
        const int32_t m = h_lengths[i];
        if (m < 1) {
            delete b;
            return MDTW_ERR_QUERY_TOO_SHORT;
        }
        if (m > MDTW_MAX_QUERY) {
            delete b;
            return MDTW_ERR_QUERY_TOO_LONG;
        }
        if (h_bands[i] < 0) {
            delete b;
            return MDTW_ERR_BAND;
        }
        const int32_t w = std::min(h_bands[i], m - 1);
        if (w > kMaxBandW) {
            delete b;
            return MDTW_ERR_UNSUPPORTED;
        }
        for (int32_t j = 0; j < m; ++j)
            if (!std::isfinite(h_zqueries[src + j])) {
                delete b;
                return MDTW_ERR_NOT_FINITE;
            }
        b->m[i] = m;
        b->w[i] = w;
        b->off[i] = total;
        total += 3 * static_cast<int64_t>(m);
        b->qx_off[i] = total;
        b->qx_len[i] = m + 2 * w + kBandChunk + 8;
        total += b->qx_len[i];
        total = (total + 63) & ~int64_t(63);
        src += m;
    }
    std::vector<float> arena(static_cast<size_t>(total), kInf);
    src = 0;
    for (int32_t i = 0; i < nq; ++i) {
        const int32_t m = b->m[i], w = b->w[i];
        float* q = arena.data() + b->off[i];
        std::memcpy(q, h_zqueries + src, sizeof(float) * m);
        host_envelope(q, w, q + m, q + 2 * m, m);
        float* qx = arena.data() + b->qx_off[i];
        for (int32_t j = 0; j < m; ++j) qx[j + w] = q[j];
        src += m;
    }
    int rc = MDTW_OK;
    if (hipMalloc(&b->d_arena, sizeof(float) * arena.size()) != hipSuccess) {
        (void)hipGetLastError();
        rc = MDTW_ERR_ALLOC;
    }
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpy(b->d_arena, arena.data(), sizeof(float) * arena.size(),
                              hipMemcpyHostToDevice));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking));
    b->streams.assign(b->P, nullptr);
    b->join.assign(b->P, nullptr);
    for (int p = 0; p < b->P && rc == MDTW_OK; ++p) {
        rc = hip_ok(hipStreamCreateWithFlags(&b->streams[p], hipStreamNonBlocking));
        if (rc == MDTW_OK) rc = hip_ok(hipEventCreateWithFlags(&b->join[p], hipEventDisableTiming));
    }
    if (rc == MDTW_OK) rc = hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    if (rc == MDTW_OK) rc = hip_ok(hipEventCreate(&b->t0));
    if (rc == MDTW_OK) rc = hip_ok(hipEventCreate(&b->t1));
    if (rc != MDTW_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return MDTW_OK;
}

int mdtw_batch_destroy(mdtw_batch* b) {
    if (!b) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard guard(b->e->device);
    (void)hipDeviceSynchronize();
    delete b;
    return MDTW_OK;
}

int mdtw_batch_shape(const mdtw_batch* b, int32_t i, int32_t* h_shape) {
    if (!b || !h_shape || i < 0 || i >= b->nq) return MDTW_ERR_INVALID_ARGUMENT;
    const LaunchShape& s = b->shape_of(i);
    h_shape[0] = s.dp_blocks_per_cu;
    h_shape[1] = s.wide_waves_per_cu;
    h_shape[2] = s.cascade_blocks_per_cu;
    h_shape[3] = s.dp_form;
    return MDTW_OK;
}

int mdtw_batch_set_shape(mdtw_batch* b, int32_t i, const int32_t* h_shape) {
    if (!b || !h_shape || i < 0 || i >= b->nq) return MDTW_ERR_INVALID_ARGUMENT;
    if (h_shape[0] < 1 || h_shape[0] > 64 || h_shape[1] < 1 || h_shape[1] > 256 ||
        h_shape[2] < 1 || h_shape[2] > 256 || h_shape[3] < kDpFormAuto || h_shape[3] > kDpFormWide)
        return MDTW_ERR_INVALID_ARGUMENT;
    LaunchShape s;
    s.dp_blocks_per_cu = h_shape[0];
    s.wide_waves_per_cu = h_shape[1];
    s.cascade_blocks_per_cu = h_shape[2];
    s.dp_form = h_shape[3];
    b->shapes[shape_bucket(b->m[i], b->w[i])] = s;
    b->shapes_version++;
    return MDTW_OK;
}

int mdtw_batch_tune(mdtw_batch* b, const float* d_series, int64_t n, int32_t metric, int32_t k,
                    int64_t cell_budget, int32_t* h_buckets_tuned) {
    int rc = check_common(b, d_series, n, metric, k);
    if (rc != MDTW_OK) return rc;
    DeviceGuard guard(b->e->device);
    if (cell_budget <= 0) cell_budget = int64_t(1) << 39;
    const std::vector<int64_t> ex = excl_of(b, nullptr, n);
    rc = ensure_ws(b, slot_bytes(b, n, k, ex));
    if (rc != MDTW_OK) return rc;
    // Queries per bucket; the median-cost one represents the bucket.
    std::map<int, std::vector<int32_t>> buckets;
    for (int32_t i = 0; i < b->nq; ++i) buckets[shape_bucket(b->m[i], b->w[i])].push_back(i);
    int32_t tuned = 0;
    for (auto& kv : buckets) {
        std::vector<int32_t>& qs = kv.second;
        std::sort(qs.begin(), qs.end(), [&](int32_t a, int32_t c) {
            return band_cells(b->m[a], b->w[a]) < band_cells(b->m[c], b->w[c]);
        });
        const int32_t i = qs[qs.size() / 2];
        const bool wide = b->w[i] > kBandMaxW;
        // Stream prefix holding about cell_budget band cells.
        int64_t ns = cell_budget / std::max<int64_t>(1, band_cells(b->m[i], b->w[i]));
        ns = std::min(n, std::max<int64_t>(ns + b->m[i] - 1, int64_t(1) << 16));
        ns = std::max<int64_t>(ns, b->m[i]);
        std::vector<LaunchShape> cands;
        for (int32_t dp : {4, 8, 16})
            for (int32_t cas : {16, 64}) {
                LaunchShape s;
                s.dp_blocks_per_cu = dp;
                s.cascade_blocks_per_cu = cas;
                if (wide) s.wide_waves_per_cu = dp * 4;
                cands.push_back(s);
            }
        float best_ms = INFINITY;
        LaunchShape best;
        auto trial = [&](const LaunchShape& s) {
            float ms = 0.0f;
            int r = run_one(b, i, d_series, ns, metric, k, ex[i], s, nullptr, nullptr, nullptr, &ms);
            if (r == MDTW_OK)
                r = run_one(b, i, d_series, ns, metric, k, ex[i], s, nullptr, nullptr, nullptr, &ms);
            if (r == MDTW_OK && ms < best_ms) {
                best_ms = ms;
                best = s;
            }
            return r;
        };
        for (const LaunchShape& s : cands) {
            rc = trial(s);
            if (rc != MDTW_OK) return rc;
        }
        // Then the DP's formulation, on the grids just chosen: with the
        // cumulative LB_Keogh bound (one-lane register bands), and for the
        // LDS-tail / split bands the anti-diagonal kernel.
        const LaunchShape grid_best = best;
        if (b->w[i] < kBandPairMinW) {
            LaunchShape s = grid_best;
            s.dp_form = kDpFormLb;
            rc = trial(s);
            if (rc != MDTW_OK) return rc;
        }
        if (wide && b->w[i] <= kBandWideMaxW) {
            for (int32_t ww : {16, 32, 64}) {
                LaunchShape s = grid_best;
                s.dp_form = kDpFormWide;
                s.wide_waves_per_cu = ww;
                rc = trial(s);
                if (rc != MDTW_OK) return rc;
            }
        }
        b->shapes[kv.first] = best;
        ++tuned;
    }
    b->shapes_version++;
    if (h_buckets_tuned) *h_buckets_tuned = tuned;
    return MDTW_OK;
}

int mdtw_batch_search(mdtw_batch* b, const float* d_series, int64_t n, int32_t metric,
                      int32_t k, const int64_t* h_excl, int64_t* h_pos, float* h_dist,
                      mdtw_search_info* h_info, float* h_elapsed_ms) {
    int rc = check_common(b, d_series, n, metric, k);
    if (rc != MDTW_OK) return rc;
    if (!h_pos || !h_dist) return MDTW_ERR_INVALID_ARGUMENT;
    mdtw_engine* e = b->e;
    DeviceGuard guard(e->device);
    const std::vector<int64_t> ex = excl_of(b, h_excl, n);
    rc = ensure_ws(b, slot_bytes(b, n, k, ex));
    if (rc == MDTW_OK) rc = ensure_outputs(b, k);
    if (rc != MDTW_OK) return rc;
    if (!b->have_graph || b->g_s != d_series || b->g_n != n || b->g_k != k ||
        b->g_metric != metric || b->g_excl != ex || b->g_shapes != b->shapes_version) {
        rc = capture(b, d_series, n, metric, k, ex);
        if (rc != MDTW_OK) return rc;
    }
    rc = hip_ok(hipEventRecord(b->t0, e->stream));
    if (rc == MDTW_OK) rc = hip_ok(hipGraphLaunch(b->exec, e->stream));
    if (rc == MDTW_OK) rc = hip_ok(hipEventRecord(b->t1, e->stream));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(e->stream));
    if (rc != MDTW_OK) return rc;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, b->t0, b->t1);
    if (h_elapsed_ms) *h_elapsed_ms = ms;
    const size_t nk = static_cast<size_t>(b->nq) * k;
    std::memcpy(h_pos, b->h_pos, sizeof(int64_t) * nk);
    std::memcpy(h_dist, b->h_dist, sizeof(float) * nk);
    for (int32_t i = 0; i < b->nq; ++i) {
        Ctrl c = b->h_ctrls[i];
        if (b->h_ovf[i]) {
            rc = run_one(b, i, d_series, n, metric, k, ex[i], b->shape_of(i),
                         h_pos + static_cast<size_t>(i) * k, h_dist + static_cast<size_t>(i) * k,
                         &c, nullptr);
            if (rc != MDTW_OK) return rc;
        }
        if (h_info) {
            mdtw_search_info& inf = h_info[i];
            inf.windows = n - b->m[i] + 1;
            inf.seeded = search_seed_count(n - b->m[i] + 1, k, ex[i]);
            inf.lb_kim_pruned = static_cast<int64_t>(c.kim);
            inf.lb_keogh_pruned = static_cast<int64_t>(c.keogh);
            inf.dp_evaluated = static_cast<int64_t>(c.dp);
            inf.dp_abandoned = static_cast<int64_t>(c.abandoned);
            int32_t found = 0;
            for (int32_t j = 0; j < k; ++j) found += h_pos[static_cast<size_t>(i) * k + j] >= 0;
            inf.found = found;
            if (k > 1) {
                inf.threshold = c.thr;
            } else {
                const uint32_t bits = static_cast<uint32_t>(c.best >> 32);
                float d;
                std::memcpy(&d, &bits, sizeof(d));
                inf.threshold = d;
            }
            inf.elapsed_ms = ms;
        }
    }
    return MDTW_OK;
}

int mdtw_batch_backward(mdtw_batch* b, const float* d_series, int64_t n, int32_t metric,
                        int32_t k, const int64_t* h_pos, const float* h_grad,
                        int32_t sqrt_dist, const float* h_query_raw, float* h_grad_out,
                        float* h_dist) {
    int rc = check_common(b, d_series, n, metric, k);
    if (rc != MDTW_OK) return rc;
    if (!h_pos || !h_grad || !h_grad_out) return MDTW_ERR_INVALID_ARGUMENT;
    mdtw_engine* e = b->e;
    DeviceGuard guard(e->device);
    const int32_t nq = b->nq;
    // Pair lists, query by query (entries with position < 0 are skipped).
    std::vector<int64_t> po(nq + 1, 0), qo(nq + 1, 0);
    std::vector<int64_t> pos64;
    std::vector<int32_t> pos32, plen;
    std::vector<float> pgrad;
    for (int32_t i = 0; i < nq; ++i) {
        for (int32_t j = 0; j < k; ++j) {
            const int64_t p = h_pos[static_cast<size_t>(i) * k + j];
            if (p < 0) continue;
            if (p > n - b->m[i]) return MDTW_ERR_INVALID_ARGUMENT;
            pos64.push_back(p);
            pos32.push_back(static_cast<int32_t>(p));
            plen.push_back(b->m[i]);
            pgrad.push_back(h_grad[static_cast<size_t>(i) * k + j]);
        }
        po[i + 1] = static_cast<int64_t>(pos64.size());
        qo[i + 1] = qo[i] + b->m[i];
    }
    const int64_t np = po[nq], nqf = qo[nq];
    // Queries whose direction words fit in LDS run in the batch kernels: one
    // launch per class (slots per lane) over all their windows, longest
    // queries first; the others run query by query on the pipelines.
    std::vector<BwdQuery> tab(static_cast<size_t>(nq));
    std::vector<int32_t> pair_query(static_cast<size_t>(np));
    int64_t rows = 0;
    int32_t max_m = 0;
    size_t wsb = 0;
    bool any_single = false;
    for (int32_t i = 0; i < nq; ++i) {
        const int32_t m = b->m[i], w = b->w[i];
        const int64_t cnt = po[i + 1] - po[i];
        const bool bt = bwd_batchable(m, w);
        tab[i] = BwdQuery{po[i], bt ? rows : 0, b->off[i], qo[i], m, w,
                          static_cast<int32_t>(cnt), bt ? 1 : 0};
        for (int64_t p = po[i]; p < po[i + 1]; ++p) pair_query[p] = i;
        if (bt) {
            rows += cnt * m;
            max_m = std::max(max_m, m);
        } else {
            any_single = true;
            wsb = std::max(wsb, bwd_ws_bytes(m, w, cnt));
        }
    }
    struct Cls {
        int C;
        int64_t begin, count;
        int32_t max_m;
    };
    std::vector<int32_t> list;
    list.reserve(static_cast<size_t>(np));
    std::vector<Cls> cls;
    for (int C = 1; C <= 16; C *= 2) {
        std::vector<int32_t> qs;
        for (int32_t i = 0; i < nq; ++i)
            if (tab[i].batched && bwd_class(b->w[i]) == C) qs.push_back(i);
        std::stable_sort(qs.begin(), qs.end(),
                         [&](int32_t x, int32_t y) { return b->m[x] > b->m[y]; });
        const int64_t begin = static_cast<int64_t>(list.size());
        int32_t cm = 0;
        for (int32_t i : qs) {
            for (int64_t p = po[i]; p < po[i + 1]; ++p) list.push_back(static_cast<int32_t>(p));
            if (po[i + 1] > po[i]) cm = std::max(cm, b->m[i]);
        }
        if (static_cast<int64_t>(list.size()) > begin)
            cls.push_back(Cls{C, begin, static_cast<int64_t>(list.size()) - begin, cm});
    }
    const size_t bytes = carve_size({sizeof(int64_t) * np, sizeof(int32_t) * np,
                                     sizeof(int32_t) * np, sizeof(float) * np,
                                     sizeof(float) * np, sizeof(float) * np, sizeof(float) * np,
                                     sizeof(float) * nqf, sizeof(float) * nqf,
                                     sizeof(BwdQuery) * nq, sizeof(int32_t) * np,
                                     sizeof(int32_t) * np, sizeof(float) * rows,
                                     sizeof(double) * nqf});
    if (bytes > b->bwd_cap) {
        if (b->bwd_buf) (void)hipFree(b->bwd_buf);
        b->bwd_buf = nullptr;
        b->bwd_cap = 0;
        if (hipMalloc(&b->bwd_buf, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return MDTW_ERR_ALLOC;
        }
        b->bwd_cap = bytes;
    }
    if (any_single && (wsb > b->bwd_ws_cap || static_cast<int>(b->bwd_ws.size()) != b->P)) {
        for (char* p : b->bwd_ws)
            if (p) (void)hipFree(p);
        b->bwd_ws.clear();
        b->bwd_ws_cap = 0;
        for (int p = 0; p < b->P; ++p) {
            char* ptr = nullptr;
            if (hipMalloc(&ptr, wsb) != hipSuccess) {
                (void)hipGetLastError();
                return MDTW_ERR_ALLOC;
            }
            b->bwd_ws.push_back(ptr);
        }
        b->bwd_ws_cap = wsb;
    }
    Carve cv{b->bwd_buf};
    int64_t* d_pos64 = cv.take<int64_t>(np);
    int32_t* d_pos32 = cv.take<int32_t>(np);
    int32_t* d_len = cv.take<int32_t>(np);
    float* d_grad = cv.take<float>(np);
    float* d_mu = cv.take<float>(np);
    float* d_sd = cv.take<float>(np);
    float* d_dist = cv.take<float>(np);
    float* d_raw = cv.take<float>(nqf);
    float* d_out = cv.take<float>(nqf);
    BwdQuery* d_tab = cv.take<BwdQuery>(nq);
    int32_t* d_pq = cv.take<int32_t>(np);
    int32_t* d_list = cv.take<int32_t>(np);
    float* d_rows = cv.take<float>(rows);
    double* d_gsum = cv.take<double>(nqf);

    hipStream_t st = b->origin;
    rc = hip_ok(hipMemcpyAsync(d_pos64, pos64.data(), sizeof(int64_t) * np,
                               hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(d_pos32, pos32.data(), sizeof(int32_t) * np,
                                   hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(d_len, plen.data(), sizeof(int32_t) * np,
                                   hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(d_grad, pgrad.data(), sizeof(float) * np,
                                   hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK && h_query_raw)
        rc = hip_ok(hipMemcpyAsync(d_raw, h_query_raw, sizeof(float) * nqf,
                                   hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(d_tab, tab.data(), sizeof(BwdQuery) * nq,
                                   hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK && np)
        rc = hip_ok(hipMemcpyAsync(d_pq, pair_query.data(), sizeof(int32_t) * np,
                                   hipMemcpyHostToDevice, st));
    if (rc == MDTW_OK && !list.empty())
        rc = hip_ok(hipMemcpyAsync(d_list, list.data(), sizeof(int32_t) * list.size(),
                                   hipMemcpyHostToDevice, st));
    // Window statistics of the matched windows, as the forward computed them.
    if (rc == MDTW_OK) rc = launch_stats_at(st, d_series, n, d_len, d_pos64, np, d_mu, d_sd);
    if (any_single) {
        if (rc == MDTW_OK) rc = hip_ok(hipEventRecord(b->fork, st));
        for (int p = 0; p < b->P && rc == MDTW_OK; ++p)
            rc = hip_ok(hipStreamWaitEvent(b->streams[p], b->fork, 0));
        int slot = 0;
        for (int32_t i = 0; i < nq && rc == MDTW_OK; ++i) {
            if (tab[i].batched) continue;
            const int p = slot++ % b->P;
            rc = enqueue_backward(b->streams[p], d_series, d_mu + po[i], d_sd + po[i], true,
                                  d_pos32 + po[i], d_grad + po[i], po[i + 1] - po[i],
                                  b->d_arena + b->off[i], b->m[i], b->w[i], metric, sqrt_dist,
                                  h_query_raw ? d_raw + qo[i] : nullptr, d_dist + po[i],
                                  d_out + qo[i], b->bwd_ws[p]);
        }
    }
    for (const Cls& c : cls) {
        if (rc != MDTW_OK) break;
        rc = launch_bwd_path_batch(st, c.C, metric == MDTW_METRIC_L2, c.max_m, d_series, d_mu,
                                   d_sd, d_pos32, d_pq, d_list + c.begin, c.count, d_tab,
                                   b->d_arena, d_grad, d_rows, d_dist, sqrt_dist);
    }
    if (rc == MDTW_OK && max_m > 0)
        rc = launch_bwd_reduce_batch(st, d_rows, d_tab, nq, max_m, h_query_raw ? d_raw : nullptr,
                                     d_gsum, d_out);
    if (any_single) {
        for (int p = 0; p < b->P && rc == MDTW_OK; ++p) {
            rc = hip_ok(hipEventRecord(b->join[p], b->streams[p]));
            if (rc == MDTW_OK) rc = hip_ok(hipStreamWaitEvent(st, b->join[p], 0));
        }
    }
    std::vector<float> dist(static_cast<size_t>(np));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(h_grad_out, d_out, sizeof(float) * nqf, hipMemcpyDeviceToHost,
                                   st));
    if (rc == MDTW_OK && h_dist && np)
        rc = hip_ok(hipMemcpyAsync(dist.data(), d_dist, sizeof(float) * np,
                                   hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(st));
    if (rc != MDTW_OK) return rc;
    if (h_dist) {
        for (int32_t i = 0; i < nq; ++i) {
            int64_t q = po[i];
            for (int32_t j = 0; j < k; ++j) {
                float& o = h_dist[static_cast<size_t>(i) * k + j];
                o = h_pos[static_cast<size_t>(i) * k + j] >= 0 ? dist[q++] : INFINITY;
            }
        }
    }
    return MDTW_OK;
}

}  // extern "C"
