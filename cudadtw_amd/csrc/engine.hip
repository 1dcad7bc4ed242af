// engine.hip -- host side of the C ABI: engine lifecycle, query state,
// bit-exact host helpers and the profile entry points.  No kernels here.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "cdtw_band.hpp"

namespace mdtw {

static thread_local hipError_t g_last_hip = hipSuccess;
void set_last_hip_error(hipError_t e) { g_last_hip = e; }

int ws_reserve(mdtw_engine* e, size_t bytes, void** out) {
    if (e->ws_limit && bytes > e->ws_limit) return MDTW_ERR_WORKSPACE;
    if (bytes <= e->ws_cap && e->d_ws) {
        *out = e->d_ws;
        return MDTW_OK;
    }
    hipError_t err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    size_t cap = std::max(bytes, e->ws_cap + e->ws_cap / 2);
    if (e->ws_limit) cap = std::min(cap, e->ws_limit);
    if (e->d_ws) (void)hipFree(e->d_ws);
    e->d_ws = nullptr;
    e->ws_cap = 0;
    cap = (cap + (1u << 20) - 1) & ~static_cast<size_t>((1u << 20) - 1);
    err = hipMalloc(&e->d_ws, cap);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        e->d_ws = nullptr;
        return MDTW_ERR_ALLOC;
    }
    e->ws_cap = cap;
    *out = e->d_ws;
    return MDTW_OK;
}

int upload(mdtw_engine* e, void* dst, const void* src, size_t bytes) {
    hipError_t err = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

// bounds.cuh:13-64 envelope: up[i] / lo[i] = max / min of q over
// [i - w, i + w] clipped to [0, m), by a monotone-deque sliding window (the
// published O(m) method the reference also uses): ub holds indices of
// decreasing values, lb of increasing ones, each at most m entries (the
// caller's arrays hold m + 1).  Exact maxima and minima, so every correct
// method gives these bits.  w >= m is clamped to m - 1 (global max/min).
void host_envelope(const float* q, int32_t w, float* lo, float* up, int32_t m, int* ub,
                   int* lb) {
    if (m <= 0) return;
    if (w > m - 1) w = m - 1;
    if (w < 0) w = 0;
    int uh = 0, ut = 0, lh = 0, lt = 0;
    for (int j = 0; j < m + w; ++j) {
        if (j < m) {
            while (ut > uh && q[ub[ut - 1]] <= q[j]) --ut;
            ub[ut++] = j;
            while (lt > lh && q[lb[lt - 1]] >= q[j]) --lt;
            lb[lt++] = j;
        }
        const int c = j - w;  // the window centred at c is complete once j = c + w
        if (c < 0) continue;
        while (ub[uh] < c - w) ++uh;
        while (lb[lh] < c - w) ++lh;
        up[c] = q[ub[uh]];
        lo[c] = q[lb[lh]];
    }
}

void host_envelope(const float* q, int32_t w, float* lo, float* up, int32_t m) {
    if (m <= 0) return;
    std::vector<int> ub(static_cast<size_t>(m) + 1), lb(static_cast<size_t>(m) + 1);
    host_envelope(q, w, lo, up, m, ub.data(), lb.data());
}

int ensure_envelope(mdtw_engine* e, int32_t w) {
    if (e->w_env == w) return MDTW_OK;
    int* dq = e->env_deques.data();
    host_envelope(e->h_q.data(), w, e->h_lo.data(), e->h_up.data(), e->m, dq,
                  dq + MDTW_MAX_QUERY + 1);
    int rc = upload(e, e->d_lo, e->h_lo.data(), sizeof(float) * e->m);
    if (rc == MDTW_OK) rc = upload(e, e->d_up, e->h_up.data(), sizeof(float) * e->m);
    if (rc == MDTW_OK) e->w_env = w;
    return rc;
}

// Padded query for the register-band kernel: qx[k] = q[k - w] for
// w <= k < m + w, +inf elsewhere, with room for the last load chunk.
int ensure_qx(mdtw_engine* e, int32_t w) {
    if (e->w_qx == w) return MDTW_OK;
    const size_t len = static_cast<size_t>(e->m) + 2 * static_cast<size_t>(w) + kBandChunk + 8;
    if (len > e->qx_cap) {
        if (e->d_qx) (void)hipFree(e->d_qx);
        e->d_qx = nullptr;
        e->qx_cap = 0;
        if (hipMalloc(&e->d_qx, sizeof(float) * len) != hipSuccess) {
            (void)hipGetLastError();
            return MDTW_ERR_ALLOC;
        }
        e->qx_cap = len;
    }
    std::vector<float> qx(len, kInf);
    for (int32_t i = 0; i < e->m; ++i) qx[static_cast<size_t>(i) + w] = e->h_q[i];
    int rc = upload(e, e->d_qx, qx.data(), sizeof(float) * len);
    if (rc == MDTW_OK) {
        e->w_qx = w;
        e->qx_len = static_cast<int32_t>(len);
    }
    return rc;
}

}  // namespace mdtw

using namespace mdtw;

extern "C" {

const char* mdtw_status_string(int status) {
    switch (status) {
        case MDTW_OK: return "ok";
        case MDTW_ERR_INVALID_ARGUMENT: return "invalid argument";
        case MDTW_ERR_QUERY_TOO_SHORT: return "query too short (LB_Kim needs M >= 11)";
        case MDTW_ERR_QUERY_TOO_LONG: return "query longer than MDTW_MAX_QUERY";
        case MDTW_ERR_BAND: return "band W out of range";
        case MDTW_ERR_STREAM_TOO_SHORT: return "stream shorter than the query";
        case MDTW_ERR_NO_QUERY: return "no query set (mdtw_set_query)";
        case MDTW_ERR_ALLOC: return "allocation failed";
        case MDTW_ERR_HIP: return hipGetErrorString(g_last_hip);
        case MDTW_ERR_NOT_FINITE: return "query contains NaN or Inf";
        case MDTW_ERR_K: return "k out of range";
        case MDTW_ERR_UNSUPPORTED: return "unsupported option";
        case MDTW_ERR_WORKSPACE: return "workspace request too large";
        default: return "unknown status";
    }
}

int mdtw_create(int device, mdtw_engine** out) {
    if (!out) return MDTW_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    hipError_t err = hipGetDeviceCount(&ndev);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    if (device < 0 || device >= ndev) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard g(device);
    err = hipSetDevice(device);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    auto* e = new (std::nothrow) mdtw_engine();
    if (!e) return MDTW_ERR_ALLOC;
    e->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        e->cus = prop.multiProcessorCount;
        std::snprintf(e->name, sizeof(e->name), "%s", prop.gcnArchName);
    }
    if (e->cus <= 0) e->cus = 256;
    const size_t qbytes = sizeof(float) * MDTW_MAX_QUERY;
    if (hipMalloc(&e->d_q, qbytes) != hipSuccess || hipMalloc(&e->d_lo, qbytes) != hipSuccess ||
        hipMalloc(&e->d_up, qbytes) != hipSuccess ||
        hipMalloc(&e->d_ctrl, sizeof(Ctrl)) != hipSuccess ||
        hipHostMalloc(&e->h_ctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
        (void)hipGetLastError();
        mdtw_destroy(e);
        return MDTW_ERR_ALLOC;
    }
    e->h_q.resize(MDTW_MAX_QUERY);
    e->h_lo.resize(MDTW_MAX_QUERY);
    e->h_up.resize(MDTW_MAX_QUERY);
    e->env_deques.resize(2 * (MDTW_MAX_QUERY + 1));
    // The fixed scratch array of mdtw_search_noalloc, looked up on this device
    // (current here) once; a failure is kept and reported by that entry point.
    (void)scratch_setup(device);
    *out = e;
    return MDTW_OK;
}

int mdtw_destroy(mdtw_engine* e) {
    if (!e) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard g(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    else (void)hipDeviceSynchronize();
    if (e->d_q) (void)hipFree(e->d_q);
    if (e->d_lo) (void)hipFree(e->d_lo);
    if (e->d_up) (void)hipFree(e->d_up);
    if (e->d_qx) (void)hipFree(e->d_qx);
    if (e->d_ws) (void)hipFree(e->d_ws);
    if (e->d_ctrl) (void)hipFree(e->d_ctrl);
    if (e->h_ctrl) (void)hipHostFree(e->h_ctrl);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    delete e;
    return MDTW_OK;
}

int mdtw_set_workspace_limit(mdtw_engine* e, size_t bytes) {
    if (!e) return MDTW_ERR_INVALID_ARGUMENT;
    e->ws_limit = bytes;
    return MDTW_OK;
}

int mdtw_set_stream(mdtw_engine* e, void* hip_stream) {
    if (!e) return MDTW_ERR_INVALID_ARGUMENT;
    e->stream = static_cast<hipStream_t>(hip_stream);
    return MDTW_OK;
}

int mdtw_synchronize(mdtw_engine* e) {
    if (!e) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard g(e->device);
    hipError_t err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int mdtw_device_info(mdtw_engine* e, char* name, int32_t name_len, int32_t* cus) {
    if (!e) return MDTW_ERR_INVALID_ARGUMENT;
    if (name && name_len > 0) std::snprintf(name, static_cast<size_t>(name_len), "%s", e->name);
    if (cus) *cus = e->cus;
    return MDTW_OK;
}

int mdtw_set_query(mdtw_engine* e, const float* h_zquery, const float* h_lower,
                   const float* h_upper, int32_t m, int32_t w) {
    if (!e || !h_zquery || m < 1 || w < 0) return m > MDTW_MAX_QUERY ? MDTW_ERR_QUERY_TOO_LONG
                                                                     : MDTW_ERR_INVALID_ARGUMENT;
    if (m > MDTW_MAX_QUERY) return MDTW_ERR_QUERY_TOO_LONG;
    if ((h_lower == nullptr) != (h_upper == nullptr)) return MDTW_ERR_INVALID_ARGUMENT;
    for (int32_t i = 0; i < m; ++i)
        if (!std::isfinite(h_zquery[i])) return MDTW_ERR_NOT_FINITE;
    DeviceGuard g(e->device);
    std::memcpy(e->h_q.data(), h_zquery, sizeof(float) * m);
    e->m = m;
    e->w_qx = -1;
    const int32_t wc = std::min(w, m - 1);
    if (h_lower) {
        std::memcpy(e->h_lo.data(), h_lower, sizeof(float) * m);
        std::memcpy(e->h_up.data(), h_upper, sizeof(float) * m);
    } else {
        host_envelope(e->h_q.data(), wc, e->h_lo.data(), e->h_up.data(), m);
    }
    int rc = upload(e, e->d_q, e->h_q.data(), sizeof(float) * m);
    if (rc == MDTW_OK) rc = upload(e, e->d_lo, e->h_lo.data(), sizeof(float) * m);
    if (rc == MDTW_OK) rc = upload(e, e->d_up, e->h_up.data(), sizeof(float) * m);
    if (rc != MDTW_OK) {
        e->has_query = false;
        return rc;
    }
    e->w_env = wc;
    e->has_query = true;
    return MDTW_OK;
}

// stats.cuh:8-28 with value_t = float: sums in index order in FP32.
int mdtw_znormalize(float* h, int32_t len) {
    if (!h || len < 1) return MDTW_ERR_INVALID_ARGUMENT;
    float avg = 0.0f, sd = 0.0f;
    for (int32_t i = 0; i < len; ++i) avg += h[i];
    avg /= static_cast<float>(len);
    for (int32_t i = 0; i < len; ++i) sd += h[i] * h[i];
    sd = std::sqrt(sd / static_cast<float>(len) - avg * avg);
    for (int32_t i = 0; i < len; ++i) {
        h[i] -= avg;
        h[i] /= sd;
    }
    return MDTW_OK;
}

int mdtw_znormalize_f64(double* h, int32_t len) {
    if (!h || len < 1) return MDTW_ERR_INVALID_ARGUMENT;
    double avg = 0.0, sd = 0.0;
    for (int32_t i = 0; i < len; ++i) avg += h[i];
    avg /= static_cast<double>(len);
    for (int32_t i = 0; i < len; ++i) sd += h[i] * h[i];
    sd = std::sqrt(sd / static_cast<double>(len) - avg * avg);
    for (int32_t i = 0; i < len; ++i) {
        h[i] -= avg;
        h[i] /= sd;
    }
    return MDTW_OK;
}

int mdtw_envelope(const float* h, int32_t w, float* lo, float* up, int32_t m) {
    if (!h || !lo || !up || m < 1 || w < 0) return MDTW_ERR_INVALID_ARGUMENT;
    host_envelope(h, w, lo, up, m);
    return MDTW_OK;
}

int mdtw_avg_std(mdtw_engine* e, const float* d_series, int64_t n, int32_t m, float* d_avg,
                 float* d_std, int32_t mode) {
    if (!e || !d_series || !d_avg || !d_std || m < 1) return MDTW_ERR_INVALID_ARGUMENT;
    if (m > MDTW_MAX_QUERY) return MDTW_ERR_QUERY_TOO_LONG;
    if (n < m) return MDTW_ERR_STREAM_TOO_SHORT;
    DeviceGuard g(e->device);
    if (mode == MDTW_STATS_FP64) return launch_stats_fp64(e->stream, d_series, n, m, d_avg, d_std);
    if (mode == MDTW_STATS_REFCOMPAT)
        return launch_stats_refcompat(e, d_series, n, m, d_avg, d_std);
    if (mode == MDTW_STATS_NONE)
        return launch_stats_none(e->stream, d_series, n, m, d_avg, d_std);
    return MDTW_ERR_INVALID_ARGUMENT;
}

int mdtw_lb_kim(mdtw_engine* e, const float* d_series, const float* d_avg, const float* d_std,
                float* d_out, const int32_t* d_indices, int64_t count, int32_t metric) {
    if (!e || !d_series || !d_avg || !d_std || !d_out || count < 0) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    if (e->m < 11) return MDTW_ERR_QUERY_TOO_SHORT;
    DeviceGuard g(e->device);
    return launch_lb_kim(e->stream, d_series, d_avg, d_std, e->d_q, e->m, d_indices, count,
                         metric, d_out);
}

int mdtw_lb_keogh(mdtw_engine* e, const float* d_series, const float* d_avg, const float* d_std,
                  float* d_out, const int32_t* d_indices, int64_t count, int32_t metric) {
    if (!e || !d_series || !d_avg || !d_std || !d_out || count < 0) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    DeviceGuard g(e->device);
    return launch_lb_keogh(e->stream, d_series, d_avg, d_std, e->d_lo, e->d_up, e->m, d_indices,
                           count, metric, d_out);
}

int mdtw_cdtw(mdtw_engine* e, const float* d_series, const float* d_avg, const float* d_std,
              float* d_out, const int32_t* d_indices, int64_t count, int32_t w, int32_t metric) {
    if (!e || !d_series || !d_avg || !d_std || !d_out || count < 0 || w < 0)
        return w < 0 ? MDTW_ERR_BAND : MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    DeviceGuard g(e->device);
    const int32_t wc = std::min(w, e->m - 1);
    if (wc > kMaxBandW) return MDTW_ERR_UNSUPPORTED;
    int rc = ensure_qx(e, wc);
    if (rc != MDTW_OK) return rc;
    BandArgs a;
    a.s = d_series;
    a.avg = d_avg;
    a.sd = d_std;
    a.qx = e->d_qx;
    a.qx_len = e->qx_len;
    a.m = e->m;
    a.idx = d_indices;
    a.count = count;
    a.mode = kModeProfile;
    a.ctrl = e->d_ctrl;
    a.out = d_out;
    return launch_dp(e->stream, e->cus, a, e->d_q, wc, metric, false);
}

// Testing/tuning hook: the same profile through the anti-diagonal kernel.
int mdtw_cdtw_wide(mdtw_engine* e, const float* d_series, const float* d_avg, const float* d_std,
                   float* d_out, const int32_t* d_indices, int64_t count, int32_t w,
                   int32_t metric) {
    if (!e || !d_series || !d_avg || !d_std || !d_out || count < 0 || w < 0)
        return w < 0 ? MDTW_ERR_BAND : MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    DeviceGuard g(e->device);
    const int32_t wc = std::min(w, e->m - 1);
    if (wc > kMaxBandW) return MDTW_ERR_UNSUPPORTED;
    BandArgs a;
    a.s = d_series;
    a.avg = d_avg;
    a.sd = d_std;
    a.m = e->m;
    a.idx = d_indices;
    a.count = count;
    a.mode = kModeProfile;
    a.ctrl = e->d_ctrl;
    a.out = d_out;
    return launch_dp(e->stream, e->cus, a, e->d_q, wc, metric, true);
}

int mdtw_ldtw(mdtw_engine* e, const float* d_series, const float* d_avg, const float* d_std,
              float* d_out, int64_t count, int32_t w, int32_t metric) {
    if (!e || !d_series || !d_out || count < 0 || w < 0)
        return w < 0 ? MDTW_ERR_BAND : MDTW_ERR_INVALID_ARGUMENT;
    if ((d_avg == nullptr) != (d_std == nullptr)) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    DeviceGuard g(e->device);
    if (!d_avg && count > 0) {  // raw windows: mean 0, sigma 1 in the workspace
        void* ws = nullptr;
        int rc = ws_reserve(e, carve_size({sizeof(float) * count, sizeof(float) * count}), &ws);
        if (rc != MDTW_OK) return rc;
        Carve cv{static_cast<char*>(ws)};
        float* avg = cv.take<float>(count);
        float* sd = cv.take<float>(count);
        rc = launch_stats_none(e->stream, d_series, count + e->m - 1, e->m, avg, sd);
        if (rc != MDTW_OK) return rc;
        d_avg = avg;
        d_std = sd;
    }
    BandArgs a;
    a.s = d_series;
    a.avg = d_avg;
    a.sd = d_std;
    a.qx = e->d_q;
    a.m = e->m;
    a.count = count;
    a.mode = kModeProfile;
    a.ctrl = e->d_ctrl;
    a.out = d_out;
    return launch_ldtw(e->stream, e->cus, a, std::min(w, e->m - 1), metric);
}

int mdtw_ed(mdtw_engine* e, const float* d_series, const float* d_avg, const float* d_std,
            float* d_out, int64_t count) {
    if (!e || !d_series || !d_avg || !d_std || !d_out || count < 0) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    DeviceGuard g(e->device);
    BandArgs a;
    a.s = d_series;
    a.avg = d_avg;
    a.sd = d_std;
    a.qx = e->d_q;
    a.m = e->m;
    a.count = count;
    a.mode = kModeProfile;
    a.ctrl = e->d_ctrl;
    a.out = d_out;
    a.ed = ed_query_sums(e->h_q.data(), e->m);
    return launch_ed(e->stream, e->cus, a);
}

int mdtw_threshold(mdtw_engine* e, const float* d_values, int32_t* d_indices, int64_t count,
                   float thr, int64_t* h_count) {
    if (!e || !d_values || !d_indices || !h_count || count < 0) return MDTW_ERR_INVALID_ARGUMENT;
    DeviceGuard g(e->device);
    *h_count = 0;
    if (count == 0) return MDTW_OK;
    void* ws = nullptr;
    int rc = ws_reserve(e, sizeof(int32_t) * threshold_ws_ints(count) + 512, &ws);
    if (rc != MDTW_OK) return rc;
    int32_t* total = nullptr;
    rc = launch_threshold(e->stream, d_values, d_indices, count, thr, static_cast<int32_t*>(ws),
                          &total);
    if (rc != MDTW_OK) return rc;
    int32_t hc = 0;
    hipError_t err = hipMemcpyAsync(&hc, total, sizeof(int32_t), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    *h_count = hc;
    return MDTW_OK;
}

int mdtw_argmin(mdtw_engine* e, const float* d_values, int64_t count, int64_t* h_index,
                float* h_value) {
    if (!e || !d_values || !h_index || !h_value || count < 1) return MDTW_ERR_INVALID_ARGUMENT;
    if (count > 0xffffffffll) return MDTW_ERR_UNSUPPORTED;
    DeviceGuard g(e->device);
    int rc = launch_ctrl_reset(e->stream, e->d_ctrl, ~0ull);
    if (rc == MDTW_OK) rc = launch_argmin(e->stream, d_values, count, e->d_ctrl);
    if (rc != MDTW_OK) return rc;
    hipError_t err =
        hipMemcpyAsync(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    const unsigned long long key = e->h_ctrl->best;
    const uint32_t ord = static_cast<uint32_t>(key >> 32);
    const uint32_t bits = (ord & 0x80000000u) ? (ord & 0x7fffffffu) : ~ord;
    float v;
    std::memcpy(&v, &bits, sizeof(v));
    *h_index = static_cast<int64_t>(key & 0xffffffffull);
    *h_value = v;
    return MDTW_OK;
}

}  // extern "C"
