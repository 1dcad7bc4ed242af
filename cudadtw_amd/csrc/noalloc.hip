// noalloc.hip -- mdtw_search_noalloc: mdtw_search for callers that cannot
// allocate (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  No reference counterpart: the reference allocates and
// frees its scratch on every call (cdtw.cuh:510-592).
//
// Everything a search needs beyond its inputs and outputs -- window
// statistics, candidate lists, kept keys, bucket minima, seeds, the padded
// query and the selected matches -- is carved from one fixed device array,
// `scratch`, of kScratchBytes.  Every GPU the library runs on holds its own
// copy of the array until the process ends.  Host code never uses the
// array's name as an address (on the host it is only a stand-in): its device
// address is looked up once per device, with that device current, when the
// first engine on the device is created, and kept per device.
//
// All calls on one GPU share that GPU's array, and it keeps the last call's
// contents: every piece that must start at zero is cleared by the search's
// own reset kernel, the rest is written before it is read.  Two calls that
// overlap on one GPU (engines on different host threads) race on the array.
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>

#include "cdtw_band.hpp"

namespace mdtw {
namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchSlot {
    bool looked_up = false;
    int status = MDTW_OK;
    char* ptr = nullptr;
};
std::mutex g_scratch_mu;
std::array<ScratchSlot, kMaxDevices> g_scratch;

// The register-band kernel's padded query: qx[i] = q[i - w] for w <= i < m + w,
// +inf elsewhere (ensure_qx builds the same array on the host).
__global__ void pad_query_kernel(const float* __restrict__ q, int32_t m, int32_t w,
                                 float* __restrict__ qx, int32_t len) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x)
        qx[i] = (i >= w && i < w + m) ? q[i - w] : kInf;
}

int scratch_of(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) return MDTW_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    const ScratchSlot& s = g_scratch[device];
    if (!s.looked_up) return MDTW_ERR_UNSUPPORTED;
    if (s.status != MDTW_OK) return s.status;
    *out = s.ptr;
    return MDTW_OK;
}

}  // namespace

int scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return MDTW_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ScratchSlot& s = g_scratch[device];
    if (s.looked_up) return s.status;
    void* p = nullptr;
    const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    s.looked_up = true;
    if (err != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        s.status = hip_status(err != hipSuccess ? err : hipErrorInvalidSymbol);
        return s.status;
    }
    s.ptr = static_cast<char*>(p);
    s.status = MDTW_OK;
    return MDTW_OK;
}

}  // namespace mdtw

using namespace mdtw;

extern "C" int mdtw_search_noalloc(mdtw_engine* e, const float* d_series, int64_t n,
                                   const float* d_avg, const float* d_std, int32_t w,
                                   int32_t algo, int32_t metric, int32_t stats_mode, int32_t k,
                                   int64_t excl, int64_t* h_pos, float* h_dist,
                                   mdtw_search_info* info) {
    int32_t wc = 0;
    int rc = check_search_args(e, d_series, n, d_avg, d_std, w, algo, metric, stats_mode, k, h_pos,
                               h_dist, &excl, &wc);
    if (rc != MDTW_OK) return rc;
    // The FFT form and the REFCOMPAT statistics size their own workspaces.
    if (algo == MDTW_ALGO_ED_FFT) return MDTW_ERR_UNSUPPORTED;
    if (d_avg == nullptr && stats_mode == MDTW_STATS_REFCOMPAT) return MDTW_ERR_UNSUPPORTED;
    for (int32_t i = 0; i < k; ++i) {
        h_pos[i] = -1;
        h_dist[i] = INFINITY;
    }
    const int32_t m = e->m;
    const int64_t L = n - m + 1;
    DeviceGuard guard(e->device);
    char* base = nullptr;
    rc = scratch_of(e->device, &base);
    if (rc != MDTW_OK) return rc;

    // Layout: search workspace, padded query (cDTW), k positions, k distances,
    // the selection's overflow flag.
    const bool cdtw = algo == MDTW_ALGO_CDTW;
    const size_t qx_len = cdtw ? static_cast<size_t>(m) + 2 * static_cast<size_t>(wc) +
                                     kBandChunk + 8
                               : 0;
    const size_t ws_bytes = search_ws_bytes(L, k, excl, d_avg == nullptr);
    const size_t need = carve_size({ws_bytes, sizeof(float) * qx_len, sizeof(int64_t) * k,
                                    sizeof(float) * k, sizeof(int32_t)});
    if (need > kScratchBytes) return MDTW_ERR_WORKSPACE;
    Carve cv{base};
    char* ws = cv.take<char>(ws_bytes);
    float* qx = cv.take<float>(qx_len);
    int64_t* d_pos = cv.take<int64_t>(k);
    float* d_dist = cv.take<float>(k);
    int32_t* d_ovf = cv.take<int32_t>(1);

    // Query-side state: the envelope for this band (host work into the
    // engine's own buffers, then a copy), the padded query on the device.
    hipStream_t st = e->stream;
    if (algo != MDTW_ALGO_ED) {
        rc = ensure_envelope(e, wc);
        if (rc != MDTW_OK) return rc;
    }
    if (cdtw) {
        const int32_t len = static_cast<int32_t>(qx_len);
        hipLaunchKernelGGL(pad_query_kernel, dim3((len + 255) / 256), dim3(256), 0, st, e->d_q, m,
                           wc, qx, len);
        rc = hip_status(hipGetLastError());
        if (rc != MDTW_OK) return rc;
    }

    SearchPlan p;
    p.st = st;
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
    p.qx = cdtw ? qx : nullptr;
    p.qx_len = static_cast<int32_t>(qx_len);
    if (algo == MDTW_ALGO_ED) p.ed = ed_query_sums(e->h_q.data(), m);
    p.m = m;
    p.w = wc;
    p.algo = algo;
    p.metric = metric;
    p.k = k;
    p.excl = excl;
    p.ctrl = e->d_ctrl;
    p.ws = ws;
    SearchBufs bufs;
    int32_t overflow = 0;
    rc = hip_status(hipEventRecord(e->ev0, st));
    if (rc == MDTW_OK) rc = enqueue_search(p, &bufs);
    if (rc == MDTW_OK)
        rc = launch_topk_select(st, e->d_ctrl, bufs.res, k, excl, metric, d_pos, d_dist, d_ovf);
    if (rc == MDTW_OK) rc = hip_status(hipEventRecord(e->ev1, st));
    if (rc == MDTW_OK)
        rc = hip_status(
            hipMemcpyAsync(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK)
        rc = hip_status(
            hipMemcpyAsync(h_pos, d_pos, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK)
        rc = hip_status(
            hipMemcpyAsync(h_dist, d_dist, sizeof(float) * k, hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK)
        rc = hip_status(
            hipMemcpyAsync(&overflow, d_ovf, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK) rc = hip_status(hipStreamSynchronize(st));
    if (rc != MDTW_OK) return rc;
    // More kept keys than one workgroup selects from: mdtw_search would sort
    // them on the host, which allocates.
    if (overflow) {
        for (int32_t i = 0; i < k; ++i) {
            h_pos[i] = -1;
            h_dist[i] = INFINITY;
        }
        return MDTW_ERR_WORKSPACE;
    }
    const Ctrl c = *e->h_ctrl;
    int32_t found = 0;
    while (found < k && h_pos[found] >= 0) ++found;
    if (info) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
        float thr = INFINITY;
        if (k > 1) {
            thr = c.thr;
        } else if ((c.best >> 32) < 0x7f800000ull) {
            const uint32_t bits = static_cast<uint32_t>(c.best >> 32);
            std::memcpy(&thr, &bits, sizeof(thr));
        }
        info->windows = L;
        info->seeded = bufs.seeded;
        info->lb_kim_pruned = static_cast<int64_t>(c.kim);
        info->lb_keogh_pruned = static_cast<int64_t>(c.keogh);
        info->dp_evaluated = static_cast<int64_t>(c.dp);
        info->dp_abandoned = static_cast<int64_t>(c.abandoned);
        info->found = found;
        info->threshold = thr;
        info->elapsed_ms = ms;
    }
    return MDTW_OK;
}
