// stats.hip -- sliding-window mean and population sigma of every window.
//
// Two modes (mdtw_avg_std, replacing stats.cuh:115-142 avg_std<double>):
//
// MDTW_STATS_FP64 (default).  One pass, no global prefix.  A workgroup owns T
// consecutive windows [w0, w0+T).  With c = x[w0] as a tile-local shift,
// window r's sums are
//     S(r) = S0 + sum_{k<r} (x[w0+m+k] - x[w0+k])         (and the same for
//                                                         squares of x - c),
// S0 being the first window's sum.  S0 is a workgroup reduction, the
// differences are a workgroup exclusive scan, all in FP64 on values shifted
// by c, so cancellation is bounded by the tile's local range instead of the
// stream's absolute level (the reference's global FP64 prefix loses ~1e-2 of
// sigma on offset data; its FP32 finalize loses more, stats.cuh:40-52).
// HBM traffic: read (T + m)/T floats per window, write 8 bytes.
//
// MDTW_STATS_REFCOMPAT.  Restates the reference arithmetic: per 10^7-sample
// epoch with M-1 overlap (stats.cuh:120-141), a global inclusive FP64 prefix
// with P[0] = 0 (cub_util.cuh:29-54), (P[i+M] - P[i]) / M in FP64 stored as
// FP32 (stats.cuh:30-38), sigma = sqrt(fmaf(-X, X, X2)) or 1 (stats.cuh:40-52,
// nvcc's default contraction of X2 - X*X).
#include <cfloat>

#include "internal.hpp"

namespace mdtw {
namespace {

constexpr int kThreads = 256;

// Inclusive scan of a double across the 256 threads of the workgroup
// (4 waves); returns the inclusive value, *total gets the workgroup total.
__device__ __forceinline__ double block_scan_incl(double v, double* sh, double* total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        double o = __shfl_up(v, off, kWave);
        if (lane >= off) v += o;
    }
    if (lane == kWave - 1) sh[wid] = v;
    __syncthreads();
    double base = 0.0, tot = 0.0;
#pragma unroll
    for (int i = 0; i < kThreads / kWave; ++i) {
        double t = sh[i];
        if (i < wid) base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return v + base;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kThreads / kWave; ++i) t += sh[i];
    __syncthreads();
    return t;
}

template <int R>
__global__ __launch_bounds__(kThreads) void stats_fp64_kernel(
    const float* __restrict__ s, int64_t L, int32_t m, float* __restrict__ avg,
    float* __restrict__ sd) {
    constexpr int T = kThreads * R;
    __shared__ double sh[kThreads / kWave];
    const int64_t w0 = static_cast<int64_t>(blockIdx.x) * T;
    const int tid = threadIdx.x;
    const int64_t nw64 = L - w0;
    const int nw = nw64 < T ? static_cast<int>(nw64) : T;
    const float* x = s + w0;
    const double c = static_cast<double>(x[0]);

    // S0: first window of the tile.
    double a1 = 0.0, a2 = 0.0, b1 = 0.0;
    for (int k = tid; k < m; k += kThreads) {
        double d = static_cast<double>(x[k]) - c;
        a1 += d;
        b1 += fabs(d);
        a2 = fma(d, d, a2);
    }
    const double S1 = block_sum(a1, sh);
    const double S2 = block_sum(a2, sh);
    const double B1 = block_sum(b1, sh);

    // Transition terms of this thread's R windows (k = tid*R + v), used by
    // windows k+1 and later.
    double d1[R], d2[R], e1[R], e2[R];
    double t1 = 0.0, t2 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
    for (int v = 0; v < R; ++v) {
        const int k = tid * R + v;
        double in1 = 0.0, in2 = 0.0, ab1 = 0.0, ab2 = 0.0;
        if (k + 1 < nw) {
            double xo = static_cast<double>(x[k]) - c;
            double xn = static_cast<double>(x[k + m]) - c;
            in1 = xn - xo;
            double xo2 = xo * xo, xn2 = xn * xn;
            in2 = xn2 - xo2;
            ab1 = fabs(xn) + fabs(xo);
            ab2 = xn2 + xo2;
        }
        // exclusive prefix within the thread
        d1[v] = t1; d2[v] = t2; e1[v] = u1; e2[v] = u2;
        t1 += in1; t2 += in2; u1 += ab1; u2 += ab2;
    }
    double tot;
    const double p1 = block_scan_incl(t1, sh, &tot) - t1;
    const double p2 = block_scan_incl(t2, sh, &tot) - t2;
    const double q1 = block_scan_incl(u1, sh, &tot) - u1;
    const double q2 = block_scan_incl(u2, sh, &tot) - u2;

    const double inv_m = 1.0 / static_cast<double>(m);
#pragma unroll
    for (int v = 0; v < R; ++v) {
        const int r = tid * R + v;
        if (r < nw) {
            const double sum1 = S1 + (p1 + d1[v]);
            const double sum2 = S2 + (p2 + d2[v]);
            const double mean = sum1 * inv_m;
            double var = fma(-mean, mean, sum2 * inv_m);
            // Rounding-level variance of an exactly constant window is noise:
            // treat it as zero (sigma := 1, the reference's guard).
            const double mag2 = (S2 + q2 + e2[v]) * inv_m;
            const double mag1 = (B1 + q1 + e1[v]) * inv_m;
            const double tol = 64.0 * DBL_EPSILON * (mag2 + 2.0 * fabs(mean) * mag1);
            if (var <= tol) var = 0.0;
            avg[w0 + r] = static_cast<float>(c + mean);
            sd[w0 + r] = var > 0.0 ? static_cast<float>(sqrt(var)) : 1.0f;
        }
    }
}

// ---- REFCOMPAT ------------------------------------------------------------
constexpr int kChunk = kThreads * 16;  // elements per workgroup in the scan

__global__ __launch_bounds__(kThreads) void ref_block_sums(
    const float* __restrict__ x, int64_t len, double* __restrict__ bs1,
    double* __restrict__ bs2) {
    __shared__ double sh[kThreads / kWave];
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x * 16;
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        int64_t k = b0 + v;
        if (k < len) {
            double d = static_cast<double>(x[k]);
            a1 += d;
            a2 += d * d;
        }
    }
    double tot;
    block_scan_incl(a1, sh, &tot);
    if (threadIdx.x == 0) bs1[blockIdx.x] = tot;
    block_scan_incl(a2, sh, &tot);
    if (threadIdx.x == 0) bs2[blockIdx.x] = tot;
}

// Exclusive scan of nb block totals in one workgroup (nb <= 256 * 64).
__global__ __launch_bounds__(kThreads) void ref_scan_blocks(double* __restrict__ bs1,
                                                            double* __restrict__ bs2,
                                                            int nb) {
    __shared__ double sh[kThreads / kWave];
    const int per = (nb + kThreads - 1) / kThreads;
    const int lo = threadIdx.x * per;
    double a1 = 0.0, a2 = 0.0;
    for (int i = 0; i < per; ++i)
        if (lo + i < nb) {
            a1 += bs1[lo + i];
            a2 += bs2[lo + i];
        }
    double tot;
    double o1 = block_scan_incl(a1, sh, &tot) - a1;
    double o2 = block_scan_incl(a2, sh, &tot) - a2;
    for (int i = 0; i < per; ++i)
        if (lo + i < nb) {
            double v1 = bs1[lo + i], v2 = bs2[lo + i];
            bs1[lo + i] = o1;
            bs2[lo + i] = o2;
            o1 += v1;
            o2 += v2;
        }
}

__global__ __launch_bounds__(kThreads) void ref_prefix(const float* __restrict__ x,
                                                       int64_t len,
                                                       const double* __restrict__ off1,
                                                       const double* __restrict__ off2,
                                                       double* __restrict__ P1,
                                                       double* __restrict__ P2) {
    __shared__ double sh[kThreads / kWave];
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kChunk + threadIdx.x * 16;
    double v1[16], v2[16];
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        int64_t k = b0 + v;
        double d = k < len ? static_cast<double>(x[k]) : 0.0;
        a1 += d;
        a2 += d * d;
        v1[v] = a1;
        v2[v] = a2;
    }
    double tot;
    const double base1 = off1[blockIdx.x] + (block_scan_incl(a1, sh, &tot) - a1);
    const double base2 = off2[blockIdx.x] + (block_scan_incl(a2, sh, &tot) - a2);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        int64_t k = b0 + v;
        if (k < len) {
            P1[k + 1] = base1 + v1[v];
            P2[k + 1] = base2 + v2[v];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P1[0] = 0.0;
        P2[0] = 0.0;
    }
}

__global__ void ref_finalize(const double* __restrict__ P1, const double* __restrict__ P2,
                             int64_t nwin, int32_t m, float* __restrict__ avg,
                             float* __restrict__ sd) {
    const double dm = static_cast<double>(m);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nwin;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float X = static_cast<float>((P1[i + m] - P1[i]) / dm);
        const float X2 = static_cast<float>((P2[i + m] - P2[i]) / dm);
        const float v = fmaf(-X, X, X2);
        avg[i] = X;
        sd[i] = v > 0.0f ? sqrtf(v) : 1.0f;
    }
}

}  // namespace

int launch_stats_fp64(hipStream_t st, const float* s, int64_t n, int32_t m, float* avg,
                      float* sd) {
    const int64_t L = n - m + 1;
    if (L <= 0) return MDTW_ERR_STREAM_TOO_SHORT;
    if (m >= 2048) {
        constexpr int T = kThreads * 16;
        const int64_t nb = (L + T - 1) / T;
        hipLaunchKernelGGL(stats_fp64_kernel<16>, dim3(static_cast<unsigned>(nb)),
                           dim3(kThreads), 0, st, s, L, m, avg, sd);
    } else {
        constexpr int T = kThreads * 8;
        const int64_t nb = (L + T - 1) / T;
        hipLaunchKernelGGL(stats_fp64_kernel<8>, dim3(static_cast<unsigned>(nb)),
                           dim3(kThreads), 0, st, s, L, m, avg, sd);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int launch_stats_refcompat(mdtw_engine* e, const float* s, int64_t n, int32_t m,
                           float* avg, float* sd) {
    // Epoch plan of stats.cuh:120-134 (a final short chunk merges into the
    // previous epoch; a single short chunk is processed as is).
    const int64_t epoch = 10000000;
    const int64_t width = epoch - m + 1;
    std::vector<int64_t> lowers, uppers;
    for (int64_t lower = 0; lower < n; lower += width) {
        int64_t upper = lower + epoch < n ? lower + epoch : n;
        if (upper - lower > m + 1 || lowers.empty()) {
            lowers.push_back(lower);
            uppers.push_back(upper);
        } else {
            uppers.back() = upper;
        }
    }
    int64_t maxlen = 0;
    for (size_t i = 0; i < lowers.size(); ++i)
        maxlen = std::max(maxlen, uppers[i] - lowers[i]);
    const int64_t nbmax = (maxlen + kChunk - 1) / kChunk;
    if (nbmax > kThreads * 64) return MDTW_ERR_UNSUPPORTED;
    void* ws = nullptr;
    const size_t bytes = carve_size({sizeof(double) * (size_t)(maxlen + 1),
                                     sizeof(double) * (size_t)(maxlen + 1),
                                     sizeof(double) * (size_t)nbmax,
                                     sizeof(double) * (size_t)nbmax});
    int rc = ws_reserve(e, bytes, &ws);
    if (rc != MDTW_OK) return rc;
    Carve cv{static_cast<char*>(ws)};
    double* P1 = cv.take<double>(maxlen + 1);
    double* P2 = cv.take<double>(maxlen + 1);
    double* bs1 = cv.take<double>(nbmax);
    double* bs2 = cv.take<double>(nbmax);
    for (size_t i = 0; i < lowers.size(); ++i) {
        const int64_t lo = lowers[i], len = uppers[i] - lowers[i];
        if (len < m) continue;
        const int nb = static_cast<int>((len + kChunk - 1) / kChunk);
        hipLaunchKernelGGL(ref_block_sums, dim3(nb), dim3(kThreads), 0, e->stream, s + lo,
                           len, bs1, bs2);
        hipLaunchKernelGGL(ref_scan_blocks, dim3(1), dim3(kThreads), 0, e->stream, bs1,
                           bs2, nb);
        hipLaunchKernelGGL(ref_prefix, dim3(nb), dim3(kThreads), 0, e->stream, s + lo, len,
                           bs1, bs2, P1, P2);
        const int64_t nwin = len - m + 1;
        const int64_t g = std::min<int64_t>((nwin + 255) / 256, 4096);
        hipLaunchKernelGGL(ref_finalize, dim3(static_cast<unsigned>(g)), dim3(256), 0,
                           e->stream, P1, P2, nwin, m, avg + lo, sd + lo);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

}  // namespace mdtw
