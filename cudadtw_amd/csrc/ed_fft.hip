// ed_fft.hip -- Euclidean distance of every window through FFT
// cross-correlation (FP64), for long queries.
//
// The reference computes ED this way for every query (ed.cuh:71-153: cuFFT
// of the stream and of the query in FP64, a product, an inverse, then
// 2M - 2 corr / sigma with FP64 window statistics).  Direct ED costs 3M
// VALU ops per window and wins for short queries; the FFT costs
// O(log N) per window whatever M is, so it wins for long ones.
//
// Overlap-save in blocks of N real samples (N = the smallest power of two
// >= 2M, at least 4096, at most 16384: M <= 8192): block b holds samples
// [bV, bV + N) with V = N - M + 1 windows, minus the block's first sample c
// (offset data loses no precision).  One workgroup of H/8 threads per
// block, all in LDS (H = N/2 complex doubles plus one pad per 8, at most
// 144 KiB):
//   - the N real samples are packed as H complex values z[n] = x[2n] +
//     i x[2n+1], stored at bit-reversed positions,
//   - forward decimation-in-time FFT of size H, in place, three radix-2
//     stages per pass in registers (one LDS round trip and one barrier per
//     three stages),
//   - real-spectrum split X[k] = E[k] + W^k O[k], E/O from Z[k] and
//     conj(Z[H-k]) (pairs k, H-k in one thread), product Y = X conj(Q)
//     with the query's spectrum Q (same transform, computed once), and the
//     inverse's packing A[k] + i B[k], A = Y[k] + conj(Y[H-k]),
//     B = (Y[k] - conj(Y[H-k])) conj(W^k),
//   - inverse decimation-in-frequency FFT, three stages per pass (natural order in,
//     bit-reversed out), so y[2n] + i y[2n+1] = z'[n] / N is the
//     cross-correlation dot_r = sum_i q_i (x_{bV+r+i} - c), r < V,
//   - ED^2 = sum q^2 + sum z^2 - 2 (dot_r - (mu_r - c) sum q) / sigma_r,
//     with the window's FP64 mean and variance (the statistics tile kernel
//     in FP64): sum z^2 = M (0 for a constant window, sigma := 1).  Raw
//     values are not offered here: sum q^2 + sum x^2 - 2 sum q x cancels
//     about eleven digits on offset data.
// Twiddles W^j = exp(-2 pi i j / N), j < N/2, come from sincospi (exact
// argument 2j/N) into a table read through the cache.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "internal.hpp"

using namespace mdtw;

namespace {

constexpr int kFftMinLog = 12;  // N >= 4096 real samples per block
constexpr int kFftMaxLog = 14;  // N <= 16384 (H = 8192 complex doubles in LDS)

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ double2 cconj(double2 a) { return {a.x, -a.y}; }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return {a.x * s, a.y * s}; }

template <int LOGH>
__device__ __forceinline__ int bitrev(int n) {
    return static_cast<int>(__brev(static_cast<unsigned>(n)) >> (32 - LOGH));
}

__global__ __launch_bounds__(256) void twiddle_kernel(double2* __restrict__ tw, int logn) {
    const int h = 1 << (logn - 1);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= h) return;
    double sn, cs;
    sincospi(static_cast<double>(2 * j) / static_cast<double>(1 << logn), &sn, &cs);
    tw[j] = {cs, -sn};
}

// LDS layout of the H complex values: one pad entry after every 8, so the
// eight-element groups of the first radix-8 pass (lane strides of 8
// entries) and the strided groups of the later ones fall in distinct banks.
__device__ __forceinline__ int fx(int i) { return i + (i >> 3); }
template <int LOGH>
constexpr int fft_lds_entries() {
    return (1 << LOGH) + (1 << LOGH) / 8;
}
// Threads per transform block: one radix-8 group each.
template <int LOGH>
constexpr int fft_threads() {
    return (1 << LOGH) / 8;
}

__device__ __forceinline__ void bfly_dit(double2& u, double2& v, double2 w) {
    const double2 t = cmul(v, w);
    v = csub(u, t);
    u = cadd(u, t);
}
__device__ __forceinline__ void bfly_dif(double2& u, double2& v, double2 w) {
    const double2 d = csub(u, v);
    u = cadd(u, v);
    v = cmul(d, cconj(w));
}

// In-place forward DIT FFT of a[0..H) (bit-reversed in, natural out).
// Stage s combines pairs (i, i + 2^(s-1)) within blocks of 2^s with the
// twiddle of i's position in its block; three consecutive stages s..s+2 only
// combine the eight entries pos + k 2^(s-1) (k < 8) of a block of 2^(s+2),
// so a thread takes those eight into registers, applies the three stages
// (the same butterflies, in the same order per entry, as stage by stage) and
// writes them back: one LDS round trip and one barrier per three stages.
template <int LOGH>
__device__ __forceinline__ void fft_dit(double2* a, const double2* __restrict__ tw) {
    constexpr int H = 1 << LOGH;
    constexpr int NT = fft_threads<LOGH>();
    int s = 1;
    for (; s + 2 <= LOGH; s += 3) {
        const int h = 1 << (s - 1);
        for (int b = threadIdx.x; b < H / 8; b += NT) {
            const int pos = b & (h - 1);
            const int base = ((b >> (s - 1)) << (s + 2)) + pos;
            double2 e[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) e[k] = a[fx(base + k * h)];
            const double2 w1 = tw[pos << (LOGH - s + 1)];
#pragma unroll
            for (int k = 0; k < 8; k += 2) bfly_dit(e[k], e[k + 1], w1);
            const double2 w2a = tw[pos << (LOGH - s)], w2b = tw[(pos + h) << (LOGH - s)];
            bfly_dit(e[0], e[2], w2a);
            bfly_dit(e[1], e[3], w2b);
            bfly_dit(e[4], e[6], w2a);
            bfly_dit(e[5], e[7], w2b);
#pragma unroll
            for (int k = 0; k < 4; ++k) bfly_dit(e[k], e[k + 4], tw[(pos + k * h) << (LOGH - s - 1)]);
#pragma unroll
            for (int k = 0; k < 8; ++k) a[fx(base + k * h)] = e[k];
        }
        __syncthreads();
    }
    for (; s <= LOGH; ++s) {  // the remaining one or two stages, radix 2
        const int half = 1 << (s - 1);
        for (int b = threadIdx.x; b < H / 2; b += NT) {
            const int pos = b & (half - 1);
            const int i = ((b >> (s - 1)) << s) + pos;
            double2 u = a[fx(i)], v = a[fx(i + half)];
            bfly_dit(u, v, tw[pos << (LOGH - s + 1)]);
            a[fx(i)] = u;
            a[fx(i + half)] = v;
        }
        __syncthreads();
    }
}

// In-place inverse DIF FFT of a[0..H) (natural in, bit-reversed out),
// unnormalised: sum_k a[k] exp(+2 pi i n k / H).  Stages s = LOGH .. 1
// (pairs (i, i + 2^(s-1)), difference times the conjugate twiddle of i's
// position), three per LDS round trip as in fft_dit.
template <int LOGH>
__device__ __forceinline__ void ifft_dif(double2* a, const double2* __restrict__ tw) {
    constexpr int H = 1 << LOGH;
    constexpr int NT = fft_threads<LOGH>();
    int s = LOGH;
    for (; s - 2 >= 1; s -= 3) {
        const int h = 1 << (s - 3);
        for (int b = threadIdx.x; b < H / 8; b += NT) {
            const int pos = b & (h - 1);
            const int base = ((b >> (s - 3)) << s) + pos;
            double2 e[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) e[k] = a[fx(base + k * h)];
#pragma unroll
            for (int k = 0; k < 4; ++k) bfly_dif(e[k], e[k + 4], tw[(pos + k * h) << (LOGH - s + 1)]);
            const double2 w2a = tw[pos << (LOGH - s + 2)], w2b = tw[(pos + h) << (LOGH - s + 2)];
            bfly_dif(e[0], e[2], w2a);
            bfly_dif(e[1], e[3], w2b);
            bfly_dif(e[4], e[6], w2a);
            bfly_dif(e[5], e[7], w2b);
            const double2 w1 = tw[pos << (LOGH - s + 3)];
#pragma unroll
            for (int k = 0; k < 8; k += 2) bfly_dif(e[k], e[k + 1], w1);
#pragma unroll
            for (int k = 0; k < 8; ++k) a[fx(base + k * h)] = e[k];
        }
        __syncthreads();
    }
    for (; s >= 1; --s) {
        const int half = 1 << (s - 1);
        for (int b = threadIdx.x; b < H / 2; b += NT) {
            const int pos = b & (half - 1);
            const int i = ((b >> (s - 1)) << s) + pos;
            double2 u = a[fx(i)], v = a[fx(i + half)];
            bfly_dif(u, v, tw[pos << (LOGH - s + 1)]);
            a[fx(i)] = u;
            a[fx(i + half)] = v;
        }
        __syncthreads();
    }
}

// X[k] and X[H-k] of the real N-point spectrum from Z[k], Z[H-k] (1 <= k < H).
__device__ __forceinline__ void real_split(double2 zk, double2 zm, double2 wk, double2 wm,
                                           double2* xk, double2* xm) {
    const double2 e = cscale(cadd(zk, cconj(zm)), 0.5);
    const double2 d = csub(zk, cconj(zm));
    const double2 o = {0.5 * d.y, -0.5 * d.x};  // d / (2i)
    *xk = cadd(e, cmul(wk, o));
    *xm = cadd(cconj(e), cmul(wm, cconj(o)));
}

// The query's real spectrum Q[0..H] (q zero-padded to N = 2H); QT is the
// query's type (FP32, or FP64 for the FP64-stream form).
template <int LOGH, class QT>
__global__ __launch_bounds__(fft_threads<LOGH>()) void query_spectrum_kernel(
    const QT* __restrict__ q, int32_t m, const double2* __restrict__ tw,
    double2* __restrict__ qs) {
    constexpr int H = 1 << LOGH;
    constexpr int NT = fft_threads<LOGH>();
    __shared__ double2 a[fft_lds_entries<LOGH>()];
    for (int n = threadIdx.x; n < H; n += NT) {
        const double v0 = 2 * n < m ? static_cast<double>(q[2 * n]) : 0.0;
        const double v1 = 2 * n + 1 < m ? static_cast<double>(q[2 * n + 1]) : 0.0;
        a[fx(bitrev<LOGH>(n))] = {v0, v1};
    }
    __syncthreads();
    fft_dit<LOGH>(a, tw);
    for (int k = threadIdx.x; k <= H / 2; k += NT) {
        if (k == 0) {
            const double2 z0 = a[fx(0)];
            qs[0] = {z0.x + z0.y, 0.0};
            qs[H] = {z0.x - z0.y, 0.0};
        } else {
            double2 xk, xm;
            real_split(a[fx(k)], a[fx(H - k)], tw[k], tw[H - k], &xk, &xm);
            qs[k] = xk;
            qs[H - k] = xm;  // k = H/2: the same value twice
        }
    }
}

struct FftScalars {
    int64_t L;        // windows of the stream
    int64_t n;        // stream samples
    int64_t V;        // windows per block
    int32_t m;
    int32_t pad;
    double q2;        // sum of q^2
    double qsum;      // sum of q
};

// S: the stream's type, O: the profile's (FP32 for FP32 streams, FP64 for
// FP64 ones).
template <int LOGH, class S, class O>
__global__ __launch_bounds__(fft_threads<LOGH>()) void ed_fft_kernel(
    const S* __restrict__ s, const double* __restrict__ avg, const double* __restrict__ var,
    const double2* __restrict__ tw, const double2* __restrict__ qs, O* __restrict__ out,
    FftScalars a) {
    constexpr int H = 1 << LOGH;
    constexpr int N = 2 * H;
    constexpr int NT = fft_threads<LOGH>();
    __shared__ double2 z[fft_lds_entries<LOGH>()];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * a.V;
    // Non-finite samples enter the transform as 0 (their windows' statistics
    // are NaN, so those windows' distances are NaN); one would otherwise
    // spread through the whole block's transform.
    const double c = isfinite(s[base]) ? static_cast<double>(s[base]) : 0.0;
    for (int n = threadIdx.x; n < H; n += NT) {
        const int64_t i0 = base + 2 * n;
        const S x0 = i0 < a.n ? s[i0] : S(0);
        const S x1 = i0 + 1 < a.n ? s[i0 + 1] : S(0);
        const double v0 = isfinite(x0) && i0 < a.n ? static_cast<double>(x0) - c : 0.0;
        const double v1 = isfinite(x1) && i0 + 1 < a.n ? static_cast<double>(x1) - c : 0.0;
        z[fx(bitrev<LOGH>(n))] = {v0, v1};
    }
    __syncthreads();
    fft_dit<LOGH>(z, tw);
    for (int k = threadIdx.x; k <= H / 2; k += NT) {
        if (k == 0) {
            const double2 z0 = z[fx(0)];
            const double2 y0 = cmul({z0.x + z0.y, 0.0}, cconj(qs[0]));
            const double2 yh = cmul({z0.x - z0.y, 0.0}, cconj(qs[H]));
            const double2 A = cadd(y0, cconj(yh));
            const double2 B = csub(y0, cconj(yh));
            z[fx(0)] = {A.x - B.y, A.y + B.x};
        } else {
            const int km = H - k;
            const double2 wk = tw[k], wm = tw[km];
            double2 xk, xm;
            real_split(z[fx(k)], z[fx(km)], wk, wm, &xk, &xm);
            const double2 yk = cmul(xk, cconj(qs[k]));
            const double2 ym = cmul(xm, cconj(qs[km]));
            const double2 Ak = cadd(yk, cconj(ym));
            const double2 Bk = cmul(csub(yk, cconj(ym)), cconj(wk));
            z[fx(k)] = {Ak.x - Bk.y, Ak.y + Bk.x};
            if (km != k) {
                const double2 Am = cadd(ym, cconj(yk));
                const double2 Bm = cmul(csub(ym, cconj(yk)), cconj(wm));
                z[fx(km)] = {Am.x - Bm.y, Am.y + Bm.x};
            }
        }
    }
    __syncthreads();
    ifft_dif<LOGH>(z, tw);
    const double inv_n = 1.0 / static_cast<double>(N);
    const double m = static_cast<double>(a.m);
    for (int r = threadIdx.x; r < a.V; r += NT) {
        const int64_t p = base + r;
        if (p >= a.L) break;
        const double2 v = z[fx(bitrev<LOGH>(r >> 1))];
        const double dot = ((r & 1) ? v.y : v.x) * inv_n;  // sum q (x - c)
        const double mu = avg[p], va = var[p];
        const double sig = va > 0.0 ? sqrt(va) : 1.0;
        const double sz2 = va > 0.0 ? m : 0.0;
        const double ed2 = a.q2 + sz2 - 2.0 * (dot - (mu - c) * a.qsum) / sig;
        // below 0 (rounding) -> +0; NaN (a non-finite sample) stays NaN
        out[p] = static_cast<O>(ed2 > 0.0 ? ed2 : (ed2 == ed2 ? 0.0 : ed2));
    }
}

// ---- selection over a distance profile --------------------------------------
__global__ __launch_bounds__(256) void profile_best_kernel(const float* __restrict__ v, int64_t n,
                                                           Ctrl* ctrl) {
    unsigned long long best = ~0ull;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long key = pack_key(v[i], static_cast<uint32_t>(i));
        best = key < best ? key : best;
    }
    best = wave_min_u64(best);
    if ((threadIdx.x & (kWave - 1)) == 0 && best != ~0ull) atomicMin(&ctrl->best, best);
}

__global__ __launch_bounds__(256) void profile_gather_kernel(const float* __restrict__ v,
                                                             int64_t count, int64_t stride,
                                                             float* __restrict__ seeds) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < count) seeds[j] = v[j * stride];
}

// Keys (distance, index) of every entry at or below ctrl->thr.
__global__ __launch_bounds__(256) void profile_collect_kernel(const float* __restrict__ v,
                                                              int64_t n, Ctrl* ctrl,
                                                              unsigned long long* __restrict__ res) {
    const float thr = load_thr(ctrl);
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        const bool keep = i < n && v[i] <= thr;
        const unsigned long long km = __ballot(keep);
        if (!km) continue;
        unsigned long long rb = 0;
        if (lane == 0) rb = atomicAdd(&ctrl->nres, static_cast<unsigned long long>(__popcll(km)));
        rb = __shfl(rb, 0, kWave);
        if (keep) {
            const uint32_t rank =
                __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(km >> 32),
                                          __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(km), 0u));
            res[rb + rank] = pack_key(v[i], static_cast<uint32_t>(i));
        }
    }
}

int hip_ok(hipError_t err) {
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int fft_log(int32_t m) {
    int lg = kFftMinLog;
    while ((int64_t(1) << lg) < 2 * static_cast<int64_t>(m)) ++lg;
    return lg;
}

template <int LOGH, class S, class O>
int launch_fft_pair(hipStream_t st, const S* s, const double* avg, const double* var,
                    const double2* tw, const S* q, double2* qs, O* out, const FftScalars& a) {
    constexpr int NT = fft_threads<LOGH>();
    hipLaunchKernelGGL((query_spectrum_kernel<LOGH, S>), dim3(1), dim3(NT), 0, st, q,
                       a.m, tw, qs);
    int rc = hip_ok(hipGetLastError());
    if (rc != MDTW_OK) return rc;
    const int64_t nb = (a.L + a.V - 1) / a.V;
    hipLaunchKernelGGL((ed_fft_kernel<LOGH, S, O>), dim3(static_cast<unsigned>(nb)),
                       dim3(NT), 0, st, s, avg, var, tw, qs, out, a);
    return hip_ok(hipGetLastError());
}

// Statistics, twiddles, query spectrum and the ED^2 profile of all L windows
// (avg, var: L FP64 each; tw: N/2; qs: N/2 + 1 entries, N = 2^fft_log(m)).
template <class S, class O>
int run_ed_fft(hipStream_t st, const S* s, int64_t n, const S* q, int32_t m, double q2,
               double qsum, double* avg, double* var, double2* tw, double2* qs, O* out) {
    const int64_t L = n - m + 1;
    const int lg = fft_log(m);
    const int64_t h = int64_t(1) << (lg - 1);
    int rc = std::is_same<S, float>::value
                 ? launch_stats_f64out(st, reinterpret_cast<const float*>(s), n, m, avg, var)
                 : launch_stats_d_var(st, reinterpret_cast<const double*>(s), n, m, avg, var);
    if (rc != MDTW_OK) return rc;
    hipLaunchKernelGGL(twiddle_kernel, dim3(static_cast<unsigned>((h + 255) / 256)), dim3(256), 0,
                       st, tw, lg);
    rc = hip_ok(hipGetLastError());
    if (rc != MDTW_OK) return rc;
    FftScalars a;
    a.L = L;
    a.n = n;
    a.V = (int64_t(1) << lg) - m + 1;
    a.m = m;
    a.pad = 0;
    a.q2 = q2;
    a.qsum = qsum;
    switch (lg) {
        case 12: return launch_fft_pair<11>(st, s, avg, var, tw, q, qs, out, a);
        case 13: return launch_fft_pair<12>(st, s, avg, var, tw, q, qs, out, a);
        default: return launch_fft_pair<13>(st, s, avg, var, tw, q, qs, out, a);
    }
}

constexpr int64_t kFftSeedTarget = 8192;

// Sum of q^2 and of q over the engine's query (its first m host entries).
void query_sums(const mdtw_engine* e, double* q2, double* qsum) {
    double a = 0.0, b = 0.0;
    for (int32_t i = 0; i < e->m; ++i) {
        const double v = static_cast<double>(e->h_q[static_cast<size_t>(i)]);
        a += v * v;
        b += v;
    }
    *q2 = a;
    *qsum = b;
}

}  // namespace

namespace mdtw {

size_t ed_fft_ws_bytes(int64_t L, int32_t m, int32_t k) {
    const int lg = fft_log(m);
    const int64_t h = int64_t(1) << (lg - 1);
    return carve_size({sizeof(double) * L, sizeof(double) * L, sizeof(float) * L,
                       sizeof(double2) * h, sizeof(double2) * (h + 1),
                       sizeof(float) * kSeedMaxCount,
                       k > 1 ? sizeof(unsigned long long) * L : 0});
}

// ED^2 profile of all L windows into out (ws: ed_fft_ws_bytes(L, m, 1)).
int enqueue_ed_fft(hipStream_t st, const float* s, int64_t n, const float* q, int32_t m,
                   double q2, double qsum, int32_t stats_mode, char* ws, float* out,
                   float** prof_out) {
    if (m < 1 || m > MDTW_MAX_QUERY) return MDTW_ERR_QUERY_TOO_LONG;
    if (stats_mode != MDTW_STATS_FP64) return MDTW_ERR_UNSUPPORTED;
    const int64_t L = n - m + 1;
    if (L <= 0) return MDTW_ERR_STREAM_TOO_SHORT;
    const int lg = fft_log(m);
    if (lg > kFftMaxLog) return MDTW_ERR_UNSUPPORTED;
    const int64_t h = int64_t(1) << (lg - 1);
    Carve cv{ws};
    double* avg = cv.take<double>(L);
    double* var = cv.take<double>(L);
    float* prof = cv.take<float>(L);
    double2* tw = cv.take<double2>(h);
    double2* qs = cv.take<double2>(h + 1);
    if (!out) out = prof;
    if (prof_out) *prof_out = out;
    return run_ed_fft<float, float>(st, s, n, q, m, q2, qsum, avg, var, tw, qs, out);
}

}  // namespace mdtw

extern "C" {

int mdtw_ed_fft(mdtw_engine* e, const float* d_series, int64_t n, int32_t stats_mode,
                float* d_out) {
    if (!e || !d_series || !d_out) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    const int32_t m = e->m;
    if (n < m) return MDTW_ERR_STREAM_TOO_SHORT;
    if (fft_log(m) > kFftMaxLog) return MDTW_ERR_UNSUPPORTED;
    DeviceGuard guard(e->device);
    const int64_t L = n - m + 1;
    void* ws = nullptr;
    int rc = ws_reserve(e, ed_fft_ws_bytes(L, m, 1), &ws);
    if (rc != MDTW_OK) return rc;
    double q2 = 0.0, qsum = 0.0;
    query_sums(e, &q2, &qsum);
    return enqueue_ed_fft(e->stream, d_series, n, e->d_q, m, q2, qsum, stats_mode,
                          static_cast<char*>(ws), d_out, nullptr);
}

}  // extern "C"

namespace mdtw {

// Best / top-k over the FFT ED profile (mdtw_search with MDTW_ALGO_ED_FFT).
int search_ed_fft(mdtw_engine* e, const float* d_series, int64_t n, int32_t stats_mode, int32_t k,
                  int64_t excl, int64_t* h_pos, float* h_dist, mdtw_search_info* info) {
    const int32_t m = e->m;
    const int64_t L = n - m + 1;
    void* wsv = nullptr;
    int rc = ws_reserve(e, ed_fft_ws_bytes(L, m, k), &wsv);
    if (rc != MDTW_OK) return rc;
    char* ws = static_cast<char*>(wsv);
    double q2 = 0.0, qsum = 0.0;
    query_sums(e, &q2, &qsum);
    hipStream_t st = e->stream;
    rc = hip_ok(hipEventRecord(e->ev0, st));
    if (rc == MDTW_OK) rc = launch_ctrl_reset(st, e->d_ctrl, kBestNone);
    float* prof = nullptr;
    if (rc == MDTW_OK)
        rc = enqueue_ed_fft(st, d_series, n, e->d_q, m, q2, qsum, stats_mode, ws, nullptr, &prof);
    if (rc != MDTW_OK) return rc;
    // The profile's scratch layout, as enqueue_ed_fft carved it.
    const int lg = fft_log(m);
    const int64_t h = int64_t(1) << (lg - 1);
    Carve cv{ws};
    (void)cv.take<double>(L);
    (void)cv.take<double>(L);
    (void)cv.take<float>(L);
    (void)cv.take<double2>(h);
    (void)cv.take<double2>(h + 1);
    float* seeds = cv.take<float>(kSeedMaxCount);
    unsigned long long* res = k > 1 ? cv.take<unsigned long long>(L) : nullptr;
    const unsigned nbl = static_cast<unsigned>(
        std::min<int64_t>((L + 255) / 256, static_cast<int64_t>(e->cus) * 16));
    int64_t seeded = 0;
    if (k == 1) {
        hipLaunchKernelGGL(profile_best_kernel, dim3(nbl), dim3(256), 0, st, prof, L, e->d_ctrl);
        rc = hip_ok(hipGetLastError());
    } else {
        // k-th smallest of entries >= 2*excl apart bounds the k-th greedy
        // pick; every entry at or below it is kept and picked on the host.
        int64_t stride = std::max<int64_t>(std::max<int64_t>(1, L / kFftSeedTarget),
                                           2 * std::max<int64_t>(excl, 0));
        int64_t count = (L - 1) / stride + 1;
        if (count > kSeedMaxCount) {
            stride = std::max<int64_t>((L + kSeedMaxCount - 1) / kSeedMaxCount,
                                       2 * std::max<int64_t>(excl, 0));
            count = (L - 1) / stride + 1;
        }
        seeded = count;
        hipLaunchKernelGGL(profile_gather_kernel, dim3(static_cast<unsigned>((count + 255) / 256)),
                           dim3(256), 0, st, prof, count, stride, seeds);
        rc = hip_ok(hipGetLastError());
        if (rc == MDTW_OK) rc = launch_seed_threshold(st, seeds, count, k, e->d_ctrl, INFINITY);
        if (rc == MDTW_OK) {
            hipLaunchKernelGGL(profile_collect_kernel, dim3(nbl), dim3(256), 0, st, prof, L,
                               e->d_ctrl, res);
            rc = hip_ok(hipGetLastError());
        }
    }
    if (rc == MDTW_OK) rc = hip_ok(hipEventRecord(e->ev1, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(e->h_ctrl, e->d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(st));
    if (rc != MDTW_OK) return rc;
    const Ctrl c = *e->h_ctrl;
    int32_t found = 0;
    float thr_out = INFINITY;
    if (k == 1) {
        if ((c.best >> 32) < 0x7f800000ull) {
            const uint32_t bits = static_cast<uint32_t>(c.best >> 32);
            float d;
            std::memcpy(&d, &bits, sizeof(d));
            h_pos[0] = static_cast<int64_t>(c.best & 0xffffffffull);
            h_dist[0] = std::sqrt(d);
            found = 1;
            thr_out = d;
        }
    } else {
        thr_out = c.thr;
        std::vector<unsigned long long> keys(static_cast<size_t>(c.nres));
        if (c.nres) {
            rc = hip_ok(hipMemcpyAsync(keys.data(), res, sizeof(unsigned long long) * c.nres,
                                       hipMemcpyDeviceToHost, st));
            if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(st));
            if (rc != MDTW_OK) return rc;
        }
        std::sort(keys.begin(), keys.end());
        std::vector<std::pair<float, int64_t>> v;
        v.reserve(keys.size());
        for (unsigned long long key : keys) {
            const uint32_t bits = static_cast<uint32_t>(key >> 32);
            float d;
            std::memcpy(&d, &bits, sizeof(d));
            v.emplace_back(d, static_cast<int64_t>(key & 0xffffffffull));
        }
        found = greedy_pick(v, k, excl, MDTW_METRIC_L2, h_pos, h_dist);
    }
    if (info) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
        info->windows = L;
        info->seeded = seeded;
        info->lb_kim_pruned = 0;
        info->lb_keogh_pruned = 0;
        info->dp_evaluated = L;
        info->dp_abandoned = 0;
        info->found = found;
        info->threshold = thr_out;
        info->elapsed_ms = ms;
    }
    return MDTW_OK;
}

}  // namespace mdtw

// ---- FP64 streams ------------------------------------------------------------
// The reference's ED driver is FP64 end to end (GPU-ed.cu:20-82: an FP64
// stream, avg_std<double>, calculate_ed on doubles, ed.cuh:71-153).  These
// entry points take the FP64 stream and the FP64 z-normalised query as they
// are: the same overlap-save transform with FP64 samples, FP64 window
// statistics of the FP64 stream, an FP64 profile, and its argmin in FP64.
namespace {

__global__ void f64_min_reset_kernel(unsigned long long* r) {
    if (threadIdx.x == 0) {
        r[0] = ~0ull;
        r[1] = ~0ull;
    }
}

// r[0] = bits of the lowest value of v[0..n) (NaN skipped; the profile is
// >= +0, where the bit patterns order like the values).
__global__ __launch_bounds__(256) void f64_min_bits_kernel(const double* __restrict__ v,
                                                           int64_t n, unsigned long long* r) {
    unsigned long long b = ~0ull;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double x = v[i];
        const unsigned long long k = static_cast<unsigned long long>(__double_as_longlong(x));
        if (x == x && k < b) b = k;
    }
    b = wave_min_u64(b);
    if ((threadIdx.x & (kWave - 1)) == 0 && b != ~0ull) atomicMin(&r[0], b);
}

// r[1] = the lowest index holding that value.
__global__ __launch_bounds__(256) void f64_min_index_kernel(const double* __restrict__ v,
                                                            int64_t n, unsigned long long* r) {
    const unsigned long long want = r[0];
    unsigned long long b = ~0ull;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const unsigned long long k = static_cast<unsigned long long>(__double_as_longlong(v[i]));
        if (k == want && static_cast<unsigned long long>(i) < b) b = static_cast<unsigned long long>(i);
    }
    b = wave_min_u64(b);
    if ((threadIdx.x & (kWave - 1)) == 0 && b != ~0ull) atomicMin(&r[1], b);
}

int check_f64_query(const double* h_zquery, int32_t m) {
    if (!h_zquery || m < 1) return MDTW_ERR_INVALID_ARGUMENT;
    if (m > MDTW_MAX_QUERY || fft_log(m) > kFftMaxLog) return MDTW_ERR_QUERY_TOO_LONG;
    for (int32_t i = 0; i < m; ++i)
        if (!std::isfinite(h_zquery[i])) return MDTW_ERR_NOT_FINITE;
    return MDTW_OK;
}

size_t ed_f64_ws_bytes(int64_t L, int32_t m) {
    const int64_t h = int64_t(1) << (fft_log(m) - 1);
    return carve_size({sizeof(double) * L, sizeof(double) * L, sizeof(double) * L,
                       sizeof(double2) * h, sizeof(double2) * (h + 1), sizeof(double) * m,
                       2 * sizeof(unsigned long long)});
}

struct F64Bufs {
    double *avg, *var, *prof;
    double2 *tw, *qs;
    double* q;
    unsigned long long* r;
};

// Uploads the query and enqueues the profile (into out, or the workspace's).
int enqueue_ed_f64(mdtw_engine* e, const double* d_series, int64_t n, const double* h_zquery,
                   int32_t m, char* ws, double* out, F64Bufs* b) {
    const int64_t L = n - m + 1;
    const int64_t h = int64_t(1) << (fft_log(m) - 1);
    Carve cv{ws};
    b->avg = cv.take<double>(L);
    b->var = cv.take<double>(L);
    b->prof = cv.take<double>(L);
    b->tw = cv.take<double2>(h);
    b->qs = cv.take<double2>(h + 1);
    b->q = cv.take<double>(m);
    b->r = cv.take<unsigned long long>(2);
    if (out) b->prof = out;
    double q2 = 0.0, qsum = 0.0;
    for (int32_t i = 0; i < m; ++i) {
        q2 += h_zquery[i] * h_zquery[i];
        qsum += h_zquery[i];
    }
    int rc = upload(e, b->q, h_zquery, sizeof(double) * m);
    if (rc != MDTW_OK) return rc;
    return run_ed_fft<double, double>(e->stream, d_series, n, b->q, m, q2, qsum, b->avg, b->var,
                                      b->tw, b->qs, b->prof);
}

}  // namespace

extern "C" {

int mdtw_avg_std_f64(mdtw_engine* e, const double* d_series, int64_t n, int32_t m,
                     double* d_avg, double* d_std) {
    if (!e || !d_series || !d_avg || !d_std || m < 1) return MDTW_ERR_INVALID_ARGUMENT;
    if (m > MDTW_MAX_QUERY) return MDTW_ERR_QUERY_TOO_LONG;
    if (n < m) return MDTW_ERR_STREAM_TOO_SHORT;
    DeviceGuard guard(e->device);
    return launch_stats_d(e->stream, d_series, n, m, d_avg, d_std, true);
}

int mdtw_ed_f64(mdtw_engine* e, const double* d_series, int64_t n, const double* h_zquery,
                int32_t m, double* d_out) {
    if (!e || !d_series || !d_out) return MDTW_ERR_INVALID_ARGUMENT;
    int rc = check_f64_query(h_zquery, m);
    if (rc != MDTW_OK) return rc;
    if (n < m) return MDTW_ERR_STREAM_TOO_SHORT;
    DeviceGuard guard(e->device);
    void* ws = nullptr;
    rc = ws_reserve(e, ed_f64_ws_bytes(n - m + 1, m), &ws);
    if (rc != MDTW_OK) return rc;
    F64Bufs b;
    return enqueue_ed_f64(e, d_series, n, h_zquery, m, static_cast<char*>(ws), d_out, &b);
}

int mdtw_ed_search_f64(mdtw_engine* e, const double* d_series, int64_t n,
                       const double* h_zquery, int32_t m, int64_t* h_pos, double* h_dist,
                       mdtw_search_info* info) {
    if (!e || !d_series || !h_pos || !h_dist) return MDTW_ERR_INVALID_ARGUMENT;
    int rc = check_f64_query(h_zquery, m);
    if (rc != MDTW_OK) return rc;
    if (n < m) return MDTW_ERR_STREAM_TOO_SHORT;
    const int64_t L = n - m + 1;
    if (L > kMaxSearchWindows) return MDTW_ERR_UNSUPPORTED;
    *h_pos = -1;
    *h_dist = INFINITY;
    DeviceGuard guard(e->device);
    void* ws = nullptr;
    rc = ws_reserve(e, ed_f64_ws_bytes(L, m), &ws);
    if (rc != MDTW_OK) return rc;
    hipStream_t st = e->stream;
    F64Bufs b;
    rc = hip_ok(hipEventRecord(e->ev0, st));
    if (rc == MDTW_OK)
        rc = enqueue_ed_f64(e, d_series, n, h_zquery, m, static_cast<char*>(ws), nullptr, &b);
    const unsigned nbl = static_cast<unsigned>(
        std::min<int64_t>((L + 255) / 256, static_cast<int64_t>(e->cus) * 16));
    if (rc == MDTW_OK) {
        hipLaunchKernelGGL(f64_min_reset_kernel, dim3(1), dim3(64), 0, st, b.r);
        hipLaunchKernelGGL(f64_min_bits_kernel, dim3(nbl), dim3(256), 0, st, b.prof, L, b.r);
        hipLaunchKernelGGL(f64_min_index_kernel, dim3(nbl), dim3(256), 0, st, b.prof, L, b.r);
        rc = hip_ok(hipGetLastError());
    }
    unsigned long long r[2] = {~0ull, ~0ull};
    if (rc == MDTW_OK) rc = hip_ok(hipEventRecord(e->ev1, st));
    if (rc == MDTW_OK)
        rc = hip_ok(hipMemcpyAsync(r, b.r, sizeof(r), hipMemcpyDeviceToHost, st));
    if (rc == MDTW_OK) rc = hip_ok(hipStreamSynchronize(st));
    if (rc != MDTW_OK) return rc;
    double best = INFINITY;
    if (r[0] != ~0ull && r[1] != ~0ull) {
        std::memcpy(&best, &r[0], sizeof(best));
        *h_pos = static_cast<int64_t>(r[1]);
        *h_dist = std::sqrt(best);
    }
    if (info) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->ev0, e->ev1);
        info->windows = L;
        info->seeded = 0;
        info->lb_kim_pruned = 0;
        info->lb_keogh_pruned = 0;
        info->dp_evaluated = L;
        info->dp_abandoned = 0;
        info->found = *h_pos >= 0 ? 1 : 0;
        info->threshold = static_cast<float>(best);
        info->elapsed_ms = ms;
    }
    return MDTW_OK;
}

}  // extern "C"
