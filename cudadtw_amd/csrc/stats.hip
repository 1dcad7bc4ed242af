// stats.hip -- sliding-window mean and population sigma of every window.
//
// Two modes (mdtw_avg_std, replacing stats.cuh:115-142 avg_std<double>):
//
// MDTW_STATS_FP64 (default).  One pass, no global prefix.  A wave owns a
// tile of consecutive windows [w0, w0 + T).  With c = x[w0] as a tile-local
// shift, window r's sums are
//     S(r) = S0 + sum_{k<r} (x[w0+m+k] - x[w0+k])         (and the same for
//                                                         squares of x - c),
// S0 being the first window's sum, all in FP64 on values shifted by c, so
// cancellation is bounded by the tile's local range instead of the stream's
// absolute level (the reference's global FP64 prefix loses ~1e-2 of sigma on
// offset data; its FP32 finalize loses more, stats.cuh:40-52).
// HBM traffic: read (2T + m)/T floats per window (the second read of each
// sample mostly hits L2), write 8 bytes.  The whole-stream kernels sweep the
// tiles with as many waves as the device holds, each wave loading its next
// tile while it computes the current one (stats_sweep).
//
// Non-finite samples (NaN, +-Inf), FP64 and raw modes: a window holding one
// gets NaN for its mean and sigma, which makes every distance and bound of it
// NaN (unordered: never a match, never kept by a top-k bound).  The tile sums
// skip such samples (counted instead), so they never reach the windows
// around them; without that, one NaN would poison every later window of its
// tile through the running sums.
//
// MDTW_STATS_REFCOMPAT.  Restates the reference arithmetic: per 10^7-sample
// epoch with M-1 overlap (stats.cuh:120-141), a global inclusive FP64 prefix
// with P[0] = 0 written over the epoch's first sample, which the in-place
// scan then reads as 0 (cub_util.cuh:29-54), (P[i+M] - P[i]) / M in FP64 stored as
// FP32 (stats.cuh:30-38), sigma = sqrt(fmaf(-X, X, X2)) or 1 (stats.cuh:40-52,
// nvcc's default contraction of X2 - X*X).
#include <algorithm>
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

// ---- MDTW_STATS_FP64: wave-local tiles -------------------------------------
// A wave owns kStatsTW consecutive windows [w0, w0 + kStatsTW) and needs no
// barrier: the first window's sums are a wave reduction (coalesced loads),
// x[w0 + i] and x[w0 + m + i] for the tile are staged through LDS with
// coalesced loads, lane l takes the kStatsR consecutive windows l*R .. l*R+R-1,
// the transitions before them are a wave exclusive scan, and the results go
// back through LDS for coalesced stores.  LDS index i is stored at
// i + i/R, so lane l's consecutive elements ((R+1)l + v) fall in distinct
// banks.
constexpr int kStatsR = 16;
constexpr int kStatsTW = kWave * kStatsR;
constexpr int kStatsWaves = 4;
constexpr int kStatsLds = kStatsTW + kStatsTW / kStatsR;

__device__ __forceinline__ int stats_skew(int i) { return i + i / kStatsR; }
// stats_skew(v * kWave + lane) written as one lane base plus a constant per
// v, so unrolled accesses share one address register (immediate offsets).
static_assert(kWave % kStatsR == 0, "a wave row must hold whole skew groups");
__device__ __forceinline__ int stats_slot(int v, int lane) {
    return v * (kWave + kWave / kStatsR) + lane + lane / kStatsR;
}

// Wave-wide inclusive scan of a double through DPP (GFX9 family): row_shr
// 1/2/4/8 inside 16-lane rows, then row_bcast:15 and row_bcast:31 carry row
// totals upward.  Register-only; lanes with no DPP source add 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROWS, 0xf, false);
    const int hi =
        __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double wave_incl_scan_f64(double v) {
    v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
    v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
    v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
    v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
    v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
    const double incl = wave_incl_scan_f64(v);
    const long long b = __double_as_longlong(incl);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), kWave - 1);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), kWave - 1);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
}

__device__ __forceinline__ double wave_excl_scan_f64(double v) {
    return wave_incl_scan_f64(v) - v;
}

__device__ __forceinline__ void stats_wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads of one tile's samples into registers: x[w0 + i] (leaving sample of
// window i) and x[w0 + m + i] (entering), i = v * 64 + lane, clamped to the
// last stream sample.  Issued together; nothing waits on them until
// stats_stage, so a sweep can have the next tile's loads in flight while it
// computes the current one.
template <class S>
__device__ __forceinline__ void stats_load(const S* __restrict__ s, int64_t L, int32_t m,
                                           int64_t w0, S (&va)[kStatsR], S (&vb)[kStatsR]) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t last = L - 1 + m - 1;  // last stream sample
    // Two wave-uniform bases and one lane offset, so the loads hold no
    // per-load 64-bit address.
    const S* __restrict__ pa = s + w0;
    const S* __restrict__ pb = pa + m;
    if (w0 + m + kStatsTW - 1 <= last) {  // whole tile inside the stream
        const S* __restrict__ qa = pa + lane;
        const S* __restrict__ qb = pb + lane;
#pragma unroll
        for (int v = 0; v < kStatsR; ++v) {  // immediate offsets from qa / qb
            va[v] = qa[v * kWave];
            vb[v] = qb[v * kWave];
        }
        return;
    }
    // The stream's last tile: index min(v * 64 + lane, last - w0), written
    // as v * 64 + min(lane, b) with the wave-uniform b = last - w0 - v * 64
    // (last - w0 < m + kStatsTW fits an int; a negative b gives last - w0).
    const int la = static_cast<int>(last - w0);
#pragma unroll
    for (int v = 0; v < kStatsR; ++v) {
        const int ca = la - v * kWave, cb = la - m - v * kWave;
        va[v] = pa[v * kWave + (lane < ca ? lane : ca)];
        vb[v] = pb[v * kWave + (lane < cb ? lane : cb)];
    }
}

// Stages the loaded samples in LDS; returns (wave-uniform) whether all of
// them are finite.
template <class T, class S>
__device__ __forceinline__ bool stats_stage(const S (&va)[kStatsR], const S (&vb)[kStatsR], T* xa,
                                            T* xb) {
    const int lane = threadIdx.x & (kWave - 1);
    int fin = 1;
#pragma unroll
    for (int v = 0; v < kStatsR; ++v) {
        xa[stats_slot(v, lane)] = va[v];
        xb[stats_slot(v, lane)] = vb[v];
        // bitwise, not short-circuit: no branch per sample
        fin &= static_cast<int>(isfinite(va[v])) & static_cast<int>(isfinite(vb[v]));
    }
    stats_wave_fence();
    return __all(fin);
}

// The windows of a staged tile from the first window's sums (S1, S2, B1 of
// x - c, F1 non-finite samples): pass 1 totals each lane's transitions, a
// wave scan offsets them, pass 2 walks the lane's windows.  CHECK keeps the
// non-finite bookkeeping (needed only when the tile holds such a sample).
// FULL (nw == kStatsTW, every tile but a stream's last): every lane has
// kStatsR windows, so the loops run without a per-lane bound.  Either way a
// lane's totals hold exactly its transitions r -> r + 1 with r + 1 < nw: the
// exclusive scan is the inclusive scan minus the lane's own total, so that
// total enters the rounding of the lane's own windows.
template <class T, class S, bool RAW, bool CHECK, bool FULL>
__device__ __forceinline__ void stats_windows(int32_t m, int nw, double c, double S1, double S2,
                                              double B1, double F1, T* xa, T* xb) {
    const int lane = threadIdx.x & (kWave - 1);
    // Pass 1: this lane's transition totals (window r -> r + 1).
    const int r0 = lane * kStatsR;
    // A non-finite sample enters the sums as 0 and the count as 1.
    auto shifted = [c](T v, double& fcount) {
        if constexpr (!CHECK) return static_cast<double>(v) - c;
        const bool fin = isfinite(v);
        fcount += fin ? 0.0 : 1.0;
        return fin ? static_cast<double>(v) - c : 0.0;
    };
    double t1 = 0.0, t2 = 0.0, u1 = 0.0, u2 = 0.0, tf = 0.0;
    // r0 + v < nw as v < nv: one per-lane bound, compared with constants
    const int nv = nw - r0;
#pragma unroll 4
    for (int v = 0; v < kStatsR; ++v) {
        const int r = r0 + v;
        // transition r -> r + 1 exists when r + 1 < nw
        if (FULL ? (v + 1 < kStatsR || lane + 1 < kWave) : v + 1 < nv) {
            double fo = 0.0, fn = 0.0;
            const double xo = shifted(xa[stats_skew(r)], fo);
            const double xn = shifted(xb[stats_skew(r)], fn);
            t1 += xn - xo;
            t2 += xn * xn - xo * xo;
            u1 += fabs(xn) + fabs(xo);
            u2 += xn * xn + xo * xo;
            if constexpr (CHECK) tf += fn - fo;
        }
    }
    double p1 = wave_excl_scan_f64(t1);
    double p2 = wave_excl_scan_f64(t2);
    double q1 = wave_excl_scan_f64(u1);
    double q2 = wave_excl_scan_f64(u2);
    double pf = CHECK ? wave_excl_scan_f64(tf) : 0.0;

    // Pass 2: the lane's windows; results overwrite their own LDS slots
    // after the slot's samples are read.
    const double inv_m = 1.0 / static_cast<double>(m);
#pragma unroll 4
    for (int v = 0; v < kStatsR; ++v) {
        const int r = r0 + v;
        if (!FULL && v >= nv) break;
        const double sum1 = S1 + p1;
        const double sum2 = S2 + p2;
        const double mean = sum1 * inv_m;
        double var = fma(-mean, mean, sum2 * inv_m);
        // Rounding-level variance of an exactly constant window is noise:
        // treat it as zero (sigma := 1, the reference's guard).
        const double mag2 = (S2 + q2) * inv_m;
        const double mag1 = (B1 + q1) * inv_m;
        const double tol = 64.0 * DBL_EPSILON * (mag2 + 2.0 * fabs(mean) * mag1);
        if (var <= tol) var = 0.0;
        const int k = stats_skew(r);
        // the window holds a non-finite sample
        const bool bad = CHECK && F1 + pf != 0.0;
        double fo = 0.0, fn = 0.0;
        const double xo = shifted(xa[k], fo);
        const double xn = shifted(xb[k], fn);
        const T nan = static_cast<T>(__builtin_nan(""));
        if constexpr (RAW) {
            xa[k] = bad ? nan : static_cast<T>(0);
            xb[k] = bad ? nan : static_cast<T>(1);
        } else if constexpr (sizeof(T) == sizeof(double)) {
            xa[k] = bad ? nan : c + mean;
            xb[k] = bad ? nan : var;
        } else {
            xa[k] = bad ? nan : static_cast<float>(c + mean);
            // FP32 square root of the FP64 variance rounded to FP32: within
            // one ulp of the FP64 root's rounding, at a fraction of its cost.
            xb[k] = bad ? nan : (var > 0.0 ? sqrtf(static_cast<float>(var)) : 1.0f);
        }
        p1 += xn - xo;
        p2 += xn * xn - xo * xo;
        q1 += fabs(xn) + fabs(xo);
        q2 += xn * xn + xo * xo;
        if constexpr (CHECK) pf += fn - fo;
    }
}

// One tile [w0, w0 + nw) of windows by one wave (w0 a multiple of kStatsTW),
// its samples already staged by stats_stage: leaves mean and sigma of window
// w0 + i in xa / xb at stats_skew(i).  Every kernel below runs exactly this,
// so a window's statistics do not depend on which of them computed it.
// T = double (the FFT ED's statistics): the same sums, leaving the FP64 mean
// and the FP64 variance (0 for a constant window) instead.  S is the sample
// type (FP32 streams, or FP64 ones for the FP64 ED; T must hold S exactly).
// RAW (MDTW_STATS_NONE): mean 0 and sigma 1, which make (x - mean) / sigma
// == x exactly in every kernel -- only the non-finite rule is computed.
// A window holding a non-finite sample gets NaN mean and sigma (variance).
// Global memory is read only for a first window longer than the tile
// (m > kStatsTW).
template <class T = float, class S = float, bool RAW = false>
__device__ __forceinline__ int stats_tile_staged(const S* __restrict__ s, int64_t L, int32_t m,
                                                 int64_t w0, T* xa, T* xb, bool staged_finite) {
    static_assert(sizeof(T) >= sizeof(S), "the LDS tile must hold samples exactly");
    const int lane = threadIdx.x & (kWave - 1);
    const int nw = L - w0 < kStatsTW ? static_cast<int>(L - w0) : kStatsTW;
    const S* x = s + w0;
    const S x0 = static_cast<S>(xa[stats_skew(0)]);  // x[w0]
    const double c = isfinite(x0) ? static_cast<double>(x0) : 0.0;

    // First window of the tile: x[w0 .. w0 + m), from the staged tile where
    // it lies inside it (x[w0 + k] = xa[k] for k < kStatsTW).
    // When the staged samples are all finite, the sums skip the non-finite
    // test for them and count only the samples read from memory (m >
    // kStatsTW); one non-finite sample there sends the window through the
    // checked sums instead.  Finite samples give the same sums either way.
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, f1 = 0.0;
    auto add = [&](S xk) {
        const double d = static_cast<double>(xk) - c;
        a1 += d;
        b1 += fabs(d);
        a2 = fma(d, d, a2);
    };
    const int mt = m < kStatsTW ? m : kStatsTW;
    bool clean = false;
    if (staged_finite) {
        int fin = 1;
        int k = lane;
        for (; k < mt; k += kWave) add(static_cast<S>(xa[stats_skew(k)]));
        // Samples past the tile come from memory: sixteen loads in flight at
        // a time (one at a time, each waited for, cost a memory latency per
        // sample), added in the same order.
        for (; k + 15 * kWave < m; k += 16 * kWave) {
            S xk[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) xk[u] = x[k + u * kWave];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                fin &= static_cast<int>(isfinite(xk[u]));
                add(xk[u]);
            }
        }
        for (; k < m; k += kWave) {
            const S xk = x[k];
            fin &= static_cast<int>(isfinite(xk));
            add(xk);
        }
        clean = __all(fin);
    }
    if (!clean) {
        a1 = a2 = b1 = 0.0;
        auto add_checked = [&](S xk) {
            const bool fin = isfinite(xk);
            add(fin ? xk : static_cast<S>(c));  // contributes 0
            f1 += fin ? 0.0 : 1.0;
        };
        int k = lane;
        for (; k < mt; k += kWave) add_checked(static_cast<S>(xa[stats_skew(k)]));
        for (; k < m; k += kWave) add_checked(x[k]);
    }
    const double S1 = wave_sum_f64(a1);
    const double S2 = wave_sum_f64(a2);
    const double B1 = wave_sum_f64(b1);
    // non-finite samples of the first window
    const double F1 = clean ? 0.0 : wave_sum_f64(f1);

    // Every sample a window of the tile reads is in the first window or
    // among the staged ones: when all of those are finite (the usual case),
    // the transitions skip the non-finite bookkeeping -- the same sums
    // otherwise, so the same results.
    const bool fin = staged_finite && F1 == 0.0;
    if (nw == kStatsTW) {
        if (fin)
            stats_windows<T, S, RAW, false, true>(m, nw, c, S1, S2, B1, F1, xa, xb);
        else
            stats_windows<T, S, RAW, true, true>(m, nw, c, S1, S2, B1, F1, xa, xb);
    } else {
        if (fin)
            stats_windows<T, S, RAW, false, false>(m, nw, c, S1, S2, B1, F1, xa, xb);
        else
            stats_windows<T, S, RAW, true, false>(m, nw, c, S1, S2, B1, F1, xa, xb);
    }
    stats_wave_fence();
    return nw;
}

// Load, stage and compute one tile (the window-list kernel's form).
template <class T = float, class S = float, bool RAW = false>
__device__ __forceinline__ int stats_tile(const S* __restrict__ s, int64_t L, int32_t m, int64_t w0,
                                          T* xa, T* xb) {
    S va[kStatsR], vb[kStatsR];
    stats_load(s, L, m, w0, va, vb);
    const bool fin = stats_stage(va, vb, xa, xb);
    return stats_tile_staged<T, S, RAW>(s, L, m, w0, xa, xb, fin);
}


// Whole-stream sweep: wave t of the grid takes tiles t, t + stride, ...
// (stride = the grid's waves, sized to what the device holds at once) and
// issues the loads of its next tile before computing the current one, so
// the tile's FP64 work overlaps the next tile's HBM reads instead of
// alternating with them.  emit(w0, nw) writes the tile's results out of
// xa / xb.
template <class T, class S, bool RAW, class Emit>
__device__ __forceinline__ void stats_sweep(const S* __restrict__ s, int64_t L, int32_t m, T* xa,
                                            T* xb, Emit emit) {
    const int64_t ntiles = (L + kStatsTW - 1) / kStatsTW;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kStatsWaves;
    // wave-uniform, so the tile bases and the loop stay in scalar registers
    const int wid = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x / kWave));
    int64_t t = static_cast<int64_t>(blockIdx.x) * kStatsWaves + wid;
    if (t >= ntiles) return;
    S va[kStatsR], vb[kStatsR];
    stats_load(s, L, m, t * kStatsTW, va, vb);
    for (;;) {
        const int64_t w0 = t * kStatsTW;
        const bool fin = stats_stage(va, vb, xa, xb);
        const int64_t tn = t + stride;
        if (tn < ntiles) stats_load(s, L, m, tn * kStatsTW, va, vb);
        const int nw = stats_tile_staged<T, S, RAW>(s, L, m, w0, xa, xb, fin);
        emit(w0, nw);
        if (tn >= ntiles) break;
        t = tn;
    }
}

__global__ __launch_bounds__(kStatsWaves * kWave) void stats_fp64_kernel(
    const float* __restrict__ s, int64_t L, int32_t m, float* __restrict__ avg,
    float* __restrict__ sd) {
    __shared__ float xa_all[kStatsWaves][kStatsLds];
    __shared__ float xb_all[kStatsWaves][kStatsLds];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    float* xa = xa_all[wid];
    float* xb = xb_all[wid];
    stats_sweep<float, float, false>(s, L, m, xa, xb, [&](int64_t w0, int nw) {
        float* pa = avg + w0 + lane;
        float* pb = sd + w0 + lane;
#pragma unroll 4
        for (int v = 0; v < kStatsR; ++v) {
            if (v * kWave + lane < nw) {  // streamed out: not re-read by this kernel
                __builtin_nontemporal_store(xa[stats_slot(v, lane)], pa + v * kWave);
                __builtin_nontemporal_store(xb[stats_slot(v, lane)], pb + v * kWave);
            }
        }
    });
}

// FP64 mean and variance of every window (the FFT ED's statistics).
__global__ __launch_bounds__(kStatsWaves * kWave) void stats_f64out_kernel(
    const float* __restrict__ s, int64_t L, int32_t m, double* __restrict__ avg,
    double* __restrict__ var) {
    __shared__ double xa_all[kStatsWaves][kStatsLds];
    __shared__ double xb_all[kStatsWaves][kStatsLds];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    double* xa = xa_all[wid];
    double* xb = xb_all[wid];
    stats_sweep<double, float, false>(s, L, m, xa, xb, [&](int64_t w0, int nw) {
        double* pa = avg + w0 + lane;
        double* pb = var + w0 + lane;
#pragma unroll 4
        for (int v = 0; v < kStatsR; ++v) {
            if (v * kWave + lane < nw) {
                __builtin_nontemporal_store(xa[stats_slot(v, lane)], pa + v * kWave);
                __builtin_nontemporal_store(xb[stats_slot(v, lane)], pb + v * kWave);
            }
        }
    });
}

// MDTW_STATS_NONE: mean 0 and sigma 1 for every window (NaN for a window
// holding a non-finite sample).
__global__ __launch_bounds__(kStatsWaves * kWave) void stats_none_kernel(
    const float* __restrict__ s, int64_t L, int32_t m, float* __restrict__ avg,
    float* __restrict__ sd) {
    __shared__ float xa_all[kStatsWaves][kStatsLds];
    __shared__ float xb_all[kStatsWaves][kStatsLds];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    float* xa = xa_all[wid];
    float* xb = xb_all[wid];
    stats_sweep<float, float, true>(s, L, m, xa, xb, [&](int64_t w0, int nw) {
        float* pa = avg + w0 + lane;
        float* pb = sd + w0 + lane;
#pragma unroll 4
        for (int v = 0; v < kStatsR; ++v) {
            if (v * kWave + lane < nw) {
                __builtin_nontemporal_store(xa[stats_slot(v, lane)], pa + v * kWave);
                __builtin_nontemporal_store(xb[stats_slot(v, lane)], pb + v * kWave);
            }
        }
    });
}

// FP64 stream (the reference's avg_std<double> for its ED driver): FP64 mean
// and population sigma of every window, sigma := 1 for a constant window
// (SIGMA), or the variance, 0 for a constant window (the FFT ED's input).
template <bool SIGMA>
__global__ __launch_bounds__(kStatsWaves * kWave) void stats_d_kernel(
    const double* __restrict__ s, int64_t L, int32_t m, double* __restrict__ avg,
    double* __restrict__ sd) {
    __shared__ double xa_all[kStatsWaves][kStatsLds];
    __shared__ double xb_all[kStatsWaves][kStatsLds];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    double* xa = xa_all[wid];
    double* xb = xb_all[wid];
    stats_sweep<double, double, false>(s, L, m, xa, xb, [&](int64_t w0, int nw) {
        double* pa = avg + w0 + lane;
        double* pb = sd + w0 + lane;
#pragma unroll 4
        for (int v = 0; v < kStatsR; ++v) {
            if (v * kWave + lane < nw) {
                const double var = xb[stats_slot(v, lane)];
                // NaN (a non-finite sample) stays NaN
                const double sg =
                    !SIGMA ? var : (var > 0.0 ? sqrt(var) : (var == 0.0 ? 1.0 : var));
                __builtin_nontemporal_store(xa[stats_slot(v, lane)], pa + v * kWave);
                __builtin_nontemporal_store(sg, pb + v * kWave);
            }
        }
    });
}

// Statistics of listed windows: entry p is window pos[p] of length len[p]
// over the stream s[0..n); one wave computes that window's tile.
__global__ __launch_bounds__(kWave) void stats_at_kernel(const float* __restrict__ s, int64_t n,
                                                         const int32_t* __restrict__ len,
                                                         const int64_t* __restrict__ pos,
                                                         int64_t count, float* __restrict__ avg,
                                                         float* __restrict__ sd) {
    __shared__ float xa[kStatsLds];
    __shared__ float xb[kStatsLds];
    const int64_t p = blockIdx.x;
    if (p >= count) return;
    const int32_t m = len[p];
    const int64_t w = pos[p];
    const int64_t L = n - m + 1;
    const int64_t w0 = (w / kStatsTW) * kStatsTW;
    stats_tile(s, L, m, w0, xa, xb);
    if (threadIdx.x == 0) {
        avg[p] = xa[stats_skew(static_cast<int>(w - w0))];
        sd[p] = xb[stats_skew(static_cast<int>(w - w0))];
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
        if (k > 0 && k < len) {  // sample 0: see ref_prefix
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
        // The reference scans in place (prefix_sum(Aprefix, Aprefix, N),
        // stats.cuh:100-101): the 0 it stores as P[0] (cub_util.cuh:35-36)
        // overwrites the epoch's first sample before the scan reads it, so
        // that sample counts as 0 -- only the epoch's first window changes.
        double d = k > 0 && k < len ? static_cast<double>(x[k]) : 0.0;
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

// Blocks of a whole-stream sweep: as many as the device holds at once for
// the kernel (at least one), fewer when the stream has fewer tiles.
template <class K>
unsigned sweep_blocks(K kernel, int64_t L) {
    constexpr int64_t per_block = static_cast<int64_t>(kStatsTW) * kStatsWaves;
    const int64_t nb = (L + per_block - 1) / per_block;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kStatsWaves * kWave, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return static_cast<unsigned>(std::max<int64_t>(
        1, std::min<int64_t>(nb, static_cast<int64_t>(cus) * per_cu)));
}

}  // namespace

int launch_stats_at(hipStream_t st, const float* s, int64_t n, const int32_t* len,
                    const int64_t* pos, int64_t count, float* avg, float* sd) {
    if (count <= 0) return MDTW_OK;
    hipLaunchKernelGGL(stats_at_kernel, dim3(static_cast<unsigned>(count)), dim3(kWave), 0, st,
                       s, n, len, pos, count, avg, sd);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int launch_stats_fp64(hipStream_t st, const float* s, int64_t n, int32_t m, float* avg,
                      float* sd) {
    const int64_t L = n - m + 1;
    if (L <= 0) return MDTW_ERR_STREAM_TOO_SHORT;
    hipLaunchKernelGGL(stats_fp64_kernel, dim3(sweep_blocks(stats_fp64_kernel, L)),
                       dim3(kStatsWaves * kWave), 0, st, s, L, m, avg, sd);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int launch_stats_f64out(hipStream_t st, const float* s, int64_t n, int32_t m, double* avg,
                        double* var) {
    const int64_t L = n - m + 1;
    if (L <= 0) return MDTW_ERR_STREAM_TOO_SHORT;
    hipLaunchKernelGGL(stats_f64out_kernel, dim3(sweep_blocks(stats_f64out_kernel, L)),
                       dim3(kStatsWaves * kWave), 0, st, s, L, m, avg, var);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int launch_stats_none(hipStream_t st, const float* s, int64_t n, int32_t m, float* avg,
                      float* sd) {
    const int64_t L = n - m + 1;
    if (L <= 0) return MDTW_ERR_STREAM_TOO_SHORT;
    hipLaunchKernelGGL(stats_none_kernel, dim3(sweep_blocks(stats_none_kernel, L)),
                       dim3(kStatsWaves * kWave), 0, st, s, L, m, avg, sd);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int launch_stats_d(hipStream_t st, const double* s, int64_t n, int32_t m, double* avg,
                   double* sd, bool sigma) {
    const int64_t L = n - m + 1;
    if (L <= 0) return MDTW_ERR_STREAM_TOO_SHORT;
    if (sigma)
        hipLaunchKernelGGL(stats_d_kernel<true>, dim3(sweep_blocks(stats_d_kernel<true>, L)),
                           dim3(kStatsWaves * kWave), 0, st, s, L, m, avg, sd);
    else
        hipLaunchKernelGGL(stats_d_kernel<false>, dim3(sweep_blocks(stats_d_kernel<false>, L)),
                           dim3(kStatsWaves * kWave), 0, st, s, L, m, avg, sd);
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
