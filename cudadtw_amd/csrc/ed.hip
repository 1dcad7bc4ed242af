// ed.hip -- z-normalised Euclidean distance of every window.
//
// The reference computes ED^2 = 2M - 2 * corr / sigma through an FP64 cuFFT
// correlation of the stream with the query (ed.cuh:71-153), FP64 statistics
// and a full radix sort for the argmin.  Here the correlation is summed
// directly, also in FP64, with the square terms kept so that the result is
// the direct form's value for the window statistics given:
//
//   ED^2 = sum_k (q_k - (x_{w+k} - mu) / sigma)^2
//        = Sqq - 2 (A - u Sq) / sigma + (S2 - 2 u S1 + M u^2) / sigma^2,
//
// with the query centred on the host (qc_k = q_k - qbar, Sq = sum qc_k,
// Sqq = sum qc_k^2, all FP64), the samples shifted by a constant c of their
// group (y = x - c, below), A = sum_k qc_k y_{w+k}, S1 = sum y, S2 = sum y^2
// over the window, and u = mu - c + qbar sigma.  Centring both sides keeps
// every term of the size of the distance itself -- no cancellation for near
// matches, also for raw windows (mu = 0, sigma = 1) far from zero -- so the
// FP64 value is the exact direct-form distance to ~1e-13 relative before it
// is rounded to FP32 (the FP32 direct form this replaces was within ~1e-6).
//
// Groups.  Windows 16g .. 16g + 15 (index relative to the stream pointer)
// form group g and share c = x[16g + min(16, M) - 1], a sample inside every
// one of them, so |y| is bounded by the group's local range.  S1 and S2 are
// summed over window 16g (k ascending) and slid to the others one window at
// a time; A is a per-window sum in k.  Every kernel that evaluates a window
// follows this sequence exactly, so a window's distance is the same bits
// whichever kernel computes it and whatever range a call covers.
//
// Kernels:
//  - ed_tile_kernel: contiguous windows, one group per lane (16 windows),
//    256 groups per workgroup; the samples of a 256-position query chunk are
//    staged in LDS as FP64 with a pad double every 16 (lane strides of 17
//    doubles: conflict-free), the query chunk as centred FP64 (a broadcast
//    read).  Per query position a lane issues one LDS read, one shift, 16
//    FMAs (its windows) and 2 operations for S1/S2: 19 FP64 operations per
//    16 window-elements.  The FP32 direct form it replaces issued 3 FP32
//    operations per element (sub, fma with the reciprocal, fma); measured on
//    this device a wave64 FP64 FMA issues in ~4.6 cycles and an FP32 one in
//    ~4.3, so the correlation form does ~2.5x less issue work per element.
//    The lane's statistics and outputs move as 16-byte vectors.
//  - ed_lane_kernel: one window per lane, any window list, the same
//    arithmetic from global loads (search seeds, strided windows).  With
//    SCREEN (search rounds with queries of kEdScreenMinM or more, where
//    abandoning pays, and only over the search's own FP64 or raw statistics):
//    an FP32 direct partial sum first, abandoned per wave once every lane is
//    above prune_limit(threshold) plus an absolute 1e-13 M -- the rounding
//    margin of that sum, which needs sum z^2 = M over the window (true for
//    those statistics, not for arbitrary ones) -- and the exact FP64
//    evaluation only for the windows that remain (the UCR_ED early abandon,
//    UCR_ED_BIN.cpp:55-65, applied per wave).
//
// Non-finite samples: a window holding one has a non-finite A and gets NaN
// (never a match); a clean window whose slid S1 / S2 picked one up from a
// neighbour of its group sums its own samples instead.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {

constexpr int kEdG = 16;                          // windows per group (per lane)
constexpr int kEdThreads = 256;                   // groups per workgroup
constexpr int kEdTileW = kEdG * kEdThreads;       // windows per workgroup
constexpr int kEdKC = 256;                        // query positions per LDS chunk
constexpr int kEdSpan = kEdKC + kEdTileW + kEdG - 1;  // samples staged per chunk
constexpr int kEdLds = kEdSpan + kEdSpan / 16 + 1;    // with one pad per 16
// Searches with queries at least this long screen with the FP32 partial sum
// (the early abandon), shorter ones evaluate every window in the tile kernel.
constexpr int32_t kEdScreenMinM = 512;

__device__ __forceinline__ int ed_pad(int i) { return i + (i >> 4); }

__device__ __forceinline__ uint32_t ed_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ double ed_y(const float* __restrict__ s, int64_t i, double c) {
    return static_cast<double>(s[i]) - c;
}

__device__ __forceinline__ double ed_shift(const float* __restrict__ s, int64_t w0, int32_t m) {
    return static_cast<double>(s[w0 + (m < kEdG ? m : kEdG) - 1]);
}

// S1 / S2 of the window starting at w, k ascending.
__device__ __forceinline__ void ed_sums(const float* __restrict__ s, int64_t w, int32_t m,
                                        double c, double& s1, double& s2) {
    s1 = 0.0;
    s2 = 0.0;
    for (int32_t k = 0; k < m; ++k) {
        const double y = ed_y(s, w + k, c);
        s1 += y;
        s2 = fma(y, y, s2);
    }
}

// One window slide: sample yo leaves, yi enters.
__device__ __forceinline__ void ed_slide(double yo, double yi, double& s1, double& s2) {
    s1 = s1 + (yi - yo);
    s2 = fma(yi, yi, fma(-yo, yo, s2));
}

// ED^2 from a window's correlation a and sums s1, s2 (all finite).  The
// reciprocal of sigma is the hardware estimate refined by two Newton steps
// (relative error far below what the FP32 result keeps).
__device__ __forceinline__ float ed_value(int32_t m, double c, double a, double s1, double s2,
                                          float mu, float sg, const EdQuerySums& eq) {
    const double sd = static_cast<double>(sg);
    const double u = fma(eq.mean, sd, static_cast<double>(mu) - c);
    const double num = fma(-u, eq.sum, a);
    const double den = fma(u, fma(static_cast<double>(m), u, -2.0 * s1), s2);
    const double r0 = __builtin_amdgcn_rcp(sd);
    double r = fma(fma(-sd, r0, 1.0), r0, r0);
    r = fma(fma(-sd, r, 1.0), r, r);
    const double e2 = fma(fma(den, r, -2.0 * num), r, eq.sumsq);
    const float v = static_cast<float>(e2);
    return v < 0.0f ? 0.0f : v;  // NaN stays NaN
}

__device__ __forceinline__ bool ed_finite(double s1, double s2) {
    return __builtin_isfinite(s1) && __builtin_isfinite(s2);
}

// The whole evaluation of window w from global memory (ed_lane_kernel, and
// the tile kernel's rare windows whose slid sums are not finite); the tile
// kernel's sequence, operation for operation: the group's first-window sums
// and this window's correlation each accumulate k ascending (three
// independent chains, one pass, loads issued 4 positions ahead), then the
// slides from the group's first window to this one.
__device__ float ed_window(const float* __restrict__ s, const float* __restrict__ q, int32_t m,
                           int64_t w, float mu, float sg, const EdQuerySums& eq) {
    const int64_t w0 = w & ~static_cast<int64_t>(kEdG - 1);
    const double c = ed_shift(s, w0, m);
    const float* __restrict__ xg = s + w0;
    const float* __restrict__ xw = s + w;
    double s1 = 0.0, s2 = 0.0, a = 0.0;
    int32_t k = 0;
    for (; k + 4 <= m; k += 4) {
        float fg[4], fw[4], fq[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            fg[u] = xg[k + u];
            fw[u] = xw[k + u];
            fq[u] = q[k + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double yg = static_cast<double>(fg[u]) - c;
            s1 += yg;
            s2 = fma(yg, yg, s2);
            a = fma(static_cast<double>(fq[u]) - eq.mean, static_cast<double>(fw[u]) - c, a);
        }
    }
    for (; k < m; ++k) {
        const double yg = ed_y(s, w0 + k, c);
        s1 += yg;
        s2 = fma(yg, yg, s2);
        a = fma(static_cast<double>(q[k]) - eq.mean, ed_y(s, w + k, c), a);
    }
    for (int64_t t = 1; t <= w - w0; ++t)
        ed_slide(ed_y(s, w0 + t - 1, c), ed_y(s, w0 + m + t - 1, c), s1, s2);
    if (!__builtin_isfinite(a)) return __builtin_nanf("");
    if (!ed_finite(s1, s2)) ed_sums(s, w, m, c, s1, s2);
    return ed_value(m, c, a, s1, s2, mu, sg, eq);
}

struct EdTileArgs {
    int64_t first, end;  // windows [first, end)
    int32_t m;
    int32_t mode;
    EdQuerySums eq;
};

// One query position u (0..15 within a block of 16) of the tile kernel:
// the new sample y[w0 + k + 15] enters ring slot (u + 15) % 16, window j
// takes slot (u + j) % 16, window 0's sample (slot u) goes into S1 / S2.
template <int U>
__device__ __forceinline__ void ed_step(const double* __restrict__ xb, const double* __restrict__ qb,
                                        double c, double (&y)[kEdG], double (&a)[kEdG],
                                        double& s1, double& s2) {
    constexpr int t = U + kEdG - 1;
    y[(U + kEdG - 1) % kEdG] = xb[t + (t >> 4)] - c;
    const double qk = qb[U];
#pragma unroll
    for (int j = 0; j < kEdG; ++j) a[j] = fma(qk, y[(U + j) % kEdG], a[j]);
    s1 += y[U];
    s2 = fma(y[U], y[U], s2);
}

template <int U>
__device__ __forceinline__ void ed_steps(const double* __restrict__ xb,
                                         const double* __restrict__ qb, double c,
                                         double (&y)[kEdG], double (&a)[kEdG], double& s1,
                                         double& s2, int rem) {
    if constexpr (U < kEdG) {
        if (U < rem) {
            ed_step<U>(xb, qb, c, y, a, s1, s2);
            ed_steps<U + 1>(xb, qb, c, y, a, s1, s2, rem);
        }
    }
}

template <int U>
__device__ __forceinline__ void ed_block(const double* __restrict__ xb,
                                         const double* __restrict__ qb, double c,
                                         double (&y)[kEdG], double (&a)[kEdG], double& s1,
                                         double& s2) {
    if constexpr (U < kEdG) {
        ed_step<U>(xb, qb, c, y, a, s1, s2);
        ed_block<U + 1>(xb, qb, c, y, a, s1, s2);
    }
}

template <int MODE>
__global__ __launch_bounds__(kEdThreads) void ed_tile_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, float* __restrict__ out, unsigned long long* __restrict__ res,
    Ctrl* __restrict__ ctrl, EdTileArgs ta) {
    __shared__ double xs[kEdLds];
    __shared__ double qs[kEdKC];
    const int tid = threadIdx.x;
    const int32_t m = ta.m;
    const int64_t tb = (ta.first / kEdG + static_cast<int64_t>(blockIdx.x) * kEdThreads) * kEdG;
    const int64_t w0 = tb + kEdG * tid;  // this lane's group
    // samples [0, nlim) are the ones the windows of [first, end) read
    const int64_t nlim = ta.end + m - 1;
    const bool lane_live = w0 < ta.end;
    const double c = lane_live ? ed_shift(s, w0, m) : 0.0;
    double a[kEdG], y[kEdG];
#pragma unroll
    for (int j = 0; j < kEdG; ++j) a[j] = 0.0;
    double s1 = 0.0, s2 = 0.0;
    const double* xl = xs + (kEdG + 1) * tid;  // ed_pad(16 tid + t) = 17 tid + ed_pad(t)
    for (int32_t k0 = 0; k0 < m; k0 += kEdKC) {
        const int kc = m - k0 < kEdKC ? m - k0 : kEdKC;
        const int span = kc + kEdTileW + kEdG - 1;
        __syncthreads();
        for (int i = tid; i < span; i += kEdThreads) {
            const int64_t p = tb + k0 + i;
            xs[ed_pad(i)] = p < nlim ? static_cast<double>(s[p]) : 0.0;
        }
        for (int i = tid; i < kc; i += kEdThreads)
            qs[i] = static_cast<double>(q[k0 + i]) - ta.eq.mean;
        __syncthreads();
        if (k0 == 0) {
#pragma unroll
            for (int t = 0; t < kEdG - 1; ++t) y[t] = xl[t] - c;
        }
        // query positions k0 + 16 b + u: LDS index 16 tid + 16 b + u + 15
        int b = 0;
        for (; 16 * b + kEdG <= kc; ++b)
            ed_block<0>(xl + (kEdG + 1) * b, qs + kEdG * b, c, y, a, s1, s2);
        if (16 * b < kc) ed_steps<0>(xl + (kEdG + 1) * b, qs + kEdG * b, c, y, a, s1, s2, kc - 16 * b);
    }
    if (!lane_live) {  // the search modes keep every lane for their wave-wide steps
        if constexpr (MODE == kModeProfile) return;
    }
    // Finish the lane's windows w0 + j, j <= jmax (its last one before end).
    // Entering samples y[w0 + m + j - 1] are in the last chunk's LDS, leaving
    // ones y[w0 + j - 1] in the first chunk's (still there when there was
    // one chunk); statistics and outputs move as 16-byte vectors when the
    // lane's 16 windows are all in range and aligned.
    const int64_t last = ta.end - 1 - w0;
    const int jmax = lane_live ? (last < kEdG - 1 ? static_cast<int>(last) : kEdG - 1) : -1;
    const bool full = lane_live && w0 >= ta.first && jmax == kEdG - 1;
    const int kcl = m - ((m - 1) / kEdKC) * kEdKC;  // the last chunk's length
    const bool one_chunk = m <= kEdKC;
    float mu[kEdG], sg[kEdG];
    if (full && ((reinterpret_cast<uintptr_t>(avg + w0) | reinterpret_cast<uintptr_t>(sd + w0)) & 15) == 0) {
#pragma unroll
        for (int v4 = 0; v4 < kEdG / 4; ++v4) {
            const float4 am = *reinterpret_cast<const float4*>(avg + w0 + 4 * v4);
            const float4 as = *reinterpret_cast<const float4*>(sd + w0 + 4 * v4);
            mu[4 * v4] = am.x; mu[4 * v4 + 1] = am.y; mu[4 * v4 + 2] = am.z; mu[4 * v4 + 3] = am.w;
            sg[4 * v4] = as.x; sg[4 * v4 + 1] = as.y; sg[4 * v4 + 2] = as.z; sg[4 * v4 + 3] = as.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kEdG; ++j) {
            const int64_t w = w0 + j;
            const bool valid = j <= jmax && w >= ta.first;
            mu[j] = valid ? avg[w] : 0.0f;
            sg[j] = valid ? sd[w] : 1.0f;
        }
    }
    float v[kEdG];
    uint32_t slow = 0;  // windows whose slid sums are not finite
#pragma unroll
    for (int j = 0; j < kEdG; ++j) {
        if (j >= 1 && j <= jmax) {
            const int ii = kcl + j - 1;  // LDS index of y[w0 + m + j - 1] is 16 tid + ii
            const double yi = xl[ii + (ii >> 4)] - c;
            const double yo = one_chunk ? xl[j - 1] - c : ed_y(s, w0 + j - 1, c);
            ed_slide(yo, yi, s1, s2);
        }
        if (!__builtin_isfinite(a[j])) {
            v[j] = __builtin_nanf("");
        } else if (!ed_finite(s1, s2)) {
            v[j] = __builtin_nanf("");
            slow |= 1u << j;
        } else {
            v[j] = ed_value(m, c, a[j], s1, s2, mu[j], sg[j], ta.eq);
        }
    }
    // rare: a window whose group neighbours hold a non-finite sample
    if (slow) {
        for (int j = 0; j < kEdG; ++j) {
            const int64_t w = w0 + j;
            if (((slow >> j) & 1u) && j <= jmax && w >= ta.first) {
                const float r = ed_window(s, q, m, w, avg[w], sd[w], ta.eq);
#pragma unroll
                for (int jj = 0; jj < kEdG; ++jj)
                    if (jj == j) v[jj] = r;
            }
        }
    }
    if constexpr (MODE == kModeProfile) {
        float* o = out + (w0 - ta.first);
        if (full && (reinterpret_cast<uintptr_t>(o) & 15) == 0) {
#pragma unroll
            for (int v4 = 0; v4 < kEdG / 4; ++v4)
                *reinterpret_cast<float4*>(o + 4 * v4) =
                    make_float4(v[4 * v4], v[4 * v4 + 1], v[4 * v4 + 2], v[4 * v4 + 3]);
        } else {
#pragma unroll
            for (int j = 0; j < kEdG; ++j)
                if (j <= jmax && w0 + j >= ta.first) o[j] = v[j];
        }
    } else if constexpr (MODE == kModeBest) {
        unsigned long long best = ~0ull;
#pragma unroll
        for (int j = 0; j < kEdG; ++j) {
            const int64_t w = w0 + j;
            if (j <= jmax && w >= ta.first) {
                const unsigned long long key = pack_key(v[j], static_cast<uint32_t>(w));
                best = key < best ? key : best;
            }
        }
        best = wave_min_u64(best);
        // keys only fall: skip the same-address atomic when it cannot win
        if ((tid & (kWave - 1)) == 0 && best != ~0ull &&
            best < __hip_atomic_load(&ctrl->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&ctrl->best, best);
    } else {
        const float thr = load_thr(ctrl);
#pragma unroll
        for (int j = 0; j < kEdG; ++j) {
            const int64_t w = w0 + j;
            const bool keep = j <= jmax && w >= ta.first && v[j] <= thr;
            if (keep) note_bucket(ctrl->bmin, ctrl->bucket_len, w, v[j]);
            const unsigned long long km = __ballot(keep);
            if (km) {
                const int lead = __builtin_ctzll(km);
                unsigned long long rb = 0;
                if ((tid & (kWave - 1)) == lead)
                    rb = atomicAdd(&ctrl->nres, static_cast<unsigned long long>(__popcll(km)));
                rb = __shfl(rb, lead, kWave);
                if (keep) res[rb + ed_rank(km)] = pack_key(v[j], static_cast<uint32_t>(w));
            }
        }
    }
}

struct EdLaneArgs {
    int32_t m;
    int32_t mode;
    int64_t count;
    int64_t base;
    int64_t stride;
    EdQuerySums eq;
};

template <bool SCREEN>
__global__ __launch_bounds__(256) void ed_lane_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, EdLaneArgs la) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    const int32_t m = la.m;
    for (int64_t base = wave * kWave; base < la.count; base += nwaves * kWave) {
        const int64_t g = base + lane;
        const bool valid = g < la.count;
        const int64_t gc = valid ? g : la.count - 1;
        const int64_t w = idx ? static_cast<int64_t>(idx[gc]) : la.base + gc * la.stride;
        const float mu = avg[w], sg = sd[w];
        float thr = kInf;
        if (la.mode == kModeBest || la.mode == kModeCollect)
            thr = la.mode == kModeBest ? load_best_dist(ctrl) : load_thr(ctrl);
        bool alive = valid;
        if constexpr (SCREEN) {
            // FP32 direct partial sums: abandon once every lane is above the
            // rounding-aware limit of the threshold
            const float lim = prune_limit(thr, m, true) + 1e-13f * static_cast<float>(m);
            const float* __restrict__ x = s + w;
            const float rs = 1.0f / sg;
            float acc = 0.0f;
            int32_t k = 0;
            bool gone = false;
            for (; k + 16 <= m; k += 16) {
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const float d = fmaf(-(x[k + u] - mu), rs, q[k + u]);
                    acc = fmaf(d, d, acc);
                }
                alive = alive && !(acc > lim);
                if (!__any(alive)) {
                    gone = true;
                    break;
                }
            }
            if (gone) continue;
            for (; k < m; ++k) {
                const float d = fmaf(-(x[k] - mu), rs, q[k]);
                acc = fmaf(d, d, acc);
            }
            alive = alive && !(acc > lim);
        }
        const float v = alive ? ed_window(s, q, m, w, mu, sg, la.eq) : kInf;
        if (SCREEN) {  // windows the screen left for the exact evaluation
            const unsigned long long am = __ballot(alive);
            if (lane == 0 && am) atomicAdd(&ctrl->dp, static_cast<unsigned long long>(__popcll(am)));
        }
        if (la.mode == kModeProfile || la.mode == kModeSeed) {
            if (valid) out[g] = v;
        }
        if (la.mode == kModeBest || la.mode == kModeSeed) {
            unsigned long long key = alive ? pack_key(v, static_cast<uint32_t>(w)) : ~0ull;
            key = wave_min_u64(key);
            if (lane == 0 && key != ~0ull) atomicMin(&ctrl->best, key);
            if (la.mode == kModeSeed && valid) note_bucket(ctrl->bmin, ctrl->bucket_len, w, v);
        } else if (la.mode == kModeCollect) {
            const bool keep = alive && v <= thr;
            if (keep) note_bucket(ctrl->bmin, ctrl->bucket_len, w, v);
            const unsigned long long km = __ballot(keep);
            if (km) {
                unsigned long long rb = 0;
                if (lane == 0)
                    rb = atomicAdd(&ctrl->nres, static_cast<unsigned long long>(__popcll(km)));
                rb = __shfl(rb, 0, kWave);
                if (keep) res[rb + ed_rank(km)] = pack_key(v, static_cast<uint32_t>(w));
            }
        }
    }
}

int ed_status() {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

}  // namespace

EdQuerySums ed_query_sums(const float* h_q, int32_t m) {
    EdQuerySums eq;
    double t = 0.0;
    for (int32_t k = 0; k < m; ++k) t += static_cast<double>(h_q[k]);
    eq.mean = m > 0 ? t / m : 0.0;
    for (int32_t k = 0; k < m; ++k) {
        const double qc = static_cast<double>(h_q[k]) - eq.mean;
        eq.sum += qc;
        eq.sumsq = std::fma(qc, qc, eq.sumsq);
    }
    return eq;
}

int launch_ed(hipStream_t st, int cus, const BandArgs& a) {
    if (a.count <= 0) return MDTW_OK;
    const bool contiguous = !a.idx && a.stride == 1;
    const bool screen = (a.mode == kModeBest || a.mode == kModeCollect) && contiguous &&
                        a.ed_screen && a.m >= kEdScreenMinM;
    const bool tile = contiguous && a.mode != kModeSeed && !screen;
    if (tile) {
        EdTileArgs ta{a.base, a.base + a.count, a.m, a.mode, a.ed};
        const int64_t g_first = a.base / kEdG;
        const int64_t g_end = (ta.end + kEdG - 1) / kEdG;
        const int64_t blocks = (g_end - g_first + kEdThreads - 1) / kEdThreads;
        const dim3 grid(static_cast<unsigned>(blocks)), block(kEdThreads);
        if (a.mode == kModeProfile)
            hipLaunchKernelGGL(ed_tile_kernel<kModeProfile>, grid, block, 0, st, a.s, a.avg, a.sd,
                               a.qx, a.out, a.res, a.ctrl, ta);
        else if (a.mode == kModeBest)
            hipLaunchKernelGGL(ed_tile_kernel<kModeBest>, grid, block, 0, st, a.s, a.avg, a.sd,
                               a.qx, a.out, a.res, a.ctrl, ta);
        else
            hipLaunchKernelGGL(ed_tile_kernel<kModeCollect>, grid, block, 0, st, a.s, a.avg, a.sd,
                               a.qx, a.out, a.res, a.ctrl, ta);
        return ed_status();
    }
    EdLaneArgs la{a.m, a.mode, a.count, a.base, a.stride, a.ed};
    const int64_t waves = (a.count + kWave - 1) / kWave;
    const int64_t blocks = std::min<int64_t>((waves + 3) / 4, static_cast<int64_t>(cus) * 64);
    if (screen)
        hipLaunchKernelGGL(ed_lane_kernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                           st, a.s, a.avg, a.sd, a.qx, a.idx, a.out, a.res, a.ctrl, la);
    else
        hipLaunchKernelGGL(ed_lane_kernel<false>, dim3(static_cast<unsigned>(blocks)), dim3(256),
                           0, st, a.s, a.avg, a.sd, a.qx, a.idx, a.out, a.res, a.ctrl, la);
    return ed_status();
}

}  // namespace mdtw
