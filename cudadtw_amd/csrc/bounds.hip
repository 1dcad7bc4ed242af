// bounds.hip -- LB_Kim and LB_Keogh lower bounds, and the fused search
// cascade that filters every window of the stream through both.
//
// Profile kernels (mdtw_lb_kim / mdtw_lb_keogh) z-normalise with an IEEE
// division and follow the operation order of bounds.cuh:111-144 and
// :283-407, so they are bit-exact with the CPU restatement.
//
// The cascade kernel is the search's HBM-streaming filter: one lane per
// window, 64 consecutive windows per wave (coalesced stream reads, the
// envelope and query read as wave-uniform scalars).  It z-normalises with a
// reciprocal (x - mu) * (1/sigma) and prunes against prune_limit(threshold),
// which absorbs that rounding; LB_Keogh abandons early when every lane of
// the wave is already above the limit.
#include "internal.hpp"

namespace mdtw {
namespace {

template <bool EXACT>
__device__ __forceinline__ float zn(float x, float mu, float sg, float rs) {
    if constexpr (EXACT)
        return (x - mu) / sg;
    else
        return (x - mu) * rs;
}

template <bool L2>
__device__ __forceinline__ float lp(float d) {
    if constexpr (L2)
        return d * d;
    else
        return fabsf(d);
}
template <bool L2>
__device__ __forceinline__ float lpa(float d, float acc) {
    if constexpr (L2)
        return fmaf(d, d, acc);
    else
        return fabsf(d) + acc;
}

// bounds.cuh:283-407: front 5x5 unconstrained DTW, minimum over its exit
// boundary, plus the back 5x5 block entered anywhere on its top row / left
// column.  Needs m >= 11 (bounds.cuh:418).
template <bool L2, bool EXACT>
__device__ __forceinline__ float lb_kim_window(const float* __restrict__ x,
                                               const float* __restrict__ q, int32_t m,
                                               float mu, float sg, float rs) {
    float q0, q1, q2, q3, q4, s0, s1, s2, s3, s4, mem;
    float p0, p1, p2, p3, p4, p5, p6, p7, p8, p9;
    q0 = q[0]; q1 = q[1]; q2 = q[2]; q3 = q[3]; q4 = q[4];
    s0 = zn<EXACT>(x[0], mu, sg, rs);
    s1 = zn<EXACT>(x[1], mu, sg, rs);
    s2 = zn<EXACT>(x[2], mu, sg, rs);
    s3 = zn<EXACT>(x[3], mu, sg, rs);
    s4 = zn<EXACT>(x[4], mu, sg, rs);

    p0 = lp<L2>(q0 - s0);
    p1 = lpa<L2>(q0 - s1, p0);
    p2 = lpa<L2>(q0 - s2, p1);
    p3 = lpa<L2>(q0 - s3, p2);
    p4 = lpa<L2>(q0 - s4, p3);
    mem = p4;

    p5 = lpa<L2>(q1 - s0, p0);
    p6 = lpa<L2>(q1 - s1, fminf(p5, fminf(p0, p1)));
    p7 = lpa<L2>(q1 - s2, fminf(p6, fminf(p1, p2)));
    p8 = lpa<L2>(q1 - s3, fminf(p7, fminf(p2, p3)));
    p9 = lpa<L2>(q1 - s4, fminf(p8, fminf(p3, p4)));
    mem = fminf(mem, p9);

    p0 = lpa<L2>(q2 - s0, p5);
    p1 = lpa<L2>(q2 - s1, fminf(p0, fminf(p5, p6)));
    p2 = lpa<L2>(q2 - s2, fminf(p1, fminf(p6, p7)));
    p3 = lpa<L2>(q2 - s3, fminf(p2, fminf(p7, p8)));
    p4 = lpa<L2>(q2 - s4, fminf(p3, fminf(p8, p9)));
    mem = fminf(mem, p4);

    p5 = lpa<L2>(q3 - s0, p0);
    p6 = lpa<L2>(q3 - s1, fminf(p5, fminf(p0, p1)));
    p7 = lpa<L2>(q3 - s2, fminf(p6, fminf(p1, p2)));
    p8 = lpa<L2>(q3 - s3, fminf(p7, fminf(p2, p3)));
    p9 = lpa<L2>(q3 - s4, fminf(p8, fminf(p3, p4)));
    mem = fminf(mem, p9);

    p0 = lpa<L2>(q4 - s0, p5);
    p1 = lpa<L2>(q4 - s1, fminf(p0, fminf(p5, p6)));
    p2 = lpa<L2>(q4 - s2, fminf(p1, fminf(p6, p7)));
    p3 = lpa<L2>(q4 - s3, fminf(p2, fminf(p7, p8)));
    p4 = lpa<L2>(q4 - s4, fminf(p3, fminf(p8, p9)));
    mem = fminf(mem, fminf(p0, fminf(p1, fminf(p2, fminf(p3, p4)))));

    q0 = q[m - 5]; q1 = q[m - 4]; q2 = q[m - 3]; q3 = q[m - 2]; q4 = q[m - 1];
    s0 = zn<EXACT>(x[m - 5], mu, sg, rs);
    s1 = zn<EXACT>(x[m - 4], mu, sg, rs);
    s2 = zn<EXACT>(x[m - 3], mu, sg, rs);
    s3 = zn<EXACT>(x[m - 2], mu, sg, rs);
    s4 = zn<EXACT>(x[m - 1], mu, sg, rs);

    p0 = lp<L2>(q0 - s0);
    p1 = lp<L2>(q0 - s1);
    p2 = lp<L2>(q0 - s2);
    p3 = lp<L2>(q0 - s3);
    p4 = lp<L2>(q0 - s4);

    p5 = lp<L2>(q1 - s0);
    p6 = lpa<L2>(q1 - s1, fminf(p5, fminf(p0, p1)));
    p7 = lpa<L2>(q1 - s2, fminf(p6, fminf(p1, p2)));
    p8 = lpa<L2>(q1 - s3, fminf(p7, fminf(p2, p3)));
    p9 = lpa<L2>(q1 - s4, fminf(p8, fminf(p3, p4)));

    p0 = lp<L2>(q2 - s0);
    p1 = lpa<L2>(q2 - s1, fminf(p0, fminf(p5, p6)));
    p2 = lpa<L2>(q2 - s2, fminf(p1, fminf(p6, p7)));
    p3 = lpa<L2>(q2 - s3, fminf(p2, fminf(p7, p8)));
    p4 = lpa<L2>(q2 - s4, fminf(p3, fminf(p8, p9)));

    p5 = lp<L2>(q3 - s0);
    p6 = lpa<L2>(q3 - s1, fminf(p5, fminf(p0, p1)));
    p7 = lpa<L2>(q3 - s2, fminf(p6, fminf(p1, p2)));
    p8 = lpa<L2>(q3 - s3, fminf(p7, fminf(p2, p3)));
    p9 = lpa<L2>(q3 - s4, fminf(p8, fminf(p3, p4)));

    p0 = lp<L2>(q4 - s0);
    p1 = lpa<L2>(q4 - s1, fminf(p0, fminf(p5, p6)));
    p2 = lpa<L2>(q4 - s2, fminf(p1, fminf(p6, p7)));
    p3 = lpa<L2>(q4 - s3, fminf(p2, fminf(p7, p8)));
    p4 = lpa<L2>(q4 - s4, fminf(p3, fminf(p8, p9)));

    return mem + p4;
}

template <bool L2>
__global__ __launch_bounds__(256) void lb_kim_kernel(const float* __restrict__ s,
                                                     const float* __restrict__ avg,
                                                     const float* __restrict__ sd,
                                                     const float* __restrict__ q,
                                                     int32_t m,
                                                     const int32_t* __restrict__ idx,
                                                     int64_t count, float* __restrict__ out) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= count) return;
    const int64_t w = idx ? idx[g] : g;
    out[g] = lb_kim_window<L2, true>(s + w, q, m, avg[w], sd[w], 0.0f);
}

// bounds.cuh:111-144.  e = z - clamp(z, lo, up) is the reference's `upper`
// (z > up) or `lower` (z < lo) or 0; at most one is non-zero.
template <bool L2>
__global__ __launch_bounds__(256) void lb_keogh_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ lo, const float* __restrict__ up, int32_t m,
    const int32_t* __restrict__ idx, int64_t count, float* __restrict__ out) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= count) return;
    const int64_t w = idx ? idx[g] : g;
    const float* x = s + w;
    const float mu = avg[w], sg = sd[w];
    float acc = 0.0f;
    for (int32_t i = 0; i < m; ++i) {
        const float z = (x[i] - mu) / sg;
        const float e = z - __builtin_amdgcn_fmed3f(z, lo[i], up[i]);
        acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
    }
    out[g] = acc;
}

__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Search cascade over windows 0..L-1, wave-strided: each wave takes 64
// consecutive windows per iteration.  Survivors go to cand (order not
// significant: the search result does not depend on it).
template <bool L2>
__global__ __launch_bounds__(256) void cascade_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, const float* __restrict__ lo, const float* __restrict__ up,
    int32_t m, int64_t L, bool use_best, Ctrl* ctrl, int32_t* __restrict__ cand) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    const float thr = use_best ? load_best_dist(ctrl) : load_thr(ctrl);
    const float lim = prune_limit(thr, m, L2);
    unsigned long long n_kim = 0, n_keogh = 0;

    for (int64_t base = wave * kWave; base < L; base += nwaves * kWave) {
        const int64_t w = base + lane;
        const bool valid = w < L;
        const int64_t wc = valid ? w : L - 1;
        const float* x = s + wc;
        const float mu = avg[wc], sg = sd[wc];
        const float rs = 1.0f / sg;
        const float kim = lb_kim_window<L2, false>(x, q, m, mu, sg, rs);
        bool alive = valid && !(kim > lim);
        n_kim += __popcll(__ballot(valid && !alive));
        if (!__any(alive)) continue;
        float acc = 0.0f;
        int32_t i = 0;
        bool gone = false;
        for (; i + 16 <= m; i += 16) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const float z = (x[i + u] - mu) * rs;
                const float e = z - __builtin_amdgcn_fmed3f(z, lo[i + u], up[i + u]);
                acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
            }
            if (!__any(alive && !(acc > lim))) {
                gone = true;
                break;
            }
        }
        for (; !gone && i < m; ++i) {
            const float z = (x[i] - mu) * rs;
            const float e = z - __builtin_amdgcn_fmed3f(z, lo[i], up[i]);
            acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
        }
        const bool keep = alive && !(acc > lim);
        const unsigned long long km = __ballot(keep);
        n_keogh += __popcll(__ballot(alive && !keep));
        if (km) {
            unsigned long long basec = 0;
            if (lane == 0)
                basec = atomicAdd(&ctrl->count, static_cast<unsigned long long>(__popcll(km)));
            basec = __shfl(basec, 0, kWave);
            if (keep) cand[basec + lane_rank(km)] = static_cast<int32_t>(w);
        }
    }
    if (lane == 0) {
        if (n_kim) atomicAdd(&ctrl->kim, n_kim);
        if (n_keogh) atomicAdd(&ctrl->keogh, n_keogh);
    }
}

inline int check_launch() {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

}  // namespace

int launch_lb_kim(hipStream_t st, const float* s, const float* avg, const float* sd,
                  const float* q, int32_t m, const int32_t* idx, int64_t count,
                  int32_t metric, float* out) {
    if (count <= 0) return MDTW_OK;
    const unsigned nb = static_cast<unsigned>((count + 255) / 256);
    if (metric)
        hipLaunchKernelGGL(lb_kim_kernel<true>, dim3(nb), dim3(256), 0, st, s, avg, sd, q, m,
                           idx, count, out);
    else
        hipLaunchKernelGGL(lb_kim_kernel<false>, dim3(nb), dim3(256), 0, st, s, avg, sd, q, m,
                           idx, count, out);
    return check_launch();
}

int launch_lb_keogh(hipStream_t st, const float* s, const float* avg, const float* sd,
                    const float* lo, const float* up, int32_t m, const int32_t* idx,
                    int64_t count, int32_t metric, float* out) {
    if (count <= 0) return MDTW_OK;
    const unsigned nb = static_cast<unsigned>((count + 255) / 256);
    if (metric)
        hipLaunchKernelGGL(lb_keogh_kernel<true>, dim3(nb), dim3(256), 0, st, s, avg, sd, lo,
                           up, m, idx, count, out);
    else
        hipLaunchKernelGGL(lb_keogh_kernel<false>, dim3(nb), dim3(256), 0, st, s, avg, sd, lo,
                           up, m, idx, count, out);
    return check_launch();
}

int launch_cascade(hipStream_t st, const float* s, const float* avg, const float* sd,
                   const float* q, const float* lo, const float* up, int32_t m, int64_t L,
                   int32_t metric, bool use_best, Ctrl* ctrl, int32_t* cand) {
    if (L <= 0) return MDTW_OK;
    // Enough waves to cover the chip several times over; each wave strides.
    const int64_t waves_needed = (L + kWave - 1) / kWave;
    const int64_t nb64 = std::min<int64_t>((waves_needed + 3) / 4, 256 * 64);
    const unsigned nb = static_cast<unsigned>(nb64);
    if (metric)
        hipLaunchKernelGGL(cascade_kernel<true>, dim3(nb), dim3(256), 0, st, s, avg, sd, q, lo,
                           up, m, L, use_best, ctrl, cand);
    else
        hipLaunchKernelGGL(cascade_kernel<false>, dim3(nb), dim3(256), 0, st, s, avg, sd, q,
                           lo, up, m, L, use_best, ctrl, cand);
    return check_launch();
}

}  // namespace mdtw
