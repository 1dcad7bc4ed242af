// cdtw_band.hpp -- banded cDTW with the whole band of one window in the
// registers of one lane (included by the cdtw_band_*.hip instantiation units).
//
// Mapping.  Lane = window (64 consecutive windows per wave, so stream reads
// coalesce).  The DP runs column by column over the window's samples j; the
// band of column j is held as E[t] = D[j - W + t][j], t = 0..2W.  In these
// coordinates the three predecessors of E[t] are E[t] and E[t+1] of the
// previous column and E[t-1] of this column, so the update is in place in
// increasing t with no data movement:
//     E[t] = fmaf(d, d, min(E[t-1], E[t], E[t+1])),   d = q[j-W+t] - z_j
// (v_sub, v_min3, v_fmac: three VALU slots per cell).  The query index
// j - W + t is wave-uniform, so the query is read with scalar loads from a
// copy padded with W +inf sentinels on each side (out-of-matrix cells cost
// +inf, which min() then ignores).  z_j = (x_j - mu) / sigma is one IEEE
// division per column per lane, amortised over 2W + 1 cells.  The answer is
// E[W] after column M - 1.
//
// The arithmetic per cell is that of cdtw.cuh:339-350 (z-normalise by
// division, q - z, fused square-add onto the minimum), so results are
// bit-exact with the CPU restatement of sequential_cdtw (cdtw.cuh:23-88).
//
// Early abandoning (search modes): every warping path crosses every column,
// and cell values only grow along a path, so once the minimum of a column
// exceeds the threshold the window cannot finish at or below it.
#pragma once

#include "internal.hpp"

namespace mdtw {

using BandLaunchFn = int (*)(hipStream_t, unsigned grid, const BandArgs&);

// Column-block unroll and abandon-check interval (in columns).
constexpr int kBandU = 4;
constexpr int kBandCheck = 16;
constexpr int kBandChunk = 16;  // query scalars per load group

__device__ __forceinline__ int64_t band_next(Ctrl* ctrl, int lane) {
    unsigned long long b = 0;
    if (lane == 0) b = atomicAdd(&ctrl->work, static_cast<unsigned long long>(kWave));
    return static_cast<int64_t>(__shfl(b, 0, kWave));
}

template <int W, bool L2>
__device__ __forceinline__ float band_dp(const float* __restrict__ x, float mu, float sg,
                                         const float* __restrict__ qx, int32_t m,
                                         bool& alive, int mode, const Ctrl* ctrl) {
    constexpr int B = 2 * W + 1;
    constexpr int U = kBandU;
    constexpr int CH = kBandChunk;
    constexpr int NC = (B + CH - 1) / CH;
    float E[B];
#pragma unroll
    for (int t = 0; t < B; ++t) E[t] = kInf;
    E[W] = 0.0f;  // D[-1][-1]

    // Samples of the next U columns; indices are clamped rather than
    // branched so the loads stay unconditional and overlap a whole block.
    float xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = x[u < m ? u : m - 1];

    const bool abandon = mode == kModeBest || mode == kModeCollect;
    for (int32_t j0 = 0; j0 < m; j0 += U) {
        float z[U];
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = (xv[u] - mu) / sg;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t c = j0 + U + u;
            xv[u] = x[c < m ? c : m - 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (j0 + u < m) {
                const float* __restrict__ qc = qx + (j0 + u);
                const float zj = z[u];
                float prev = kInf;
                // The query band is read CH scalars at a time, one chunk
                // ahead; the scheduling barrier keeps the scalar loads from
                // being hoisted together (that would exceed the SGPR file).
                float qa[CH];
#pragma unroll
                for (int i = 0; i < CH; ++i) qa[i] = qc[i];
#pragma unroll
                for (int ch = 0; ch < NC; ++ch) {
                    float qb[CH];
                    if (ch + 1 < NC) {
#pragma unroll
                        for (int i = 0; i < CH; ++i) qb[i] = qc[(ch + 1) * CH + i];
                    }
                    __builtin_amdgcn_sched_barrier(0x96);  // VALU, SALU, VMEM, DS may cross
#pragma unroll
                    for (int i = 0; i < CH; ++i) {
                        const int t = ch * CH + i;
                        if (t < B) {
                            const float nxt = t + 1 < B ? E[t + 1] : kInf;
                            const float mn = fminf(fminf(prev, E[t]), nxt);
                            const float d = qa[i] - zj;
                            const float e = L2 ? fmaf(d, d, mn) : fabsf(d) + mn;
                            E[t] = e;
                            prev = e;
                        }
                    }
#pragma unroll
                    for (int i = 0; i < CH; ++i) qa[i] = qb[i];
                }
            }
        }
        if (abandon && ((j0 / U) % (kBandCheck / U)) == (kBandCheck / U) - 1) {
            float cm = E[0];
#pragma unroll
            for (int t = 1; t < B; ++t) cm = fminf(cm, E[t]);
            const float thr = mode == kModeBest ? load_best_dist(ctrl) : load_thr(ctrl);
            alive = alive && !(cm > thr);
            if (!__any(alive)) return kInf;
        }
    }
    return E[W];
}

__device__ __forceinline__ uint32_t band_lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <int W, bool L2>
__global__ __launch_bounds__(64) void band_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ qx, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, BandScalars a) {
    const int lane = threadIdx.x;
    int64_t count = a.count;
    if (a.idx_count_from_ctrl)
        count = static_cast<int64_t>(
            __hip_atomic_load(&ctrl->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    unsigned long long n_dp = 0, n_ab = 0;
    int64_t base = a.persistent ? band_next(ctrl, lane)
                                : static_cast<int64_t>(blockIdx.x) * kWave;
    while (base < count) {
        const int64_t g = base + lane;
        const bool valid = g < count;
        const int64_t gc = valid ? g : count - 1;
        const int64_t w = idx ? static_cast<int64_t>(idx[gc]) : a.base + gc * a.stride;
        const float mu = avg[w], sg = sd[w];
        bool alive = valid;
        const float d = band_dp<W, L2>(s + w, mu, sg, qx, a.m, alive, a.mode, ctrl);
        if (a.mode == kModeProfile || a.mode == kModeSeed) {
            if (valid) out[g] = d;
        }
        if (a.mode == kModeBest || a.mode == kModeSeed) {
            unsigned long long key = alive ? pack_key(d, static_cast<uint32_t>(w)) : ~0ull;
            key = wave_min_u64(key);
            if (lane == 0 && key != ~0ull) atomicMin(&ctrl->best, key);
        } else if (a.mode == kModeCollect) {
            const bool keep = alive && d <= load_thr(ctrl);
            const unsigned long long km = __ballot(keep);
            if (km) {
                unsigned long long rb = 0;
                if (lane == 0)
                    rb = atomicAdd(&ctrl->nres, static_cast<unsigned long long>(__popcll(km)));
                rb = __shfl(rb, 0, kWave);
                if (keep) res[rb + band_lane_rank(km)] = pack_key(d, static_cast<uint32_t>(w));
            }
        }
        n_dp += __popcll(__ballot(valid));
        n_ab += __popcll(__ballot(valid && !alive));
        if (!a.persistent) break;
        base = band_next(ctrl, lane);
    }
    if (lane == 0 && a.mode != kModeProfile) {
        if (n_dp) atomicAdd(&ctrl->dp, n_dp);
        if (n_ab) atomicAdd(&ctrl->abandoned, n_ab);
    }
}

template <int W, bool L2>
int band_launch(hipStream_t st, unsigned grid, const BandArgs& a) {
    BandScalars sc{a.m, a.idx_count_from_ctrl, a.count, a.base, a.stride, a.mode, a.persistent};
    hipLaunchKernelGGL((band_kernel<W, L2>), dim3(grid), dim3(kWave), 0, st, a.s, a.avg, a.sd,
                       a.qx, a.idx, a.out, a.res, a.ctrl, sc);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

// Instantiation units provide these for their W ranges.
BandLaunchFn band_table_0(int w, bool l2);  // W 0..29
BandLaunchFn band_table_1(int w, bool l2);  // W 30..59
BandLaunchFn band_table_2(int w, bool l2);  // W 60..89
BandLaunchFn band_table_3(int w, bool l2);  // W 90..kBandMaxW

constexpr int kBandMaxW = 112;

}  // namespace mdtw
