// cdtw_band.hpp -- banded cDTW with the whole band of one window in the
// registers of one lane (included by the cdtw_band_*.hip instantiation units).
//
// Mapping.  Lane = window (64 consecutive windows per wave, so stream reads
// coalesce; 4 waves per workgroup).  The DP runs column by column over the
// window's samples j; the band of column j is held as E[t] = D[j - W + t][j],
// t = 0..2W.  In these coordinates the three predecessors of E[t] are E[t]
// and E[t+1] of the previous column and E[t-1] of this column, so the update
// is in place in increasing t with no data movement:
//     E[t] = fmaf(d, d, min(E[t-1], E[t], E[t+1])),   d = q[j-W+t] - z_j
// (v_sub, v_min3, v_fmac: three VALU slots per cell).  The query index
// j - W + t is wave-uniform, so the query -- padded with W +inf sentinels on
// each side (out-of-matrix cells cost +inf, which min() then ignores) -- is
// read with scalar loads into SGPRs: no per-lane copies, no LDS bandwidth (a
// broadcast LDS read still moves 64 lanes' worth of data).  z_j =
// (x_j - mu) / sigma is computed once per column per lane (see band_z),
// amortised over 2W + 1 cells.  The answer is E[W] after column M - 1.
//
// The arithmetic per cell is that of cdtw.cuh:339-350 (z-normalise by
// division, q - z, fused square-add onto the minimum), so results are
// bit-exact with the CPU restatement of sequential_cdtw (cdtw.cuh:23-88).
//
// Early abandoning (search modes): every warping path crosses every column,
// and cell values only grow along a path, so once the minimum of a column
// exceeds the threshold the window cannot finish at or below it.
#pragma once

#include <type_traits>

#include "internal.hpp"

namespace mdtw {

using BandLaunchFn = int (*)(hipStream_t, unsigned grid, const BandArgs&);

// Column-block unroll, abandon-check interval (columns), query chunk size.
constexpr int kBandU = 4;
constexpr int kBandCheck = 16;
constexpr int kBandChunk = 16;
constexpr int kBandWaves = 4;  // waves per workgroup
// Narrow-band limit: query window plus prefetch held in scalar registers
// (up to W = 23; measured over W = 0..112 against the chunked loads, which
// stall on their scalar waits for W = 18..23 and are faster from W = 24).
constexpr int kBandSlideMax = 54;
// Bands up to kBandMaxW unroll kBandU columns per block (W = 108..110: two,
// W = 111..112: one, which keeps them within 256 registers, i.e. two waves
// per SIMD); W = 113..114 keep the 2W + 1 band values in vector registers,
// one column at a time (two waves per SIMD), W = 115..kBandTailMaxW in vector
// registers plus an LDS tail (below), and the bands from kBandPairMinW to
// kBandWideMaxW are split over two or three waves (cdtw_pair.hpp).
constexpr int kBandMaxW = 112;
constexpr int kBandWideMaxW = 240;
// Bands W = 115..kBandTailMaxW keep band values 0..kBandTailR-1 in vector
// registers and the rest (the tail, t >= kBandTailR) in LDS, 16 bytes per
// lane per chunk of four values, so two workgroups (two waves per SIMD, the
// register budget of 256) share a CU's 160 KiB.  Their kernels are compiled
// for two waves per SIMD (band_kernel_tail); the compiler then spills a few
// band values per column to scratch memory, a number that grows with the
// tail: measured at M = 1024 (2^22 windows), L2, the tail kernel runs W = 115
// at 17.7 T cells/s, 121 at 16.8, 124 at 12.9 and 136 at 10.7, where the
// two-wave split runs 16.5..16.9 over the whole range; L1 alike (18.0 and
// 17.2 at W = 115 and 120, 14.0 at 125, the split 15.3, 16.9, 16.9).
constexpr int kBandTailR = 212;
constexpr int kBandTailMaxW = 121;
__host__ __device__ constexpr int band_tail_len(int w) {
    return (w > 114 && w <= kBandTailMaxW) ? 2 * w + 1 - kBandTailR : 0;
}
constexpr int kBandPairMinW = kBandTailMaxW + 1;
constexpr int kBandTailStride = 4 * 64;  // float4 slots between a lane's chunks

template <int NCHUNK>
__device__ __forceinline__ float4* band_tail_lds() {
    __shared__ float4 buf[NCHUNK * kBandTailStride];
    return buf + threadIdx.x;
}
__device__ __forceinline__ float f4_get(const float4& v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
__device__ __forceinline__ void f4_set(float4& v, int i, float f) {
    if (i == 0) v.x = f;
    else if (i == 1) v.y = f;
    else if (i == 2) v.z = f;
    else v.w = f;
}

// Next 64 windows from the shared cursor (one-wave workgroups).
__device__ __forceinline__ int64_t band_next(Ctrl* ctrl, int lane) {
    unsigned long long b = 0;
    if (lane == 0) b = atomicAdd(&ctrl->work, static_cast<unsigned long long>(kWave));
    return static_cast<int64_t>(__shfl(b, 0, kWave));
}

// Next 64 windows from the shared cursor, which hands out the windows from
// `first` on (the ones before were assigned by grid position).  Once a wave
// of the block has seen the cursor past the end (it only grows), the block's
// other waves stop without another same-address atomic: at the end of a
// launch every wave would otherwise queue one on the cursor at once.
__device__ __forceinline__ int64_t band_next(Ctrl* ctrl, int lane, int64_t first, int64_t count,
                                             int* drained) {
    unsigned long long b = static_cast<unsigned long long>(count);
    if (lane == 0 && first < count &&
        !__hip_atomic_load(drained, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
        b = static_cast<unsigned long long>(first) +
            atomicAdd(&ctrl->work, static_cast<unsigned long long>(kWave));
        if (b >= static_cast<unsigned long long>(count))
            __hip_atomic_store(drained, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    // lane 0's value, kept in scalar registers (the whole wave calls this)
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
    return static_cast<int64_t>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// One column of the band update with the query values of this column in
// q[0..B) (wave-uniform scalars), over cells T0..T1 (the others are outside
// the matrix in the columns this is called for; see band_dp).
template <int B, bool L2, int T0, int T1, class Q>
__device__ __forceinline__ void band_column(float (&E)[B], const Q& q, float zj) {
    float prev = kInf;
#pragma unroll
    for (int t = T0; t <= T1; ++t) {
        const float nxt = t + 1 < B ? E[t + 1] : kInf;
        const float mn = fminf(fminf(prev, E[t]), nxt);
        const float d = q(t) - zj;
        const float e = L2 ? fmaf(d, d, mn) : fabsf(d) + mn;
        E[t] = e;
        prev = e;
    }
}

// LB: cumulative-bound abandon (see suffix_limit): the DP also sums the
// window's LB_Keogh terms over the columns it has done, with the envelope
// held lane-distributed (lane l: lo/up[64c + l]) and broadcast per column by
// readlane, the next 64 values loaded a chunk ahead.
template <int W, bool L2, bool LB>
__device__ __forceinline__ float band_dp(const float* __restrict__ x, float mu, float sg,
                                         const float* __restrict__ qx, int32_t m, bool& alive,
                                         int mode, const Ctrl* ctrl,
                                         const float* __restrict__ elo,
                                         const float* __restrict__ eup, float lbt) {
    static_assert(W < kBandPairMinW, "wider bands run split over two waves (cdtw_pair.hpp)");
    constexpr int B = 2 * W + 1;
    constexpr int U = W > 110 ? 1 : (W > 107 ? 2 : kBandU);
    constexpr int CH = kBandChunk;
    constexpr int NC = (B + CH - 1) / CH;
    // Narrow bands keep the query window of a whole block of U columns in
    // scalar registers and slide it by U per block; the next U values are
    // requested a block ahead, so the only scalar-memory wait is on data
    // that has had a block of compute to arrive.  (Scalar loads complete out
    // of order, so every wait is for all of them.)  Wider bands read the
    // query in chunks of CH scalars per column.
    constexpr bool kSlide = B + 2 * U - 1 <= kBandSlideMax;
    // KT band values in the LDS tail (0 for most bands), BR in registers.
    constexpr int KT = band_tail_len(W);
    constexpr int KT4 = (KT + 3) / 4;
    constexpr int BR = B - KT;
    static_assert(KT == 0 || (U == 1 && !kSlide && W < BR), "tail layout");
    float E[BR];
#pragma unroll
    for (int t = 0; t < BR; ++t) E[t] = kInf;
    E[W] = 0.0f;  // D[-1][-1]
    [[maybe_unused]] float4* tl = nullptr;
    // minimum of the tail's values after the last column (for the abandon
    // check, which then needs no LDS reads)
    [[maybe_unused]] float tail_min = kInf;
    const float4 inf4 = make_float4(kInf, kInf, kInf, kInf);
    if constexpr (KT4 > 0) {
        tl = band_tail_lds<KT4>();
#pragma unroll
        for (int c = 0; c < KT4; ++c) tl[c * kBandTailStride] = inf4;
    }

    const float y = 1.0f / sg;
    const bool wfast = band_z_window_fast(mu, sg);

    // Samples of the next U columns; indices are clamped rather than
    // branched so the loads stay unconditional and overlap a whole block.
    float xv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xv[u] = x[u < m ? u : m - 1];

    constexpr int R = kSlide ? B + U - 1 : 1;
    float qs[R];
    if constexpr (kSlide) {
#pragma unroll
        for (int k = 0; k < R; ++k) qs[k] = qx[k];
    }

    const bool abandon = mode == kModeBest || mode == kModeCollect;
    const int lane = threadIdx.x & (kWave - 1);
    // Bands that fit in vector registers read the abandon bound once per
    // batch of windows (load_abandon_bound); the wider ones re-read it at each
    // check: their band already fills the 256 registers, one more value live
    // across the DP made the compiler spill it to scratch memory, and their
    // checks are hundreds of cells apart anyway.
    constexpr bool kBoundOnce = W <= kBandMaxW;
    const float thr_once = kBoundOnce && abandon ? load_abandon_bound(ctrl, mode) : kInf;
    auto bound = [&]() {
        return kBoundOnce ? thr_once
                          : (mode == kModeBest ? load_best_dist(ctrl) : load_thr(ctrl));
    };
    float lo_c = 0.0f, up_c = 0.0f, lo_n = 0.0f, up_n = 0.0f, P = 0.0f;
    if constexpr (LB) {
        // The bound may have fallen since the cascade kept this window.
        alive = alive && !(lbt > prune_limit(bound(), m, L2));
        if (!__any(alive)) return kInf;
        lo_c = elo[lane < m ? lane : m - 1];
        up_c = eup[lane < m ? lane : m - 1];
        lo_n = elo[kWave + lane < m ? kWave + lane : m - 1];
        up_n = eup[kWave + lane < m ? kWave + lane : m - 1];
    }
    // One block of U columns from column j0, computing cells T0..T1 of each
    // (all of them but in the first and last columns, below); returns true
    // once every window of the wave is abandoned.
    auto block = [&](int32_t j0, auto t0c, auto t1c) -> bool {
        constexpr int T0 = decltype(t0c)::value;
        constexpr int T1 = decltype(t1c)::value;
        float z[U];
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = band_z_w(xv[u], mu, sg, y, wfast);
        if constexpr (LB) {
            // U divides 64, so a block of U columns never straddles a chunk.
            if (j0 > 0 && (j0 & (kWave - 1)) == 0) {
                lo_c = lo_n;
                up_c = up_n;
                const int32_t c = j0 + kWave + lane;
                lo_n = elo[c < m ? c : m - 1];
                up_n = eup[c < m ? c : m - 1];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j0 + u < m) {
                    const int sl = (j0 + u) & (kWave - 1);
                    const float l = __int_as_float(
                        __builtin_amdgcn_readlane(__float_as_int(lo_c), sl));
                    const float h = __int_as_float(
                        __builtin_amdgcn_readlane(__float_as_int(up_c), sl));
                    const float e = z[u] - __builtin_amdgcn_fmed3f(z[u], l, h);
                    P = L2 ? fmaf(e, e, P) : P + fabsf(e);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t c = j0 + U + u;
            xv[u] = x[c < m ? c : m - 1];
        }
        if constexpr (kSlide) {
            // next block's new values: qx[j0 + U + R - U .. + U)
            float pf[U];
#pragma unroll
            for (int u = 0; u < U; ++u) pf[u] = qx[j0 + R + u];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j0 + u < m)
                    band_column<B, L2, T0, T1>(E, [&](int t) { return qs[u + t]; }, z[u]);
            }
#pragma unroll
            for (int k = 0; k + U < R; ++k) qs[k] = qs[k + U];
#pragma unroll
            for (int u = 0; u < U; ++u) qs[R - U + u] = pf[u];
        } else if constexpr (KT4 > 0) {
            // One column (U == 1): cells 0..BR-1 in registers, then the tail
            // in chunks of four, each read from LDS two chunks ahead (its
            // first value is the last register cell's right neighbour, or the
            // previous chunk's), updated in registers and written back.
            const float* __restrict__ qc = qx + j0;
            const float zj = z[0];
            float prev = kInf;
            float4 T0 = tl[0];
            float4 T1 = KT4 > 1 ? tl[kBandTailStride] : inf4;
            float4 T2 = KT4 > 2 ? tl[2 * kBandTailStride] : inf4;
            __builtin_amdgcn_sched_barrier(0x16);
            tail_min = kInf;
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
                __builtin_amdgcn_sched_barrier(0x96);
#pragma unroll
                for (int i = 0; i < CH; ++i) {
                    const int t = ch * CH + i;
                    if (t < BR) {
                        const float nxt = t + 1 < BR ? E[t + 1] : T0.x;
                        const float mn = fminf(fminf(prev, E[t]), nxt);
                        const float d = qa[i] - zj;
                        const float e = L2 ? fmaf(d, d, mn) : fabsf(d) + mn;
                        E[t] = e;
                        prev = e;
                    } else if (t < B) {
                        const int k = t - BR, e4 = k & 3, c4 = k >> 2;
                        const float nxt = e4 < 3 ? f4_get(T0, e4 + 1) : T1.x;
                        const float mn = fminf(fminf(prev, f4_get(T0, e4)), nxt);
                        const float d = qa[i] - zj;
                        const float e = L2 ? fmaf(d, d, mn) : fabsf(d) + mn;
                        f4_set(T0, e4, e);
                        prev = e;
                        if (e4 == 3 || t == B - 1) {
                            tail_min = fminf(tail_min, fminf(fminf(T0.x, T0.y), fminf(T0.z, T0.w)));
                            tl[c4 * kBandTailStride] = T0;
                            T0 = T1;
                            T1 = T2;
                            T2 = c4 + 3 < KT4 ? tl[(c4 + 3) * kBandTailStride] : inf4;
                            // LDS traffic stays where it is written: hoisted
                            // reads would hold more of the tail in registers
                            __builtin_amdgcn_sched_barrier(0x16);
                        }
                    }
                }
#pragma unroll
                for (int i = 0; i < CH; ++i) qa[i] = qb[i];
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (j0 + u < m) {
                    const float* __restrict__ qc = qx + (j0 + u);
                    const float zj = z[u];
                    float prev = kInf;
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
                        // keep the scalar loads in chunks (hoisting them all
                        // would exceed the scalar register file)
                        __builtin_amdgcn_sched_barrier(0x96);
#pragma unroll
                        for (int i = 0; i < CH; ++i) {
                            const int t = ch * CH + i;
                            if (t >= T0 && t <= T1 && t < B) {
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
        }
        if (abandon && ((j0 / U) % (kBandCheck / U)) == (kBandCheck / U) - 1) {
            float cm = E[0];
#pragma unroll
            for (int t = 1; t < BR; ++t) cm = fminf(cm, E[t]);
            if constexpr (KT4 > 0) {
                cm = fminf(cm, tail_min);
            }
            const float thr = bound();
            if constexpr (LB) {
                const float rest = lbt - P;
                const float bound = cm + (rest > 0.0f ? rest : 0.0f);
                alive = alive && !(cm > thr) && !(bound > suffix_limit(thr, lbt, m, L2));
            } else {
                alive = alive && !(cm > thr);
            }
            if (!__any(alive)) return true;
        }
        return false;
    };
    // Cell t of column j is row j - W + t, inside the matrix for
    // W - j <= t <= W + m - 1 - j; the cells outside cost +inf (their query
    // values are the +inf padding), so computing them changes nothing.  The
    // first P1 columns (P1 <= W/2, a multiple of U) have no cell inside below
    // T0a = W - P1 + 1, and the columns from J1 (>= m - W/2) none above
    // T1b = W + W/2 - 1: those blocks skip them (2.5% of the cells at
    // W = 10% of M; measured against the full band on one device, M = 256:
    // W = 32 +3%, W = 51 +3.8%; M = 1024: W = 80..112 +1.5..2%).  A skipped cell below the matrix still holds its initial
    // +inf when it enters the matrix (its first computation reads it as the
    // diagonal neighbour, as +inf); one above keeps a stale value, which no
    // cell inside reads (a cell's left neighbour is its own row one column
    // back, inside the computed range there) and which can only lower an
    // abandon check's column minimum, delaying that window's abandon.  The
    // results are the full band's bit for bit.
    // (From W = 32: at W = 24..28 the extra loops' registers cost a wave
    // per SIMD and the net measured 0..0.3%; above kBandMaxW the band already
    // fills the registers.)
    constexpr int P1 = (KT4 == 0 && W >= 32 && W <= kBandMaxW) ? ((W / 2) / U) * U : 0;
    constexpr int T0a = W - P1 + 1;
    constexpr int T1b = W + W / 2 - 1;
    int32_t j0 = 0;
    if (P1 > 0 && m >= 2 * W + 2 * U) {
        const int32_t J1 = ((m - W / 2 + U - 1) / U) * U;
        for (; j0 < P1; j0 += U)
            if (block(j0, std::integral_constant<int, T0a>{}, std::integral_constant<int, B - 1>{}))
                return kInf;
        for (; j0 < J1; j0 += U)
            if (block(j0, std::integral_constant<int, 0>{}, std::integral_constant<int, B - 1>{}))
                return kInf;
        for (; j0 < m; j0 += U)
            if (block(j0, std::integral_constant<int, 0>{}, std::integral_constant<int, T1b>{}))
                return kInf;
    } else {
        for (; j0 < m; j0 += U)
            if (block(j0, std::integral_constant<int, 0>{}, std::integral_constant<int, B - 1>{}))
                return kInf;
    }
    return E[W];
}

__device__ __forceinline__ uint32_t band_lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

template <int W, bool L2, bool LB>
__device__ __forceinline__ void band_kernel_body(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ qx, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, BandScalars a,
    const float* __restrict__ lbt, const float* __restrict__ elo,
    const float* __restrict__ eup) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    int64_t count = a.count;
    if (a.idx_count_from_ctrl)
        count = static_cast<int64_t>(
            __hip_atomic_load(&ctrl->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    unsigned long long n_dp = 0, n_ab = 0;
    uint32_t* const bmin = ctrl->bmin;
    const long long bucket_len = ctrl->bucket_len;
    __shared__ int drained;
    __shared__ unsigned long long part[kBandWaves][2];
    if (a.persistent) {  // uniform: the flag is read by persistent waves only
        if (threadIdx.x == 0) drained = 0;
        __syncthreads();
    }
    // Every wave's first 64 windows are fixed by its position in the grid;
    // after that, persistent waves take 64 windows at a time from an atomic
    // cursor over the windows beyond the grid's first pass, and the others
    // stride by the grid (every wave the same amount of work).  A persistent
    // wave whose first chunk is already past the end, or a grid that covers
    // every window in one pass, touches the cursor not at all: same-address
    // atomics serialise (~15 ns each), which for a short candidate list and a
    // grid sized for the chip would otherwise cost more than the DP.
    const int64_t step = static_cast<int64_t>(gridDim.x) * kBandWaves * kWave;
    int64_t base = (static_cast<int64_t>(blockIdx.x) * kBandWaves + wid) * kWave;
    while (base < count) {
        const int64_t g = base + lane;
        const bool valid = g < count;
        const int64_t gc = valid ? g : count - 1;
        const int64_t w = idx ? static_cast<int64_t>(idx[gc]) : a.base + gc * a.stride;
        const float mu = avg[w], sg = sd[w];
        bool alive = valid;
        const float d = band_dp<W, L2, LB>(s + w, mu, sg, qx, a.m, alive, a.mode, ctrl, elo,
                                           eup, LB ? lbt[gc] : 0.0f);
        if (a.mode == kModeProfile || a.mode == kModeSeed) {
            if (valid) out[g] = d;
        }
        if (a.mode == kModeBest || a.mode == kModeSeed) {
            unsigned long long key = alive ? pack_key(d, static_cast<uint32_t>(w)) : ~0ull;
            key = wave_min_u64(key);
            if (lane == 0 && key != ~0ull) atomicMin(&ctrl->best, key);
            if (a.mode == kModeSeed && valid) note_bucket(bmin, bucket_len, w, d);
        } else if (a.mode == kModeCollect) {
            const bool keep = alive && d <= load_thr(ctrl);
            if (keep) note_bucket(bmin, bucket_len, w, d);
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
        base = a.persistent ? band_next(ctrl, lane, step, count, &drained) : base + step;
    }
    // Counters summed per block: two atomics per block, not per wave.
    if (a.mode != kModeProfile) {
        if (lane == 0) {
            part[wid][0] = n_dp;
            part[wid][1] = n_ab;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long td = 0, ta = 0;
            for (int v = 0; v < kBandWaves; ++v) {
                td += part[v][0];
                ta += part[v][1];
            }
            if (td) atomicAdd(&ctrl->dp, td);
            if (ta) atomicAdd(&ctrl->abandoned, ta);
        }
    }
}

#define MDTW_BAND_KERNEL_ARGS                                                                 \
    const float *__restrict__ s, const float *__restrict__ avg, const float *__restrict__ sd,   \
        const float *__restrict__ qx, const int32_t *__restrict__ idx, float *__restrict__ out, \
        unsigned long long *__restrict__ res, Ctrl *__restrict__ ctrl, BandScalars a,            \
        const float *__restrict__ lbt, const float *__restrict__ elo, const float *__restrict__ eup

template <int W, bool L2, bool LB>
__global__ __launch_bounds__(256) void band_kernel(MDTW_BAND_KERNEL_ARGS) {
    band_kernel_body<W, L2, LB>(s, avg, sd, qx, idx, out, res, ctrl, a, lbt, elo, eup);
}

// Bands with an LDS tail, compiled for two waves per SIMD (see band_tail_len).
template <int W, bool L2, bool LB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void band_kernel_tail(
    MDTW_BAND_KERNEL_ARGS) {
    band_kernel_body<W, L2, LB>(s, avg, sd, qx, idx, out, res, ctrl, a, lbt, elo, eup);
}
#undef MDTW_BAND_KERNEL_ARGS

template <int W, bool L2, bool LB>
constexpr auto band_kernel_fn() {
    if constexpr (band_tail_len(W) > 0)
        return &band_kernel_tail<W, L2, LB>;
    else
        return &band_kernel<W, L2, LB>;
}

}  // namespace mdtw

#include "cdtw_pair.hpp"

namespace mdtw {

// Waves per SIMD a kernel can keep resident, from its register count
// (512 registers per lane per SIMD, allocated in granules of 8).
inline int resident_waves(const void* fn) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, fn) != hipSuccess || attr.numRegs <= 0) {
        (void)hipGetLastError();
        return 0;
    }
    const int regs = (attr.numRegs + 7) & ~7;
    return regs > 0 ? 512 / regs : 0;
}

// Bytes of scratch memory (register spills) a kernel needs per lane, or -1.
inline int spill_bytes(const void* fn) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, fn) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return static_cast<int>(attr.localSizeBytes);
}

// The cumulative-bound variant holds a few more registers; it is used unless
// that drops residency below min(2, the plain variant's) waves per SIMD, where
// the lost latency hiding costs more than the earlier abandon saves, or makes
// it spill registers to scratch memory where the plain variant does not.
template <int W, bool L2>
bool band_lb_pays() {
    static const bool pays = [] {
        const void* fw = reinterpret_cast<const void*>(band_kernel_fn<W, L2, true>());
        const void* fo = reinterpret_cast<const void*>(band_kernel_fn<W, L2, false>());
        const int with = resident_waves(fw);
        const int without = resident_waves(fo);
        const int sw = spill_bytes(fw), so = spill_bytes(fo);
        return with > 0 && with >= (without < 2 ? without : 2) && sw >= 0 && sw <= so;
    }();
    return pays;
}

// `grid` counts workgroups of kBandWaves one-lane waves (256 windows per
// pass); a split band takes four pairs or trios of waves for each (the same
// windows per pass, twice or three times the waves: two or three of them fit
// a SIMD, against one or two one-lane waves of these widths).
template <int W, bool L2>
int band_launch(hipStream_t st, unsigned grid, const BandArgs& a) {
    BandScalars sc{a.m, a.idx_count_from_ctrl, a.count, a.base, a.stride, a.mode, a.persistent,
                   0, a.qx_len};
    if constexpr (band_trio(W)) {
        const dim3 g(grid * kBandWaves), b(kTrioWaves * kWave);
        if constexpr (band_trio_three(W))
            hipLaunchKernelGGL((band_trio_kernel3<W, L2>), g, b, 0, st, a.s, a.avg, a.sd, a.qx,
                               a.idx, a.out, a.res, a.ctrl, sc);
        else
            hipLaunchKernelGGL((band_trio_kernel<W, L2>), g, b, 0, st, a.s, a.avg, a.sd, a.qx,
                               a.idx, a.out, a.res, a.ctrl, sc);
    } else if constexpr (W >= kBandPairMinW) {
        const dim3 g(grid * kBandWaves), b(kPairWaves * kWave);
        if constexpr (band_pair_tight(W))
            hipLaunchKernelGGL((band_pair_kernel2<W, L2>), g, b, 0, st, a.s, a.avg, a.sd, a.qx,
                               a.idx, a.out, a.res, a.ctrl, sc);
        else if constexpr (band_pair_three(W))
            hipLaunchKernelGGL((band_pair_kernel3<W, L2>), g, b, 0, st, a.s, a.avg, a.sd, a.qx,
                               a.idx, a.out, a.res, a.ctrl, sc);
        else
            hipLaunchKernelGGL((band_pair_kernel<W, L2>), g, b, 0, st, a.s, a.avg, a.sd, a.qx,
                               a.idx, a.out, a.res, a.ctrl, sc);
    } else {
        const bool lb = a.lbt && a.idx && a.lo && a.up &&
                        (a.mode == kModeBest || a.mode == kModeCollect) && band_lb_pays<W, L2>();
        if (lb)
            hipLaunchKernelGGL((band_kernel_fn<W, L2, true>()), dim3(grid),
                               dim3(kBandWaves * kWave), 0, st, a.s, a.avg, a.sd, a.qx, a.idx,
                               a.out, a.res, a.ctrl, sc, a.lbt, a.lo, a.up);
        else
            hipLaunchKernelGGL((band_kernel_fn<W, L2, false>()), dim3(grid),
                               dim3(kBandWaves * kWave), 0, st, a.s, a.avg, a.sd, a.qx, a.idx,
                               a.out, a.res, a.ctrl, sc, nullptr, nullptr, nullptr);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

// Instantiation units provide these for their W ranges.
BandLaunchFn band_table_0(int w, bool l2);   // W 0..29
BandLaunchFn band_table_1(int w, bool l2);   // W 30..39
BandLaunchFn band_table_12(int w, bool l2);  // W 40..49
BandLaunchFn band_table_13(int w, bool l2);  // W 50..59
BandLaunchFn band_table_2(int w, bool l2);   // W 60..69
BandLaunchFn band_table_14(int w, bool l2);  // W 70..79
BandLaunchFn band_table_15(int w, bool l2);  // W 80..89
BandLaunchFn band_table_3(int w, bool l2);   // W 90..97
BandLaunchFn band_table_16(int w, bool l2);  // W 98..105
BandLaunchFn band_table_17(int w, bool l2);  // W 106..kBandMaxW
BandLaunchFn band_table_4(int w);           // L2, W 113..144
BandLaunchFn band_table_5(int w);           // L2, W 145..176
BandLaunchFn band_table_6(int w);           // L2, W 177..208
BandLaunchFn band_table_7(int w);           // L2, W 209..kBandWideMaxW
BandLaunchFn band_table_8(int w);           // L1, W 113..144
BandLaunchFn band_table_9(int w);           // L1, W 145..176
BandLaunchFn band_table_10(int w);          // L1, W 177..208
BandLaunchFn band_table_11(int w);          // L1, W 209..kBandWideMaxW

}  // namespace mdtw
