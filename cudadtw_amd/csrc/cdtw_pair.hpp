// cdtw_pair.hpp -- banded cDTW for wide bands, one window's band split over
// two waves of a workgroup (a pair) or, for W = 159..225, three (a trio; see
// trio_dp below) (included by the wide-band instantiation units).
//
// The register-band kernel (cdtw_band.hpp) holds the 2W + 1 band values of a
// window in one lane.  Beyond W ~ 114 that no longer fits the 256 registers a
// lane has at two waves per SIMD, and at one wave per SIMD the SIMD issues at
// ~60% of its two-wave rate.  Here the band is split by rows: the lower wave
// of a pair holds t = 0..W (rows j-W..j of column j), the upper wave
// t = W+1..2W, with the same lane owning the same window in both, so each
// wave needs about W + 1 registers for its half and two (up to W ~ 230) or
// three waves share a SIMD.
//
// Dependencies across the split.  E[t] of column j needs E[t-1] of column j
// (new), E[t] and E[t+1] of column j-1 (old).  Upper cell t = W+1 of column j
// needs lower cell t = W of column j; lower cell t = W of column j needs
// upper cell t = W+1 of column j-1.  So at step s the lower wave updates
// column s and the upper wave column s-1:
//     lower: cells t = 0..W-1 of column s          upper: cells W+1..2W of
//                                                   column s-1 (its first
//                                                   cell from the lower's
//                                                   t = W of column s-1)
//     ---- barrier X ----
//     lower: cell t = W of column s (its right neighbour is the upper's first
//            cell of column s-1, written before X)
//     ---- barrier Y ----
// with the two boundary values passed through LDS (one float per lane each
// way per step).  Each wave's instruction stream is the register-band
// kernel's (v_sub with the query as a wave-uniform scalar, v_min3, v_fmac per
// cell); both waves run the same cell loop, with their own query offset,
// column and first-cell input.  The barriers are wave-level (s_barrier in
// code both waves execute), so the two waves stay one column apart.
//
// Early abandoning.  At step s the lower wave's column s (rows <= s) and the
// upper wave's column s-1 (rows >= s) form a cut every warping path crosses:
// a path leaving column s-1 from a row below s enters column s at a row <= s.
// Their combined minimum above the bound abandons the window, exactly as the
// column minimum does in the one-lane kernel.  The lower wave decides (it
// holds the answer, E[W] after column M-1) and passes the decision to the
// upper wave through LDS, so both leave the loop at the same barrier.
//
// The arithmetic per cell is the one-lane kernel's, in the same order, so
// the distances are bit-identical to it and to the oracle's DP.
//
// Part of cdtw_band.hpp (included there ahead of the launch code, whose
// band_launch picks these kernels for W >= kBandPairMinW).
#pragma once

namespace mdtw {

constexpr int kPairWaves = 2;  // one pair per workgroup
constexpr int kTrioWaves = 3;  // or one trio
// Up to W = 225 the compiler fits a half band and the rest in 256 registers
// (two waves per SIMD; three to W ~ 137); beyond, left to itself, it takes
// a few more and one wave per SIMD, so those widths are compiled for two
// waves (band_pair_kernel2), spilling 12..60 bytes per lane outside the
// cell loop: measured at M = 1024, 15.6..15.8 T cells/s at W = 228..240
// (13.1 at 236) against 8 at one wave.
constexpr bool band_pair_tight(int w) { return w > 225; }
// From W = 138 a half band and the rest take a few more than 168 registers,
// and left to itself the compiler drops to two waves per SIMD (W = 137 at
// 17.4 T cells/s, 138 at 14.5).  Up to W = 158 those widths are compiled for
// three waves (band_pair_kernel3), spilling a few values outside the cell
// loop: measured at M = 1024, 2^22 windows, W = 138 / 145 / 150 / 155 / 158
// at 17.6 / 17.7 / 17.8 / 16.6 / 16.5 T cells/s against 14.2 / 14.3 / 15.1 /
// 14.9 / 15.4 at two (W = 159 even, beyond slower).
constexpr bool band_pair_three(int w) { return w >= 138 && w <= 158; }
// From W = 159 to 225 the band is split over three waves (band_trio_kernel):
// a third of the band and the rest fit 168 registers (three waves per SIMD)
// up to W = 204, where a pair's half runs two; W = 205..225 are compiled for
// three waves (band_trio_kernel3), spilling 3..14 values outside the cell
// loop.  Measured at M = 1024, 2^22 windows, interleaved on one device,
// against the pair: W = 165 / 180 / 200 15.4 / 15.0 / 15.8 -> 15.5-15.7 /
// 15.7-15.9 / 16.1 T cells/s, W = 205 / 210 / 215 / 225 14.9-15.4 /
// 14.6-15.4 / 15.1-15.7 / 15.4-16.2 -> 15.9 / 15.9 / 16.1 / 15.9; at W = 230
// both run 15.6, and beyond the trio's spills grow (W = 240: 14.7).
constexpr bool band_trio(int w) { return w >= 159 && w <= 225; }
constexpr bool band_trio_three(int w) { return w >= 205 && w <= 225; }

struct PairLds {
    float lo_last[kWave];   // lower: E[W] of its column (the upper's next first-cell input)
    float up_first[kWave];  // upper: E[W+1] of its column (the lower's last-cell input)
    float up_min[kWave];    // upper: its cells' minimum at a check step
    int stop;               // lower: every window abandoned at a check step
    int drained;            // persistent cursor seen past the end
    long long next;         // persistent: the pair's next 64 windows
};

__device__ __forceinline__ void pair_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Cells t = 0..W-1 of one half (lower: rows j-W.., upper: rows j+1..) with
// the query values qc[0..W) (wave-uniform scalars, loaded in chunks a chunk
// ahead), z of the column, and the first cell's upper neighbour `prev`.
// E[W] is the lower wave's cell t = W of the previous column (the right
// neighbour of its cell W-1), or +inf in the upper wave.  No scalar load is
// left in flight at the end: the barrier that follows waits for every
// outstanding scalar and LDS access (one counter covers both), so a chunk
// requested for the next step here measured slower (W = 138..240: 14.2-15.9
// -> 12.8-15.3 T cells/s at M = 1024).
template <int W, bool L2>
__device__ __forceinline__ void pair_cells(float (&E)[W + 1], const float* __restrict__ qc,
                                           float zj, float prev) {
    constexpr int CH = kBandChunk;
    constexpr int NC = (W + CH - 1) / CH;
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
        // keep the scalar loads in chunks (hoisting them all would exceed
        // the scalar register file)
        __builtin_amdgcn_sched_barrier(0x96);
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int t = ch * CH + i;
            if (t < W) {
                const float mn = fminf(fminf(prev, E[t]), E[t + 1]);
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

template <int W, bool L2>
__device__ __forceinline__ float pair_min(const float (&E)[W + 1]) {
    float cm = E[0];
#pragma unroll
    for (int t = 1; t <= W; ++t) cm = fminf(cm, E[t]);
    return cm;
}

// The DP of one window per lane; both waves of the pair call it with the
// same window.  Returns the distance in the lower wave (meaningless in the
// upper); `alive` is updated in the lower wave.
template <int W, bool L2>
__device__ __forceinline__ float pair_dp(PairLds& sh, bool lower, const float* __restrict__ x,
                                         float mu, float sg, const float* __restrict__ qx,
                                         int32_t m, bool& alive, int mode, const Ctrl* ctrl) {
    const int lane = threadIdx.x & (kWave - 1);
    float E[W + 1];
#pragma unroll
    for (int t = 0; t <= W; ++t) E[t] = kInf;
    if (lower) E[W] = 0.0f;  // D[-1][-1]

    const float y = 1.0f / sg;
    const bool wfast = band_z_window_fast(mu, sg);
    const bool abandon = mode == kModeBest || mode == kModeCollect;
    // lower: column s at step s; upper: column s - 1
    const int32_t lag = lower ? 0 : 1;
    // the query rows of cell t = 0 of each half: lower j - W, upper j + 1;
    // in the padded query (W sentinels first) qx + j and qx + j + W + 1
    const int32_t qoff = lower ? 0 : W + 1;
    float xv = x[0];
    for (int32_t s = 0; s < m; ++s) {
        const int32_t j = s - lag;
        const bool check = abandon && (s % kBandCheck) == kBandCheck - 1;
        float thr = kInf;
        if (check && lower) thr = mode == kModeBest ? load_best_dist(ctrl) : load_thr(ctrl);
        if (j >= 0) {
            // the lower's cell t = W: requested before the cells, so that
            // its wait is not between the two barriers (where the upper
            // wave idles)
            const float qw = qx[j + W];
            const float zj = band_z_w(xv, mu, sg, y, wfast);
            xv = x[j + 1 < m ? j + 1 : m - 1];
            const float prev = lower ? kInf : sh.lo_last[lane];
            pair_cells<W, L2>(E, qx + j + qoff, zj, prev);
            if (!lower) {
                sh.up_first[lane] = E[0];
                if (check) sh.up_min[lane] = pair_min<W, L2>(E);
            } else {
                // the last cell needs the upper's first cell of column j - 1
                pair_sync();
                const float r = j > 0 ? sh.up_first[lane] : kInf;
                const float mn = fminf(fminf(E[W - 1], E[W]), r);
                const float d = qw - zj;
                E[W] = L2 ? fmaf(d, d, mn) : fabsf(d) + mn;
                sh.lo_last[lane] = E[W];
                if (check) {
                    const float cm = fminf(pair_min<W, L2>(E), sh.up_min[lane]);
                    alive = alive && !(cm > thr);
                    // the vote over the whole wave, then one lane stores it
                    const bool any = __any(alive);
                    if (lane == 0) sh.stop = !any;
                }
                pair_sync();
            }
        }
        if (!lower || j < 0) {
            pair_sync();
            pair_sync();
        }
        if (check && __builtin_amdgcn_readfirstlane(sh.stop)) return kInf;
    }
    return E[W];
}

// Three waves per window (a trio).  The 2W + 1 band values are split by rows
// into three parts: wave k holds global cells k(P+1) .. k(P+1)+P, P + 1 of
// them, except the top wave, which holds the last TP = 2W - 1 - 2P (<= P).
// Each wave holds P + 1 registers of band, about two thirds of a pair's half,
// so three waves share a SIMD to larger W.  Wave k updates column s - k at
// step s, and the pair's two phases apply between every two neighbours:
//     phase 1: every wave computes its cells 0..P-1 (the wave above k takes
//              its first cell's lower neighbour from k's cell P of the
//              same column, written at the previous step) and the waves
//              above the bottom publish their cell 0
//     ---- barrier X ----
//     phase 2: the bottom and middle waves compute their cell P from the
//              published cell 0 of the wave above (its column is one less)
//     ---- barrier Y ----
// The top wave runs the same P cells in phase 1; its cells TP..P-1 lie past
// the band's end and are set back to +inf after each column (their inputs
// are finite query rows), so its real cells see +inf above them as in the
// one-lane kernel.  The band centre, cell W, is the middle wave's cell
// W - P - 1 < P, so the distance is the middle wave's after column M - 1,
// at step M, and passes through LDS to the bottom wave, which writes the
// results.
//
// Abandoning: at a check step s the bottom's column s, the middle's column
// s - 1 and the top's column s - 2 form a cut that every warping path
// crosses (as the pair's two columns do: a path that stays below one part's
// rows in its column enters the next part's column at or below that part's
// first row).  The middle and bottom parts are complete only after phase 2,
// so each wave stores its part's minimum at step s and the bottom decides at
// step s + 1 before barrier Y; all three waves leave after that barrier.  A
// window abandoned a column later than it could be is still one whose
// distance exceeds the bound, so the results are unchanged.
template <int W>
struct TrioShape {
    static constexpr int P = (2 * W - 1 + 2) / 3;  // cells per wave in phase 1
    static constexpr int TP = 2 * W - 1 - 2 * P;   // the top wave's band cells
    static constexpr int C = W - (P + 1);          // cell W in the middle wave
    static_assert(TP >= 1 && TP <= P && C >= 0 && C < P, "trio split");
};

struct TrioLds {
    float lo_last[2][kWave];   // wave k: its cell P (wave k + 1's first-cell input)
    float up_first[3][kWave];  // wave k: its cell 0 (wave k - 1's cell P input; row 0 unread)
    float part_min[3][kWave];  // each wave's minimum at a check step
    float dist[kWave];         // the middle wave's distance
    int stop;
    int drained;
    long long next;
};

template <int W, bool L2>
__device__ __forceinline__ float trio_dp(TrioLds& sh, int k, const float* __restrict__ x,
                                         float mu, float sg, const float* __restrict__ qx,
                                         int32_t m, bool& alive, int mode, const Ctrl* ctrl) {
    using S = TrioShape<W>;
    constexpr int P = S::P;
    const int lane = threadIdx.x & (kWave - 1);
    const bool top = k == 2;
    float E[P + 1];
#pragma unroll
    for (int t = 0; t <= P; ++t) E[t] = kInf;
    if (k == 1) E[S::C] = 0.0f;  // D[-1][-1]

    const float y = 1.0f / sg;
    const bool wfast = band_z_window_fast(mu, sg);
    const bool abandon = mode == kModeBest || mode == kModeCollect;
    // cell 0 of wave k is band cell k (P + 1): row j - W + k (P + 1), in the
    // padded query (W sentinels first) qx + j + k (P + 1)
    const int32_t qoff = k * (P + 1);
    float xv = x[0];
    float thr = kInf;
    for (int32_t s = 0; s <= m; ++s) {
        const int32_t j = s - k;
        const bool check = abandon && s < m && (s % kBandCheck) == kBandCheck - 1;
        const bool decide = abandon && s > 0 && ((s - 1) % kBandCheck) == kBandCheck - 1;
        if (check && k == 0) thr = mode == kModeBest ? load_best_dist(ctrl) : load_thr(ctrl);
        const bool act = j >= 0 && j < m;
        float zj = 0.0f, qw = 0.0f;
        if (act) {
            qw = qx[j + qoff + P];
            zj = band_z_w(xv, mu, sg, y, wfast);
            xv = x[j + 1 < m ? j + 1 : m - 1];
            const float prev = k == 0 ? kInf : sh.lo_last[k - 1][lane];
            pair_cells<P, L2>(E, qx + j + qoff, zj, prev);
#pragma unroll
            for (int t = S::TP; t < P; ++t) E[t] = top ? kInf : E[t];
            // every wave stores (the bottom's row is never read): a store
            // under a wave condition here splits the cell loop's block, and
            // the compiler then hoists every chunk's query loads above it
            sh.up_first[k][lane] = E[0];
            if (check && top) sh.part_min[2][lane] = pair_min<P, L2>(E);
        }
        pair_sync();
        if (act && !top) {
            // cell P: its right neighbour is the upper wave's cell 0 of column j - 1
            const float r = j > 0 ? sh.up_first[k + 1][lane] : kInf;
            const float mn = fminf(fminf(E[P - 1], E[P]), r);
            const float d = qw - zj;
            E[P] = L2 ? fmaf(d, d, mn) : fabsf(d) + mn;
            sh.lo_last[k][lane] = E[P];
            if (check) sh.part_min[k][lane] = pair_min<P, L2>(E);
        }
        if (decide && k == 0) {
            const float cm =
                fminf(fminf(sh.part_min[0][lane], sh.part_min[1][lane]), sh.part_min[2][lane]);
            alive = alive && !(cm > thr);
            const bool any = __any(alive);
            if (lane == 0) sh.stop = !any;
        }
        pair_sync();
        if (decide && __builtin_amdgcn_readfirstlane(sh.stop)) return kInf;
    }
    if (k == 1) sh.dist[lane] = E[S::C];
    pair_sync();
    return sh.dist[lane];
}

template <int W, bool L2, int NW>
__device__ __forceinline__ void pair_kernel_body(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ qx, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, BandScalars a) {
    static_assert(NW == 2 || NW == 3, "a pair or a trio");
    using Lds = std::conditional_t<NW == 2, PairLds, TrioLds>;
    __shared__ Lds sh;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const bool lower = wave == 0;
    int64_t count = a.count;
    if (a.idx_count_from_ctrl)
        count = static_cast<int64_t>(
            __hip_atomic_load(&ctrl->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    unsigned long long n_dp = 0, n_ab = 0;
    uint32_t* const bmin = ctrl->bmin;
    const long long bucket_len = ctrl->bucket_len;
    if (threadIdx.x == 0) {
        sh.drained = 0;
        sh.stop = 0;
    }
    pair_sync();
    // As the one-lane kernel: the first 64 windows by grid position, then
    // the cursor (persistent) or a grid stride.  The lower wave takes from
    // the cursor and passes the base to the upper wave.
    const int64_t step = static_cast<int64_t>(gridDim.x) * kWave;
    int64_t base = static_cast<int64_t>(blockIdx.x) * kWave;
    while (base < count) {
        const int64_t g = base + lane;
        const bool valid = g < count;
        const int64_t gc = valid ? g : count - 1;
        const int64_t w = idx ? static_cast<int64_t>(idx[gc]) : a.base + gc * a.stride;
        const float mu = avg[w], sg = sd[w];
        bool alive = valid;
        float d;
        if constexpr (NW == 2)
            d = pair_dp<W, L2>(sh, lower, s + w, mu, sg, qx, a.m, alive, a.mode, ctrl);
        else
            d = trio_dp<W, L2>(sh, wave, s + w, mu, sg, qx, a.m, alive, a.mode, ctrl);
        if (lower) {
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
            if (a.persistent) {
                const int64_t nb = band_next(ctrl, lane, step, count, &sh.drained);
                if (lane == 0) sh.next = nb;
            }
            // sh.stop is not cleared here: every step that reads it writes
            // it first (the lower or bottom wave, before the barrier the
            // reads follow), and a clear here could land before a slower
            // wave's read of the previous decision
        }
        pair_sync();
        base = a.persistent ? static_cast<int64_t>(sh.next) : base + step;
        base = static_cast<int64_t>(
            (static_cast<unsigned long long>(__builtin_amdgcn_readfirstlane(
                 static_cast<uint32_t>(static_cast<unsigned long long>(base) >> 32)))
             << 32) |
            __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(base)));
        // the next write of sh.next is after this window batch's barriers
    }
    if (lower && a.mode != kModeProfile && lane == 0) {
        if (n_dp) atomicAdd(&ctrl->dp, n_dp);
        if (n_ab) atomicAdd(&ctrl->abandoned, n_ab);
    }
}

#define MDTW_PAIR_KERNEL_ARGS                                                                 \
    const float *__restrict__ s, const float *__restrict__ avg, const float *__restrict__ sd,   \
        const float *__restrict__ qx, const int32_t *__restrict__ idx, float *__restrict__ out, \
        unsigned long long *__restrict__ res, Ctrl *__restrict__ ctrl, BandScalars a

template <int W, bool L2>
__global__ __launch_bounds__(kPairWaves * kWave) void band_pair_kernel(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2, 2>(s, avg, sd, qx, idx, out, res, ctrl, a);
}

// Compiled for two waves per SIMD (256 registers per lane).
template <int W, bool L2>
__global__ __launch_bounds__(kPairWaves * kWave) __attribute__((amdgpu_waves_per_eu(2, 2))) void
band_pair_kernel2(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2, 2>(s, avg, sd, qx, idx, out, res, ctrl, a);
}
// Compiled for three waves per SIMD (168 registers per lane).
template <int W, bool L2>
__global__ __launch_bounds__(kPairWaves * kWave) __attribute__((amdgpu_waves_per_eu(3, 3))) void
band_pair_kernel3(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2, 2>(s, avg, sd, qx, idx, out, res, ctrl, a);
}
// The trio: three waves per window; and compiled for three waves per SIMD
// (168 registers per lane).
template <int W, bool L2>
__global__ __launch_bounds__(kTrioWaves * kWave) void band_trio_kernel(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2, 3>(s, avg, sd, qx, idx, out, res, ctrl, a);
}
template <int W, bool L2>
__global__ __launch_bounds__(kTrioWaves * kWave) __attribute__((amdgpu_waves_per_eu(3, 3))) void
band_trio_kernel3(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2, 3>(s, avg, sd, qx, idx, out, res, ctrl, a);
}
#undef MDTW_PAIR_KERNEL_ARGS

}  // namespace mdtw
