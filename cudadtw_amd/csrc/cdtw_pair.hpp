// cdtw_pair.hpp -- banded cDTW for wide bands, one window's band split over
// two waves of a workgroup (included by the wide-band instantiation units).
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

template <int W, bool L2>
__device__ __forceinline__ void pair_kernel_body(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ qx, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, BandScalars a) {
    __shared__ PairLds sh;
    const int lane = threadIdx.x & (kWave - 1);
    const bool lower = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave) == 0;
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
        const float d = pair_dp<W, L2>(sh, lower, s + w, mu, sg, qx, a.m, alive, a.mode, ctrl);
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
            if (lane == 0) sh.stop = 0;
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
    pair_kernel_body<W, L2>(s, avg, sd, qx, idx, out, res, ctrl, a);
}

// Compiled for two waves per SIMD (256 registers per lane).
template <int W, bool L2>
__global__ __launch_bounds__(kPairWaves * kWave) __attribute__((amdgpu_waves_per_eu(2, 2))) void
band_pair_kernel2(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2>(s, avg, sd, qx, idx, out, res, ctrl, a);
}
// Compiled for three waves per SIMD (168 registers per lane).
template <int W, bool L2>
__global__ __launch_bounds__(kPairWaves * kWave) __attribute__((amdgpu_waves_per_eu(3, 3))) void
band_pair_kernel3(MDTW_PAIR_KERNEL_ARGS) {
    pair_kernel_body<W, L2>(s, avg, sd, qx, idx, out, res, ctrl, a);
}
#undef MDTW_PAIR_KERNEL_ARGS

}  // namespace mdtw
