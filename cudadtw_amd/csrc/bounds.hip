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

// z-normalisation modes: kZnRecip (x - mu) * (1/sigma), the cascade's
// (pruning margin); kZnDiv the IEEE quotient; kZnMark the same quotient
// through band_z_fast (rs = 1/sigma), for windows band_z_window_fast admits.
enum ZnMode { kZnRecip = 0, kZnDiv = 1, kZnMark = 2 };
template <int Z>
__device__ __forceinline__ float zn(float x, float mu, float sg, float rs) {
    if constexpr (Z == kZnDiv)
        return (x - mu) / sg;
    else if constexpr (Z == kZnMark)
        return band_z_fast(x, mu, sg, rs);
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

// bounds.cuh:283-407, the 10-point LB_Kim: an unconstrained DTW over the
// front 5x5 corner (query rows q[0..5), stream columns z[0..5)) entered at
// (0, 0), whose minimum over its exit boundary (last column, last row) starts
// the bound, plus the back 5x5 corner entered anywhere on its top row or left
// column, read at (4, 4).  Needs m >= 11 (bounds.cuh:418).  Cells are
// cost + min(left, min(diagonal, up)), cost then min in that order, so the
// rounding matches the CPU restatement; the tile is fully unrolled into
// registers (five stream values, one row of cells).
template <bool L2, int EXACT, bool FRONT>
__device__ __forceinline__ float kim_corner(const float* __restrict__ x,
                                            const float* __restrict__ q, float mu, float sg,
                                            float rs, float* exit_min) {
    float z[5], row[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) z[c] = zn<EXACT>(x[c], mu, sg, rs);
    float ex = 0.0f;
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const float qr = q[r];
        float left = 0.0f, diag = 0.0f;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const float up = row[c];  // row r - 1 (unused in row 0)
            const float d = qr - z[c];
            float v;
            if (r == 0)
                v = (FRONT && c > 0) ? lpa<L2>(d, left) : lp<L2>(d);
            else if (c == 0)
                v = FRONT ? lpa<L2>(d, up) : lp<L2>(d);
            else
                v = lpa<L2>(d, fminf(left, fminf(diag, up)));
            diag = up;
            left = v;
            row[c] = v;
        }
        if constexpr (FRONT) {
            if (r == 0)
                ex = row[4];
            else if (r < 4)
                ex = fminf(ex, row[4]);
            else
                ex = fminf(ex, fminf(row[0], fminf(row[1], fminf(row[2], fminf(row[3], row[4])))));
        }
    }
    if constexpr (FRONT) *exit_min = ex;
    return row[4];
}

template <bool L2, int EXACT>
__device__ __forceinline__ float lb_kim_window(const float* __restrict__ x,
                                               const float* __restrict__ q, int32_t m,
                                               float mu, float sg, float rs) {
    float front = 0.0f;
    (void)kim_corner<L2, EXACT, true>(x, q, mu, sg, rs, &front);
    const float back = kim_corner<L2, EXACT, false>(x + (m - 5), q + (m - 5), mu, sg, rs, nullptr);
    return front + back;
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
    const float mu = avg[w], sg = sd[w];
    // the ten quotients without a division when the whole wave's windows allow it
    if (__all(band_z_window_fast(mu, sg)))
        out[g] = lb_kim_window<L2, kZnMark>(s + w, q, m, mu, sg, 1.0f / sg);
    else
        out[g] = lb_kim_window<L2, kZnDiv>(s + w, q, m, mu, sg, 0.0f);
}

// LB_Keogh of one window, z = (x - mu) / sigma as the exact IEEE quotient
// without a division per element: band_z, or band_z_fast (no per-element
// test) when FASTZ.  The envelope index is wave-uniform, so blocks of 16
// read it with scalar loads.
template <bool L2, bool FASTZ>
__device__ __forceinline__ float keogh_window(const float* __restrict__ x, float mu, float sg,
                                              const float* __restrict__ lo,
                                              const float* __restrict__ up, int32_t m) {
    const float y = 1.0f / sg;
    const bool fast = sg >= 0x1p-100f && sg <= 0x1p100f;
    auto zof = [&](float v) { return FASTZ ? band_z_fast(v, mu, sg, y) : band_z(v, mu, sg, y, fast); };
    float acc = 0.0f;
    int32_t i = 0;
    for (; i + 16 <= m; i += 16) {
        const float* __restrict__ lq = lo + i;
        const float* __restrict__ uq = up + i;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const float z = zof(x[i + t]);
            const float e = z - __builtin_amdgcn_fmed3f(z, lq[t], uq[t]);
            acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
        }
    }
    for (; i < m; ++i) {
        const float z = zof(x[i]);
        const float e = z - __builtin_amdgcn_fmed3f(z, lo[i], up[i]);
        acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
    }
    return acc;
}

template <bool L2>
__device__ __forceinline__ float keogh_window_any(const float* __restrict__ x, float mu, float sg,
                                                  const float* __restrict__ lo,
                                                  const float* __restrict__ up, int32_t m) {
    // wave-uniform choice: the loop without per-element tests when every
    // active lane's window admits it (almost always)
    if (__all(band_z_window_fast(mu, sg))) return keogh_window<L2, true>(x, mu, sg, lo, up, m);
    return keogh_window<L2, false>(x, mu, sg, lo, up, m);
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
    out[g] = keogh_window_any<L2>(s + w, avg[w], sd[w], lo, up, m);
}

// All windows: a workgroup's 256 consecutive windows read samples
// [g0, g0 + 255 + M) of the stream, staged once in LDS with coalesced loads
// (dynamic shared memory, (255 + M) floats); each lane then reads its
// window from LDS instead of issuing one global load per element.
template <bool L2>
__global__ __launch_bounds__(256) void lb_keogh_all_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ lo, const float* __restrict__ up, int32_t m, int64_t count,
    float* __restrict__ out) {
    extern __shared__ float seg[];
    const int64_t g0 = static_cast<int64_t>(blockIdx.x) * 256;
    const int64_t last = count + m - 2;  // last stream sample any window reads
    const int32_t len = 255 + m;
    for (int32_t k = threadIdx.x; k < len; k += 256) {
        const int64_t gi = g0 + k;
        seg[k] = s[gi < last ? gi : last];
    }
    __syncthreads();
    const int64_t g = g0 + threadIdx.x;
    if (g >= count) return;
    out[g] = keogh_window_any<L2>(seg + threadIdx.x, avg[g], sd[g], lo, up, m);
}

__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Search cascade over windows wbase..wbase+L-1, wave-strided: each wave takes
// 64 consecutive windows per iteration.  Survivors go to cand (order not
// significant: the search result does not depend on it).
//
// The 64 windows of a wave share all but 63 of their samples, so LB_Keogh
// reads the stream from a per-wave ring in LDS instead of one vector load
// per lane and element: lane l at step i needs sample wb + l + i.  The ring
// holds 64-sample chunk c at slots (64c mod 128) + l and, duplicated, at
// + 128, so the steps 64c .. 64c + 63 (which read chunks c and c + 1) read
// slots (64c mod 128) + l + u for u < 64 with no wrap: one ds_read with an
// immediate offset per step.  After those steps chunk c + 2 replaces chunk
// c.  The query envelope is wave-uniform, so each block of 16 steps reads it
// with scalar loads (two 16-value loads) and uses it as a scalar operand:
// 6 VALU per element (sub, mul, a move for med3's second scalar, med3, sub,
// fma).  Broadcasting a lane-distributed copy with readlane instead cost two
// more VALU per element and 13% of the C2 search.  The steps past the last
// whole block of 16 still use the lane-distributed copy (lane l holds
// lo/up[64c + l]) and readlane.  A wave reads only its own ring and LDS
// operations of one wave complete in order; a wave-level fence orders each
// refill before the reads that follow.
constexpr int kCascadeWaves = 4;
constexpr int kRing = 256;
// Survivors are staged per wave in LDS and appended kCandBuf - 64 or more at
// a time: one atomicAdd on the shared candidate counter per flush instead of
// one per 64 windows (same-address atomics serialise in L2, ~15 ns each).
// The last flush and the prune counters are summed per block first: three
// atomics per block, not per wave.  A wave takes kCascadeChunks chunks of 64
// windows where the range fills the chip twice over with that many, so a
// medium range (a streaming epoch) does not pay the atomics of a grid sized
// for a long one; a shorter range gets fewer chunks per wave (down to one),
// because there the waves' serial latency, not the atomics, sets the time.
constexpr int kCandBuf = 512;
constexpr int64_t kCascadeChunks = 16;

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <bool L2>
__device__ __forceinline__ float keogh_step(float xv, float mu, float rs, float l, float u,
                                            float acc) {
    const float z = (xv - mu) * rs;
    const float e = z - __builtin_amdgcn_fmed3f(z, l, u);
    return L2 ? fmaf(e, e, acc) : acc + fabsf(e);
}

template <bool L2>
__global__ __launch_bounds__(kCascadeWaves * kWave) void cascade_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, const float* __restrict__ lo, const float* __restrict__ up,
    int32_t m, int64_t wbase, int64_t L, bool use_best, Ctrl* ctrl,
    int32_t* __restrict__ cand, float* __restrict__ cand_lb) {
    __shared__ float ring_all[kCascadeWaves][kRing];
    __shared__ int32_t cbuf_all[kCascadeWaves][kCandBuf];
    __shared__ float lbuf_all[kCascadeWaves][kCandBuf];
    const int lane = threadIdx.x & (kWave - 1);
    float* ring = ring_all[threadIdx.x / kWave];
    int32_t* cbuf = cbuf_all[threadIdx.x / kWave];
    float* lbuf = lbuf_all[threadIdx.x / kWave];
    int32_t nbuf = 0;
    auto flush = [&]() {
        wave_lds_fence();
        unsigned long long basec = 0;
        if (lane == 0)
            basec = atomicAdd(&ctrl->count, static_cast<unsigned long long>(nbuf));
        basec = __shfl(basec, 0, kWave);
        for (int32_t t = lane; t < nbuf; t += kWave) {
            cand[basec + t] = cbuf[t];
            if (cand_lb) cand_lb[basec + t] = lbuf[t];
        }
        wave_lds_fence();
        nbuf = 0;
    };
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    const float thr = use_best ? load_best_dist(ctrl) : load_thr(ctrl);
    const float lim = prune_limit(thr, m, L2);
    // Last stream sample any window of this range reads.
    const int64_t last = wbase + L - 1 + m - 1;
    unsigned long long n_kim = 0, n_keogh = 0;

    for (int64_t base = wave * kWave; base < L; base += nwaves * kWave) {
        const bool valid = base + lane < L;
        const int64_t wb = wbase + base;  // first window of the wave
        const int64_t w = wb + lane;
        const int64_t wc = valid ? w : wbase + L - 1;
        const float* x = s + wc;
        const float mu = avg[wc], sg = sd[wc];
        const float rs = 1.0f / sg;
        const float kim = lb_kim_window<L2, kZnRecip>(x, q, m, mu, sg, rs);
        bool alive = valid && !(kim > lim);
        n_kim += __popcll(__ballot(valid && !alive));
        if (!__any(alive)) continue;

        auto fill = [&](int32_t c) {  // samples wb + 64c .. wb + 64c + 63
            const int64_t gi = wb + static_cast<int64_t>(c) * kWave + lane;
            const float v = s[gi < last ? gi : last];
            const int slot = (c & 1) * kWave + lane;
            ring[slot] = v;
            ring[slot + 2 * kWave] = v;
        };
        fill(0);
        fill(1);
        wave_lds_fence();
        float acc = 0.0f;
        bool gone = false;
        for (int32_t i0 = 0; i0 < m && !gone; i0 += kWave) {
            const int32_t ie = lane + i0;
            const float lv = ie < m ? lo[ie] : 0.0f;
            const float uv = ie < m ? up[ie] : 0.0f;
            const float* rp = ring + (i0 & kWave) + lane;  // (64c mod 128) + lane
            const int32_t nstep = m - i0 < kWave ? m - i0 : kWave;
            int32_t u = 0;
            for (; u + 16 <= nstep; u += 16) {
                // envelope of these 16 steps: wave-uniform scalar loads
                const float* __restrict__ lq = lo + (i0 + u);
                const float* __restrict__ uq = up + (i0 + u);
#pragma unroll
                for (int t = 0; t < 16; ++t)
                    acc = keogh_step<L2>(rp[u + t], mu, rs, lq[t], uq[t], acc);
                if (!__any(alive && !(acc > lim))) {
                    gone = true;
                    break;
                }
            }
            if (gone) break;
            for (; u < nstep; ++u) acc = keogh_step<L2>(rp[u], mu, rs, bcast(lv, u), bcast(uv, u), acc);
            if (i0 + kWave < m) {
                fill(i0 / kWave + 2);  // replaces the chunk just consumed
                wave_lds_fence();
            }
        }
        const bool keep = alive && !gone && !(acc > lim);
        const unsigned long long km = __ballot(keep);
        n_keogh += __popcll(__ballot(alive && !keep));
        if (km) {
            if (keep) {
                cbuf[nbuf + lane_rank(km)] = static_cast<int32_t>(w);
                lbuf[nbuf + lane_rank(km)] = acc;
            }
            nbuf += __popcll(km);
            if (nbuf > kCandBuf - kWave) flush();
        }
    }
    // Block epilogue: one reservation for every wave's remaining survivors.
    __shared__ unsigned long long part[kCascadeWaves][3];
    __shared__ unsigned long long bbase;
    const int wv = threadIdx.x / kWave;
    if (lane == 0) {
        part[wv][0] = static_cast<unsigned long long>(nbuf);
        part[wv][1] = n_kim;
        part[wv][2] = n_keogh;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tc = 0, tk = 0, te = 0;
        for (int v = 0; v < kCascadeWaves; ++v) {
            tc += part[v][0];
            tk += part[v][1];
            te += part[v][2];
        }
        bbase = tc ? atomicAdd(&ctrl->count, tc) : 0ull;
        if (tk) atomicAdd(&ctrl->kim, tk);
        if (te) atomicAdd(&ctrl->keogh, te);
    }
    __syncthreads();
    unsigned long long off = bbase;
    for (int v = 0; v < wv; ++v) off += part[v][0];
    for (int32_t t = lane; t < nbuf; t += kWave) {
        cand[off + t] = cbuf[t];
        if (cand_lb) cand_lb[off + t] = lbuf[t];
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
    if (!idx) {
        const size_t lds = sizeof(float) * static_cast<size_t>(255 + m);
        if (metric)
            hipLaunchKernelGGL(lb_keogh_all_kernel<true>, dim3(nb), dim3(256), lds, st, s, avg,
                               sd, lo, up, m, count, out);
        else
            hipLaunchKernelGGL(lb_keogh_all_kernel<false>, dim3(nb), dim3(256), lds, st, s, avg,
                               sd, lo, up, m, count, out);
        return check_launch();
    }
    if (metric)
        hipLaunchKernelGGL(lb_keogh_kernel<true>, dim3(nb), dim3(256), 0, st, s, avg, sd, lo,
                           up, m, idx, count, out);
    else
        hipLaunchKernelGGL(lb_keogh_kernel<false>, dim3(nb), dim3(256), 0, st, s, avg, sd, lo,
                           up, m, idx, count, out);
    return check_launch();
}

int launch_cascade(hipStream_t st, const float* s, const float* avg, const float* sd,
                   const float* q, const float* lo, const float* up, int32_t m, int64_t base,
                   int64_t L, int32_t metric, bool use_best, Ctrl* ctrl, int32_t* cand,
                   float* cand_lb, int64_t max_blocks, int cus) {
    if (L <= 0) return MDTW_OK;
    // Chunks of 64 windows per wave: kCascadeChunks once the range gives two
    // waves per SIMD with that many, fewer on shorter ranges; the grid is
    // capped at max_blocks and each wave strides.
    const int64_t chunks = (L + kWave - 1) / kWave;
    const int64_t target_waves = 2 * kCascadeWaves * static_cast<int64_t>(cus > 0 ? cus : 1);
    const int64_t cpw = std::min<int64_t>(
        kCascadeChunks, std::max<int64_t>(1, (chunks + target_waves - 1) / target_waves));
    const int64_t waves_needed = (chunks + cpw - 1) / cpw;
    const int64_t nb64 = std::max<int64_t>(
        1, std::min<int64_t>((waves_needed + kCascadeWaves - 1) / kCascadeWaves, max_blocks));
    const unsigned nb = static_cast<unsigned>(nb64);
    if (metric)
        hipLaunchKernelGGL(cascade_kernel<true>, dim3(nb), dim3(256), 0, st, s, avg, sd, q, lo,
                           up, m, base, L, use_best, ctrl, cand, cand_lb);
    else
        hipLaunchKernelGGL(cascade_kernel<false>, dim3(nb), dim3(256), 0, st, s, avg, sd, q,
                           lo, up, m, base, L, use_best, ctrl, cand, cand_lb);
    return check_launch();
}

}  // namespace mdtw
