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
    // z = (x - mu) / sigma as the exact IEEE quotient without a division
    // per element (band_z); the envelope index is wave-uniform, so blocks of
    // 16 read it with scalar loads.
    const float y = 1.0f / sg;
    const bool fast = sg >= 0x1p-100f && sg <= 0x1p100f;
    float acc = 0.0f;
    int32_t i = 0;
    for (; i + 16 <= m; i += 16) {
        const float* __restrict__ lq = lo + i;
        const float* __restrict__ uq = up + i;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const float z = band_z(x[i + t], mu, sg, y, fast);
            const float e = z - __builtin_amdgcn_fmed3f(z, lq[t], uq[t]);
            acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
        }
    }
    for (; i < m; ++i) {
        const float z = band_z(x[i], mu, sg, y, fast);
        const float e = z - __builtin_amdgcn_fmed3f(z, lo[i], up[i]);
        acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
    }
    out[g] = acc;
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
    const float* x = seg + threadIdx.x;
    const float mu = avg[g], sg = sd[g];
    const float y = 1.0f / sg;
    const bool fast = sg >= 0x1p-100f && sg <= 0x1p100f;
    float acc = 0.0f;
    int32_t i = 0;
    for (; i + 16 <= m; i += 16) {
        const float* __restrict__ lq = lo + i;
        const float* __restrict__ uq = up + i;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const float z = band_z(x[i + t], mu, sg, y, fast);
            const float e = z - __builtin_amdgcn_fmed3f(z, lq[t], uq[t]);
            acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
        }
    }
    for (; i < m; ++i) {
        const float z = band_z(x[i], mu, sg, y, fast);
        const float e = z - __builtin_amdgcn_fmed3f(z, lo[i], up[i]);
        acc = L2 ? fmaf(e, e, acc) : acc + fabsf(e);
    }
    out[g] = acc;
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
        const float kim = lb_kim_window<L2, false>(x, q, m, mu, sg, rs);
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
