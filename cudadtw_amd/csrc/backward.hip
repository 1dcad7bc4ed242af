// backward.hip -- query gradient of banded cDTW distances.
//
// No reference counterpart: the reference computes distances only
// (cdtw.cuh:23-88 is the recurrence differentiated here).  For each window
// p with upstream gradient g[p] the banded DP is recomputed (no checkpoint
// from the forward), the argmin predecessor of every cell is kept as two
// bits, and the optimal warping path is walked back from (M-1, M-1); every
// path cell (i, j) adds d cost / d q_i = 2 (q_i - z_j) (L2) or
// sign(q_i - z_j) (L1) to row i of the query gradient.  The per-window rows
// are summed in a fixed order (deterministic) and, when the raw query is
// given, carried through its z-normalisation:
//   dL/dq = (G - mean(G) - qhat * mean(G * qhat)) / sigma_q.
//
// Kernel layout.  One wave per window, the anti-diagonal layout of the wide
// cDTW kernel (cdtw.hip): diagonal k = i + j, slot p holds offset
// d = i - j = -W + 2p + par(k), par(k) = (k + W) & 1; lane l owns slots
// l*C .. l*C+C-1.  The query and the window's z-normalised values sit in
// LDS (2M floats), so each slot reads its own q_i and z_j.  The cell
// arithmetic is the forward's (q - z, fmaf(d, d, min)), so the recomputed
// distance is bit-identical to mdtw_cdtw's.  Direction codes (0 = (i-1,j-1),
// 1 = (i-1,j), 2 = (i,j-1); ties: diagonal if not above either other, then
// (i-1,j) if not above (i,j-1)) of 16 consecutive diagonals are packed into
// one 32-bit word per slot, word (k/16, p) at ((k/16) * 64C + (p % C) * 64 +
// p / C): coalesced stores.  The words live in LDS behind the query and the
// window when all of them fit in kBwdLdsBytes (about 32 C M + 8 M bytes: M up
// to 1638 at W < 64, 910 at W < 128, 481 at W < 256), else in a scratch
// buffer in HBM.  A batch of queries runs every window of the queries that
// fit in one launch per class C (bwd_path_batch_kernel), so thousands of
// one-wave walks proceed at once.  The walk back is serial, one
// dependent word read per path cell, so an LDS read (about 100 cycles) in
// place of an L2 round trip is what sets its speed.
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace mdtw {
namespace {

constexpr int kDirGroup = 16;  // diagonals per direction word (2 bits each)
constexpr size_t kBwdChunkBytes = size_t(256) << 20;  // scratch per launch
constexpr size_t kBwdLdsBytes = size_t(64) << 10;     // q, z and direction words in LDS

struct BwdScalars {
    int32_t m;
    int32_t w;
    int32_t sqrt_dist;
    int32_t pair_stats;  // 1: avg/sd hold one entry per window of the list
    int64_t words;       // direction words per window
};

// One window: recompute the banded DP, walk the optimal path back and write
// the window's gradient row go[0..m).  lds holds q and z (2m floats) and,
// when LDIR, the direction words behind them; otherwise dw (HBM) does.
template <int C, bool L2, bool LDIR>
__device__ __forceinline__ void bwd_window(const float* __restrict__ x, float mu, float sg,
                                           const float* __restrict__ q, int m, int W, float gin,
                                           int sqrt_dist, uint32_t* __restrict__ dw,
                                           float* __restrict__ go, float* __restrict__ dist,
                                           float* lds) {
    const int lane = threadIdx.x;
    float* lq = lds;
    float* lz = lds + m;
    for (int t = lane; t < m; t += kWave) {
        lq[t] = q[t];
        lz[t] = (x[t] - mu) / sg;
    }
    __syncthreads();

    if (LDIR) dw = reinterpret_cast<uint32_t*>(lz + m);
    const int P = W + 1;
    const int pfin = W >> 1;
    float pad0[C], pad1[C], A1[C], A2[C];
    uint32_t bits[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int p = lane * C + c;
        pad0[c] = p < P ? 0.0f : kInf;
        pad1[c] = p < P - 1 ? 0.0f : kInf;
        A1[c] = kInf;
        A2[c] = p == pfin ? 0.0f : kInf;  // D[-1][-1] = 0 on diagonal -2
        bits[c] = 0u;
    }
    const int K = 2 * m - 1;
    for (int k = 0; k < K; ++k) {
        const int par = (k + W) & 1;
        float nw[C];
        float nb_in;
        if (par == 0) {
            nb_in = __shfl_up(A1[C - 1], 1, kWave);
            if (lane == 0) nb_in = kInf;
        } else {
            nb_in = __shfl_down(A1[0], 1, kWave);
            if (lane == kWave - 1) nb_in = kInf;
        }
        const int sh = 2 * (k & (kDirGroup - 1));
#pragma unroll
        for (int c = 0; c < C; ++c) {
            // par 0: (i-1, j) is slot p-1 and (i, j-1) slot p of diagonal k-1;
            // par 1: (i-1, j) is slot p and (i, j-1) slot p+1.
            float up, lf;
            if (par == 0) {
                up = c > 0 ? A1[c - 1] : nb_in;
                lf = A1[c];
            } else {
                up = A1[c];
                lf = c + 1 < C ? A1[c + 1] : nb_in;
            }
            const float dg = A2[c];
            const int p = lane * C + c;
            const int d = -W + 2 * p + par;
            const int i = (k + d) >> 1;
            const int j = (k - d) >> 1;
            const bool iq = i >= 0 && i < m;
            const bool jz = j >= 0 && j < m;
            const float qv = iq ? lq[iq ? i : 0] : kInf;
            const float zv = jz ? lz[jz ? j : 0] : -kInf;
            const float dd = qv - zv;
            const float mn = fminf(fminf(dg, up), lf);
            const float v = L2 ? fmaf(dd, dd, mn) : fabsf(dd) + mn;
            nw[c] = fmaxf(v, par ? pad1[c] : pad0[c]);
            const uint32_t dr = (dg <= up && dg <= lf) ? 0u : (up <= lf ? 1u : 2u);
            bits[c] |= dr << sh;
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            A2[c] = A1[c];
            A1[c] = nw[c];
        }
        if ((k & (kDirGroup - 1)) == kDirGroup - 1 || k == K - 1) {
            uint32_t* row = dw + static_cast<int64_t>(k / kDirGroup) * (kWave * C);
#pragma unroll
            for (int c = 0; c < C; ++c) {
                row[c * kWave + lane] = bits[c];
                bits[c] = 0u;
            }
        }
    }
    // D[M-1][M-1]: diagonal 2M-2, offset 0 -> slot W >> 1.
    float r = A1[0];
#pragma unroll
    for (int c = 1; c < C; ++c)
        if (c == pfin % C) r = A1[c];
    const float D = __shfl(r, pfin / C, kWave);
    // Direction words stored above are read by lane 0 below.
    if (LDIR)
        __syncthreads();
    else
        __threadfence();
    if (lane != 0) return;

    float scale = gin;
    if (sqrt_dist) scale = D > 0.0f ? scale * (0.5f / sqrtf(D)) : 0.0f;
    if (dist) *dist = sqrt_dist ? sqrtf(D) : D;
    int i = m - 1, j = m - 1;
    float acc = 0.0f;
    // The path has at most 2M - 1 cells; every step lowers i + j.
    for (int step = 0; step < K; ++step) {
        const float dd = lq[i] - lz[j];
        acc += L2 ? 2.0f * dd : (dd > 0.0f ? 1.0f : (dd < 0.0f ? -1.0f : 0.0f));
        if (i == 0 && j == 0) break;
        const int k = i + j;
        const int par = (k + W) & 1;
        const int p = min(max((i - j + W - par) >> 1, 0), W);
        const int64_t at =
            static_cast<int64_t>(k / kDirGroup) * (kWave * C) + (p % C) * kWave + p / C;
        const uint32_t word =
            LDIR ? dw[at] : __hip_atomic_load(dw + at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int dr = static_cast<int>((word >> (2 * (k & (kDirGroup - 1)))) & 3u);
        if (i == 0) dr = 2;
        else if (j == 0) dr = 1;
        if (dr != 2) {  // leaving row i
            go[i] = acc * scale;
            acc = 0.0f;
            --i;
        }
        if (dr != 1) --j;
    }
    go[0] = acc * scale;
}

template <int C, bool L2, bool LDIR>
__global__ __launch_bounds__(64) void bwd_path_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const int32_t* __restrict__ idx, const float* __restrict__ q, const float* __restrict__ gin,
    uint32_t* __restrict__ dirs, float* __restrict__ gpair, float* __restrict__ dist,
    BwdScalars a) {
    extern __shared__ float lds[];
    const int64_t pr = blockIdx.x;
    const int64_t pos = idx[pr];
    const float mu = a.pair_stats ? avg[pr] : avg[pos];
    const float sg = a.pair_stats ? sd[pr] : sd[pos];
    bwd_window<C, L2, LDIR>(s + pos, mu, sg, q, a.m, a.w, gin[pr], a.sqrt_dist,
                            LDIR ? nullptr : dirs + pr * a.words, gpair + pr * a.m,
                            dist ? dist + pr : nullptr, lds);
}

// Every window of a batch of queries, one block each: block b takes pair
// list[b] of query pair_query[list[b]]; the pair's statistics are mu/sd[pr].
// All listed queries keep their direction words in LDS.
template <int C, bool L2>
__global__ __launch_bounds__(64) void bwd_path_batch_kernel(
    const float* __restrict__ s, const float* __restrict__ mu, const float* __restrict__ sd,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ pair_query,
    const int32_t* __restrict__ list, const BwdQuery* __restrict__ tab,
    const float* __restrict__ qarena, const float* __restrict__ gin, float* __restrict__ gpair,
    float* __restrict__ dist, int32_t sqrt_dist) {
    extern __shared__ float lds[];
    const int64_t pr = list[blockIdx.x];
    const BwdQuery Q = tab[pair_query[pr]];
    bwd_window<C, L2, true>(s + pos[pr], mu[pr], sd[pr], qarena + Q.q_off, Q.m, Q.w, gin[pr],
                            sqrt_dist, nullptr, gpair + Q.row_begin + (pr - Q.pair_begin) * Q.m,
                            dist + pr, lds);
}

__global__ __launch_bounds__(256) void bwd_fill_kernel(double* g, int32_t m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) g[i] = 0.0;
}

// gsum[i] += sum over the chunk's windows of gpair[p][i], p ascending.
__global__ __launch_bounds__(256) void bwd_accum_kernel(const float* __restrict__ gpair,
                                                        int64_t count, int32_t m,
                                                        double* __restrict__ gsum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double acc = gsum[i];
    for (int64_t p = 0; p < count; ++p) acc += static_cast<double>(gpair[p * m + i]);
    gsum[i] = acc;
}

constexpr int kFinThreads = 1024;

__device__ double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int off = kFinThreads / 2; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// out = G (q_raw NULL) or the z-normalisation Jacobian applied to G.
__device__ __forceinline__ void bwd_final(const double* __restrict__ gsum,
                                          const float* __restrict__ q_raw, int32_t m,
                                          float* __restrict__ out, double* red) {
    const int t = threadIdx.x;
    if (!q_raw) {
        for (int i = t; i < m; i += kFinThreads) out[i] = static_cast<float>(gsum[i]);
        return;
    }
    double s1 = 0.0;
    for (int i = t; i < m; i += kFinThreads) s1 += q_raw[i];
    const double mu = block_sum(s1, red) / m;
    double s2 = 0.0;
    for (int i = t; i < m; i += kFinThreads) {
        const double d = q_raw[i] - mu;
        s2 += d * d;
    }
    const double sig = sqrt(block_sum(s2, red) / m);
    if (!(sig > 0.0)) {  // a constant query has no z-normalisation
        for (int i = t; i < m; i += kFinThreads) out[i] = 0.0f;
        return;
    }
    double sg = 0.0, sgz = 0.0;
    for (int i = t; i < m; i += kFinThreads) {
        const double z = (q_raw[i] - mu) / sig;
        sg += gsum[i];
        sgz += gsum[i] * z;
    }
    const double mg = block_sum(sg, red) / m;
    const double mgz = block_sum(sgz, red) / m;
    for (int i = t; i < m; i += kFinThreads) {
        const double z = (q_raw[i] - mu) / sig;
        out[i] = static_cast<float>((gsum[i] - mg - z * mgz) / sig);
    }
}

__global__ __launch_bounds__(kFinThreads) void bwd_final_kernel(const double* __restrict__ gsum,
                                                                const float* __restrict__ q_raw,
                                                                int32_t m, float* __restrict__ out) {
    __shared__ double red[kFinThreads];
    bwd_final(gsum, q_raw, m, out, red);
}

// Per query of a batch (blockIdx.x): gsum = sum of its rows, pairs ascending
// (the order of bwd_accum_kernel, so the sums are the per-query ones).
__global__ __launch_bounds__(256) void bwd_accum_batch_kernel(const float* __restrict__ gpair,
                                                              const BwdQuery* __restrict__ tab,
                                                              double* __restrict__ gsum) {
    const BwdQuery Q = tab[blockIdx.x];
    const int i = blockIdx.y * blockDim.x + threadIdx.x;
    if (!Q.batched || i >= Q.m) return;
    const float* row = gpair + Q.row_begin + i;
    double acc = 0.0;
    for (int64_t p = 0; p < Q.pairs; ++p) acc += static_cast<double>(row[p * Q.m]);
    gsum[Q.r_off + i] = acc;
}

__global__ __launch_bounds__(kFinThreads) void bwd_final_batch_kernel(
    const double* __restrict__ gsum, const float* __restrict__ q_raw,
    const BwdQuery* __restrict__ tab, float* __restrict__ out) {
    __shared__ double red[kFinThreads];
    const BwdQuery Q = tab[blockIdx.x];
    if (!Q.batched) return;
    bwd_final(gsum + Q.r_off, q_raw ? q_raw + Q.r_off : nullptr, Q.m, out + Q.r_off, red);
}

using PathFn = void (*)(const float*, const float*, const float*, const int32_t*, const float*,
                        const float*, uint32_t*, float*, float*, BwdScalars);

template <int C>
PathFn path_fn(bool l2, bool ldir) {
    if (ldir) return l2 ? bwd_path_kernel<C, true, true> : bwd_path_kernel<C, false, true>;
    return l2 ? bwd_path_kernel<C, true, false> : bwd_path_kernel<C, false, false>;
}

int slots_per_lane(int32_t w) {
    const int slots = w + 1;
    if (slots <= 64) return 1;
    if (slots <= 128) return 2;
    if (slots <= 256) return 4;
    if (slots <= 512) return 8;
    return 16;
}

PathFn pick_path(int C, bool l2, bool ldir) {
    switch (C) {
        case 1: return path_fn<1>(l2, ldir);
        case 2: return path_fn<2>(l2, ldir);
        case 4: return path_fn<4>(l2, ldir);
        case 8: return path_fn<8>(l2, ldir);
        default: return path_fn<16>(l2, ldir);
    }
}

// LDS bytes of the path kernel with the direction words in LDS.
size_t lds_dir_bytes(int32_t m, int64_t words) {
    return sizeof(float) * 2 * static_cast<size_t>(m) + sizeof(uint32_t) * words;
}

}  // namespace

size_t bwd_ws_bytes(int32_t m, int32_t w, int64_t count) {
    const int C = slots_per_lane(w);
    const int64_t groups = (2 * static_cast<int64_t>(m) - 1 + kDirGroup - 1) / kDirGroup;
    const int64_t words = groups * kWave * C;
    const int64_t gwords = lds_dir_bytes(m, words) <= kBwdLdsBytes ? 0 : words;
    const size_t per_pair = sizeof(uint32_t) * gwords + sizeof(float) * m;
    const int64_t chunk = std::max<int64_t>(
        1, std::min<int64_t>(std::max<int64_t>(count, 1),
                             static_cast<int64_t>(kBwdChunkBytes / per_pair)));
    return carve_size({sizeof(double) * m, sizeof(uint32_t) * gwords * chunk,
                       sizeof(float) * m * chunk});
}

int enqueue_backward(hipStream_t st, const float* s, const float* avg, const float* sd,
                     bool pair_stats, const int32_t* idx, const float* grad, int64_t count,
                     const float* q, int32_t m, int32_t w, int32_t metric, int32_t sqrt_dist,
                     const float* q_raw, float* dist, float* out, char* ws) {
    const int C = slots_per_lane(w);
    const int64_t groups = (2 * static_cast<int64_t>(m) - 1 + kDirGroup - 1) / kDirGroup;
    const int64_t words = groups * kWave * C;
    const bool ldir = lds_dir_bytes(m, words) <= kBwdLdsBytes;
    const int64_t gwords = ldir ? 0 : words;
    const size_t per_pair = sizeof(uint32_t) * gwords + sizeof(float) * m;
    const int64_t chunk = std::max<int64_t>(
        1, std::min<int64_t>(std::max<int64_t>(count, 1),
                             static_cast<int64_t>(kBwdChunkBytes / per_pair)));
    Carve cv{ws};
    double* gsum = cv.take<double>(m);
    uint32_t* dirs = cv.take<uint32_t>(gwords * chunk);
    float* gpair = cv.take<float>(static_cast<size_t>(m) * chunk);

    const unsigned mblocks = static_cast<unsigned>((m + 255) / 256);
    hipLaunchKernelGGL(bwd_fill_kernel, dim3(mblocks), dim3(256), 0, st, gsum, m);
    int rc = hip_status(hipGetLastError());
    if (rc != MDTW_OK) return rc;
    const PathFn fn = pick_path(C, metric == MDTW_METRIC_L2, ldir);
    const BwdScalars sc{m, w, sqrt_dist ? 1 : 0, pair_stats ? 1 : 0, words};
    const size_t lds = ldir ? lds_dir_bytes(m, words) : sizeof(float) * 2 * static_cast<size_t>(m);
    for (int64_t off = 0; off < count; off += chunk) {
        const int64_t n = std::min(chunk, count - off);
        hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(n)), dim3(kWave), lds, st, s,
                           pair_stats ? avg + off : avg, pair_stats ? sd + off : sd, idx + off, q,
                           grad + off, dirs, gpair, dist ? dist + off : nullptr, sc);
        hipLaunchKernelGGL(bwd_accum_kernel, dim3(mblocks), dim3(256), 0, st, gpair, n, m, gsum);
        rc = hip_status(hipGetLastError());
        if (rc != MDTW_OK) return rc;
    }
    hipLaunchKernelGGL(bwd_final_kernel, dim3(1), dim3(kFinThreads), 0, st, gsum, q_raw, m, out);
    return hip_status(hipGetLastError());
}

bool bwd_batchable(int32_t m, int32_t w) {
    const int C = slots_per_lane(w);
    const int64_t groups = (2 * static_cast<int64_t>(m) - 1 + kDirGroup - 1) / kDirGroup;
    return lds_dir_bytes(m, groups * kWave * C) <= kBwdLdsBytes;
}

int bwd_class(int32_t w) { return slots_per_lane(w); }

namespace {

using BatchFn = void (*)(const float*, const float*, const float*, const int32_t*,
                         const int32_t*, const int32_t*, const BwdQuery*, const float*,
                         const float*, float*, float*, int32_t);

template <int C>
BatchFn batch_fn(bool l2) {
    return l2 ? bwd_path_batch_kernel<C, true> : bwd_path_batch_kernel<C, false>;
}

}  // namespace

int launch_bwd_path_batch(hipStream_t st, int C, bool l2, int32_t max_m, const float* s,
                          const float* mu, const float* sd, const int32_t* pos,
                          const int32_t* pair_query, const int32_t* list, int64_t count,
                          const BwdQuery* tab, const float* qarena, const float* gin,
                          float* gpair, float* dist, int32_t sqrt_dist) {
    if (count <= 0) return MDTW_OK;
    if (count > 0x7fffffffll) return MDTW_ERR_UNSUPPORTED;
    BatchFn fn = nullptr;
    switch (C) {
        case 1: fn = batch_fn<1>(l2); break;
        case 2: fn = batch_fn<2>(l2); break;
        case 4: fn = batch_fn<4>(l2); break;
        case 8: fn = batch_fn<8>(l2); break;
        case 16: fn = batch_fn<16>(l2); break;
        default: return MDTW_ERR_INVALID_ARGUMENT;
    }
    const int64_t groups = (2 * static_cast<int64_t>(max_m) - 1 + kDirGroup - 1) / kDirGroup;
    const size_t lds = lds_dir_bytes(max_m, groups * kWave * C);
    if (lds > kBwdLdsBytes) return MDTW_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(fn, dim3(static_cast<unsigned>(count)), dim3(kWave), lds, st, s, mu, sd,
                       pos, pair_query, list, tab, qarena, gin, gpair, dist, sqrt_dist);
    return hip_status(hipGetLastError());
}

int launch_bwd_reduce_batch(hipStream_t st, const float* gpair, const BwdQuery* tab, int32_t nq,
                            int32_t max_m, const float* q_raw, double* gsum, float* out) {
    if (nq <= 0 || max_m <= 0) return MDTW_OK;
    const unsigned mblocks = static_cast<unsigned>((max_m + 255) / 256);
    hipLaunchKernelGGL(bwd_accum_batch_kernel, dim3(static_cast<unsigned>(nq), mblocks), dim3(256),
                       0, st, gpair, tab, gsum);
    int rc = hip_status(hipGetLastError());
    if (rc != MDTW_OK) return rc;
    hipLaunchKernelGGL(bwd_final_batch_kernel, dim3(static_cast<unsigned>(nq)), dim3(kFinThreads),
                       0, st, gsum, q_raw, tab, out);
    return hip_status(hipGetLastError());
}

}  // namespace mdtw

using namespace mdtw;

extern "C" int mdtw_cdtw_backward(mdtw_engine* e, const float* d_series, const float* d_avg,
                                  const float* d_std, const int32_t* d_indices,
                                  const float* d_grad, int64_t count, int32_t w, int32_t metric,
                                  int32_t sqrt_dist, const float* d_query_raw, float* d_dist,
                                  float* d_grad_query) {
    if (!e || !d_grad_query || count < 0 || w < 0)
        return w < 0 ? MDTW_ERR_BAND : MDTW_ERR_INVALID_ARGUMENT;
    if (count > 0 && (!d_series || !d_avg || !d_std || !d_indices || !d_grad))
        return MDTW_ERR_INVALID_ARGUMENT;
    if (metric != MDTW_METRIC_L1 && metric != MDTW_METRIC_L2) return MDTW_ERR_INVALID_ARGUMENT;
    if (!e->has_query) return MDTW_ERR_NO_QUERY;
    const int32_t m = e->m;
    const int32_t wc = std::min(w, m - 1);
    if (wc > kMaxBandW) return MDTW_ERR_UNSUPPORTED;
    if (count > 0x7fffffffll) return MDTW_ERR_UNSUPPORTED;
    DeviceGuard guard(e->device);
    void* ws = nullptr;
    int rc = ws_reserve(e, bwd_ws_bytes(m, wc, count), &ws);
    if (rc != MDTW_OK) return rc;
    return enqueue_backward(e->stream, d_series, d_avg, d_std, false, d_indices, d_grad, count,
                            e->d_q, m, wc, metric, sqrt_dist, d_query_raw, d_dist, d_grad_query,
                            static_cast<char*>(ws));
}
