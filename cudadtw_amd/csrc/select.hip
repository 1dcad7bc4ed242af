// select.hip -- reductions and selections over candidate arrays: argmin with
// lowest-index ties, stable threshold compaction, and the k-th smallest of
// the seed distances.  These replace the reference's CUB wrappers
// (cub_util.cuh:7-27 reduce, :29-54 prefix_sum, :79-114 pair_sort,
// :154-172 threshold) with single-purpose kernels that allocate nothing.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {

__global__ void ctrl_reset_kernel(Ctrl* ctrl, unsigned long long best_init, uint32_t* bmin,
                                  long long bucket_len) {
    if (threadIdx.x == 0) {
        ctrl->bmin = bmin;
        ctrl->bucket_len = bucket_len;
        ctrl->best = best_init;
        ctrl->count = 0;
        ctrl->work = 0;
        ctrl->nres = 0;
        ctrl->kim = 0;
        ctrl->keogh = 0;
        ctrl->dp = 0;
        ctrl->abandoned = 0;
        ctrl->thr = kInf;
    }
}

__global__ __launch_bounds__(256) void search_reset_kernel(Ctrl* ctrl,
                                                           unsigned long long best_init,
                                                           uint32_t* bmin, long long bucket_len,
                                                           int64_t nb, uint32_t* tscratch,
                                                           int64_t ntw) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    if (g == 0) {
        ctrl->bmin = bmin;
        ctrl->bucket_len = bucket_len;
        ctrl->best = best_init;
        ctrl->count = 0;
        ctrl->work = 0;
        ctrl->nres = 0;
        ctrl->kim = 0;
        ctrl->keogh = 0;
        ctrl->dp = 0;
        ctrl->abandoned = 0;
        ctrl->thr = kInf;
    }
    for (int64_t i = g; i < nb; i += stride) bmin[i] = 0x7f800000u;
    for (int64_t i = g; i < ntw; i += stride) tscratch[i] = 0u;
}

__global__ void ctrl_set_thr_from_best(Ctrl* ctrl) {
    if (threadIdx.x == 0) ctrl->thr = key_dist(ctrl->best);
}

__global__ __launch_bounds__(256) void argmin_kernel(const float* __restrict__ v, int64_t n,
                                                     Ctrl* ctrl) {
    unsigned long long best = ~0ull;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const float x = v[i];
        // Keys order non-negative floats; map -0 and negatives below +0 so
        // the minimum is still found for signed data.
        const uint32_t b = __float_as_uint(x);
        const uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
        const unsigned long long key =
            (static_cast<unsigned long long>(ord) << 32) | static_cast<uint32_t>(i);
        best = key < best ? key : best;
    }
    best = wave_min_u64(best);
    if ((threadIdx.x & (kWave - 1)) == 0 && best != ~0ull) atomicMin(&ctrl->best, best);
}

constexpr int kTb = 1024;  // elements per block in the compaction

__global__ __launch_bounds__(256) void thr_count_kernel(const float* __restrict__ v, int64_t n,
                                                        float thr, int32_t* __restrict__ bc) {
    __shared__ int sh[4];
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kTb;
    int c = 0;
    for (int e = threadIdx.x; e < kTb; e += 256) {
        const int64_t i = b0 + e;
        if (i < n && v[i] <= thr) ++c;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, kWave);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) bc[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Exclusive scan of nb block counts in one workgroup; bc[nb] = total.
__global__ __launch_bounds__(256) void thr_scan_kernel(int32_t* __restrict__ bc, int nb) {
    __shared__ int sh[256];
    const int per = (nb + 255) / 256;
    const int lo = threadIdx.x * per;
    int s = 0;
    for (int i = 0; i < per; ++i)
        if (lo + i < nb) s += bc[lo + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int t = static_cast<int>(threadIdx.x) >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - s;
    for (int i = 0; i < per; ++i)
        if (lo + i < nb) {
            const int c = bc[lo + i];
            bc[lo + i] = run;
            run += c;
        }
    if (threadIdx.x == 255) bc[nb] = sh[255];
}

__global__ __launch_bounds__(256) void thr_scatter_kernel(const float* __restrict__ v,
                                                          const int32_t* __restrict__ in,
                                                          int64_t n, float thr,
                                                          const int32_t* __restrict__ bc,
                                                          int32_t* __restrict__ outi) {
    __shared__ int sh[4];
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kTb;
    int base = bc[blockIdx.x];
    // Four passes of 256 keep the output in input order.
    for (int e0 = 0; e0 < kTb; e0 += 256) {
        const int64_t i = b0 + e0 + threadIdx.x;
        const bool keep = i < n && v[i] <= thr;
        const unsigned long long mask = __ballot(keep);
        const uint32_t r = __builtin_amdgcn_mbcnt_hi(
            static_cast<uint32_t>(mask >> 32),
            __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x / 64] = __popcll(mask);
        __syncthreads();
        int off = base;
        for (int wv = 0; wv < static_cast<int>(threadIdx.x / 64); ++wv) off += sh[wv];
        if (keep) outi[off + r] = in[i];
        base += sh[0] + sh[1] + sh[2] + sh[3];
        __syncthreads();
    }
}

// Order-preserving map of float bits to unsigned keys (negative values
// reversed, the sign bit set on the rest), and its inverse.
__device__ __forceinline__ uint32_t ord_key(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord_key(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// One wave finds the bin of rank krem (1-based) in hist[0..64*PER): lane l
// sums bins [PER*l, PER*l + PER), a wave prefix sum picks the lane, the lane
// walks its bins.  Returns (bin, rank inside the bin) to every lane, or bin
// 0xffffffff when the bins hold fewer than krem values.
template <int PER>
__device__ __forceinline__ void wave_find_rank(const uint32_t* hist, uint32_t krem,
                                               uint32_t& bin, uint32_t& rank_in) {
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) s += hist[lane * PER + i];
    uint32_t incl = s;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += o;
    }
    const unsigned long long hit = __ballot(incl >= krem);
    bin = 0xffffffffu;
    rank_in = 0;
    if (hit == 0ull) return;
    const int L = __ffsll(static_cast<long long>(hit)) - 1;
    uint32_t b = 0xffffffffu, r = 0;
    if (lane == L) {
        uint32_t run = incl - s;
        for (int i = 0; i < PER; ++i) {
            const uint32_t c = hist[lane * PER + i];
            if (run + c >= krem) {
                b = static_cast<uint32_t>(lane * PER + i);
                r = krem - run;
                break;
            }
            run += c;
        }
    }
    bin = __shfl(b, L, kWave);
    rank_in = __shfl(r, L, kWave);
}

// k-th smallest of n (<= kSeedMax) values, exactly: radix selection over the
// order-preserving keys, 8 bits per pass from the top, with the n values held
// in registers (kSeedMax / 1024 per thread) and a 256-bin LDS histogram of
// the values that still match the selected prefix.
constexpr int kSeedMax = 16384;
constexpr int kSeedThreads = 1024;
constexpr int kSeedPer = kSeedMax / kSeedThreads;

__global__ __launch_bounds__(kSeedThreads) void seed_thr_kernel(const float* __restrict__ v,
                                                                int n, int k, float cap,
                                                                Ctrl* ctrl) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sel[2];
    if (k > n || k < 1) {
        if (threadIdx.x == 0) ctrl->thr = fminf(kInf, cap);
        return;
    }
    uint32_t key[kSeedPer];
#pragma unroll
    for (int i = 0; i < kSeedPer; ++i) {
        const int j = static_cast<int>(threadIdx.x) + i * kSeedThreads;
        key[i] = j < n ? ord_key(v[j]) : 0u;
    }
    uint32_t prefix = 0, krem = static_cast<uint32_t>(k);
    for (int shift = 24; shift >= 0; shift -= 8) {
        if (threadIdx.x < 256) hist[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t hmask = shift == 24 ? 0u : (0xffffffffu << (shift + 8));
#pragma unroll
        for (int i = 0; i < kSeedPer; ++i) {
            const int j = static_cast<int>(threadIdx.x) + i * kSeedThreads;
            if (j < n && (key[i] & hmask) == prefix)
                atomicAdd(&hist[(key[i] >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x < kWave) {
            uint32_t bin, r;
            wave_find_rank<4>(hist, krem, bin, r);
            if (threadIdx.x == 0) {
                sel[0] = prefix | (bin << shift);  // a bin exists: krem <= matching count
                sel[1] = r;
            }
        }
        __syncthreads();
        prefix = sel[0];
        krem = sel[1];
    }
    if (threadIdx.x == 0) ctrl->thr = fminf(unord_key(prefix), cap);
}

inline int check_launch() {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

__global__ void round_reset_kernel(Ctrl* ctrl) {
    if (threadIdx.x == 0) {
        ctrl->count = 0;
        ctrl->work = 0;
    }
}

// ---- top-k bound: two-pass radix selection over the bucket minima ----------
// Pass 0 histograms bits 31..20 of every bucket minimum (non-negative float
// bits order like the values), per parity class; the k-th smallest lies in
// the first bin where the running count reaches k.  Pass 1 histograms bits
// 19..8 of the minima inside that bin.  The bound is the largest value with
// the 24 selected bits, so it is >= the k-th smallest (within 2^-15 of it).
// Scratch layout (uint32): hist[2][4096], then state: prefix[2], krem[2].
constexpr int kSelBins = 4096;
constexpr int kSelState = 2 * kSelBins;

__global__ __launch_bounds__(256) void tighten_hist_kernel(const uint32_t* __restrict__ bmin,
                                                           int64_t nb, int split, int pass,
                                                           uint32_t* __restrict__ ws) {
    __shared__ uint32_t h[2 * kSelBins];
    for (int i = threadIdx.x; i < 2 * kSelBins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint32_t p0 = ws[kSelState + 0], p1 = ws[kSelState + 1];
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < nb;
         b += stride) {
        const uint32_t v = bmin[b];
        const int par = split ? static_cast<int>(b & 1) : 0;
        if (pass == 0) {
            atomicAdd(&h[par * kSelBins + (v >> 20)], 1u);
        } else if ((v >> 20) == (par ? p1 : p0)) {
            atomicAdd(&h[par * kSelBins + ((v >> 8) & 0xfffu)], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * kSelBins; i += blockDim.x)
        if (h[i]) atomicAdd(&ws[i], h[i]);
}

// One block: find the bin of the remaining rank per parity (one wave per
// parity), clear the histogram for the next pass; after pass 1 lower
// ctrl->thr.
__global__ __launch_bounds__(256) void tighten_find_kernel(uint32_t* __restrict__ ws, int32_t k,
                                                           int split, int pass,
                                                           Ctrl* __restrict__ ctrl,
                                                           int reset_round) {
    __shared__ uint32_t found[2], rank[2];
    const int t = threadIdx.x;
    const int wv = t / kWave;
    if (wv < (split ? 2 : 1)) {
        const uint32_t krem = pass == 0 ? static_cast<uint32_t>(k) : ws[kSelState + 2 + wv];
        uint32_t bin, r;
        wave_find_rank<kSelBins / kWave>(ws + wv * kSelBins, krem, bin, r);
        if ((t & (kWave - 1)) == 0) {
            found[wv] = bin;
            rank[wv] = r;
        }
    }
    __syncthreads();
    for (int i = t; i < 2 * kSelBins; i += blockDim.x) ws[i] = 0;
    if (t != 0) return;
    uint32_t bound[2] = {0x7f800000u, 0x7f800000u};
    for (int par = 0; par < (split ? 2 : 1); ++par) {
        ws[kSelState + 2 + par] = rank[par];  // rank inside the bin
        if (found[par] == 0xffffffffu) {      // fewer than k known: no bound
            ws[kSelState + par] = 0xffffffffu;
            continue;
        }
        if (pass == 0) {
            ws[kSelState + par] = found[par];
        } else {
            const uint32_t hi = (ws[kSelState + par] << 20) | (found[par] << 8) | 0xffu;
            bound[par] = hi >= 0x7f800000u ? 0x7f800000u : hi;
        }
    }
    if (pass == 1) {
        uint32_t b = bound[0];
        if (split && bound[1] < b) b = bound[1];
        const float tb = __uint_as_float(b);
        if (tb < ctrl->thr) ctrl->thr = tb;
        if (reset_round) {
            ctrl->count = 0;
            ctrl->work = 0;
        }
    }
}

// Both passes in one block, histograms in LDS, for up to kTightenSmall
// buckets (where the four launches of the general path cost more than one
// block reading the minima twice).  Same bins, same bound.
constexpr int64_t kTightenSmall = 1 << 15;
constexpr int kTightenThreads = 1024;
constexpr int kTightenUnroll = 8;

__global__ __launch_bounds__(kTightenThreads) void tighten_small_kernel(
    const uint32_t* __restrict__ bmin, int64_t nb, int32_t k, int split, Ctrl* __restrict__ ctrl,
    int reset_round) {
    __shared__ uint32_t h[2 * kSelBins];
    __shared__ uint32_t prefix[2], rank[2], sub[2];
    const int t = threadIdx.x;
    const int wv = t / kWave;
    const int npar = split ? 2 : 1;
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = t; i < 2 * kSelBins; i += kTightenThreads) h[i] = 0;
        __syncthreads();
        const uint32_t p0 = pass ? prefix[0] : 0u, p1 = pass ? prefix[1] : 0u;
        for (int64_t b0 = t; b0 < nb; b0 += static_cast<int64_t>(kTightenThreads) * kTightenUnroll) {
            uint32_t v[kTightenUnroll];
#pragma unroll
            for (int u = 0; u < kTightenUnroll; ++u) {
                const int64_t b = b0 + static_cast<int64_t>(u) * kTightenThreads;
                v[u] = b < nb ? bmin[b] : 0u;
            }
#pragma unroll
            for (int u = 0; u < kTightenUnroll; ++u) {
                const int64_t b = b0 + static_cast<int64_t>(u) * kTightenThreads;
                if (b >= nb) break;
                const int par = split ? static_cast<int>(b & 1) : 0;
                if (pass == 0)
                    atomicAdd(&h[par * kSelBins + (v[u] >> 20)], 1u);
                else if ((v[u] >> 20) == (par ? p1 : p0))
                    atomicAdd(&h[par * kSelBins + ((v[u] >> 8) & 0xfffu)], 1u);
            }
        }
        __syncthreads();
        if (wv < npar) {
            const uint32_t krem = pass == 0 ? static_cast<uint32_t>(k) : rank[wv];
            uint32_t bin, r;
            wave_find_rank<kSelBins / kWave>(h + wv * kSelBins, krem, bin, r);
            if ((t & (kWave - 1)) == 0) {
                if (pass == 0) {
                    prefix[wv] = bin;  // 0xffffffff: fewer than k buckets known
                    rank[wv] = r;
                } else {
                    sub[wv] = bin;
                }
            }
        }
        __syncthreads();
    }
    if (t != 0) return;
    uint32_t b = 0x7f800000u;
    for (int par = 0; par < npar; ++par) {
        if (prefix[par] == 0xffffffffu || sub[par] == 0xffffffffu) continue;
        uint32_t hi = (prefix[par] << 20) | (sub[par] << 8) | 0xffu;
        if (hi > 0x7f800000u) hi = 0x7f800000u;
        if (hi < b) b = hi;
    }
    const float tb = __uint_as_float(b);
    if (tb < ctrl->thr) ctrl->thr = tb;
    if (reset_round) {
        ctrl->count = 0;
        ctrl->work = 0;
    }
}

// ---- greedy top-k selection on the device ---------------------------------
constexpr int kSelThreads = 1024;

__global__ __launch_bounds__(kSelThreads) void topk_select_kernel(
    const Ctrl* __restrict__ ctrl, const unsigned long long* __restrict__ res, int32_t k,
    long long excl, int32_t metric, int64_t* __restrict__ out_pos, float* __restrict__ out_dist,
    int32_t* __restrict__ overflow) {
    __shared__ unsigned long long keys[kSelectMaxKeys];
    __shared__ int32_t nk;
    const int t = threadIdx.x;
    const bool l2 = metric == MDTW_METRIC_L2;
    if (k == 1) {
        if (t == 0) {
            const unsigned long long b = ctrl->best;
            if ((b >> 32) < 0x7f800000ull) {
                const float d = key_dist(b);
                out_pos[0] = static_cast<int64_t>(b & 0xffffffffull);
                out_dist[0] = l2 ? sqrtf(d) : d;
            } else {
                out_pos[0] = -1;
                out_dist[0] = kInf;
            }
            *overflow = 0;
        }
        return;
    }
    // Keys above the final bound cannot be among the first k picks: the
    // bound is >= the k-th pick's distance.
    const float thr = ctrl->thr;
    const long long nres = static_cast<long long>(ctrl->nres);
    if (t == 0) nk = 0;
    __syncthreads();
    for (long long i = t; i < nres; i += kSelThreads) {
        const unsigned long long key = res[i];
        if (key_dist(key) <= thr) {
            const int32_t slot = atomicAdd(&nk, 1);
            if (slot < kSelectMaxKeys) keys[slot] = key;
        }
    }
    __syncthreads();
    const int32_t n = nk;
    if (n > kSelectMaxKeys) {
        for (int32_t i = t; i < k; i += kSelThreads) {
            out_pos[i] = -1;
            out_dist[i] = kInf;
        }
        if (t == 0) *overflow = 1;
        return;
    }
    int32_t P = 1;
    while (P < n) P <<= 1;
    for (int32_t i = n + t; i < P; i += kSelThreads) keys[i] = ~0ull;
    __syncthreads();
    // Bitonic sort of keys[0..P), ascending.
    for (int32_t size = 2; size <= P; size <<= 1) {
        for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (int32_t i = t; i < P / 2; i += kSelThreads) {
                const int32_t lo = 2 * i - (i & (stride - 1));
                const int32_t hi = lo + stride;
                const bool up = (lo & size) == 0;
                const unsigned long long a = keys[lo], b = keys[hi];
                if ((a > b) == up) {
                    keys[lo] = b;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
    if (t != 0) return;
    int32_t found = 0;
    for (int32_t i = 0; i < n && found < k; ++i) {
        const unsigned long long key = keys[i];
        const int64_t pos = static_cast<int64_t>(key & 0xffffffffull);
        bool ok = true;
        for (int32_t p = 0; p < found; ++p) {
            const int64_t dd = pos > out_pos[p] ? pos - out_pos[p] : out_pos[p] - pos;
            if (dd < excl) {
                ok = false;
                break;
            }
        }
        if (!ok) continue;
        const float d = key_dist(key);
        out_pos[found] = pos;
        out_dist[found] = l2 ? sqrtf(d) : d;
        ++found;
    }
    for (int32_t i = found; i < k; ++i) {
        out_pos[i] = -1;
        out_dist[i] = kInf;
    }
    *overflow = 0;
}

}  // namespace

int launch_topk_select(hipStream_t st, const Ctrl* ctrl, const unsigned long long* res,
                       int32_t k, int64_t excl, int32_t metric, int64_t* out_pos,
                       float* out_dist, int32_t* overflow) {
    hipLaunchKernelGGL(topk_select_kernel, dim3(1), dim3(kSelThreads), 0, st, ctrl, res, k,
                       static_cast<long long>(excl), metric, out_pos, out_dist, overflow);
    return check_launch();
}

int launch_ctrl_reset(hipStream_t st, Ctrl* ctrl, unsigned long long best_init, uint32_t* bmin,
                      int64_t bucket_len) {
    hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, st, ctrl, best_init, bmin,
                       static_cast<long long>(bucket_len));
    return check_launch();
}

int launch_search_reset(hipStream_t st, Ctrl* ctrl, unsigned long long best_init,
                        uint32_t* bmin, int64_t bucket_len, int64_t nb, uint32_t* tscratch,
                        int64_t ntw) {
    const int64_t n = std::max<int64_t>(nb, ntw);
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(search_reset_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                       ctrl, best_init, bmin, static_cast<long long>(bucket_len), nb, tscratch,
                       ntw);
    return check_launch();
}

int launch_round_reset(hipStream_t st, Ctrl* ctrl) {
    hipLaunchKernelGGL(round_reset_kernel, dim3(1), dim3(64), 0, st, ctrl);
    return check_launch();
}

size_t tighten_ws_words() { return kSelState + 4; }

// scratch must hold zeros in its histogram part on entry (launch_search_reset
// clears it); every pass leaves it cleared again.
int launch_tighten(hipStream_t st, const uint32_t* bmin, int64_t nb, int32_t k, bool split,
                   uint32_t* scratch, Ctrl* ctrl, bool reset_round) {
    if (nb <= 0) return reset_round ? launch_round_reset(st, ctrl) : MDTW_OK;
    if (nb <= kTightenSmall) {
        hipLaunchKernelGGL(tighten_small_kernel, dim3(1), dim3(kTightenThreads), 0, st, bmin, nb,
                           k, split ? 1 : 0, ctrl, reset_round ? 1 : 0);
        return check_launch();
    }
    const int64_t blocks = std::min<int64_t>((nb + 255) / 256, 1024);
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(tighten_hist_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                           st, bmin, nb, split ? 1 : 0, pass, scratch);
        hipLaunchKernelGGL(tighten_find_kernel, dim3(1), dim3(256), 0, st, scratch, k,
                           split ? 1 : 0, pass, ctrl, reset_round && pass == 1 ? 1 : 0);
    }
    return check_launch();
}

int launch_thr_from_best(hipStream_t st, Ctrl* ctrl) {
    hipLaunchKernelGGL(ctrl_set_thr_from_best, dim3(1), dim3(64), 0, st, ctrl);
    return check_launch();
}

int launch_argmin(hipStream_t st, const float* v, int64_t count, Ctrl* ctrl) {
    if (count <= 0) return MDTW_OK;
    const int64_t nb = std::min<int64_t>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(argmin_kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, st, v, count,
                       ctrl);
    return check_launch();
}

size_t threshold_ws_ints(int64_t count) {
    const int64_t nb = (count + kTb - 1) / kTb;
    return static_cast<size_t>(nb + 1 + count);
}

// Stable in-place compaction; *d_total (device) gets the survivor count.
int launch_threshold(hipStream_t st, const float* v, int32_t* idx, int64_t count, float thr,
                     int32_t* tmp, int32_t** d_total) {
    const int64_t nb = (count + kTb - 1) / kTb;
    if (nb > 256 * 4096) return MDTW_ERR_UNSUPPORTED;
    int32_t* bc = tmp;
    int32_t* outi = tmp + nb + 1;
    hipLaunchKernelGGL(thr_count_kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, st, v,
                       count, thr, bc);
    hipLaunchKernelGGL(thr_scan_kernel, dim3(1), dim3(256), 0, st, bc, static_cast<int>(nb));
    hipLaunchKernelGGL(thr_scatter_kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, st, v,
                       idx, count, thr, bc, outi);
    hipError_t err = hipMemcpyAsync(idx, outi, sizeof(int32_t) * static_cast<size_t>(count),
                                    hipMemcpyDeviceToDevice, st);
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    *d_total = bc + nb;
    return check_launch();
}

int launch_seed_threshold(hipStream_t st, const float* seeds, int64_t nseeds, int32_t k,
                          Ctrl* ctrl, float cap) {
    if (nseeds > kSeedMax) return MDTW_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(seed_thr_kernel, dim3(1), dim3(1024), 0, st, seeds,
                       static_cast<int>(nseeds), k, cap, ctrl);
    return check_launch();
}

}  // namespace mdtw
