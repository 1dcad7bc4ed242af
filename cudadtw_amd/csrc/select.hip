// select.hip -- reductions and selections over candidate arrays: argmin with
// lowest-index ties, stable threshold compaction, and the k-th smallest of
// the seed distances.  These replace the reference's CUB wrappers
// (cub_util.cuh:7-27 reduce, :29-54 prefix_sum, :79-114 pair_sort,
// :154-172 threshold) with single-purpose kernels that allocate nothing.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {

__global__ void ctrl_reset_kernel(Ctrl* ctrl, unsigned long long best_init) {
    if (threadIdx.x == 0) {
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
        int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
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
        for (int wv = 0; wv < threadIdx.x / 64; ++wv) off += sh[wv];
        if (keep) outi[off + r] = in[i];
        base += sh[0] + sh[1] + sh[2] + sh[3];
        __syncthreads();
    }
}

// k-th smallest of n (<= kSeedMax) values: bitonic sort in LDS.
constexpr int kSeedMax = 16384;

__global__ __launch_bounds__(1024) void seed_thr_kernel(const float* __restrict__ v, int n, int k,
                                                        Ctrl* ctrl) {
    __shared__ float a[kSeedMax];
    int size = 1;
    while (size < n) size <<= 1;
    for (int i = threadIdx.x; i < size; i += blockDim.x) a[i] = i < n ? v[i] : kInf;
    __syncthreads();
    for (int len = 2; len <= size; len <<= 1) {
        for (int st = len >> 1; st > 0; st >>= 1) {
            for (int i = threadIdx.x; i < size; i += blockDim.x) {
                const int j = i ^ st;
                if (j > i) {
                    const bool up = (i & len) == 0;
                    const float x = a[i], y = a[j];
                    if ((x > y) == up) {
                        a[i] = y;
                        a[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) ctrl->thr = (k <= n) ? a[k - 1] : kInf;
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

int launch_ctrl_reset(hipStream_t st, Ctrl* ctrl, unsigned long long best_init) {
    hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, st, ctrl, best_init);
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
                          Ctrl* ctrl) {
    if (nseeds > kSeedMax) return MDTW_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(seed_thr_kernel, dim3(1), dim3(1024), 0, st, seeds,
                       static_cast<int>(nseeds), k, ctrl);
    return check_launch();
}

}  // namespace mdtw
