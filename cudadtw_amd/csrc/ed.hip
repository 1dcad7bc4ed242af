// ed.hip -- z-normalised Euclidean distance of every window.
//
// The reference computes ED^2 = 2M - 2 * corr / sigma through an FP64 cuFFT
// correlation (ed.cuh:71-153) and takes the argmin with a full radix sort.
// Here each lane owns one window and accumulates (q_k - (x_k - mu) / sigma)^2
// directly in FP32 (64 consecutive windows per wave: coalesced stream reads,
// the query as wave-uniform scalars); the division is a reciprocal multiply.
// The direct form has no cancellation when the distance is small (the
// correlation form does), so FP32 accumulation stays within ~1e-6 of the
// FP64 value.  Search mode seeds a threshold and abandons a wave once every
// lane's partial sum is above it (the UCR_ED early abandon, UCR_ED_BIN.cpp:
// 55-65, applied per wave).
#include "cdtw_band.hpp"

namespace mdtw {
namespace {

constexpr int kEdChunk = 16;

__device__ __forceinline__ uint32_t ed_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__global__ __launch_bounds__(256) void ed_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, BandScalars a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    const int32_t m = a.m;
    const bool abandon = a.mode == kModeBest || a.mode == kModeCollect;
    for (int64_t base = wave * kWave; base < a.count; base += nwaves * kWave) {
        const int64_t g = base + lane;
        const bool valid = g < a.count;
        const int64_t gc = valid ? g : a.count - 1;
        const int64_t w = idx ? static_cast<int64_t>(idx[gc]) : a.base + gc * a.stride;
        const float* __restrict__ x = s + w;
        const float mu = avg[w];
        const float rs = 1.0f / sd[w];
        float thr = kInf;
        if (abandon) thr = a.mode == kModeBest ? load_best_dist(ctrl) : load_thr(ctrl);
        bool alive = valid;
        float acc = 0.0f;
        int32_t k = 0;
        bool gone = false;
        for (; k + kEdChunk <= m; k += kEdChunk) {
#pragma unroll
            for (int u = 0; u < kEdChunk; ++u) {
                const float z = (x[k + u] - mu) * rs;
                const float d = q[k + u] - z;
                acc = fmaf(d, d, acc);
            }
            if (abandon) {
                alive = alive && !(acc > thr);
                if (!__any(alive)) {
                    gone = true;
                    break;
                }
            }
        }
        for (; !gone && k < m; ++k) {
            const float z = (x[k] - mu) * rs;
            const float d = q[k] - z;
            acc = fmaf(d, d, acc);
        }
        if (gone) acc = kInf;
        if (a.mode == kModeProfile || a.mode == kModeSeed) {
            if (valid) out[g] = acc;
        }
        if (a.mode == kModeBest || a.mode == kModeSeed) {
            unsigned long long key = alive ? pack_key(acc, static_cast<uint32_t>(w)) : ~0ull;
            key = wave_min_u64(key);
            if (lane == 0 && key != ~0ull) atomicMin(&ctrl->best, key);
            if (a.mode == kModeSeed && valid) note_bucket(ctrl->bmin, ctrl->bucket_len, w, acc);
        } else if (a.mode == kModeCollect) {
            const bool keep = alive && acc <= thr;
            if (keep) note_bucket(ctrl->bmin, ctrl->bucket_len, w, acc);
            const unsigned long long km = __ballot(keep);
            if (km) {
                unsigned long long rb = 0;
                if (lane == 0)
                    rb = atomicAdd(&ctrl->nres, static_cast<unsigned long long>(__popcll(km)));
                rb = __shfl(rb, 0, kWave);
                if (keep) res[rb + ed_rank(km)] = pack_key(acc, static_cast<uint32_t>(w));
            }
        }
    }
}

}  // namespace

int launch_ed(hipStream_t st, int cus, const BandArgs& a) {
    if (a.count <= 0) return MDTW_OK;
    BandScalars sc{a.m, 0, a.count, a.base, a.stride, a.mode, 0};
    const int64_t waves = (a.count + kWave - 1) / kWave;
    const int64_t blocks = std::min<int64_t>((waves + 3) / 4, static_cast<int64_t>(cus) * 64);
    hipLaunchKernelGGL(ed_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, a.s,
                       a.avg, a.sd, a.qx, a.idx, a.out, a.res, a.ctrl, sc);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

}  // namespace mdtw
