// ed.hip -- z-normalised Euclidean distance of every window.
//
// The reference computes ED^2 = 2M - 2 * corr / sigma through an FP64 cuFFT
// correlation (ed.cuh:71-153) and takes the argmin with a full radix sort.
// Here each lane owns one window and accumulates (q_k - (x_k - mu) / sigma)^2
// directly in FP32 (64 consecutive windows per wave: coalesced stream reads,
// the query as wave-uniform scalars), three operations per element:
// u = x_k - mu, d = fma(-u, 1/sigma, q_k) (one rounding where a multiply and
// a subtract would round twice), acc = fma(d, d, acc).  Measured on this
// device, that sequence issues in 8.0 SIMD cycles per wave-element against
// 9.7 for the four-operation form with a separate multiply.
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
                const float d = fmaf(-(x[k + u] - mu), rs, q[k + u]);
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
            const float d = fmaf(-(x[k] - mu), rs, q[k]);
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


// Contiguous windows (stride 1, no index list), four per lane.  Every lane of
// ed_kernel loads its own copy of each sample (64 lanes x 4 bytes per element,
// 63 of them shared with the neighbours): the vector-memory path, not the
// VALU, bounded it (4 VALU per element, ~40% of the VALU's issue rate).  Here
// lane l of a wave owns windows w0 + 4l + j (j < 4): for query positions
// k..k+3 its four windows need samples x[w0 + 4l + k .. + 7), two aligned
// 16-byte vectors, of which the second is the next step's first -- one
// 16-byte load per lane per four query positions serves sixteen
// window-elements, and a wave's loads are one contiguous kilobyte.  The
// arithmetic per element is ed_kernel's (sub, fma with the reciprocal, fma),
// so the distances are bit-identical.
constexpr int kEd4 = 4;                      // windows per lane
constexpr int kEd4Tile = kWave * kEd4;       // windows per wave

__device__ __forceinline__ float4 ed4_load(const float* __restrict__ s, int64_t i) {
    return *reinterpret_cast<const float4*>(s + i);
}

__device__ __forceinline__ float ed4_pick(const float4& a, const float4& b, int i) {
    switch (i) {
        case 0: return a.x;
        case 1: return a.y;
        case 2: return a.z;
        case 3: return a.w;
        case 4: return b.x;
        case 5: return b.y;
        case 6: return b.z;
        default: return b.w;
    }
}

// Four query positions k..k+3 of the lane's four windows: window j's samples
// are x[w0 + j + k + i], i < 4, the elements i + j of the vector pair (a, b).
__device__ __forceinline__ void ed4_chunk(const float4& a, const float4& b,
                                          const float* __restrict__ qk, const float (&mu)[kEd4],
                                          const float (&rs)[kEd4], float (&acc)[kEd4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float qi = qk[i];
#pragma unroll
        for (int j = 0; j < kEd4; ++j) {
            const float d = fmaf(-(ed4_pick(a, b, i + j) - mu[j]), rs[j], qi);
            acc[j] = fmaf(d, d, acc[j]);
        }
    }
}

__device__ __forceinline__ void ed4_windows(const float* __restrict__ s, int64_t w0,
                                            const float* __restrict__ q,
                                            int32_t m, const float (&mu)[kEd4],
                                            const float (&rs)[kEd4], float (&acc)[kEd4]) {
    // v(c) = x[w0 + 4c .. w0 + 4c + 4); chunk c needs v(c) and v(c + 1).
    // Three chunks per iteration rotate three vectors through r0, r1, r2 so
    // no vector is copied between registers; each load is issued a chunk
    // before its first use.
    auto v = [&](int32_t c) { return ed4_load(s, w0 + 4 * static_cast<int64_t>(c)); };
    float4 r0 = v(0), r1 = v(1);
    int32_t k = 0;
    for (; k + 12 <= m; k += 12) {
        const int32_t c = k / 4;
        const float4 r2 = v(c + 2);
        ed4_chunk(r0, r1, q + k, mu, rs, acc);
        r0 = v(c + 3);
        ed4_chunk(r1, r2, q + k + 4, mu, rs, acc);
        r1 = v(c + 4);
        ed4_chunk(r2, r0, q + k + 8, mu, rs, acc);
    }
    for (; k + 4 <= m; k += 4) {
        const float4 r2 = v(k / 4 + 2);
        ed4_chunk(r0, r1, q + k, mu, rs, acc);
        r0 = r1;
        r1 = r2;
    }
    // last m % 4 query positions (wave-uniform count)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (k + i >= m) break;
        const float qi = q[k + i];
#pragma unroll
        for (int j = 0; j < kEd4; ++j) {
            const float d = fmaf(-(ed4_pick(r0, r1, i + j) - mu[j]), rs[j], qi);
            acc[j] = fmaf(d, d, acc[j]);
        }
    }
}

// Profile only (every window's distance, no abandoning).
__global__ __launch_bounds__(256) void ed4_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, float* __restrict__ out, BandScalars a) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / kWave;
    const int32_t m = a.m;
    // windows [first, end); the host launches only tiles whose 16-byte loads
    // stay inside the stream (a.stride = their count), the rest go to ed_kernel
    const int64_t first = a.base, end = a.base + a.count;
    const int64_t tile0 = (first / kEd4Tile) * kEd4Tile;
    const int64_t ntiles = a.stride;
    for (int64_t t = wave; t < ntiles; t += nwaves) {
        const int64_t w0 = tile0 + t * kEd4Tile + kEd4 * lane;
        float mu[kEd4], rs[kEd4], acc[kEd4];
#pragma unroll
        for (int j = 0; j < kEd4; ++j) {
            const int64_t w = w0 + j;
            const int64_t wc = w < first ? first : (w < end ? w : end - 1);
            mu[j] = avg[wc];
            rs[j] = 1.0f / sd[wc];
            acc[j] = 0.0f;
        }
        ed4_windows(s, w0, q, m, mu, rs, acc);
#pragma unroll
        for (int j = 0; j < kEd4; ++j) {
            const int64_t w = w0 + j;
            if (w >= first && w < end) out[w - first] = acc[j];
        }
    }
}

}  // namespace

int launch_ed(hipStream_t st, int cus, const BandArgs& a) {
    if (a.count <= 0) return MDTW_OK;
    BandScalars sc{a.m, 0, a.count, a.base, a.stride, a.mode, 0};
    // Profile of contiguous windows of a 16-byte aligned stream: four windows
    // per lane for every tile of 256 whose vector loads (up to 12 samples
    // past its last window's end) stay within the samples the windows read;
    // ed_kernel takes the windows after those tiles.  The abandoning modes
    // stay on ed_kernel: a wave abandons only when all of its windows are
    // above the threshold, and 256 neighbouring windows reach that point
    // later than 64 do (at |Q| = 8192 the four-window form took 2.4x longer
    // per search, at |Q| = 256 about 5% less).
    const bool contiguous = !a.idx && a.stride == 1 && a.mode == kModeProfile &&
                            (reinterpret_cast<uintptr_t>(a.s) & 15) == 0;
    int64_t done = 0;  // windows [a.base, a.base + done) handled by ed4_kernel
    if (contiguous) {
        const int64_t first = a.base, end = a.base + a.count;
        const int64_t samples_end = end + a.m - 1;  // one past the last sample read
        const int64_t tile0 = (first / kEd4Tile) * kEd4Tile;
        // tile t is fast when tile0 + (t + 1) * 256 + m + 12 <= samples_end
        const int64_t room = samples_end - a.m - 12 - tile0;
        const int64_t nfast = room >= kEd4Tile ? room / kEd4Tile : 0;
        if (nfast > 0) {
            BandScalars s4 = sc;
            s4.stride = nfast;  // tile count, in the stride slot
            const int64_t blocks = std::min<int64_t>((nfast + 3) / 4, static_cast<int64_t>(cus) * 64);
            hipLaunchKernelGGL(ed4_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st,
                               a.s, a.avg, a.sd, a.qx, a.out, s4);
            done = std::min<int64_t>(a.count, tile0 + nfast * kEd4Tile - first);
        }
    }
    if (done < a.count) {
        BandScalars rest = sc;
        rest.count = a.count - done;
        rest.base = a.base + done * a.stride;
        float* out = a.out && done ? a.out + done : a.out;
        const int32_t* idx = a.idx ? a.idx + done : nullptr;
        const int64_t waves = (rest.count + kWave - 1) / kWave;
        const int64_t blocks = std::min<int64_t>((waves + 3) / 4, static_cast<int64_t>(cus) * 64);
        hipLaunchKernelGGL(ed_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, a.s,
                           a.avg, a.sd, a.qx, idx, out, a.res, a.ctrl, rest);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

}  // namespace mdtw
