// ldtw.hip -- greedy "locally best" band-constrained DTW (LDTW), one lane
// per window.
//
// Restates ztsh2_ldtw / z_serial_ldtw (ldtw.cuh:74-127, :294-391): from cell
// (0, 0) the path repeatedly steps to the cheapest of diagonal, down and right
// (band |i - j| <= W checked for down/right, INFTY = 1e10 for moves that are
// not allowed), with the reference's tie rules: diagonal only if strictly
// below both others, then down if not above diagonal and strictly below
// right, else right.  The cost is the sum of the visited cells' local costs.
// z-normalisation is (x - mu) / sigma (the IEEE quotient), squares and sums
// are separately rounded (no contraction), so the value is bit-exact with the
// CPU restatement.  Each z-normalised stream value is computed once and carried
// while the path stays on its column.
//
// Search modes abandon a wave once every lane's partial cost (which only
// grows) exceeds the threshold.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {

constexpr float kLdtwInf = 1e10f;  // ldtw.cuh:8 INFTY
constexpr int kLdtwCheck = 32;
constexpr int kLdtwLdsMaxM = 2048;

template <bool L2>
__device__ __forceinline__ float lcost(float t) {
    if constexpr (L2)
        return t * t;
    else
        return t < 0 ? -t : t;
}

// FASTZ: z = (x - mu) / sigma through band_z_fast (every window of the wave
// admits it, band_z_window_fast), else the IEEE division; the same quotient.
template <bool FASTZ>
__device__ __forceinline__ float ldtw_z(float x, float mu, float sg, float y) {
    if constexpr (FASTZ)
        return band_z_fast(x, mu, sg, y);
    else
        return (x - mu) / sg;
}

template <bool L2, bool QL, bool FASTZ>
__device__ __forceinline__ float ldtw_window(const float* __restrict__ x, float mu, float sg,
                             const float* __restrict__ q, int32_t M, int32_t W, int mode,
                             const Ctrl* ctrl, bool& alive) {
    int32_t i = 0, j = 0;
    const float y = 1.0f / sg;
    float qi = q[0], qi1 = M > 1 ? q[1] : 0.0f;
    float zj = ldtw_z<FASTZ>(x[0], mu, sg, y);
    float zj1 = M > 1 ? ldtw_z<FASTZ>(x[1], mu, sg, y) : 0.0f;
    float d = lcost<L2>(qi - zj);
    const bool abandon = mode == kModeBest || mode == kModeCollect;
    const float thr = abandon ? load_abandon_bound(ctrl, mode) : kInf;
    int step = 0;
    // A valid greedy path reaches (M-1, M-1) in at most 2M-2 moves (every
    // move advances i or j and the path never leaves the band); the cap and
    // the clamps below only bound the loop and the indices.
    const int max_steps = 2 * M - 2;
    while ((i + 1 != M || j + 1 != M) && step < max_steps) {
        // The three candidate costs are computed unconditionally and the
        // disallowed ones replaced by INFTY with selects: no divergent
        // branches (their exec-mask bookkeeping ran on the scalar unit).
        // |i+1 - j| <= W and |j+1 - i| <= W, with the path inside the band
        // (|i - j| <= W), reduce to i - j < W and j - i < W.
        const bool vi = i + 1 < M, vj = j + 1 < M;
        const int32_t dij = i - j;
        const float cd = lcost<L2>(qi1 - zj1);
        const float cn = lcost<L2>(qi1 - zj);
        const float cr = lcost<L2>(qi - zj1);
        const float diag = vi && vj ? cd : kLdtwInf;
        const float down = vi && dij < W ? cn : kLdtwInf;
        const float rght = vj && -dij < W ? cr : kLdtwInf;
        const bool dsd = diag < down;
        const bool dsr = diag < rght;
        const bool dnr = down < rght;
        bool adv_i, adv_j;
        if (dsd && dsr) {
            d += diag;
            adv_i = true;
            adv_j = true;
        } else if (!dsd && dnr) {
            d += down;
            adv_i = true;
            adv_j = false;
        } else {
            d += rght;
            adv_i = false;
            adv_j = true;
        }
        if constexpr (QL) {
            // query in LDS: read the next value every step and select, no
            // divergent branch
            const bool ai = adv_i && i + 1 < M;
            const float qn = q[i + 2 < M ? i + 2 : M - 1];
            const float qn1 = i + 2 < M ? qn : 0.0f;
            qi = ai ? qi1 : qi;
            qi1 = ai ? qn1 : qi1;
            i = ai ? i + 1 : i;
        } else if (adv_i && i + 1 < M) {
            ++i;
            qi = qi1;
            qi1 = i + 1 < M ? q[i + 1] : 0.0f;
        }
        if constexpr (FASTZ) {
            // stream advance as selects: the next value is loaded and
            // normalised every step (four operations through band_z_fast),
            // which costs less than the divergent branch's exec-mask work
            const bool aj = adv_j && j + 1 < M;
            const float zn = ldtw_z<FASTZ>(x[j + 2 < M ? j + 2 : M - 1], mu, sg, y);
            const float zn1 = j + 2 < M ? zn : 0.0f;
            zj = aj ? zj1 : zj;
            zj1 = aj ? zn1 : zj1;
            j = aj ? j + 1 : j;
        } else if (adv_j && j + 1 < M) {
            ++j;
            zj = zj1;
            zj1 = j + 1 < M ? ldtw_z<FASTZ>(x[j + 1], mu, sg, y) : 0.0f;
        }
        ++step;
        if (abandon && (step % kLdtwCheck) == 0) {
            alive = alive && !(d > thr);
            if (!__any(alive)) return kInf;
        }
    }
    return d;
}

__device__ __forceinline__ uint32_t rank_in(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// QL: the query is staged in LDS (dynamic shared memory, 4M bytes): the
// path's per-lane query reads are gathers, and from global memory they
// queue in the vector-memory address unit behind the stream reads.
template <bool L2, bool QL>
__global__ __launch_bounds__(64) void ldtw_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ qg, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, BandScalars a, int32_t W) {
    const int lane = threadIdx.x;
    extern __shared__ float ldtw_q[];
    const float* __restrict__ q = qg;
    if constexpr (QL) {
        for (int k = lane; k < a.m; k += kWave) ldtw_q[k] = qg[k];
        __syncthreads();
        q = ldtw_q;
    }
    int64_t count = a.count;
    if (a.idx_count_from_ctrl)
        count = static_cast<int64_t>(
            __hip_atomic_load(&ctrl->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    unsigned long long n_dp = 0, n_ab = 0;
    // First chunk by grid position; persistent waves then take chunks beyond
    // the grid's first pass from the cursor (none when the grid covers all).
    const int64_t pass = static_cast<int64_t>(gridDim.x) * kWave;
    int64_t base = static_cast<int64_t>(blockIdx.x) * kWave;
    while (base < count) {
        const int64_t g = base + lane;
        const bool valid = g < count;
        const int64_t gc = valid ? g : count - 1;
        const int64_t w = idx ? static_cast<int64_t>(idx[gc]) : a.base + gc * a.stride;
        bool alive = valid;
        const float mu = avg[w], sg = sd[w];
        const float d =
            __all(band_z_window_fast(mu, sg))
                ? ldtw_window<L2, QL, true>(s + w, mu, sg, q, a.m, W, a.mode, ctrl, alive)
                : ldtw_window<L2, QL, false>(s + w, mu, sg, q, a.m, W, a.mode, ctrl, alive);
        if (a.mode == kModeProfile || a.mode == kModeSeed) {
            if (valid) out[g] = d;
        }
        if (a.mode == kModeBest || a.mode == kModeSeed) {
            unsigned long long key = alive ? pack_key(d, static_cast<uint32_t>(w)) : ~0ull;
            key = wave_min_u64(key);
            if (lane == 0 && key != ~0ull) atomicMin(&ctrl->best, key);
            if (a.mode == kModeSeed && valid) note_bucket(ctrl->bmin, ctrl->bucket_len, w, d);
        } else if (a.mode == kModeCollect) {
            const bool keep = alive && d <= load_thr(ctrl);
            if (keep) note_bucket(ctrl->bmin, ctrl->bucket_len, w, d);
            const unsigned long long km = __ballot(keep);
            if (km) {
                unsigned long long rb = 0;
                if (lane == 0)
                    rb = atomicAdd(&ctrl->nres, static_cast<unsigned long long>(__popcll(km)));
                rb = __shfl(rb, 0, kWave);
                if (keep) res[rb + rank_in(km)] = pack_key(d, static_cast<uint32_t>(w));
            }
        }
        n_dp += __popcll(__ballot(valid));
        n_ab += __popcll(__ballot(valid && !alive));
        if (!a.persistent) break;
        base = pass >= count ? count : pass + band_next(ctrl, lane);
    }
    if (lane == 0 && a.mode != kModeProfile) {
        if (n_dp) atomicAdd(&ctrl->dp, n_dp);
        if (n_ab) atomicAdd(&ctrl->abandoned, n_ab);
    }
}

}  // namespace

int launch_ldtw(hipStream_t st, int cus, const BandArgs& a, int32_t w, int32_t metric) {
    if (!a.idx_count_from_ctrl && a.count <= 0) return MDTW_OK;
    BandScalars sc{a.m, a.idx_count_from_ctrl, a.count, a.base, a.stride, a.mode, a.persistent, 0, 0};
    // Persistent: 32 one-wave workgroups per CU, 8 waves per SIMD (the
    // kernel's register occupancy), which hides the per-step stream gathers
    // of one wave behind the others' walks (C3 |Q| = 512 / 1024 searches
    // 67 / 111 -> 60 / 91 ms against 16 per CU; 64 per CU is no faster).
    const unsigned grid = a.persistent ? static_cast<unsigned>(cus * 32)
                                       : static_cast<unsigned>((a.count + kWave - 1) / kWave);
    // Query in LDS up to kLdtwLdsMaxM (8 KiB per one-wave workgroup keeps
    // LDS from limiting residency); longer queries read it from global memory.
    const bool ql = a.m <= kLdtwLdsMaxM;
    const size_t lds = ql ? sizeof(float) * static_cast<size_t>(a.m) : 0;
    auto k = metric ? (ql ? ldtw_kernel<true, true> : ldtw_kernel<true, false>)
                    : (ql ? ldtw_kernel<false, true> : ldtw_kernel<false, false>);
    hipLaunchKernelGGL(k, dim3(grid), dim3(kWave), lds, st, a.s, a.avg, a.sd, a.qx, a.idx, a.out,
                       a.res, a.ctrl, sc, w);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

}  // namespace mdtw
