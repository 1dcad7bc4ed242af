// cdtw.hip -- banded cDTW dispatch and the wide-band kernel.
//
// W <= kBandWideMaxW (both metrics): the register-band kernel
// (cdtw_band.hpp), one lane per window.  Wider bands (up to kMaxBandW) use
// the anti-diagonal kernel below.
//
// Anti-diagonal kernel.  One wave per window.  Diagonal k = i + j holds the
// cells with offset d = i - j in [-W, W], d = k (mod 2); slot p of diagonal k
// has d = -W + 2p + par(k), par(k) = (k + W) & 1, so p = 0..W.  Lane l owns
// slots l*C .. l*C + C - 1 (slots past W are padding, forced to +inf).  With
// this layout the predecessors of slot p are: diagonal k-2 slot p (i-1, j-1);
// diagonal k-1 slots p-1 and p (par 0) or p and p+1 (par 1), so a step needs
// one cross-lane value.  Query values move one slot down and stream values
// one slot up every other step; fresh query values enter at the top physical
// slot, fresh z-normalised stream values at slot 0.  Out-of-matrix query
// values are +inf and stream values -inf, so q - z is +inf there and never
// NaN.  The reference runs the same diagonal recurrence with a barrier per
// diagonal and shared-memory exchange (cdtw.cuh:261-363); here a diagonal is
// one wave, exchanged through cross-lane shuffles.
#include "cdtw_band.hpp"

namespace mdtw {
namespace {

struct WideScalars {
    int32_t m;
    int32_t w;
    int32_t idx_count_from_ctrl;
    int64_t count;
    int64_t base;
    int64_t stride;
    int32_t mode;
    int32_t persistent;
};

constexpr int kWideCheck = 32;  // steps between abandon checks

// Whole-wave shifts by one lane through DPP (wave_shr:1 / wave_shl:1, GFX9
// family): register-only, no LDS round trip.  The lane with no source gets
// `fill`.
__device__ __forceinline__ float lane_from_below(float v, float fill) {  // lane l <- l-1
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v),
                                                      0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_from_above(float v, float fill) {  // lane l <- l+1
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v),
                                                      0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
    return v;
}

// One diagonal of parity PAR: N = the new diagonal written over A2 (the
// diagonal two back, whose slot c is the only one of it slot c reads), A1
// the previous diagonal.
template <int C, bool L2, int PAR>
__device__ __forceinline__ void wide_step(float (&A2)[C], const float (&A1)[C],
                                          const float (&Qv)[C], const float (&Zv)[C],
                                          const float (&pad)[C]) {
    if constexpr (PAR == 0) {
        const float in = lane_from_below(A1[C - 1], kInf);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float up = c > 0 ? A1[c - 1] : in;
            const float mn = fminf(fminf(A2[c], up), A1[c]);
            const float dd = Qv[c] - Zv[c];
            const float v = L2 ? fmaf(dd, dd, mn) : fabsf(dd) + mn;
            A2[c] = fmaxf(v, pad[c]);
        }
    } else {
        const float in = lane_from_above(A1[0], kInf);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float lf = c + 1 < C ? A1[c + 1] : in;
            const float mn = fminf(fminf(A2[c], A1[c]), lf);
            const float dd = Qv[c] - Zv[c];
            const float v = L2 ? fmaf(dd, dd, mn) : fabsf(dd) + mn;
            A2[c] = fmaxf(v, pad[c]);
        }
    }
}

// Diagonals are processed in pairs, the first of parity FIRST = W & 1 (the
// parity of diagonal 0), so each position's parity is fixed at compile time
// and the two newest diagonals alternate between the arrays X and Y instead
// of being copied.  Query values enter on parity-0 steps and stream values on
// parity-1 steps; their indices grow by one per pair (query: never below 0,
// since 64C > W), and values past the end come from the padded 64-value
// buffers (+inf / -inf), so a pair costs only a few scalar instructions: the
// one-diagonal form spent ~41 scalar instructions per diagonal (parity,
// index arithmetic and bounds branches), and the scalar unit, shared by a
// CU's four SIMDs, bounded the kernel.
template <int C, bool L2, int FIRST>
__device__ float wide_dp_pairs(const float* __restrict__ x, float mu, float sg,
                               const float* __restrict__ q, int32_t m, int32_t W, int mode,
                               const Ctrl* ctrl, bool& alive) {
    const int lane = threadIdx.x;
    const int P = W + 1;
    float pad0[C], pad1[C], X[C], Y[C], Qv[C], Zv[C];
    const int pfin = W >> 1;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int p = lane * C + c;
        pad0[c] = p < P ? 0.0f : kInf;
        pad1[c] = p < P - 1 ? 0.0f : kInf;
        Y[c] = kInf;                       // diagonal -1
        X[c] = p == pfin ? 0.0f : kInf;    // diagonal -2: D[-1][-1]
        const int d = -W + 2 * p + FIRST;  // diagonal 0: i = d/2, j = -d/2
        const int i = d / 2, j = -i;
        Qv[c] = (i >= 0 && i < m) ? q[i] : kInf;
        Zv[c] = (j >= 0 && j < m) ? (x[j] - mu) / sg : -kInf;
    }
    // Index of the query value entering at the next parity-0 step and of the
    // stream value entering at the next parity-1 step (diagonal k: (k - W)/2
    // + 64C and (k + W + 1)/2).
    int iq = (FIRST == 0 ? 0 : 1) - W;
    iq = iq / 2 + kWave * C;
    int iz = (FIRST == 1 ? 0 : 1) + W + 1;
    iz = iz / 2;
    int ib = iq & ~(kWave - 1);
    float qb = ib + lane < m ? q[ib + lane] : kInf;
    int jb = iz & ~(kWave - 1);
    float zb = jb + lane < m ? (x[jb + lane] - mu) / sg : -kInf;
    const bool top = lane == kWave - 1, bottom = lane == 0;

    auto shift_q = [&]() {  // i advances: slot p takes slot p+1's query value
        const float in = lane_from_above(Qv[0], kInf);
#pragma unroll
        for (int c = 0; c + 1 < C; ++c) Qv[c] = Qv[c + 1];
        if (iq >= ib + kWave) {
            ib = iq & ~(kWave - 1);
            qb = ib + lane < m ? q[ib + lane] : kInf;
        }
        const float fresh = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qb), iq - ib));
        Qv[C - 1] = top ? fresh : in;
        ++iq;
    };
    auto shift_z = [&]() {  // j advances: slot p takes slot p-1's stream value
        const float in = lane_from_below(Zv[C - 1], -kInf);
#pragma unroll
        for (int c = C - 1; c > 0; --c) Zv[c] = Zv[c - 1];
        if (iz >= jb + kWave) {
            jb = iz & ~(kWave - 1);
            zb = jb + lane < m ? (x[jb + lane] - mu) / sg : -kInf;
        }
        const float fresh = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(zb), iz - jb));
        Zv[0] = bottom ? fresh : in;
        ++iz;
    };
    auto step_first = [&]() {  // diagonal 2p (parity FIRST) into X
        if constexpr (FIRST == 0) {
            wide_step<C, L2, 0>(X, Y, Qv, Zv, pad0);
            shift_q();
        } else {
            wide_step<C, L2, 1>(X, Y, Qv, Zv, pad1);
            shift_z();
        }
    };
    auto step_second = [&]() {  // diagonal 2p + 1 into Y
        if constexpr (FIRST == 0) {
            wide_step<C, L2, 1>(Y, X, Qv, Zv, pad1);
            shift_z();
        } else {
            wide_step<C, L2, 0>(Y, X, Qv, Zv, pad0);
            shift_q();
        }
    };

    const bool abandon = mode == kModeBest || mode == kModeCollect;
    const float thr = abandon ? load_abandon_bound(ctrl, mode) : kInf;
    // 2M - 1 diagonals: M - 1 pairs, then diagonal 2M - 2 alone.
    for (int32_t pr = 0; pr + 1 < m; ++pr) {
        step_first();
        step_second();
        if (abandon && (pr % (kWideCheck / 2)) == kWideCheck / 2 - 1) {
            float cm = kInf;
#pragma unroll
            for (int c = 0; c < C; ++c) cm = fminf(cm, fminf(X[c], Y[c]));
            cm = wave_min_f32(cm);
            if (cm > thr) {
                alive = false;
                return kInf;
            }
        }
    }
    step_first();
    // D[M-1][M-1]: diagonal 2M-2 (in X), d = 0 -> slot W >> 1.
    float r = X[0];
#pragma unroll
    for (int c = 1; c < C; ++c)
        if (c == pfin % C) r = X[c];
    return __shfl(r, pfin / C, kWave);
}

template <int C, bool L2>
__device__ float wide_dp(const float* __restrict__ x, float mu, float sg,
                         const float* __restrict__ q, int32_t m, int32_t W, int mode,
                         const Ctrl* ctrl, bool& alive) {
    return (W & 1) ? wide_dp_pairs<C, L2, 1>(x, mu, sg, q, m, W, mode, ctrl, alive)
                   : wide_dp_pairs<C, L2, 0>(x, mu, sg, q, m, W, mode, ctrl, alive);
}

template <int C, bool L2>
__global__ __launch_bounds__(64) void wide_kernel(
    const float* __restrict__ s, const float* __restrict__ avg, const float* __restrict__ sd,
    const float* __restrict__ q, const int32_t* __restrict__ idx, float* __restrict__ out,
    unsigned long long* __restrict__ res, Ctrl* __restrict__ ctrl, WideScalars a) {
    const int lane = threadIdx.x;
    int64_t count = a.count;
    if (a.idx_count_from_ctrl)
        count = static_cast<int64_t>(
            __hip_atomic_load(&ctrl->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    unsigned long long n_dp = 0, n_ab = 0;
    auto next = [&]() -> int64_t {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&ctrl->work, 1ull);
        return static_cast<int64_t>(__shfl(b, 0, kWave));
    };
    // First window by grid position; persistent waves then take windows
    // beyond the grid's first pass from the cursor (none when it covers all).
    const int64_t pass = static_cast<int64_t>(gridDim.x);
    int64_t g = static_cast<int64_t>(blockIdx.x);
    while (g < count) {
        const int64_t w = idx ? static_cast<int64_t>(idx[g]) : a.base + g * a.stride;
        bool alive = true;
        const float mu = avg[w];
        float d = wide_dp<C, L2>(s + w, mu, sd[w], q, a.m, a.w, a.mode, ctrl, alive);
        // A window holding a non-finite sample has NaN statistics and so a
        // NaN distance; the band padding (fmaxf with -inf) would turn NaN
        // cells into -inf, so the NaN is restored here.
        if (!(mu == mu)) d = mu;
        if (lane == 0) {
            if (a.mode == kModeProfile || a.mode == kModeSeed) out[g] = d;
            if ((a.mode == kModeBest || a.mode == kModeSeed) && alive)
                atomicMin(&ctrl->best, pack_key(d, static_cast<uint32_t>(w)));
            if (a.mode == kModeSeed) note_bucket(ctrl->bmin, ctrl->bucket_len, w, d);
            if (a.mode == kModeCollect && alive && d <= load_thr(ctrl)) {
                const unsigned long long r = atomicAdd(&ctrl->nres, 1ull);
                res[r] = pack_key(d, static_cast<uint32_t>(w));
                note_bucket(ctrl->bmin, ctrl->bucket_len, w, d);
            }
        }
        n_dp += 1;
        n_ab += alive ? 0 : 1;
        if (!a.persistent) break;
        g = pass >= count ? count : pass + next();
    }
    if (lane == 0 && a.mode != kModeProfile) {
        if (n_dp) atomicAdd(&ctrl->dp, n_dp);
        if (n_ab) atomicAdd(&ctrl->abandoned, n_ab);
    }
}

template <int C, bool L2>
int wide_launch(hipStream_t st, unsigned grid, const BandArgs& a, int32_t w) {
    WideScalars sc{a.m, w, a.idx_count_from_ctrl, a.count, a.base, a.stride, a.mode,
                   a.persistent};
    hipLaunchKernelGGL((wide_kernel<C, L2>), dim3(grid), dim3(kWave), 0, st, a.s, a.avg, a.sd,
                       a.qx, a.idx, a.out, a.res, a.ctrl, sc);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_hip_error(err);
        return MDTW_ERR_HIP;
    }
    return MDTW_OK;
}

int wide_dispatch(hipStream_t st, unsigned grid, const BandArgs& a, int32_t w, bool l2) {
    const int slots = w + 1;
    if (slots <= 64)
        return l2 ? wide_launch<1, true>(st, grid, a, w) : wide_launch<1, false>(st, grid, a, w);
    if (slots <= 128)
        return l2 ? wide_launch<2, true>(st, grid, a, w) : wide_launch<2, false>(st, grid, a, w);
    if (slots <= 256)
        return l2 ? wide_launch<4, true>(st, grid, a, w) : wide_launch<4, false>(st, grid, a, w);
    if (slots <= 512)
        return l2 ? wide_launch<8, true>(st, grid, a, w) : wide_launch<8, false>(st, grid, a, w);
    return l2 ? wide_launch<16, true>(st, grid, a, w) : wide_launch<16, false>(st, grid, a, w);
}

BandLaunchFn band_lookup(int w, bool l2) {
    if (w < 0) return nullptr;
    if (w > kBandMaxW) {
        if (w > kBandWideMaxW) return nullptr;
        if (w <= 144) return l2 ? band_table_4(w) : band_table_8(w);
        if (w <= 176) return l2 ? band_table_5(w) : band_table_9(w);
        if (w <= 208) return l2 ? band_table_6(w) : band_table_10(w);
        return l2 ? band_table_7(w) : band_table_11(w);
    }
    if (w < 30) return band_table_0(w, l2);
    if (w < 40) return band_table_1(w, l2);
    if (w < 50) return band_table_12(w, l2);
    if (w < 60) return band_table_13(w, l2);
    if (w < 70) return band_table_2(w, l2);
    if (w < 80) return band_table_14(w, l2);
    if (w < 90) return band_table_15(w, l2);
    if (w < 98) return band_table_3(w, l2);
    if (w < 106) return band_table_16(w, l2);
    return band_table_17(w, l2);
}

}  // namespace

// Band kernels need the padded query (qx); the wide kernel the plain query.
// The caller passes both through BandArgs::qx / q_plain.
int launch_dp(hipStream_t st, int cus, const BandArgs& a, const float* q_plain, int32_t w,
              int32_t metric, bool force_wide, const LaunchShape& shape) {
    const bool l2 = metric != 0;
    const int64_t count_hint = a.idx_count_from_ctrl ? -1 : a.count;
    if (!a.idx_count_from_ctrl && a.count <= 0) return MDTW_OK;
    BandLaunchFn fn = force_wide || a.qx == nullptr ? nullptr : band_lookup(w, l2);
    if (fn) {
        const int64_t per_block = static_cast<int64_t>(kBandWaves) * kWave;
        int64_t grid = static_cast<int64_t>(cus) * shape.dp_blocks_per_cu;
        if (!a.persistent) grid = std::min(grid, (count_hint + per_block - 1) / per_block);
        return fn(st, static_cast<unsigned>(grid), a);
    }
    if (w > kMaxBandW) return MDTW_ERR_UNSUPPORTED;
    BandArgs b = a;
    b.qx = q_plain;
    unsigned grid;
    if (a.persistent)
        grid = static_cast<unsigned>(cus * shape.wide_waves_per_cu);
    else
        grid = static_cast<unsigned>(count_hint);
    return wide_dispatch(st, grid, b, w, l2);
}

}  // namespace mdtw
