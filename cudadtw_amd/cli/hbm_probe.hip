// hbm_probe -- the HBM rates this device reaches for the access patterns of
// the path's memory-bound kernels: read-only (profile argmin, ED inputs),
// write-only, copy, and one read to two writes (the window statistics read 4
// and write 8 bytes per window).  Each pattern runs over 2 GiB buffers with
// 16-byte accesses, plain and nontemporal, at several grid sizes; the best
// rate per pattern is the ceiling a kernel of that pattern can be measured
// against (the datasheet's 8 TB/s is not reached by any of them).
//
// Usage: hbm_probe            (prints one line per pattern, grid and form)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

typedef float v4f __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "hbm_probe: %s (line %d)\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

template <bool NT>
__global__ __launch_bounds__(256) void read_kernel(const v4f* __restrict__ a, int64_t n,
                                                   float* out) {
    float s = 0.0f;
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const v4f v = NT ? __builtin_nontemporal_load(a + i) : a[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == 1234.5f) out[0] = s;  // keeps the loads; never true for the zero-filled input
}

template <bool NT>
__global__ __launch_bounds__(256) void write_kernel(v4f* __restrict__ a, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const v4f v = {1.0f, 2.0f, 3.0f, float(i)};
        if (NT)
            __builtin_nontemporal_store(v, a + i);
        else
            a[i] = v;
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const v4f* __restrict__ a,
                                                   v4f* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        if (NT)
            __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
        else
            b[i] = a[i];
    }
}

template <bool NT>
__global__ __launch_bounds__(256) void read1_write2_kernel(const v4f* __restrict__ a,
                                                           v4f* __restrict__ b,
                                                           v4f* __restrict__ c, int64_t n) {
    for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        const v4f v = a[i];
        if (NT) {
            __builtin_nontemporal_store(v, b + i);
            __builtin_nontemporal_store(v, c + i);
        } else {
            b[i] = v;
            c[i] = v;
        }
    }
}

int main() {
    const int64_t bytes = 2ll << 30;
    const int64_t n = bytes / 16;
    v4f *a, *b, *c;
    float* o;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMalloc(&c, bytes));
    CHECK(hipMalloc(&o, 64));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    CHECK(hipMemset(c, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int reps = 5;
    for (int grid : {2048, 8192, 16384, 65536}) {
        auto run = [&](const char* name, double moved, auto launch) -> int {
            launch();
            CHECK(hipDeviceSynchronize());
            CHECK(hipEventRecord(e0));
            for (int r = 0; r < reps; ++r) launch();
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            std::printf("grid %6d  %-14s %8.3f ms  %5.2f TB/s\n", grid, name, ms / reps,
                        moved / (ms / reps * 1e-3) / 1e12);
            return 0;
        };
        const dim3 g(grid), t(256);
        if (run("read", bytes, [&] { read_kernel<false><<<g, t>>>(a, n, o); }) ||
            run("read_nt", bytes, [&] { read_kernel<true><<<g, t>>>(a, n, o); }) ||
            run("write", bytes, [&] { write_kernel<false><<<g, t>>>(b, n); }) ||
            run("write_nt", bytes, [&] { write_kernel<true><<<g, t>>>(b, n); }) ||
            run("copy", 2.0 * bytes, [&] { copy_kernel<false><<<g, t>>>(a, b, n); }) ||
            run("copy_nt", 2.0 * bytes, [&] { copy_kernel<true><<<g, t>>>(a, b, n); }) ||
            run("read1_write2", 3.0 * bytes,
                [&] { read1_write2_kernel<false><<<g, t>>>(a, b, c, n); }) ||
            run("read1_write2_nt", 3.0 * bytes,
                [&] { read1_write2_kernel<true><<<g, t>>>(a, b, c, n); }))
            return 1;
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(c));
    CHECK(hipFree(o));
    return 0;
}
