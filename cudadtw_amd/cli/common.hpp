// common.hpp -- shared pieces of the GPU-DTW / GPU-ED / GPU-LDTW drivers:
// raw little-endian binary input (GPU-dtw.cu:59-64), HIP allocation with
// checks, elapsed-time lines.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "mdtw_compat.hpp"

#define HIP_OK(call)                                                                  \
    do {                                                                              \
        hipError_t err_ = (call);                                                     \
        if (err_ != hipSuccess) {                                                     \
            std::cerr << "HIP error " << hipGetErrorString(err_) << " at " << __FILE__ \
                      << ":" << __LINE__ << std::endl;                                \
            std::exit(1);                                                             \
        }                                                                             \
    } while (0)

template <class T>
std::vector<T> read_bin(const char* path, long count) {
    std::vector<T> v(static_cast<size_t>(count));
    std::ifstream f(path, std::ios::binary | std::ios::in);
    if (!f) {
        std::cerr << "cannot open " << path << std::endl;
        std::exit(1);
    }
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(sizeof(T) * count));
    if (f.gcount() != static_cast<std::streamsize>(sizeof(T) * count)) {
        std::cerr << path << ": expected " << count << " values" << std::endl;
        std::exit(1);
    }
    return v;
}

struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
            .count();
    }
};
