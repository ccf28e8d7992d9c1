// Shared device/host helpers for the gfx950 SPAI-GMRES-IR library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "hypot.hpp"

namespace bspai_hip {

// ---------------------------------------------------------------- host errors
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what, const char *file, int line) {
    if (e != hipSuccess) {
        char buf[512];
        std::snprintf(buf, sizeof buf, "%s failed at %s:%d: %s", what, file, line,
                      hipGetErrorString(e));
        throw HipError(buf);
    }
}
#define HIP_CHECK(expr) ::bspai_hip::hip_check((expr), #expr, __FILE__, __LINE__)
#define HIP_LAUNCH_CHECK() ::bspai_hip::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

constexpr int kWave = 64; // CDNA wavefront width

// ---------------------------------------------------------------- reductions
template <typename T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

template <typename T> __device__ inline T wave_max(T v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const T o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Sum over a group of G consecutive lanes (G a power of two <= 64); every lane
// of the group gets the result.  Fixed butterfly order: deterministic.
template <int G, typename T> __device__ inline T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, G);
    return v;
}

// Block-wide sum (blockDim.x a multiple of 64, <= 1024) in a fixed order.
template <typename T> __device__ inline T block_sum(T v, T *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0)
        lds[wid] = v;
    __syncthreads();
    T r = 0;
    const int nw = blockDim.x / kWave;
    if (threadIdx.x == 0) {
        for (int i = 0; i < nw; ++i)
            r += lds[i];
        lds[0] = r;
    }
    __syncthreads();
    r = lds[0];
    __syncthreads();
    return r;
}

template <typename T> __device__ inline T block_max(T v, T *lds) {
    v = wave_max(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (lane == 0)
        lds[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T r = lds[0];
        for (int i = 1; i < (int)(blockDim.x / kWave); ++i)
            r = lds[i] > r ? lds[i] : r;
        lds[0] = r;
    }
    __syncthreads();
    const T r = lds[0];
    __syncthreads();
    return r;
}

inline unsigned ceil_div(std::int64_t a, std::int64_t b) {
    return (unsigned)((a + b - 1) / b);
}

} // namespace bspai_hip
