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

// ---------------------------------------------------------------- global memory
// Loads and stores through pointers known to address device memory (HBM):
// a pointer that reaches a kernel through memory (a table of basis-vector
// pointers, a struct) is a generic pointer to the compiler, which then emits
// flat instructions; flat loads count against both the vector-memory and
// the LDS counters, so every wait for an LDS value also waits for them.
// Through the global address space they are global_load/global_store, waited
// for on the vector-memory counter alone.
#define BSPAI_GLOBAL_AS __attribute__((address_space(1)))
template <typename T> __device__ inline T gld(const T *p) {
    return *reinterpret_cast<const BSPAI_GLOBAL_AS T *>(reinterpret_cast<std::uintptr_t>(p));
}
template <typename T> __device__ inline void gst(T *p, const T &v) {
    *reinterpret_cast<BSPAI_GLOBAL_AS T *>(reinterpret_cast<std::uintptr_t>(p)) = v;
}
// The same for HIP's vector types (float4, double2), through a native
// vector of the same elements.
template <typename V> struct NativeOf {
    using E = decltype(V{}.x);
    typedef E type __attribute__((ext_vector_type(sizeof(V) / sizeof(E))));
};
template <typename V> __device__ inline V gld_vec(const V *p) {
    using N = typename NativeOf<V>::type;
    const N n = *reinterpret_cast<const BSPAI_GLOBAL_AS N *>(reinterpret_cast<std::uintptr_t>(p));
    V v;
    __builtin_memcpy(&v, &n, sizeof(V));
    return v;
}
template <typename V> __device__ inline void gst_vec(V *p, const V &v) {
    using N = typename NativeOf<V>::type;
    N n;
    __builtin_memcpy(&n, &v, sizeof(V));
    *reinterpret_cast<BSPAI_GLOBAL_AS N *>(reinterpret_cast<std::uintptr_t>(p)) = n;
}

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

// ---------------------------------------------------------------- row sums
// Accumulators of the sparse products.  Acc<TA> forms each product and the
// running sum in TA.  Acc<Comp32> is compensated FP32 (Ogita, Rump and Oishi,
// "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005,
// Algorithm Dot2): every operation is an FP32 operation; a product is split
// exactly into p + e by one fused multiply-add, the sum by TwoSum, and the
// errors gather in lo.  The row sum comes out as if accumulated in twice the
// FP32 precision and rounded once, which keeps the digits that cancelling
// rows (a stencil's interior rows sum to zero) lose in a plain FP32 sum.
struct Comp32 {
    float hi, lo;
};

template <typename TA> struct Acc {
    using Prod = TA;
    template <typename A, typename B> __device__ static Prod mul(A a, B b) {
        return (TA)a * (TA)b;
    }
    __device__ static void add(TA &s, Prod p) { s += p; }
    __device__ static TA zero() { return (TA)0; }
    template <int G> __device__ static TA group(TA v) { return group_sum<G>(v); }
    template <typename TY> __device__ static TY out(TA v) { return (TY)v; }
};

__device__ inline void two_sum(float a, float b, float &s, float &t) {
    s = a + b;
    const float bb = s - a;
    t = (a - (s - bb)) + (b - bb);
}

template <> struct Acc<Comp32> {
    using Prod = Comp32; // p = fl(a b), e = a b - p exactly
    __device__ static Prod mul(float a, float b) {
        const float p = a * b;
        return Comp32{p, __builtin_fmaf(a, b, -p)};
    }
    __device__ static void add(Comp32 &s, Prod q) {
        float h, t;
        two_sum(s.hi, q.hi, h, t);
        s.hi = h;
        s.lo = s.lo + (t + q.lo);
    }
    __device__ static Comp32 zero() { return Comp32{0.f, 0.f}; }
    template <int G> __device__ static Comp32 group(Comp32 v) {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            const Comp32 o{__shfl_xor(v.hi, off, G), __shfl_xor(v.lo, off, G)};
            float h, t;
            two_sum(v.hi, o.hi, h, t);
            v.lo = (v.lo + o.lo) + t;
            v.hi = h;
        }
        return v;
    }
    template <typename TY> __device__ static TY out(Comp32 v) { return (TY)(v.hi + v.lo); }
};

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
