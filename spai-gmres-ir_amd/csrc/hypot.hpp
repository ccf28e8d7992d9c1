// hypot matching the host libm bit for bit; usable from host and device code.
#pragma once

#include <cmath>

#ifndef __HIPCC__
#define BSPAI_HD
#else
#define BSPAI_HD __host__ __device__
#endif

namespace bspai_hip {

// ---------------------------------------------------------------- hypot
// hypot(x, y) as computed by glibc 2.35's binary64 hypot: the corrected
// algorithm of C. F. Borges, "An Improved Algorithm for hypot(a,b)" (2019),
// non-FMA correction form, with binary-power rescaling outside
// [2^-511, 2^511].  Every SPAI norm in the reference is a chain of std::hypot
// calls (src/spai.cpp:22-27), so the GPU must reproduce that function bit for
// bit; the sequence below was checked against the host libm on 3e7 inputs
// (tests/test_hypot.py repeats the check on every build).  It needs IEEE
// add/sub/mul/div/sqrt without contraction (-ffp-contract=off).
BSPAI_HD inline double hypot_kernel(double ax, double ay) {
    double h = std::sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double delta = h - ay;
        t1 = ax * (2.0 * delta - ax);
        t2 = (delta - 2.0 * (ax - ay)) * delta;
    } else {
        const double delta = h - ax;
        t1 = 2.0 * delta * (ax - 2.0 * ay);
        t2 = (4.0 * delta - ay) * ay + delta * delta;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

BSPAI_HD inline double hypot_ref(double x, double y) {
    if (!std::isfinite(x) || !std::isfinite(y)) {
        if (std::isinf(x) || std::isinf(y))
            return INFINITY;
        return x + y;
    }
    x = std::fabs(x);
    y = std::fabs(y);
    const double ax = x < y ? y : x;
    const double ay = x < y ? x : y;
    const double kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-511, kEps = 0x1p-54;
    if (ax > kLarge) {
        if (ay <= ax * kEps)
            return ax + ay;
        return hypot_kernel(ax * kScale, ay * kScale) / kScale;
    }
    if (ay < kTiny) {
        if (ax >= ay / kEps)
            return ax + ay;
        return hypot_kernel(ax / kScale, ay / kScale) * kScale;
    }
    if (ay <= ax * kEps)
        return ax + ay;
    return hypot_kernel(ax, ay);
}

// binary32 hypot for the FP32 SPAI: squares are exact in binary64, so the
// double-precision root rounded once to float is within one rounding of the
// correctly rounded result.
BSPAI_HD inline float hypot_ref(float x, float y) {
    const double dx = x, dy = y;
    return (float)std::sqrt(dx * dx + dy * dy);
}

} // namespace bspai_hip
