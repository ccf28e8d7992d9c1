// Adaptive SPAI construction on gfx950.
//
// Method (reference src/spai.cpp:100-374): scale rows by d_i = 1/max_j|a_ij|,
// set B = A^T D, and fit each column k of N ~ B^-1 by adaptive Householder
// least squares; the preconditioner is P = N^T D.  Because B = A^T D, column j
// of B is row j of A scaled by d_j, so the kernel reads A's CSR directly and
// needs B's CSR only as a structure (row i of B lists the rows of A holding
// column i).
//
// Execution: one lane owns one column k and runs its whole adaptive loop
// (shadow, dense gather, QR, residual, candidates, scores, selection) in the
// reference's floating-point order, so FP64 results are bit-identical to the
// CPU's: every dot product and hypot chain is sequential in the same order, and
// hypot_ref reproduces the host libm bit for bit.  Columns are independent, so
// 64 lanes of a wavefront run 64 columns in SIMT; on stencil matrices the
// interior columns share one control path.  Per-lane work arrays live in global
// memory interleaved by lane (element e of lane l at [(wave*E + e)*64 + l]) so
// that the lanes of a wave access one contiguous 512-byte run together.
// Columns whose shadow, pattern or candidate list outgrow the launch's caps
// stop with kOverflow and are re-run by the host with larger caps (size
// buckets); re-running from scratch is exact because the method is
// deterministic.
#include "common.hpp"
#include "kernels.hpp"

namespace bspai_hip {

namespace {

template <typename T> struct Unit;
template <> struct Unit<double> {
    __device__ static double u() { return 0x1p-53; }
};
template <> struct Unit<float> {
    __device__ static float u() { return 0x1p-24f; }
};

// src/spai.cpp:273-306
template <typename T>
__device__ inline int score_of(T rr, T dot, T den, T &score, bool &clamped) {
    clamped = false;
    if (!isfinite(rr) || !isfinite(dot) || !isfinite(den))
        return kSpaiNumericalFailure;
    if (den == 0)
        return kSpaiZeroDenominator;
    if (rr < 0 || den < 0)
        return kSpaiInvalidScore;
    const T quotient = dot * dot / den;
    T delta = rr - quotient;
    const T bound = T(32) * Unit<T>::u() * (rr + fabs(quotient));
    if (!isfinite(quotient) || !isfinite(delta) || !isfinite(bound))
        return kSpaiNumericalFailure;
    if (delta < 0) {
        if (-delta > bound)
            return kSpaiInvalidScore;
        delta = 0;
        clamped = true;
    }
    score = sqrt(delta);
    return kSpaiSuccess;
}

struct Caps {
    int mcap, qcap, ccap;
};

constexpr int kFuse = 4; // trailing columns reflected together (register budget)

template <typename T> struct Lane {
    T *d;          // float work area
    int *i;        // index work area
    std::int64_t dbase, ibase;
    __device__ T &D(int e) const { return d[dbase + (std::int64_t)e * kWave]; }
    __device__ int &I(int e) const { return i[ibase + (std::int64_t)e * kWave]; }
};

// Householder vector from x (held in v[0..len), 2-norm `length` > 0):
// v = x / length; v_0 += sign(v_0); v /= ||v|| (src/spai.cpp:48-56).  Returns
// the diagonal entry -copysign(length, x_0).
// The sequential loops below load kLB elements ahead of their dependent
// chains (hypot, accumulation), so a lane keeps several loads in flight.
constexpr int kLB = 3;

template <typename T>
__device__ T householder_vector(const Lane<T> &L, int ov, int len, T length) {
    const T diagonal = -copysign(length, L.D(ov));
    T vnorm = 0;
    for (int r0 = 0; r0 < len; r0 += kLB) {
        T x[kLB];
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            x[u] = L.D(ov + min(r0 + u, len - 1));
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            if (r0 + u < len) {
                T y = x[u] / length;
                if (r0 + u == 0)
                    y += copysign(T(1), y);
                L.D(ov + r0 + u) = y;
                vnorm = hypot_ref(vnorm, y);
            }
    }
    for (int r0 = 0; r0 < len; r0 += kLB) {
        T x[kLB];
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            x[u] = L.D(ov + min(r0 + u, len - 1));
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            if (r0 + u < len)
                L.D(ov + r0 + u) = x[u] / vnorm;
    }
    return diagonal;
}

// hypot chain over a lane array, loads batched ahead of the chain.
template <typename T> __device__ T lane_hypot_chain(const Lane<T> &L, int o, int len) {
    T acc = 0;
    for (int r0 = 0; r0 < len; r0 += kLB) {
        T x[kLB];
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            x[u] = L.D(o + min(r0 + u, len - 1));
#pragma unroll
        for (int u = 0; u < kLB; ++u)
            if (r0 + u < len)
                acc = hypot_ref(acc, x[u]);
    }
    return acc;
}

} // namespace

// Elements per lane for the given caps.
std::int64_t spai_lane_doubles(int mcap, int qcap, int ccap) {
    return (std::int64_t)mcap * qcap + 3LL * mcap + ccap + qcap;
}
std::int64_t spai_lane_ints(int mcap, int qcap, int ccap) {
    return 2LL * mcap + 2LL * qcap + 2LL * ccap + qcap;
}

// Six waves per SIMD: the register budget (80 VGPRs) spills a few values to
// scratch, but more lanes in flight hide more of the work areas' load latency
// (measured at 128^3, FP64: 141.9 ms at four waves, 130.4 at five, 125.2 at
// six, 125.4 at eight).
template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
spai_columns_kernel(int n, const int *__restrict__ aptr, const int *__restrict__ acol,
                    const T *__restrict__ aval, const T *__restrict__ dscale,
                    const int *__restrict__ bptr, const int *__restrict__ bcol, T eps,
                    long long alpha, int beta, const int *__restrict__ worklist, int nwork,
                    int *__restrict__ cursor, Caps caps, T *__restrict__ dwork,
                    int *__restrict__ iwork, std::int64_t lane_d, std::int64_t lane_i,
                    SpaiColumnOut out, int *__restrict__ out_j, T *__restrict__ out_c,
                    std::int64_t out_stride) {
    const std::int64_t g = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const std::int64_t wave = g / kWave;
    const int lane = (int)(g % kWave);
    Lane<T> L{dwork, iwork, wave * lane_d * kWave + lane, wave * lane_i * kWave + lane};
    const int M = caps.mcap, Q = caps.qcap, CC = caps.ccap;
    // double area: R [Q][M] | rhs [M] | v [M] | res [M] | scores [CC] | coef [Q]
    const int oR = 0, oRhs = M * Q, oV = oRhs + M, oRes = oV + M, oSc = oRes + M, oCoef = oSc + CC;
    // int area: I [M] | Itmp [M] | J [Q] | Jtmp [Q] | cand [CC] | ctmp [CC] | sel [Q]
    int oI = 0, oIt = M;
    const int oJ = 2 * M, oJt = 2 * M + Q, oSel = 2 * M + 2 * Q + 2 * CC;
    int oC = 2 * M + 2 * Q, oCt = oC + CC;

    for (;;) {
        int slot = atomicAdd(cursor, 1);
        if (slot >= nwork)
            break;
        const int k = worklist[slot];
        int status = kSpaiNotAttempted;
        int fits = 0, growths = 0, clamps = 0, q = 1, qdone = 0;
        double qr_flops = 0;
        double last_res = __builtin_nan(""), last_piv = __builtin_nan("");
        L.I(oJ) = k;
        int *jout = out_j + (std::int64_t)slot * out_stride;
        T *cout_ = out_c + (std::int64_t)slot * out_stride;
        for (;;) {
            // ---- shadow I = union of cols(A row j), j in J (ascending)
            int m = 0;
            bool over = false;
            for (int jj = 0; jj < q && !over; ++jj) {
                const int j = L.I(oJ + jj);
                int a = 0, e = aptr[j];
                const int e1 = aptr[j + 1];
                int t = 0;
                while (a < m || e < e1) {
                    int x;
                    if (e >= e1 || (a < m && L.I(oI + a) < acol[e]))
                        x = L.I(oI + a++);
                    else if (a >= m || acol[e] < L.I(oI + a))
                        x = acol[e++];
                    else {
                        x = acol[e++];
                        ++a;
                    }
                    if (t >= M) {
                        over = true;
                        break;
                    }
                    L.I(oIt + t++) = x;
                }
                const int tmp = oI;
                oI = oIt;
                oIt = tmp;
                m = t;
            }
            if (over) {
                status = kSpaiOverflow;
                break;
            }
            if (m < q) {
                status = kSpaiRankDeficient;
                break;
            }
            // ---- dense gather R(i, jj) = a_{J[jj], I[i]} * d_{J[jj]}, rhs = e_k(I)
            for (int c = 0; c < q; ++c)
                for (int r = 0; r < m; ++r)
                    L.D(oR + c * M + r) = T(0);
            int pk = -1;
            for (int r = 0; r < m; ++r) {
                const int ir = L.I(oI + r);
                L.D(oRhs + r) = ir == k ? T(1) : T(0);
                if (ir == k)
                    pk = r;
            }
            if (pk < 0) { // k not in I: impossible after the diagonal check
                status = kSpaiNumericalFailure;
                break;
            }
            for (int c = 0; c < q; ++c) {
                const int j = L.I(oJ + c);
                const T dj = dscale[j];
                int p = 0;
                for (int e = aptr[j]; e < aptr[j + 1]; ++e) {
                    const int col = acol[e];
                    while (L.I(oI + p) < col)
                        ++p;
                    L.D(oR + c * M + p) = aval[e] * dj;
                }
            }
            // ---- Householder least squares (src/spai.cpp:30-98)
            T inf_norm = 0;
            for (int r = 0; r < m; ++r) {
                T s = 0;
                for (int c = 0; c < q; ++c)
                    s += fabs(L.D(oR + c * M + r));
                inf_norm = s > inf_norm ? s : inf_norm;
            }
            T min_pivot = T(INFINITY);
            int fit_status = kSpaiSuccess;
            // Step j reflects column j+1 first: its updated rows j+1.. are the
            // next Householder vector, so the next norm chain runs in the same
            // pass.  The other trailing columns (and the right side, slot q)
            // are then reflected in one pass that also accumulates their dot
            // products with the next vector.  Every value sees the
            // reference's operations in the reference's order
            // (src/spai.cpp:41-78); only the number of passes over the work
            // area changes.  Status precedence is kept: a failure while
            // reflecting step j is reported before the next vector's norm.
            int oVc = oV, oVn = oRes; // Householder vectors of steps j, j+1
            const int oDot = oSc;     // per trailing slot: dot with the current v
            auto slot_base = [&](int c) { return c < q ? oR + c * M : oRhs; };
            for (int r = 0; r < m; ++r)
                L.D(oVc + r) = L.D(oR + r);
            T length = lane_hypot_chain(L, oVc, m);
            T diagonal = 0;
            if (!isfinite(length))
                fit_status = kSpaiNumericalFailure;
            else if (length == 0)
                fit_status = kSpaiRankDeficient;
            else
                diagonal = householder_vector(L, oVc, m, length);
            if (fit_status == kSpaiSuccess) {
                // dots of step 0: slots 1..q (q = right side), groups of four
                for (int c0 = 1; c0 <= q; c0 += 4) {
                    const int nb = q + 1 - c0 < 4 ? q + 1 - c0 : 4;
                    T d[4] = {0, 0, 0, 0};
                    for (int r = 0; r < m; ++r) {
                        const T vr = L.D(oVc + r);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (u < nb)
                                d[u] += vr * L.D(slot_base(c0 + u) + r);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        if (u < nb)
                            L.D(oDot + c0 + u) = d[u];
                }
            }
            for (int j = 0; j < q && fit_status == kSpaiSuccess; ++j) {
                const int len = m - j;
                const bool next = j + 1 < q;
                bool ok = true;
                T nlength = 0, ndiagonal = 0;
                int nstatus = kSpaiSuccess;
                if (next) {
                    const T dt = L.D(oDot + j + 1);
                    const int base = oR + (j + 1) * M + j;
                    T &x0 = L.D(base);
                    x0 -= T(2) * L.D(oVc) * dt;
                    ok = isfinite(x0);
                    for (int r0 = 1; r0 < len; r0 += kLB) {
                        T xb[kLB], vb[kLB];
#pragma unroll
                        for (int u = 0; u < kLB; ++u) {
                            const int r = min(r0 + u, len - 1);
                            xb[u] = L.D(base + r);
                            vb[u] = L.D(oVc + r);
                        }
#pragma unroll
                        for (int u = 0; u < kLB; ++u)
                            if (r0 + u < len) {
                                const T x = xb[u] - T(2) * vb[u] * dt;
                                L.D(base + r0 + u) = x;
                                ok = ok && isfinite(x);
                                L.D(oVn + r0 + u - 1) = x;
                                nlength = hypot_ref(nlength, x);
                            }
                    }
                    if (!isfinite(nlength))
                        nstatus = kSpaiNumericalFailure;
                    else if (nlength == 0)
                        nstatus = kSpaiRankDeficient;
                    else if (ok)
                        ndiagonal = householder_vector(L, oVn, len - 1, nlength);
                }
                const bool fuse = next && ok && nstatus == kSpaiSuccess;
                // the remaining slots (j+2..q, or just the right side on the
                // last step): reflect with v_j; fused dots with v_{j+1}, in
                // groups of kFuse slots that share each load of v_j and v_{j+1}
                for (int c0 = next ? j + 2 : j + 1; c0 <= q && ok; c0 += kFuse) {
                    const int nb = q + 1 - c0 < kFuse ? q + 1 - c0 : kFuse;
                    T dt[kFuse], dn[kFuse];
                    int base[kFuse];
#pragma unroll
                    for (int u = 0; u < kFuse; ++u) {
                        const int c = c0 + (u < nb ? u : 0);
                        dt[u] = L.D(oDot + c);
                        dn[u] = 0;
                        base[u] = slot_base(c) + j;
                    }
                    {
                        const T v2 = T(2) * L.D(oVc);
#pragma unroll
                        for (int u = 0; u < kFuse; ++u)
                            if (u < nb) {
                                T &y = L.D(base[u]);
                                y -= v2 * dt[u];
                                ok = ok && isfinite(y);
                            }
                    }
                    for (int r = 1; r < len; ++r) {
                        const T w2 = T(2) * L.D(oVc + r);
                        const T vn = fuse ? L.D(oVn + r - 1) : T(0);
#pragma unroll
                        for (int u = 0; u < kFuse; ++u)
                            if (u < nb) {
                                T &x = L.D(base[u] + r);
                                x -= w2 * dt[u];
                                ok = ok && isfinite(x);
                                dn[u] += vn * x;
                            }
                    }
#pragma unroll
                    for (int u = 0; u < kFuse; ++u)
                        if (u < nb)
                            L.D(oDot + c0 + u) = dn[u];
                }
                if (!ok) {
                    fit_status = kSpaiNumericalFailure;
                    break;
                }
                L.D(oR + j * M + j) = diagonal;
                const T ad = fabs(diagonal);
                min_pivot = ad < min_pivot ? ad : min_pivot;
                if (next && nstatus != kSpaiSuccess) {
                    fit_status = nstatus;
                    break;
                }
                diagonal = ndiagonal;
                const int t = oVc;
                oVc = oVn;
                oVn = t;
            }
            if (fit_status == kSpaiSuccess) {
                for (int j = q - 1; j >= 0; --j) {
                    T value = L.D(oRhs + j);
                    for (int c = j + 1; c < q; ++c)
                        value -= L.D(oR + c * M + j) * L.D(oCoef + c);
                    const T dg = L.D(oR + j * M + j);
                    if (!isfinite(dg)) {
                        fit_status = kSpaiNumericalFailure;
                        break;
                    }
                    if (dg == 0) {
                        fit_status = kSpaiRankDeficient;
                        break;
                    }
                    const T s = value / dg;
                    if (!isfinite(s)) {
                        fit_status = kSpaiNumericalFailure;
                        break;
                    }
                    L.D(oCoef + j) = s;
                }
            }
            if (fit_status != kSpaiSuccess) {
                status = fit_status;
                break;
            }
            const T pivot_ratio = inf_norm == 0 ? T(0) : min_pivot / inf_norm;
            // ---- fitted residual s = B(I,J) m - e_k(I) (stored entries only)
            for (int r = 0; r < m; ++r)
                L.D(oRes + r) = T(0);
            for (int c = 0; c < q; ++c) {
                const int j = L.I(oJ + c);
                const T dj = dscale[j], cf = L.D(oCoef + c);
                int p = 0;
                for (int e = aptr[j]; e < aptr[j + 1]; ++e) {
                    const int col = acol[e];
                    while (L.I(oI + p) < col)
                        ++p;
                    L.D(oRes + p) += (aval[e] * dj) * cf;
                }
            }
            L.D(oRes + pk) -= T(1);
            // (a non-finite entry makes the chain non-finite: hypot returns
            // inf for an infinite argument and NaN for a NaN one otherwise)
            const T rnorm = lane_hypot_chain(L, oRes, m);
            if (!isfinite(rnorm)) {
                status = kSpaiNumericalFailure;
                break;
            }
            if (fits != 0)
                ++growths;
            ++fits;
            {
                const double fm = m, fq = q;
                qr_flops += 2 * fm * fq * fq - 2 * fq * fq * fq / 3 + 4 * fm * fq + fq * fq;
            }
            for (int c = 0; c < q; ++c) {
                jout[c] = L.I(oJ + c);
                cout_[c] = L.D(oCoef + c);
            }
            qdone = q;
            last_res = (double)rnorm;
            last_piv = (double)pivot_ratio;
            if (rnorm <= eps) {
                status = kSpaiSuccess;
                break;
            }
            if ((long long)growths == alpha) {
                status = kSpaiIterationLimit;
                break;
            }
            // ---- candidates: union of B rows i in I, minus J (ascending)
            int c = 0;
            for (int r = 0; r < m && !over; ++r) {
                const int i = L.I(oI + r);
                int a = 0, e = bptr[i];
                const int e1 = bptr[i + 1];
                int t = 0;
                while (a < c || e < e1) {
                    int x;
                    if (e >= e1 || (a < c && L.I(oC + a) < bcol[e]))
                        x = L.I(oC + a++);
                    else if (a >= c || bcol[e] < L.I(oC + a))
                        x = bcol[e++];
                    else {
                        x = bcol[e++];
                        ++a;
                    }
                    if (t >= CC) {
                        over = true;
                        break;
                    }
                    L.I(oCt + t++) = x;
                }
                const int tmp = oC;
                oC = oCt;
                oCt = tmp;
                c = t;
            }
            if (over) {
                status = kSpaiOverflow;
                break;
            }
            {
                int t = 0, a = 0;
                for (int r = 0; r < c; ++r) {
                    const int x = L.I(oC + r);
                    while (a < q && L.I(oJ + a) < x)
                        ++a;
                    if (a < q && L.I(oJ + a) == x)
                        continue;
                    L.I(oC + t++) = x;
                }
                c = t;
            }
            if (c == 0) {
                status = kSpaiEmptyCandidates;
                break;
            }
            // ---- restricted scores (src/spai.cpp:198-215)
            const T rr = rnorm * rnorm;
            int score_status = kSpaiSuccess;
            for (int t = 0; t < c; ++t) {
                const int j = L.I(oC + t);
                const T dj = dscale[j];
                T dot = 0, den = 0;
                int lo = 0;
                for (int e = aptr[j]; e < aptr[j + 1]; ++e) {
                    const int col = acol[e];
                    int hi = m;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (L.I(oI + mid) < col)
                            lo = mid + 1;
                        else
                            hi = mid;
                    }
                    if (lo < m && L.I(oI + lo) == col) {
                        const T value = aval[e] * dj;
                        dot += L.D(oRes + lo) * value;
                        den += value * value;
                    }
                }
                T s = 0;
                bool cl = false;
                const int st = score_of(rr, dot, den, s, cl);
                if (st != kSpaiSuccess) {
                    score_status = st;
                    break;
                }
                clamps += cl ? 1 : 0;
                L.D(oSc + t) = s;
            }
            if (score_status != kSpaiSuccess) {
                status = score_status;
                break;
            }
            T mean = 0;
            for (int t = 0; t < c; ++t)
                mean += L.D(oSc + t);
            mean = mean / T(c);
            if (!isfinite(mean)) {
                status = kSpaiNumericalFailure;
                break;
            }
            // ---- select <= beta smallest (score, index) with score <= mean
            int nsel = 0;
            T prev_s = T(-INFINITY);
            int prev_j = -1;
            for (; nsel < beta; ++nsel) {
                T best_s = T(INFINITY);
                int best_j = -1;
                for (int t = 0; t < c; ++t) {
                    const T s = L.D(oSc + t);
                    const int j = L.I(oC + t);
                    if (s > mean)
                        continue;
                    const bool after = s > prev_s || (s == prev_s && j > prev_j);
                    const bool before = best_j < 0 || s < best_s || (s == best_s && j < best_j);
                    if (after && before) {
                        best_s = s;
                        best_j = j;
                    }
                }
                if (best_j < 0)
                    break;
                if (q + nsel + 1 > Q) { // the grown pattern would not fit
                    nsel = -1;
                    break;
                }
                L.I(oSel + nsel) = best_j;
                prev_s = best_s;
                prev_j = best_j;
            }
            if (nsel < 0) {
                status = kSpaiOverflow;
                break;
            }
            if (nsel == 0) {
                status = kSpaiNoGrowth;
                break;
            }
            // sort the selection by index, then merge into J
            for (int a = 1; a < nsel; ++a) {
                const int x = L.I(oSel + a);
                int b = a - 1;
                while (b >= 0 && L.I(oSel + b) > x) {
                    L.I(oSel + b + 1) = L.I(oSel + b);
                    --b;
                }
                L.I(oSel + b + 1) = x;
            }
            {
                int a = 0, b = 0, t = 0;
                while (a < q || b < nsel) {
                    if (b >= nsel || (a < q && L.I(oJ + a) < L.I(oSel + b)))
                        L.I(oJt + t++) = L.I(oJ + a++);
                    else
                        L.I(oJt + t++) = L.I(oSel + b++);
                }
                for (int r = 0; r < t; ++r)
                    L.I(oJ + r) = L.I(oJt + r);
                q = t;
            }
        }
        out.status[k] = status;
        out.fits[k] = fits;
        out.growths[k] = growths;
        out.clamps[k] = clamps;
        out.q[k] = status == kSpaiOverflow ? 0 : qdone;
        out.residual[k] = last_res;
        out.pivot[k] = last_piv;
        out.flops[k] = qr_flops;
        out.off[k] = (std::int64_t)slot * out_stride;
        out.pass[k] = out.pass_id;
    }
}

template <typename T>
void launch_spai_columns(hipStream_t s, int blocks, int threads, int n, const int *aptr,
                         const int *acol, const T *aval, const T *dscale, const int *bptr,
                         const int *bcol, T eps, long long alpha, int beta, const int *worklist,
                         int nwork, int *cursor, int mcap, int qcap, int ccap, T *dwork,
                         int *iwork, SpaiColumnOut out, int *out_j, T *out_c,
                         std::int64_t out_stride) {
    Caps caps{mcap, qcap, ccap};
    hipLaunchKernelGGL(spai_columns_kernel<T>, dim3(blocks), dim3(threads), 0, s, n, aptr, acol,
                       aval, dscale, bptr, bcol, eps, alpha, beta, worklist, nwork, cursor, caps,
                       dwork, iwork, spai_lane_doubles(mcap, qcap, ccap),
                       spai_lane_ints(mcap, qcap, ccap), out, out_j, out_c, out_stride);
    HIP_LAUNCH_CHECK();
}

template void launch_spai_columns<double>(hipStream_t, int, int, int, const int *, const int *,
                                          const double *, const double *, const int *,
                                          const int *, double, long long, int, const int *, int,
                                          int *, int, int, int, double *, int *, SpaiColumnOut,
                                          int *, double *, std::int64_t);
template void launch_spai_columns<float>(hipStream_t, int, int, int, const int *, const int *,
                                         const float *, const float *, const int *, const int *,
                                         float, long long, int, const int *, int, int *, int,
                                         int, int, float *, int *, SpaiColumnOut, int *, float *,
                                         std::int64_t);

// ================================================================ wave kernel
// FP32 fits of the first-bucket columns (the mixed path), one wavefront per
// column, the dense problem in registers and LDS: lane r owns row r of the
// least-squares matrix (m <= 64).
//   * shadow: the rows of A in J are loaded cooperatively, then (column,
//     entry) keys are bitonic-sorted across the wave; adjacent compares give
//     the sorted, duplicate-free set I and every entry's position in it;
//   * QR: each lane keeps its row in registers; the dot products of the
//     Householder vector with the trailing columns come from one
//     reduce-scatter over the wave (lane 4c ends with column c);
//   * candidates: the columns of B's rows i in I go into an open-addressing
//     set in LDS, whose distinct members are sorted (so the sort sees about
//     a quarter of the entries at 128^3) and J removed; the entries, with
//     their values, are grouped by candidate in ascending i;
//   * candidate scores: the terms res_i * b_ij d_j and (b_ij d_j)^2 are
//     summed per candidate over its group, ascending i as the lane kernel
//     sums A's row j, with no further loads;
//   * the beta smallest (score, index) pairs are found by argmin rounds
//     over the wave.
// The method and its decisions are the lane kernel's; the FP32 summation
// order of the QR and norms differs, which the FP32 build allows (its
// patterns are compared with the FP64 oracle's and differences counted).
// Columns beyond the caps stop with kOverflow and go to the workgroup kernel.
constexpr int kWM = 64;      // rows (one per lane)
constexpr int kWQ = 16;      // pattern columns
constexpr int kWC = 128;     // distinct candidate rows including J
constexpr int kWS = 512;     // staged entries (9-bit entry numbers in the sort keys)
constexpr int kWKeyColBits = 23; // column bits of a 32-bit sort key (n < 2^23)
constexpr int kWW = kWQ + 1; // LDS row stride of R
constexpr int kWH = 2 * kWC; // candidate set slots

struct WaveLds {
    // R is read last by the residual, before the candidate stage; the
    // candidate stage's arrays are read last by the scores, before the next
    // fit's gather.  9,152 bytes in all: 16 waves fit one CU's LDS.
    union {
        float R[kWM][kWW]; // gathered B(I, J) (unchanged by the QR, which runs in registers)
        struct {
            int ht[kWH]; // open-addressing set of candidate columns (-1: empty)
            union {
                int ulist[kWH];            // the set's members, (column << 8) | slot
                unsigned short grp[kWS];   // entries grouped by slot, ascending in a group
            };
            int cur[kWH];                  // entries per slot, then the groups' cursors
            unsigned char es[kWS];         // entry -> slot
            unsigned char cslot[kWC];      // candidate -> slot
        };
    };
    alignas(16) float bq[kWQ]; // wave-uniform values of a step, broadcast by LDS reads
    float sval[kWS];     // staged values by entry (A's values in the shadow, B's for candidates)
    signed char sc[kWS]; // entry -> staging row
    int sbeg[kWM];       // staging row r holds entries e of source index sbeg[r] + e
    float sd[kWM];       // staging row scale
    int I[kWM];
    int J[kWQ];
    int sel[kWQ];
    int cand[kWC];
    float score[kWC];
    float res[kWM];
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}
// Exclusive prefix sum over the wave by DPP (no LDS crossbar): shifts by 1,
// 2, 4 and 8 lanes within each row of 16 (lanes without a source add the
// old value 0), then row 0's total into row 1 and row 2's into row 3 (row
// broadcast of lane 15), then lane 31's total into rows 2 and 3.
__device__ inline int wave_scan_excl(int v, int lane) {
    (void)lane;
    int x = v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false); // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false); // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false); // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false); // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31
    return x - v;
}
// Cross-lane moves without the LDS crossbar (gfx950; every lane active):
//   DPP quad_perm gives lane i the value of lane i^1, i^2 or i^3; DPP
//   row_half_mirror lane i^7 (within 8 lanes); DPP row_ror:8 lane i^8 (within
//   16 lanes);
//   permlane16_swap(x, y) trades rows 1 and 3 of x (16 lanes each) for rows 0
//   and 2 of y; permlane32_swap(x, y) trades lanes 32..63 of x for lanes 0..31
//   of y.
constexpr int kDppXor1 = 0xB1; // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E; // quad_perm [2, 3, 0, 1]
constexpr int kDppXor3 = 0x1B; // quad_perm [3, 2, 1, 0]
constexpr int kDppXor7 = 0x141; // row_half_mirror
constexpr int kDppXor8 = 0x128; // row_ror:8
template <int kCtrl> __device__ inline unsigned dpp_u(unsigned v) {
    return (unsigned)__builtin_amdgcn_mov_dpp((int)v, kCtrl, 0xf, 0xf, true);
}
__device__ inline void swap16(unsigned &x, unsigned &y) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}
__device__ inline void swap32(unsigned &x, unsigned &y) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    x = r[0];
    y = r[1];
}
// The value of lane (lane ^ m), m a power of two below 64 (wave-uniform).
__device__ inline unsigned xor_lane(unsigned v, int m, int lane) {
    switch (m) {
    case 1:
        return dpp_u<kDppXor1>(v);
    case 2:
        return dpp_u<kDppXor2>(v);
    case 4:
        return dpp_u<kDppXor3>(dpp_u<kDppXor7>(v)); // i ^ 7 ^ 3
    case 8:
        return dpp_u<kDppXor8>(v);
    case 16: {
        unsigned x = v, y = v;
        swap16(x, y); // x: rows 0 0 2 2, y: rows 1 1 3 3
        return (lane & 16) ? x : y;
    }
    default: {
        unsigned x = v, y = v;
        swap32(x, y); // x: lanes 0..31 twice, y: lanes 32..63 twice
        return (lane & 32) ? x : y;
    }
    }
}
template <int kCtrl> __device__ inline float dpp_f(float v) {
    return __uint_as_float(dpp_u<kCtrl>(__float_as_uint(v)));
}
// Sums over the wave, every lane ending with the same bits: each step adds a
// lane's value and its partner's, which both lanes of a pair form alike.
__device__ inline float wave_sum_f(float v) {
    v += dpp_f<kDppXor1>(v);
    v += dpp_f<kDppXor2>(v);
    v += dpp_f<kDppXor7>(v);
    v += dpp_f<kDppXor8>(v);
    unsigned x = __float_as_uint(v), y = x;
    swap16(x, y);
    v = __uint_as_float(x) + __uint_as_float(y);
    x = __float_as_uint(v);
    y = x;
    swap32(x, y);
    return __uint_as_float(x) + __uint_as_float(y);
}
template <int kCtrl> __device__ inline double dpp_d(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = dpp_u<kCtrl>((unsigned)b), hi = dpp_u<kCtrl>((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double wave_sum_d(double v) {
    v += dpp_d<kDppXor1>(v);
    v += dpp_d<kDppXor2>(v);
    v += dpp_d<kDppXor7>(v);
    v += dpp_d<kDppXor8>(v);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        unsigned xl_ = (unsigned)b, yl = xl_, xh = (unsigned)(b >> 32), yh = xh;
        if (s == 0) {
            swap16(xl_, yl);
            swap16(xh, yh);
        } else {
            swap32(xl_, yl);
            swap32(xh, yh);
        }
        v = __longlong_as_double((long long)(((unsigned long long)xh << 32) | xl_)) +
            __longlong_as_double((long long)(((unsigned long long)yh << 32) | yl));
    }
    return v;
}
__device__ inline int wave_sum_int(int v) {
    v += (int)dpp_u<kDppXor1>((unsigned)v);
    v += (int)dpp_u<kDppXor2>((unsigned)v);
    v += (int)dpp_u<kDppXor7>((unsigned)v);
    v += (int)dpp_u<kDppXor8>((unsigned)v);
    unsigned x = (unsigned)v, y = x;
    swap16(x, y);
    v = (int)(x + y);
    x = (unsigned)v;
    y = x;
    swap32(x, y);
    return (int)(x + y);
}
// Maximum over the wave (operands finite or +inf and >= +0: every lane ends
// with the same bits).
__device__ inline float wave_max_f(float v) {
    float o = dpp_f<kDppXor1>(v);
    v = o > v ? o : v;
    o = dpp_f<kDppXor2>(v);
    v = o > v ? o : v;
    o = dpp_f<kDppXor7>(v);
    v = o > v ? o : v;
    o = dpp_f<kDppXor8>(v);
    v = o > v ? o : v;
    unsigned x = __float_as_uint(v), y = x;
    swap16(x, y);
    v = __uint_as_float(x) > __uint_as_float(y) ? __uint_as_float(x) : __uint_as_float(y);
    x = __float_as_uint(v);
    y = x;
    swap32(x, y);
    return __uint_as_float(x) > __uint_as_float(y) ? __uint_as_float(x) : __uint_as_float(y);
}
// Minimum over the wave of 32-bit keys (every lane ends with the same key).
__device__ inline unsigned wave_min_u32(unsigned v) {
    unsigned o = dpp_u<kDppXor1>(v);
    v = o < v ? o : v;
    o = dpp_u<kDppXor2>(v);
    v = o < v ? o : v;
    o = dpp_u<kDppXor7>(v);
    v = o < v ? o : v;
    o = dpp_u<kDppXor8>(v);
    v = o < v ? o : v;
    unsigned x = v, y = x;
    swap16(x, y);
    v = x < y ? x : y;
    x = v;
    y = x;
    swap32(x, y);
    return x < y ? x : y;
}
__device__ inline float lane_value(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Lane r < nr stages row `row` of a CSR structure: entries [off_r, off_r +
// len_r) map to row r (sc) with source index sbeg[r] + e and scale sd[r]
// (dscale[row], or 1).  Returns the entry count, or -1 beyond kWS.
__device__ int wave_rows(WaveLds &W, const int *__restrict__ ptr,
                         const float *__restrict__ dscale, int row, int nr, int lane) {
    const bool has = lane < nr;
    const int e0 = has ? ptr[row] : 0;
    const int e1 = has ? ptr[row + 1] : 0;
    const float d = has && dscale ? dscale[row] : 1.f;
    const int len = e1 - e0;
    const int off = wave_scan_excl(len, lane);
    const int total = __builtin_amdgcn_readlane(off + len, kWave - 1);
    if (total > kWS)
        return -1;
    W.sbeg[lane] = e0 - off;
    W.sd[lane] = d;
    for (int t = 0; t < len; ++t)
        W.sc[off + t] = (signed char)lane;
    wave_sync();
    return total;
}

// Bitonic sort of 64 P keys, element e = lane * P + i held in key[i].  For
// P <= 2 (the shadows and candidate sets of the stencil problems) every stage
// is unrolled, so each cross-lane partner is a compile-time lane ^ m moved by
// DPP or a permlane swap (a few cycles, against a round trip through the LDS
// crossbar).  Wider sorts keep the stage loops rolled (k and j wave-uniform)
// so that the sorts of all widths together stay small in the instruction
// cache; only the P keys of a lane are unrolled.
template <int P> __device__ inline void wave_bitonic(unsigned (&key)[P], int lane) {
    if constexpr (P <= 2) {
#pragma unroll
        for (int k = 2; k <= kWave * P; k <<= 1) {
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1) {
                if (j >= P) { // partner lane ^ (j / P), same slot
                    const bool lower = (lane & (j / P)) == 0;
#pragma unroll
                    for (int i = 0; i < P; ++i) {
                        const unsigned o = xor_lane(key[i], j / P, lane);
                        const bool up = ((lane * P + i) & k) == 0;
                        const bool take_min = lower == up;
                        const bool less = o < key[i];
                        key[i] = take_min == less ? o : key[i];
                    }
                } else { // P == 2, j == 1: slots 0 and 1 of the lane
                    const bool up = ((lane * P) & k) == 0;
                    const unsigned x = key[0], y = key[P - 1];
                    const bool sw = up ? x > y : x < y;
                    key[0] = sw ? y : x;
                    key[P - 1] = sw ? x : y;
                }
            }
        }
        return;
    }
#pragma unroll 1
    for (int k = 2; k <= kWave * P; k <<= 1) {
#pragma unroll 1
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= P) { // partner in another lane: lane ^ (j / P), same slot
                const int lj = j / P;
                const bool lower = (lane & lj) == 0;
#pragma unroll
                for (int i = 0; i < P; ++i) {
                    const unsigned o = __shfl_xor(key[i], lj, kWave);
                    const bool up = ((lane * P + i) & k) == 0;
                    const bool take_min = lower == up;
                    const bool less = o < key[i];
                    key[i] = take_min == less ? o : key[i];
                }
            } else { // partner in the same lane: slot i ^ j
#pragma unroll
                for (int jj = P / 2; jj > 0; jj >>= 1) {
                    if (jj == j) {
#pragma unroll
                        for (int i = 0; i < P; ++i) {
                            const int p = i ^ jj;
                            if (p > i) {
                                const bool up = ((lane * P + i) & k) == 0;
                                const unsigned x = key[i], y = key[p];
                                const bool sw = up ? x > y : x < y;
                                key[i] = sw ? y : x;
                                key[p] = sw ? x : y;
                            }
                        }
                    }
                }
            }
        }
    }
}

// The shadow: the staged entries e < total of (col, val), sorted by (col, e)
// and numbered by distinct column; I <- the sorted set and R(position,
// staging row) <- val * sd.  Returns the set size; outputs are written only
// when it is within the cap kWM, which the caller checks.
template <int P>
__device__ int wave_sort_set(WaveLds &W, const int *__restrict__ col,
                             const float *__restrict__ val, int total, int lane) {
    unsigned key[P]; // (column << 9) | entry; the empty key ~0u sorts last
    {
        int src[P];
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int e = lane * P + i;
            const int ec = e < total ? e : 0;
            src[i] = W.sbeg[W.sc[ec]] + ec;
        }
        int c[P];
        float v[P];
#pragma unroll
        for (int i = 0; i < P; ++i) {
            c[i] = gld(col + src[i]);
            v[i] = gld(val + src[i]);
        }
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int e = lane * P + i;
            if (e < total)
                W.sval[e] = v[i] * W.sd[W.sc[e]];
            key[i] = e < total ? ((unsigned)c[i] << 9) | (unsigned)e : ~0u;
        }
    }
    wave_bitonic<P>(key, lane);
    wave_sync(); // sval of every entry written
    int kv[P], ev[P], pos[P];
    bool first[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        kv[i] = key[i] == ~0u ? -1 : (int)(key[i] >> 9);
        ev[i] = (int)(key[i] & 511u);
    }
    // the previous lane's last key (DPP wave_shr:1; lane 0 keeps -2)
    const int prev = __builtin_amdgcn_update_dpp(-2, kv[P - 1], 0x138, 0xf, 0xf, false);
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const int pv = i ? kv[i - 1] : prev;
        first[i] = kv[i] >= 0 && kv[i] != pv;
        cnt += first[i] ? 1 : 0;
    }
    const int nset = wave_sum_int(cnt);
    {
        int run = wave_scan_excl(cnt, lane) - 1;
#pragma unroll
        for (int i = 0; i < P; ++i) {
            if (first[i])
                ++run;
            pos[i] = run;
        }
    }
    if (nset > kWM)
        return nset;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        if (first[i])
            W.I[pos[i]] = kv[i];
        if (kv[i] >= 0)
            W.R[pos[i]][W.sc[ev[i]]] = W.sval[ev[i]];
    }
    wave_sync();
    return nset;
}

// (kept out of line: the sorts' registers then do not add to the kernel's)
__device__ __attribute__((noinline)) int wave_shadow(WaveLds &W, const int *col, const float *val,
                                                     int total, int lane) {
    if (total <= 0) // (rows of A in this stage always hold an entry)
        return 0;
    if (total <= kWave)
        return wave_sort_set<1>(W, col, val, total, lane);
    if (total <= 2 * kWave)
        return wave_sort_set<2>(W, col, val, total, lane);
    if (total <= 4 * kWave)
        return wave_sort_set<4>(W, col, val, total, lane);
    return wave_sort_set<8>(W, col, val, total, lane);
}

// The distinct columns of the list ((column << 8) | slot), ascending, minus
// J -> cand and cslot; returns their count.  P keys per lane hold the
// n <= 64 P list entries.
template <int P>
__device__ int wave_sorted_candidates(WaveLds &W, int n, int q, int lane) {
    unsigned key[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
        const int e = lane * P + i;
        key[i] = e < n ? (unsigned)W.ulist[e] : ~0u;
    }
    wave_bitonic<P>(key, lane);
    bool keep[P];
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        bool in_j = false;
        if (key[i] != ~0u) {
            const int x = (int)(key[i] >> 8);
            int lo = 0, hi = q;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (W.J[mid] < x)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            in_j = lo < q && W.J[lo] == x;
        }
        keep[i] = key[i] != ~0u && !in_j;
        cnt += keep[i] ? 1 : 0;
    }
    const int nc = wave_sum_int(cnt);
    int at = wave_scan_excl(cnt, lane);
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (keep[i]) {
            W.cand[at] = (int)(key[i] >> 8);
            W.cslot[at] = (unsigned char)(key[i] & 255u);
            ++at;
        }
    wave_sync();
    return nc;
}

// The candidates of a fit: the distinct columns of B's rows i in I (staged
// entries e < total), minus J, ascending in cand; returns their count, or
// kWC + 1 when the set (J included) exceeds kWC.
//   * Each entry's column goes into an open-addressing set in LDS
//     (multiplicative hash, linear probing bounded by the table size; a
//     compare-and-swap inserts or finds it); the entry keeps its slot and
//     its value b_ij, and the slot counts its entries.
//   * The set's members are compacted and sorted (the sort sees the distinct
//     columns only, not every entry), J removed.
//   * The entries are grouped by slot, each group in ascending entry order
//     (a stable counting sort: group starts by a scan of the counts, ranks
//     within a 64-entry round from eight ballots on the slot bits), so that
//     candidate j finds its terms in ascending i with no further loads.
__device__ __attribute__((noinline)) int wave_candidates(WaveLds &W, const int *__restrict__ bcol,
                                                         const float *__restrict__ bval,
                                                         int total, int q, int lane) {
    constexpr int kPer = kWH / kWave;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        W.ht[lane + u * kWave] = -1;
        W.cur[lane + u * kWave] = 0;
    }
    wave_sync();
    int added = 0;
    bool full = false;
    for (int e0 = 0; e0 < total; e0 += kWave) {
        const int e = e0 + lane;
        if (e < total) {
            const int src = W.sbeg[W.sc[e]] + e;
            const int x = gld(bcol + src);
            W.sval[e] = gld(bval + src);
            unsigned h = ((unsigned)x * 2654435761u) >> 24; // kWH = 256 slots
            int probe = 0;
            for (; probe < kWH; ++probe) {
                const int old = atomicCAS(&W.ht[h], -1, x);
                if (old == -1) {
                    ++added;
                    break;
                }
                if (old == x)
                    break;
                h = (h + 1) & (kWH - 1);
            }
            full = full || probe == kWH;
            W.es[e] = (unsigned char)h;
            atomicAdd(&W.cur[h], 1);
        }
    }
    wave_sync();
    const int nu = wave_sum_int(added);
    if (__any(full) || nu > kWC)
        return kWC + 1;
    // compaction: slots lane * kPer .. in slot order, (column << 8) | slot
    {
        int v[kPer], cnt = 0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            v[u] = W.ht[lane * kPer + u];
            cnt += v[u] >= 0 ? 1 : 0;
        }
        int at = wave_scan_excl(cnt, lane);
#pragma unroll
        for (int u = 0; u < kPer; ++u)
            if (v[u] >= 0)
                W.ulist[at++] = (v[u] << 8) | (lane * kPer + u);
    }
    wave_sync();
    const int nc = nu <= kWave ? wave_sorted_candidates<1>(W, nu, q, lane)
                               : wave_sorted_candidates<2>(W, nu, q, lane);
    // group starts: exclusive scan of the slot counts, in slot order
    {
        int c[kPer], sum = 0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            c[u] = W.cur[lane * kPer + u];
            sum += c[u];
        }
        int at = wave_scan_excl(sum, lane);
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            W.cur[lane * kPer + u] = at;
            at += c[u];
        }
    }
    wave_sync(); // (the sort has read the list: its space now takes grp)
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int e0 = 0; e0 < total; e0 += kWave) {
        const int e = e0 + lane;
        const bool valid = e < total;
        const int sl = valid ? (int)W.es[e] : 0;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (sl >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        const int pos = valid ? W.cur[sl] + rank : 0;
        wave_sync(); // every lane has read its group's cursor
        if (valid) {
            W.grp[pos] = (unsigned short)e;
            if (rank == 0)
                W.cur[sl] += __popcll(peers);
        }
        wave_sync();
    }
    // each cursor now ends its group; a group starts where the previous
    // slot's ends
    return nc;
}

// dot = sum res_i b_ij d_j and den = sum (b_ij d_j)^2 over candidate t's
// entries (one per row i of I holding column j), ascending i: the group of
// its slot, in entry order.
__device__ inline void candidate_terms(const WaveLds &W, int t, float dj, float &dot,
                                       float &den) {
    const int sl = W.cslot[t];
    const int p1 = W.cur[sl], p0 = sl > 0 ? W.cur[sl - 1] : 0;
    dot = 0.f;
    den = 0.f;
    for (int p = p0; p < p1; ++p) {
        const int e = W.grp[p];
        const float value = W.sval[e] * dj;
        dot += W.res[W.sc[e]] * value;
        den += value * value;
    }
}

// Sums a[c] (c < 16) over the wave: lane l ends with the sum of column
// (l >> 2) & 15.  Halving exchanges: lanes 32 apart, then 16 apart, trade
// halves through the permlane swaps (after a swap both halves add alike);
// then lanes 8 apart (DPP row_ror:8) and lanes i and i^7 (row_half_mirror)
// each form both candidate sums and keep theirs; then the quad is summed.
__device__ inline float wave_reduce_scatter16(float (&a)[kWQ], int lane) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        unsigned x = __float_as_uint(a[i]), y = __float_as_uint(a[i + 8]);
        swap32(x, y);
        a[i] = __uint_as_float(x) + __uint_as_float(y);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        unsigned x = __float_as_uint(a[i]), y = __float_as_uint(a[i + 4]);
        swap16(x, y);
        a[i] = __uint_as_float(x) + __uint_as_float(y);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const float u = a[i] + dpp_f<kDppXor8>(a[i]);
        const float w = a[i + 2] + dpp_f<kDppXor8>(a[i + 2]);
        a[i] = (lane & 8) ? w : u;
    }
    {
        const float u = a[0] + dpp_f<kDppXor7>(a[0]);
        const float w = a[1] + dpp_f<kDppXor7>(a[1]);
        a[0] = (lane & 4) ? w : u;
    }
    float x = a[0];
    x += dpp_f<kDppXor1>(x);
    x += dpp_f<kDppXor2>(x);
    return x;
}

// Four waves per SIMD: 128 VGPRs, and 16 waves' LDS fit one CU.
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(4)))
spai_wave_kernel(const int *__restrict__ aptr, const int *__restrict__ acol,
                 const float *__restrict__ aval, const float *__restrict__ dscale,
                 const int *__restrict__ bptr, const int *__restrict__ bcol,
                 const float *__restrict__ bval, float eps, long long alpha, int beta,
                 const int *__restrict__ worklist, int nwork, int *__restrict__ cursor,
                 SpaiColumnOut out, int *__restrict__ out_j, float *__restrict__ out_c,
                 std::int64_t out_stride) {
    __shared__ WaveLds W;
    const int lane = threadIdx.x;
    for (;;) {
        int slot = 0;
        if (lane == 0)
            slot = atomicAdd(cursor, 1);
        slot = __builtin_amdgcn_readfirstlane(slot);
        if (slot >= nwork)
            break;
        const int k = worklist[slot];
        int status = kSpaiNotAttempted;
        int fits = 0, growths = 0, clamps = 0, q = 1, qdone = 0;
        double last_res = __builtin_nan(""), last_piv = __builtin_nan(""), qr_flops = 0;
        if (lane == 0)
            W.J[0] = k;
        wave_sync();
        int *jout = out_j + (std::int64_t)slot * out_stride;
        float *cout_ = out_c + (std::int64_t)slot * out_stride;
        for (;;) {
            // ---- shadow I = union of cols(A row j), j in J, and the gather
#pragma unroll
            for (int c = 0; c < kWW; ++c)
                W.R[lane][c] = 0.f;
            int m;
            if (q == 1) {
                // the first fit: I is row k of A itself (a canonical CSR row
                // is sorted and free of duplicates), R(r, 0) = a_k,I[r] d_k
                const int e0 = gld(aptr + k), len = gld(aptr + k + 1) - e0;
                m = len;
                if (len <= kWM) {
                    const float dk = gld(dscale + k);
                    if (lane < len) {
                        W.I[lane] = gld(acol + e0 + lane);
                        W.R[lane][0] = gld(aval + e0 + lane) * dk;
                    }
                    wave_sync();
                }
            } else {
                const int na = __builtin_amdgcn_readfirstlane(
                    wave_rows(W, aptr, dscale, lane < q ? W.J[lane] : 0, q, lane));
                if (na < 0) {
                    status = kSpaiOverflow;
                    break;
                }
                m = __builtin_amdgcn_readfirstlane(wave_shadow(W, acol, aval, na, lane));
            }
            if (m > kWM) {
                status = kSpaiOverflow;
                break;
            }
            if (m < q) {
                status = kSpaiRankDeficient;
                break;
            }
            const int irow = lane < m ? W.I[lane] : -1;
            const unsigned long long kmask = __ballot(irow == k);
            if (kmask == 0ull) {
                status = kSpaiNumericalFailure;
                break;
            }
            const int pk = __ffsll((long long)kmask) - 1;
            // ---- Householder least squares, lane r holding row r
            float r[kWQ];
#pragma unroll
            for (int c = 0; c < kWQ; ++c)
                r[c] = W.R[lane][c];
            float rh = lane == pk ? 1.f : 0.f;
            float rowsum = 0.f;
#pragma unroll
            for (int c = 0; c < kWQ; ++c)
                rowsum += fabsf(r[c]);
            const float inf_norm = wave_max_f(rowsum);
            float min_pivot = INFINITY, dg = 0.f;
            int fit_status = kSpaiSuccess;
            // One rolled loop over the steps; column j of the row is read from
            // r[] by its (wave-uniform) index, so r[] stays in registers and
            // the code holds one step (an unrolled loop over 16 steps made the
            // kernel larger than the instruction cache).  Every value tested
            // below is wave-uniform (butterfly sums give every lane the same
            // bits), so the loop exits are uniform.
#pragma unroll 1
            for (int j = 0; j < q; ++j) {
                const bool act = lane >= j && lane < m;
                const float x = act ? r[j] : 0.f;
                // The squared norm in FP32 (entries of B(I, J) are at most 1
                // in magnitude and reflections keep column norms, so the sum
                // cannot overflow); a sum below 2^-100, where the squares of
                // the small entries lose bits, and a non-finite one take FP64.
                float length;
                {
                    const float s2 = wave_sum_f(x * x);
                    if (s2 >= 0x1p-100f && s2 <= 0x1p+100f)
                        length = __builtin_amdgcn_sqrtf(s2);
                    else
                        length = (float)sqrt(wave_sum_d((double)x * x));
                    length = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(length)));
                }
                if (!isfinite(length)) {
                    fit_status = kSpaiNumericalFailure;
                    break;
                }
                if (length == 0.f) {
                    fit_status = kSpaiRankDeficient;
                    break;
                }
                const float xj = lane_value(x, j);
                const float diagonal = -copysignf(length, xj);
                // reciprocals and root of the step's (wave-uniform) scales by
                // the hardware rcp and sqrt (1 ulp; the FP32 build's order and
                // rounding are its own) in place of per-lane divisions
                const float inv_len = __builtin_amdgcn_rcpf(length);
                const float vnorm = __builtin_amdgcn_sqrtf(2.f + 2.f * fabsf(xj) * inv_len);
                float v = act ? x * inv_len : 0.f;
                if (lane == j)
                    v += copysignf(1.f, v);
                v = v * __builtin_amdgcn_rcpf(vnorm);
                // v is 0 outside rows j..m-1, so neither the products nor the
                // reflection need a mask: rows outside keep their bits (x - 0
                // = x).  Columns c <= j are reflected too; their rows >= j are
                // never read again, and the reflection keeps them bounded.
                // Columns c >= q hold zeros and stay zero.
                float p[kWQ];
#pragma unroll
                for (int c = 0; c < kWQ; ++c)
                    p[c] = v * r[c];
                const float d = wave_reduce_scatter16(p, lane);
                // column c's sum (lanes 4c..4c+3) to every lane through LDS:
                // four broadcast reads instead of 15 readlanes (the kernel is
                // bound by VALU issue; LDS reads are not VALU instructions)
                if ((lane & 3) == 0)
                    W.bq[lane >> 2] = d;
                wave_sync();
                float dq[kWQ];
#pragma unroll
                for (int c4 = 0; c4 < kWQ / 4; ++c4) {
                    const float4 t4 = reinterpret_cast<const float4 *>(W.bq)[c4];
                    dq[4 * c4] = t4.x;
                    dq[4 * c4 + 1] = t4.y;
                    dq[4 * c4 + 2] = t4.z;
                    dq[4 * c4 + 3] = t4.w;
                }
                const float dr = wave_sum_f(v * rh);
                const float v2 = 2.f * v;
                // fused multiply-adds (one rounding, one VALU instruction
                // each; the FP32 build's rounding is its own)
#pragma unroll
                for (int c = 1; c < kWQ; ++c)
                    r[c] = __builtin_fmaf(-v2, dq[c], r[c]);
                wave_sync(); // every lane has read bq before the next step writes it
                rh = __builtin_fmaf(-v2, dr, rh);
                if (lane == j)
                    dg = diagonal;
                min_pivot = fminf(min_pivot, fabsf(diagonal));
            }
            // A non-finite value left by a reflection stays non-finite through
            // every later step (and makes a later column's norm non-finite),
            // so one check after the loop finds what a check per step would,
            // with the same status.
            if (fit_status == kSpaiSuccess) {
                bool bad = !isfinite(rh);
#pragma unroll
                for (int c = 1; c < kWQ; ++c)
                    bad = bad || (c < q && !isfinite(r[c]));
                if (__any(bad))
                    fit_status = kSpaiNumericalFailure;
            }
            // back substitution, one coefficient per lane (lane j: coef_j)
            // The status is that of the first failing step (from j = q-1
            // down), as a check per step would give: each lane j records its
            // own, and the largest failing j decides after the loop (a failed
            // step's non-finite values reach only later, smaller j).
            float coef = 0.f;
            if (fit_status == kSpaiSuccess) {
                float acc = rh;
                int st = kSpaiSuccess;
#pragma unroll 1
                for (int j = q - 1; j >= 0; --j) {
                    float cj = 0.f;
                    if (lane == j) {
                        if (!isfinite(dg))
                            st = kSpaiNumericalFailure;
                        else if (dg == 0.f)
                            st = kSpaiRankDeficient;
                        else {
                            cj = acc / dg;
                            if (!isfinite(cj))
                                st = kSpaiNumericalFailure;
                        }
                        coef = cj;
                    }
                    cj = lane_value(cj, j);
                    if (lane < j)
                        acc = __builtin_fmaf(-r[j], cj, acc);
                }
                const unsigned long long badm = __ballot(st != kSpaiSuccess);
                if (badm != 0ull)
                    fit_status = __builtin_amdgcn_readlane(st, 63 - __clzll((long long)badm));
            }
            if (fit_status != kSpaiSuccess) {
                status = fit_status;
                break;
            }
            const float pivot_ratio = inf_norm == 0.f ? 0.f : min_pivot / inf_norm;
            // ---- residual s = B(I, J) m - e_k(I), from the gathered rows
            // (the coefficients broadcast through LDS, as the step sums are)
            if (lane < kWQ)
                W.bq[lane] = coef;
            wave_sync();
            float sres = 0.f;
#pragma unroll
            for (int c4 = 0; c4 < kWQ / 4; ++c4) {
                const float4 t4 = reinterpret_cast<const float4 *>(W.bq)[c4];
                const float cq[4] = {t4.x, t4.y, t4.z, t4.w};
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (4 * c4 + u < q)
                        sres = __builtin_fmaf(W.R[lane][4 * c4 + u], cq[u], sres);
            }
            if (lane == pk)
                sres -= 1.f;
            if (lane >= m)
                sres = 0.f;
            W.res[lane] = sres;
            const float rnorm = (float)sqrt(wave_sum_d((double)sres * sres));
            if (!isfinite(rnorm)) {
                status = kSpaiNumericalFailure;
                break;
            }
            if (fits != 0)
                ++growths;
            ++fits;
            {
                const double fm = m, fq = q;
                qr_flops += 2 * fm * fq * fq - 2 * fq * fq * fq / 3 + 4 * fm * fq + fq * fq;
            }
            if (lane < q) {
                jout[lane] = W.J[lane];
                cout_[lane] = coef;
            }
            qdone = q;
            last_res = (double)rnorm;
            last_piv = (double)pivot_ratio;
            if (rnorm <= eps) {
                status = kSpaiSuccess;
                break;
            }
            if ((long long)growths == alpha) {
                status = kSpaiIterationLimit;
                break;
            }
            // ---- candidates: union of B rows i in I, minus J (the union
            // holds J: row j of A lies in the shadow, so j is in B(I))
            wave_sync(); // residual visible to the candidate stage
            const int nb = __builtin_amdgcn_readfirstlane(
                wave_rows(W, bptr, nullptr, irow < 0 ? 0 : irow, m, lane));
            if (nb < 0) {
                status = kSpaiOverflow;
                break;
            }
            const int nc =
                __builtin_amdgcn_readfirstlane(wave_candidates(W, bcol, bval, nb, q, lane));
            if (nc + q > kWC) {
                status = kSpaiOverflow;
                break;
            }
            if (nc == 0) {
                status = kSpaiEmptyCandidates;
                break;
            }
            // ---- restricted scores
            const float rr = rnorm * rnorm;
            int first_bad = INT_MAX, bad_status = kSpaiSuccess, cl_before[kWC / kWave];
            float msum_l = 0.f; // the lane's scores, in its order
            float cdj[kWC / kWave];
#pragma unroll
            for (int u = 0; u < kWC / kWave; ++u) {
                const int t = lane + u * kWave;
                cdj[u] = t < nc ? gld(dscale + W.cand[t]) : 0.f;
            }
#pragma unroll
            for (int u = 0; u < kWC / kWave; ++u) {
                const int t = lane + u * kWave;
                cl_before[u] = 0;
                if (t < nc) {
                    float dt, den;
                    candidate_terms(W, t, cdj[u], dt, den);
                    float sv = 0.f;
                    bool cl = false;
                    const int st = score_of(rr, dt, den, sv, cl);
                    W.score[t] = sv;
                    if (st != kSpaiSuccess) {
                        if (t < first_bad) {
                            first_bad = t;
                            bad_status = st;
                        }
                    } else {
                        cl_before[u] = cl ? 1 : 0;
                        msum_l += sv;
                    }
                }
            }
            const int fb = -wave_max(-first_bad);
            {
                int cl = 0;
#pragma unroll
                for (int u = 0; u < kWC / kWave; ++u)
                    if (lane + u * kWave < fb)
                        cl += cl_before[u];
                clamps += wave_sum_int(cl);
            }
            if (fb != INT_MAX) {
                status = __shfl(bad_status, __ffsll((long long)__ballot(first_bad == fb)) - 1,
                                kWave);
                break;
            }
            wave_sync();
            // the mean: each lane's partial sum, then a fixed butterfly
            const float mean = wave_sum_f(msum_l) / (float)nc;
            if (!isfinite(mean)) {
                status = kSpaiNumericalFailure;
                break;
            }
            // ---- the beta smallest (score, index) with score <= mean, by
            // argmin rounds: scores are finite and >= 0, so their bits order
            // as their values; candidates ascend with t, so among equal
            // scores the smallest index is the smallest t (slot 0's lanes
            // before slot 1's)
            unsigned skey[kWC / kWave];
            int elig = 0;
#pragma unroll
            for (int u = 0; u < kWC / kWave; ++u) {
                const int t = lane + u * kWave;
                skey[u] = ~0u;
                if (t < nc) {
                    const float st_ = W.score[t];
                    if (st_ <= mean) {
                        skey[u] = st_ == 0.f ? 0u : __float_as_uint(st_);
                        ++elig;
                    }
                }
            }
            int nsel = __builtin_amdgcn_readfirstlane(wave_sum_int(elig));
            nsel = nsel < beta ? nsel : beta;
            if (nsel == 0) {
                status = kSpaiNoGrowth;
                break;
            }
            if (q + nsel > kWQ) { // the grown pattern would not fit
                status = kSpaiOverflow;
                break;
            }
            static_assert(kWC / kWave == 2, "two candidate slots per lane");
            for (int r = 0; r < nsel; ++r) {
                const unsigned lo = wave_min_u32(skey[0] < skey[1] ? skey[0] : skey[1]);
                const unsigned long long m0 = __ballot(skey[0] == lo);
                const unsigned long long m1 = __ballot(skey[1] == lo);
                const int u = m0 != 0ull ? 0 : 1;
                const int l = __ffsll((long long)(u == 0 ? m0 : m1)) - 1;
                if (lane == l) {
                    if (u == 0)
                        skey[0] = ~0u;
                    else
                        skey[1] = ~0u;
                }
                if (lane == 0)
                    W.sel[r] = W.cand[l + u * kWave];
            }
            wave_sync();
            // J <- sorted(J u selection)
            {
                const int nq = q + nsel;
                int mine = INT_MAX;
                if (lane < q)
                    mine = W.J[lane];
                else if (lane < nq)
                    mine = W.sel[lane - q];
                int rank = 0;
                for (int t = 0; t < nq; ++t)
                    rank += __builtin_amdgcn_readlane(mine, t) < mine ? 1 : 0;
                wave_sync();
                if (lane < nq)
                    W.J[rank] = mine;
                q = nq;
                wave_sync();
            }
        }
        if (lane == 0) {
            out.status[k] = status;
            out.fits[k] = fits;
            out.growths[k] = growths;
            out.clamps[k] = clamps;
            out.q[k] = status == kSpaiOverflow ? 0 : qdone;
            out.residual[k] = last_res;
            out.pivot[k] = last_piv;
            out.flops[k] = qr_flops;
            out.off[k] = (std::int64_t)slot * out_stride;
            out.pass[k] = out.pass_id;
        }
        wave_sync();
    }
}

void launch_spai_wave(hipStream_t s, int blocks, const int *aptr, const int *acol,
                      const float *aval, const float *dscale, const int *bptr, const int *bcol,
                      const float *bval, float eps, long long alpha, int beta,
                      const int *worklist, int nwork, int *cursor, SpaiColumnOut out,
                      int *out_j, float *out_c, std::int64_t out_stride) {
    hipLaunchKernelGGL(spai_wave_kernel, dim3(blocks), dim3(kWave), 0, s, aptr, acol, aval,
                       dscale, bptr, bcol, bval, eps, alpha, beta, worklist, nwork, cursor, out,
                       out_j, out_c, out_stride);
    HIP_LAUNCH_CHECK();
}
int spai_wave_qcap() {
    return kWQ;
}
int spai_wave_max_n() {
    return (1 << kWKeyColBits) - 1;
}

// ================================================================ block kernel
// Columns that outgrow the lane kernel's caps (long rows, wide patterns,
// thousands of candidates: the power-law matrices) are fitted by one
// 256-thread workgroup each.  The arithmetic is the lane kernel's, in the
// same order, so FP64 stays bit-identical to the CPU:
//   * every hypot chain, dot product, the back substitution and the mean of
//     the scores run sequentially (one thread, or one thread per column);
//   * elementwise work runs in parallel: gathers, Householder vector scaling,
//     the reflection of every (row, column) entry, per-row residual sums in
//     pattern order, per-candidate scores;
//   * index sets are formed with a per-workgroup bitmap over the n columns
//     (mark, then an ordered compaction that also clears it), replacing the
//     repeated sorted merges of the lane kernel, and the beta smallest
//     (score, index) pairs are found by block-wide argmin rounds.
// Per workgroup: T area  R[Q][M] | G[Q][M] | rhs[M] | v[M] | res[M] | score[C] | coef[Q] |
//                         dot[Q+1] | vhead[Q] | pv[P]
//                int area I[M] | J[Q] | Jt[Q] | cand[C] | clamp[C] | sel[Q] | Jf[Q] | vlen[Q] |
//                         cnt[C+1] | pr[P] | wpre[W] | bitmap[W] | inI[W] | pos[n]
// G keeps the gathered B(I,J) for the residual; pos maps a column index to its
// position in I (-1 elsewhere), replacing binary searches, and the bitmap inI
// marks the same rows: the scores test membership there first (62.5 KB per
// workgroup at n = 500,000, against pos's 2 MB), so that only the few entries
// of a candidate's row that lie in I read pos.
//
// FP32 (the mixed path's build) refits incrementally.  The reference refits
// every growth from scratch (src/spai.cpp:134: least_squares on the whole
// B(I,J)), which for the power-law matrices' columns (q in the hundreds, I in
// the thousands) is O(m q^2) per fit and O(m q^3 / beta) per column.  Here a
// row keeps the slot it got when it entered I (new rows are appended), a
// column of B(I,J) keeps its position in the factorization (Jf: new columns
// appended), and the Householder reflectors stay in R below the diagonal
// (v_j[j] in vhead[j], support slots j..vlen[j]).  Because a new row of I is
// zero in every old column (I is the union of the old columns' rows), the old
// reflectors leave it alone: a growth applies the old reflectors to the new
// columns only -- one wavefront per new column, all reflectors in order, no
// workgroup barrier -- and factors the new columns.  The least-squares
// problem, and so the coefficients, are the reference's up to FP32 rounding;
// the summation order differs, which the FP32 build allows.
constexpr int kBlk = 256;
constexpr int kWyG = 16; // (FP32) reflectors per group of the blocked application

// pcap: (FP32) room for the (row, candidate) pairs of one fit
struct BlockCaps {
    int mcap, qcap, ccap, words;
    std::int64_t pcap;
};

// T area per workgroup, rounded to 64 entries so that every workgroup's area
// (and, with M a multiple of 16, every column of R) starts 16-byte aligned;
// the FP32 build's group T factors ((Q/16 + 1) * 256) and staged pair values
// (P) follow the pair values.
std::int64_t spai_block_doubles(int M, int Q, int C, std::int64_t P) {
    const std::int64_t t = 2LL * Q * M + 3LL * M + C + 3LL * Q + 1 + 2 * P +
                           (std::int64_t)(Q / kWyG + 1) * kWyG * kWyG;
    return (t + 63) / 64 * 64;
}
std::int64_t spai_block_ints(int M, int Q, int C, std::int64_t P, int words, int n) {
    return (std::int64_t)M + 5LL * Q + 3LL * C + 1 + 3 * P + 3LL * words + n;
}

// Block-wide helpers (blockDim.x == kBlk).
__device__ inline int block_scan_excl(int v, int *sh, int &total) {
    // sh: kBlk + 1 ints
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kBlk; off <<= 1) {
        const int add = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    total = sh[kBlk - 1];
    const int ex = sh[tid] - v;
    __syncthreads();
    return ex;
}

// Ordered compaction of the marked bits in words [w0, w1] into out[0..); the
// words are cleared.  Returns the number of set bits (entries past `cap` are
// counted, not written).
__device__ int bitmap_compact(unsigned *bm, int w0, int w1, int *out, int cap, int *sh) {
    constexpr int kW = 8; // words loaded together
    const int tid = threadIdx.x;
    const int nw = w1 >= w0 ? w1 - w0 + 1 : 0;
    const int per = (nw + kBlk - 1) / kBlk;
    const int a = w0 + tid * per, b = min(w0 + nw, a + per);
    int cnt = 0;
    for (int w0b = a; w0b < b; w0b += kW) {
        unsigned x[kW];
#pragma unroll
        for (int u = 0; u < kW; ++u)
            x[u] = w0b + u < b ? bm[w0b + u] : 0u;
#pragma unroll
        for (int u = 0; u < kW; ++u)
            cnt += __popc(x[u]);
    }
    int total;
    int pos = block_scan_excl(cnt, sh, total);
    for (int w0b = a; w0b < b; w0b += kW) {
        unsigned x[kW];
#pragma unroll
        for (int u = 0; u < kW; ++u)
            x[u] = w0b + u < b ? bm[w0b + u] : 0u;
#pragma unroll
        for (int u = 0; u < kW; ++u) {
            unsigned bits = x[u];
            while (bits) {
                const int bit = __ffs(bits) - 1;
                bits &= bits - 1;
                if (pos < cap)
                    out[pos] = (w0b + u) * 32 + bit;
                ++pos;
            }
            if (w0b + u < b && x[u])
                bm[w0b + u] = 0u;
        }
    }
    __syncthreads();
    return total;
}

// As bitmap_compact, but the words keep their bits and wpre[w - w0] gets the
// number of set bits in the words before w, so that the position of a set
// bit x in out[] is wpre[(x >> 5) - w0] + popcount of the lower bits of its
// word (a rank dictionary).  The caller clears the words (bitmap_clear).
__device__ int bitmap_rank_list(const unsigned *bm, int w0, int w1, int *out, int cap, int *sh,
                                int *wpre) {
    const int tid = threadIdx.x;
    const int nw = w1 >= w0 ? w1 - w0 + 1 : 0;
    const int per = (nw + kBlk - 1) / kBlk;
    const int a = w0 + tid * per, b = min(w0 + nw, a + per);
    int cnt = 0;
    for (int w = a; w < b; ++w)
        cnt += __popc(bm[w]);
    int total;
    int pos = block_scan_excl(cnt, sh, total);
    for (int w = a; w < b; ++w) {
        wpre[w - w0] = pos;
        unsigned bits = bm[w];
        while (bits) {
            const int bit = __ffs(bits) - 1;
            bits &= bits - 1;
            if (pos < cap)
                out[pos] = w * 32 + bit;
            ++pos;
        }
    }
    __syncthreads();
    return total;
}

__device__ void bitmap_clear(unsigned *bm, int w0, int w1) {
    for (int w = w0 + (int)threadIdx.x; w <= w1; w += kBlk)
        bm[w] = 0u;
    __syncthreads();
}

// Marks column x; the caller keeps the word range [lo, hi] it touched and
// publishes it once with publish_range.
__device__ inline void mark(unsigned *bm, int x, int &lo, int &hi) {
    atomicOr(bm + (x >> 5), 1u << (x & 31));
    lo = min(lo, x >> 5);
    hi = max(hi, x >> 5);
}
__device__ inline void publish_range(int lo, int hi, int *wlo, int *whi) {
    if (hi >= 0) {
        atomicMin(wlo, lo);
        atomicMax(whi, hi);
    }
}

// Block sum in a fixed tree order (every thread gets the same bits).
__device__ inline double block_sum_fixed(double v, double *sh) {
    // wave sums by a fixed butterfly, then the four waves' in a fixed order
    const int tid = threadIdx.x;
    v = wave_sum_d(v);
    if ((tid & (kWave - 1)) == 0)
        sh[tid / kWave] = v;
    __syncthreads();
    static_assert(kBlk == 4 * kWave, "four waves per workgroup");
    const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

constexpr int kBatch = 16; // loads issued together in the sequential loops
constexpr int kRowB = 4;   // rows (or entries) per lane loaded together
constexpr int kPairB = 4;  // (FP32 scores) entries, pairs or candidates per thread at a time
constexpr int kTop = 8;    // (FP32 selection) smallest scores each thread keeps
constexpr int kPairS = 4;  // (FP32 scores) pairs of a candidate ordered in registers
constexpr int kCandB = 2;  // (FP32 scores) candidates per thread at a time in the sums

// hypot_ref(...hypot_ref(hypot_ref(0, x0), x1)..., x_{len-1}): the reference's
// norm2 chain, with the loads issued in batches ahead of the chain.
template <typename T> __device__ T hypot_chain(const T *x, int len) {
    T acc = 0;
    for (int r0 = 0; r0 < len; r0 += kBatch) {
        T b[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            b[u] = x[min(r0 + u, len - 1)];
#pragma unroll
        for (int u = 0; u < kBatch; ++u)
            if (r0 + u < len)
                acc = hypot_ref(acc, b[u]);
    }
    return acc;
}

// First position p in sorted a[0..m) with a[p] >= x.
__device__ inline int lower_bound(const int *a, int m, int x) {
    int lo = 0, hi = m;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// ---- (FP32) blocked Householder application on the matrix cores.
// Reflectors j of the incremental QR are grouped by 16 (group g holds
// j = 16g .. 16g+15); group g's product H_16g ... H_16g+15 is I - V T V^T with
// T upper triangular (16 x 16, kept in the work area), so applying the group
// to columns X in order is X - V T^T (V^T X): two tall-skinny products on
// v_mfma_f32_16x16x4f32 plus a 16 x 16 triangular one.  The instruction's
// lane layout (one wave, l = lane): a = A[l % 16][l / 16],
// b = B[l / 16][l % 16], d[i] = D[4 (l / 16) + i][l % 16].
typedef float wy_f4 __attribute__((ext_vector_type(4)));

__device__ inline wy_f4 wy_mfma(float a, float b, wy_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Rows r .. r+3 (r % 4 == 0, r + 3 < the column stride) of reflector j: zero
// above its diagonal and from row hi on, its head on the diagonal, the stored
// tail below; all zero when j >= jend (not a reflector yet).
__device__ inline wy_f4 wy_refl4(const float *R, int M, const float *vhead, int j, int jend,
                                 int r, int hi) {
    wy_f4 x = {0.f, 0.f, 0.f, 0.f};
    if (j >= jend || r >= hi || r + 3 < j)
        return x;
    x = *reinterpret_cast<const wy_f4 *>(R + (std::int64_t)j * M + r);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int row = r + e;
        x[e] = row < j || row >= hi ? 0.f : (row == j ? vhead[j] : x[e]);
    }
    return x;
}
// One entry of reflector j at row `row` (same conventions).
__device__ inline float wy_refl1(const float *R, int M, const float *vhead, int j, int jend,
                                 int row, int hi) {
    if (j >= jend || row < j || row >= hi)
        return 0.f;
    return row == j ? vhead[j] : R[(std::int64_t)j * M + row];
}

// kOcc: waves per SIMD the register allocation is held to: 4 for the first
// workgroup pass (moderate columns: more of them in flight pays), 3 beyond
// (columns with work areas of megabytes).  Measured: a batch of 256
// log-normal systems builds its FP64 SPAI in 919 instead of 1,016 ms (mixed:
// 335 instead of 367 ms); the power-law build at n = 50,000 took 8.0 s with
// every pass at 4 and 7.6 s this way, within its run-to-run range of 6.7 to
// 7.6 s at 3.
template <typename T, int kOcc>
__global__ void __launch_bounds__(kBlk) __attribute__((amdgpu_waves_per_eu(kOcc)))
spai_block_kernel(int n, const int *__restrict__ aptr, const int *__restrict__ acol,
                  const T *__restrict__ aval, const T *__restrict__ dscale,
                  const int *__restrict__ bptr, const int *__restrict__ bcol,
                  const T *__restrict__ bval, T eps,
                  long long alpha, int beta, const int *__restrict__ worklist, int nwork,
                  int *__restrict__ cursor, BlockCaps caps, T *__restrict__ dwork,
                  int *__restrict__ iwork, std::int64_t blk_d, std::int64_t blk_i,
                  SpaiColumnOut out, int *__restrict__ out_j, T *__restrict__ out_c,
                  std::int64_t out_stride) {
    __shared__ int sh[kBlk + 1];
    __shared__ int s_slot, s_pk, s_flag, s_wlo, s_whi, s_fail, s_np;
    __shared__ T s_red[kBlk];
    __shared__ int s_redi[kBlk];
    __shared__ double s_dred[kBlk];
    __shared__ T s_scal[4];
    constexpr bool kExact = sizeof(T) == sizeof(double);
    // (FP32) the back substitution's working vector, when q fits
    constexpr int kYLds = kExact ? 1 : 2048;
    __shared__ T s_y[kYLds];
    // (FP32) blocked reflector application: the waves' partial 16 x 16 sums, a
    // group's T factor and T^T W
    constexpr int kNW = kBlk / kWave;
    static_assert(kNW == 4, "the chunk dealing assumes four waves");
    constexpr int kWyLds = kExact ? 1 : (kNW + 2) * 256;
    __shared__ float s_wy[kWyLds];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const int M = caps.mcap, Q = caps.qcap, C = caps.ccap;
    T *D = dwork + (std::int64_t)blockIdx.x * blk_d;
    int *Iw = iwork + (std::int64_t)blockIdx.x * blk_i;
    T *R = D, *G = D + (std::int64_t)Q * M, *rhs = G + (std::int64_t)Q * M, *v = rhs + M,
      *res = v + M, *sc = res + M, *coef = sc + C, *dot = coef + Q, *vhead = dot + Q + 1;
    T *pv = vhead + Q; // (FP32) values of the (row, candidate) pairs
    T *Tg = pv + caps.pcap; // (FP32) T factors of the reflector groups, 16 x 16 each
    T *spv = Tg + (Q / kWyG + 1) * kWyG * kWyG; // (FP32) staged pairs' scaled values
    int *I = Iw, *J = I + M, *Jt = J + Q, *cand = Jt + Q, *clampf = cand + C, *sel = clampf + C;
    int *Jf = sel + Q, *vlen = Jf + Q; // (FP32) factorization order, reflector supports
    // (FP32) pairs per candidate (counts, then offsets), their rows, and the
    // candidate bitmap's rank dictionary (set bits before each word)
    // and, before bucketing, the staged pairs' candidate ranks and slots
    int *cnt = vlen + Q, *pr = cnt + C + 1, *sk = pr + caps.pcap, *sr = sk + caps.pcap;
    int *wpre = sr + caps.pcap;
    unsigned *bm = reinterpret_cast<unsigned *>(wpre + caps.words);
    unsigned *inI = bm + caps.words;                       // rows of I; clear between columns
    int *pos = reinterpret_cast<int *>(inI + caps.words); // all -1 between columns
    int mapped = 0; // entries of I currently recorded in pos

    for (;;) {
        if (tid == 0)
            s_slot = atomicAdd(cursor, 1);
        __syncthreads();
        const int slot = s_slot;
        __syncthreads();
        if (slot >= nwork)
            break;
        const int k = worklist[slot];
        int status = kSpaiNotAttempted;
        int fits = 0, growths = 0, clamps = 0, q = 1, qdone = 0;
        double last_res = __builtin_nan(""), last_piv = __builtin_nan(""), qr_flops = 0;
        int mslots = 0, qf = 0;   // (FP32) slots of I and columns factored so far
        T piv_min = T(INFINITY);  // (FP32) smallest |R_jj| so far
        if (tid == 0) {
            J[0] = k;
            Jf[0] = k;
        }
        int *jout = out_j + (std::int64_t)slot * out_stride;
        T *cout_ = out_c + (std::int64_t)slot * out_stride;
        __syncthreads();
        for (;;) {
            int m = 0, pk = -1;
            T inf_norm = 0, min_pivot = T(INFINITY);
            int fit_status = kSpaiSuccess;
            if constexpr (kExact) {
                // ---- shadow I = union of cols(A row j), j in J (bitmap, ordered)
                if (tid == 0) {
                    s_wlo = INT_MAX;
                    s_whi = -1;
                }
                __syncthreads();
                {
                    int lo = INT_MAX, hi = -1;
                    for (int jj = wave; jj < q; jj += kBlk / kWave) {
                        const int j = J[jj];
                        for (int e = aptr[j] + lane; e < aptr[j + 1]; e += kWave)
                            mark(bm, acol[e], lo, hi);
                    }
                    publish_range(lo, hi, &s_wlo, &s_whi);
                }
                __syncthreads();
                for (int r = tid; r < mapped; r += kBlk) { // the previous fit's shadow
                    pos[I[r]] = -1;
                    atomicAnd(inI + (I[r] >> 5), ~(1u << (I[r] & 31)));
                }
                __syncthreads();
                mapped = 0;
                m = bitmap_compact(bm, s_wlo, s_whi, I, M, sh);
                if (m > M) {
                    status = kSpaiOverflow;
                    break;
                }
                if (m < q) {
                    status = kSpaiRankDeficient;
                    break;
                }
                for (int r = tid; r < m; r += kBlk) {
                    pos[I[r]] = r;
                    atomicOr(inI + (I[r] >> 5), 1u << (I[r] & 31));
                }
                mapped = m;
                __syncthreads();
                // ---- gather R(i, c) = a_{J[c], I[i]} d_{J[c]}, rhs = e_k(I)
                if (tid == 0)
                    s_pk = -1;
                for (int c = wave; c < q; c += kBlk / kWave)
                    for (int r = lane; r < m; r += kWave) {
                        R[(std::int64_t)c * M + r] = T(0);
                        G[(std::int64_t)c * M + r] = T(0);
                    }
                __syncthreads();
                for (int r = tid; r < m; r += kBlk) {
                    rhs[r] = I[r] == k ? T(1) : T(0);
                    if (I[r] == k)
                        s_pk = r;
                }
                for (int c = wave; c < q; c += kBlk / kWave) {
                    const int j = J[c];
                    const T dj = dscale[j];
                    for (int e = aptr[j] + lane; e < aptr[j + 1]; e += kWave) {
                        const std::int64_t at = (std::int64_t)c * M + pos[acol[e]];
                        const T x = aval[e] * dj;
                        R[at] = x;
                        G[at] = x;
                    }
                }
                __syncthreads();
                pk = s_pk;
                if (pk < 0) {
                    status = kSpaiNumericalFailure;
                    break;
                }
                // ---- Householder least squares (src/spai.cpp:30-98)
                T rmax = 0;
                for (int r = tid; r < m; r += kBlk) {
                    T sum = 0;
                    for (int c0 = 0; c0 < q; c0 += kBatch) {
                        T xb[kBatch];
    #pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            xb[u] = R[(std::int64_t)min(c0 + u, q - 1) * M + r];
    #pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            if (c0 + u < q)
                                sum += fabs(xb[u]);
                    }
                    rmax = sum > rmax ? sum : rmax;
                }
                s_red[tid] = rmax;
                __syncthreads();
                if (tid == 0) {
                    T mx = 0;
                    for (int t = 0; t < kBlk; ++t)
                        mx = s_red[t] > mx ? s_red[t] : mx;
                    s_scal[3] = mx;
                }
                __syncthreads();
                inf_norm = s_scal[3];
                for (int j = 0; j < q; ++j) {
                    const int len = m - j;
                    T *colj = R + (std::int64_t)j * M + j;
                    T length;
                    if (kExact) { // the reference's hypot chain (norm2)
                        for (int r = tid; r < len; r += kBlk)
                            v[r] = colj[r];
                        __syncthreads();
                        if (tid == 0)
                            s_scal[0] = hypot_chain(v, len);
                        __syncthreads();
                        length = s_scal[0];
                    } else { // FP32: sum of squares in FP64, fixed-order tree
                        double acc = 0;
                        for (int r = tid; r < len; r += kBlk) {
                            const T x = colj[r];
                            v[r] = x;
                            acc += (double)x * (double)x;
                        }
                        length = (T)sqrt(block_sum_fixed(acc, s_dred));
                    }
                    if (!isfinite(length)) {
                        fit_status = kSpaiNumericalFailure;
                        break;
                    }
                    if (length == 0) {
                        fit_status = kSpaiRankDeficient;
                        break;
                    }
                    const T diagonal = -copysign(length, v[0]);
                    __syncthreads(); // v[0] read before it is rescaled
                    for (int r = tid; r < len; r += kBlk) {
                        T x = v[r] / length;
                        if (r == 0)
                            x += copysign(T(1), x);
                        v[r] = x;
                    }
                    __syncthreads();
                    T vnorm;
                    if (kExact) {
                        if (tid == 0)
                            s_scal[1] = hypot_chain(v, len);
                        __syncthreads();
                        vnorm = s_scal[1];
                    } else {
                        double acc = 0;
                        for (int r = tid; r < len; r += kBlk)
                            acc += (double)v[r] * (double)v[r];
                        vnorm = (T)sqrt(block_sum_fixed(acc, s_dred));
                    }
                    for (int r = tid; r < len; r += kBlk)
                        v[r] = v[r] / vnorm;
                    __syncthreads();
                    // slots j+1..q (q = right side)
                    if (kExact) {
                        // one thread per slot, rows in order; loads in batches
                        for (int c = j + 1 + tid; c <= q; c += kBlk) {
                            const T *x = (c < q ? R + (std::int64_t)c * M : rhs) + j;
                            T d = 0;
                            for (int r0 = 0; r0 < len; r0 += kBatch) {
                                T vb[kBatch], xb[kBatch];
    #pragma unroll
                                for (int u = 0; u < kBatch; ++u) {
                                    const int r = min(r0 + u, len - 1);
                                    vb[u] = v[r];
                                    xb[u] = x[r];
                                }
    #pragma unroll
                                for (int u = 0; u < kBatch; ++u)
                                    if (r0 + u < len)
                                        d += vb[u] * xb[u];
                            }
                            dot[c] = d;
                        }
                    } else {
                        // FP32: one wave per slot, lanes over rows (loads batched),
                        // FP64 partials combined by a fixed butterfly
                        for (int c = j + 1 + wave; c <= q; c += kBlk / kWave) {
                            const T *x = (c < q ? R + (std::int64_t)c * M : rhs) + j;
                            double d = 0;
                            for (int r0 = lane; r0 < len; r0 += kWave * kRowB) {
                                T vb[kRowB], xb[kRowB];
    #pragma unroll
                                for (int u = 0; u < kRowB; ++u) {
                                    const int r = min(r0 + u * kWave, len - 1);
                                    vb[u] = v[r];
                                    xb[u] = x[r];
                                }
    #pragma unroll
                                for (int u = 0; u < kRowB; ++u)
                                    if (r0 + u * kWave < len)
                                        d += (double)vb[u] * (double)xb[u];
                            }
                            d = wave_sum(d);
                            if (lane == 0)
                                dot[c] = (T)d;
                        }
                    }
                    if (tid == 0)
                        s_flag = 0;
                    __syncthreads();
                    bool bad = false;
                    for (int c = j + 1 + wave; c <= q; c += kBlk / kWave) {
                        T *x = (c < q ? R + (std::int64_t)c * M : rhs) + j;
                        const T dc = dot[c];
                        for (int r0 = lane; r0 < len; r0 += kWave * kRowB) {
                            T vb[kRowB], xb[kRowB];
    #pragma unroll
                            for (int u = 0; u < kRowB; ++u) {
                                const int r = min(r0 + u * kWave, len - 1);
                                vb[u] = v[r];
                                xb[u] = x[r];
                            }
    #pragma unroll
                            for (int u = 0; u < kRowB; ++u)
                                if (r0 + u * kWave < len) {
                                    const T y = xb[u] - T(2) * vb[u] * dc;
                                    x[r0 + u * kWave] = y;
                                    bad = bad || !isfinite(y);
                                }
                        }
                    }
                    if (bad)
                        s_flag = 1;
                    __syncthreads();
                    if (s_flag) {
                        fit_status = kSpaiNumericalFailure;
                        break;
                    }
                    if (tid == 0)
                        colj[0] = diagonal;
                    const T ad = fabs(diagonal);
                    min_pivot = ad < min_pivot ? ad : min_pivot;
                    __syncthreads();
                }
                if (fit_status == kSpaiSuccess) {
                    if (tid == 0) {
                        int st = kSpaiSuccess;
                        for (int j = q - 1; j >= 0; --j) {
                            T value = rhs[j];
                            for (int c0 = j + 1; c0 < q; c0 += kBatch) {
                                T rb[kBatch], cb[kBatch];
    #pragma unroll
                                for (int u = 0; u < kBatch; ++u) {
                                    const int c = min(c0 + u, q - 1);
                                    rb[u] = R[(std::int64_t)c * M + j];
                                    cb[u] = coef[c];
                                }
    #pragma unroll
                                for (int u = 0; u < kBatch; ++u)
                                    if (c0 + u < q)
                                        value -= rb[u] * cb[u];
                            }
                            const T dg = R[(std::int64_t)j * M + j];
                            if (!isfinite(dg)) {
                                st = kSpaiNumericalFailure;
                                break;
                            }
                            if (dg == 0) {
                                st = kSpaiRankDeficient;
                                break;
                            }
                            const T sj = value / dg;
                            if (!isfinite(sj)) {
                                st = kSpaiNumericalFailure;
                                break;
                            }
                            coef[j] = sj;
                        }
                        s_fail = st;
                    }
                    __syncthreads();
                    fit_status = s_fail;
                }
                if (fit_status != kSpaiSuccess) {
                    status = fit_status;
                    break;
                }
            } else {
                // ---- (FP32) incremental fit: see the layout notes above.  New
                // rows: those of the columns added since the last fit that have
                // no slot yet, appended in ascending order.
                if (tid == 0) {
                    s_wlo = INT_MAX;
                    s_whi = -1;
                }
                __syncthreads();
                {
                    int lo = INT_MAX, hi = -1;
                    for (int jj = qf + wave; jj < q; jj += kBlk / kWave) {
                        const int j = Jf[jj];
                        for (int e = aptr[j] + lane; e < aptr[j + 1]; e += kWave) {
                            const int row = acol[e];
                            if (pos[row] < 0)
                                mark(bm, row, lo, hi);
                        }
                    }
                    publish_range(lo, hi, &s_wlo, &s_whi);
                }
                __syncthreads();
                const int mold = mslots;
                m = mold + bitmap_compact(bm, s_wlo, s_whi, I + mold, M - mold, sh);
                if (m > M) {
                    status = kSpaiOverflow;
                    break;
                }
                if (m < q) {
                    status = kSpaiRankDeficient;
                    break;
                }
                for (int r = mold + tid; r < m; r += kBlk) {
                    pos[I[r]] = r;
                    atomicOr(inI + (I[r] >> 5), 1u << (I[r] & 31));
                }
                mapped = m;
                mslots = m;
                // Old columns are zero at the new slots; new columns are gathered
                // whole; the transformed right side is e_k at the new slots (the
                // old reflectors do not reach them).
                for (int c = wave; c < q; c += kBlk / kWave)
                    for (int r = (c < qf ? mold : 0) + lane; r < m; r += kWave) {
                        R[(std::int64_t)c * M + r] = T(0);
                        G[(std::int64_t)c * M + r] = T(0);
                    }
                for (int r = mold + tid; r < m; r += kBlk)
                    rhs[r] = I[r] == k ? T(1) : T(0);
                __syncthreads();
                for (int c = qf + wave; c < q; c += kBlk / kWave) {
                    const int j = Jf[c];
                    const T dj = dscale[j];
                    for (int e = aptr[j] + lane; e < aptr[j + 1]; e += kWave) {
                        const std::int64_t at = (std::int64_t)c * M + pos[acol[e]];
                        const T x = aval[e] * dj;
                        R[at] = x;
                        G[at] = x;
                    }
                }
                __syncthreads();
                pk = pos[k];
                if (pk < 0) {
                    status = kSpaiNumericalFailure;
                    break;
                }
                // ||B(I,J)||_inf from the gathered copy (every column, every slot)
                {
                    T rmax = 0;
                    for (int r = tid; r < m; r += kBlk) {
                        T sum = 0;
                        for (int c0 = 0; c0 < q; c0 += kBatch) {
                            T xb[kBatch];
    #pragma unroll
                            for (int u = 0; u < kBatch; ++u)
                                xb[u] = G[(std::int64_t)min(c0 + u, q - 1) * M + r];
    #pragma unroll
                            for (int u = 0; u < kBatch; ++u)
                                if (c0 + u < q)
                                    sum += fabs(xb[u]);
                        }
                        rmax = sum > rmax ? sum : rmax;
                    }
                    s_red[tid] = rmax;
                    __syncthreads();
                    if (tid == 0) {
                        T mx = 0;
                        for (int t = 0; t < kBlk; ++t)
                            mx = s_red[t] > mx ? s_red[t] : mx;
                        s_scal[3] = mx;
                    }
                    __syncthreads();
                    inf_norm = s_scal[3];
                }
                // The old reflectors on the new columns, group after group in
                // their order, in the blocked form above.  16-row chunks of the
                // slots are dealt to the waves (chunk c to wave c % 4) alike for
                // every group, so the sweep that applies group g to a wave's rows
                // also forms that wave's share of group g+1's V^T X from the rows
                // it has just rewritten.  No old reflector reaches a slot >= mold.
                {
                    float *wp = s_wy, *tt = s_wy + kNW * 256, *yy = tt + 256;
                    const int cl = lane & 15, quad = lane >> 4;
                    const int ngroups = (qf + kWyG - 1) / kWyG;
                    const int nchunks = (mold + 15) / 16;
                    bool bad = false;
                    for (int c0 = qf; c0 < q; c0 += 16) {
                        const int xc = c0 + cl;
                        const bool xok = xc < q;
                        float *X = R + (std::int64_t)(xok ? xc : 0) * M;
                        wy_f4 acc = {0.f, 0.f, 0.f, 0.f};
                        for (int ch = wave; ch < nchunks; ch += kNW) {
                            const int r = 16 * ch + 4 * quad;
                            const wy_f4 va = wy_refl4(R, M, vhead, cl, qf, r, mold);
                            wy_f4 xb = {0.f, 0.f, 0.f, 0.f};
                            if (xok && r < mold)
                                xb = *reinterpret_cast<const wy_f4 *>(X + r);
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                acc = wy_mfma(va[e], xb[e], acc);
                        }
                        for (int g = 0; g < ngroups; ++g) {
                            const int g0 = g * kWyG;
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                wp[wave * 256 + (4 * quad + e) * 16 + cl] = acc[e];
                            tt[tid] = Tg[(std::int64_t)g * 256 + tid];
                            __syncthreads();
                            {
                                // Y = T^T W (W = the waves' partial sums)
                                const int i = tid >> 4, c = tid & 15;
                                float y = 0.f;
                                if (g0 + i < qf)
                                    for (int l = 0; l <= i; ++l) {
                                        const int z = l * 16 + c;
                                        y += tt[l * 16 + i] * (wp[z] + wp[256 + z] + wp[512 + z] +
                                                               wp[768 + z]);
                                    }
                                yy[tid] = y;
                            }
                            __syncthreads();
                            float yb[4];
#pragma unroll
                            for (int kk = 0; kk < 4; ++kk)
                                yb[kk] = yy[(4 * kk + quad) * 16 + cl];
                            const bool next = g + 1 < ngroups;
                            acc = wy_f4{0.f, 0.f, 0.f, 0.f};
                            // the next chunk's loads are issued before this
                            // chunk's stores (other rows: no overlap)
                            int ch = g + ((wave - g) & (kNW - 1));
                            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
                            wy_f4 xv = {0.f, 0.f, 0.f, 0.f}, vn = {0.f, 0.f, 0.f, 0.f};
                            if (ch < nchunks) {
                                const int r0 = 16 * ch, r = r0 + 4 * quad;
                                a0 = wy_refl1(R, M, vhead, g0 + quad, qf, r0 + cl, mold);
                                a1 = wy_refl1(R, M, vhead, g0 + 4 + quad, qf, r0 + cl, mold);
                                a2 = wy_refl1(R, M, vhead, g0 + 8 + quad, qf, r0 + cl, mold);
                                a3 = wy_refl1(R, M, vhead, g0 + 12 + quad, qf, r0 + cl, mold);
                                if (xok && r < mold)
                                    xv = *reinterpret_cast<const wy_f4 *>(X + r);
                                if (next)
                                    vn = wy_refl4(R, M, vhead, g0 + kWyG + cl, qf, r, mold);
                            }
                            for (; ch < nchunks; ch += kNW) {
                                wy_f4 d = {0.f, 0.f, 0.f, 0.f};
                                d = wy_mfma(a0, yb[0], d);
                                d = wy_mfma(a1, yb[1], d);
                                d = wy_mfma(a2, yb[2], d);
                                d = wy_mfma(a3, yb[3], d);
                                wy_f4 xc = xv;
                                const wy_f4 vc = vn;
                                const int r = 16 * ch + 4 * quad;
                                if (ch + kNW < nchunks) {
                                    const int r0n = 16 * (ch + kNW), rn = r0n + 4 * quad;
                                    a0 = wy_refl1(R, M, vhead, g0 + quad, qf, r0n + cl, mold);
                                    a1 = wy_refl1(R, M, vhead, g0 + 4 + quad, qf, r0n + cl, mold);
                                    a2 = wy_refl1(R, M, vhead, g0 + 8 + quad, qf, r0n + cl, mold);
                                    a3 = wy_refl1(R, M, vhead, g0 + 12 + quad, qf, r0n + cl, mold);
                                    xv = wy_f4{0.f, 0.f, 0.f, 0.f};
                                    if (xok && rn < mold)
                                        xv = *reinterpret_cast<const wy_f4 *>(X + rn);
                                    if (next)
                                        vn = wy_refl4(R, M, vhead, g0 + kWyG + cl, qf, rn, mold);
                                }
                                if (xok && r < mold) {
                                    xc -= d; // zero at rows >= mold
                                    *reinterpret_cast<wy_f4 *>(X + r) = xc;
#pragma unroll
                                    for (int e = 0; e < 4; ++e)
                                        bad = bad || (r + e < mold && !isfinite(xc[e]));
                                }
                                if (next) {
#pragma unroll
                                    for (int e = 0; e < 4; ++e)
                                        acc = wy_mfma(vc[e], xc[e], acc);
                                }
                            }
                        }
                    }
                    if (tid == 0)
                        s_flag = 0;
                    __syncthreads();
                    if (bad)
                        s_flag = 1;
                    __syncthreads();
                    if (s_flag) {
                        status = kSpaiNumericalFailure;
                        break;
                    }
                }
                // Factor the new columns j = qf .. q-1 over slots j .. m-1; the
                // trailing updates reach the new columns and the right side only.
                for (int j = qf; j < q; ++j) {
                    const int len = m - j;
                    T *colj = R + (std::int64_t)j * M + j;
                    double acc = 0;
                    for (int r = tid; r < len; r += kBlk) {
                        const T x = colj[r];
                        v[r] = x;
                        acc += (double)x * (double)x;
                    }
                    const T length = (T)sqrt(block_sum_fixed(acc, s_dred));
                    if (!isfinite(length)) {
                        fit_status = kSpaiNumericalFailure;
                        break;
                    }
                    if (length == 0) {
                        fit_status = kSpaiRankDeficient;
                        break;
                    }
                    const T diagonal = -copysign(length, v[0]);
                    __syncthreads(); // v[0] read before it is rescaled
                    double acc2 = 0;
                    for (int r = tid; r < len; r += kBlk) {
                        T x = v[r] / length;
                        if (r == 0)
                            x += copysign(T(1), x);
                        v[r] = x;
                        acc2 += (double)x * (double)x;
                    }
                    const T vnorm = (T)sqrt(block_sum_fixed(acc2, s_dred));
                    for (int r = tid; r < len; r += kBlk)
                        v[r] = v[r] / vnorm;
                    __syncthreads();
                    for (int c = j + 1 + wave; c <= q; c += kBlk / kWave) {
                        const T *x = (c < q ? R + (std::int64_t)c * M : rhs) + j;
                        double d = 0;
                        for (int r0 = lane; r0 < len; r0 += kWave * kRowB) {
                            T vb[kRowB], xb[kRowB];
    #pragma unroll
                            for (int u = 0; u < kRowB; ++u) {
                                const int r = min(r0 + u * kWave, len - 1);
                                vb[u] = v[r];
                                xb[u] = x[r];
                            }
    #pragma unroll
                            for (int u = 0; u < kRowB; ++u)
                                if (r0 + u * kWave < len)
                                    d += (double)vb[u] * (double)xb[u];
                        }
                        d = wave_sum(d);
                        if (lane == 0)
                            dot[c] = (T)d;
                    }
                    if (tid == 0)
                        s_flag = 0;
                    __syncthreads();
                    bool bad = false;
                    for (int c = j + 1 + wave; c <= q; c += kBlk / kWave) {
                        T *x = (c < q ? R + (std::int64_t)c * M : rhs) + j;
                        const T dc = dot[c];
                        for (int r0 = lane; r0 < len; r0 += kWave * kRowB) {
                            T vb[kRowB], xb[kRowB];
    #pragma unroll
                            for (int u = 0; u < kRowB; ++u) {
                                const int r = min(r0 + u * kWave, len - 1);
                                vb[u] = v[r];
                                xb[u] = x[r];
                            }
    #pragma unroll
                            for (int u = 0; u < kRowB; ++u)
                                if (r0 + u * kWave < len) {
                                    const T y = xb[u] - T(2) * vb[u] * dc;
                                    x[r0 + u * kWave] = y;
                                    bad = bad || !isfinite(y);
                                }
                        }
                    }
                    if (bad)
                        s_flag = 1;
                    // keep the reflector: v_j[0] aside, the rest below the diagonal
                    for (int r = 1 + tid; r < len; r += kBlk)
                        colj[r] = v[r];
                    __syncthreads();
                    if (s_flag) {
                        fit_status = kSpaiNumericalFailure;
                        break;
                    }
                    if (tid == 0) {
                        colj[0] = diagonal;
                        vhead[j] = v[0];
                        vlen[j] = m;
                    }
                    const T ad = fabs(diagonal);
                    piv_min = ad < piv_min ? ad : piv_min;
                    __syncthreads();
                }
                // T factors of the groups the new reflectors joined: the Gram
                // matrix Z = V_g^T V_g on the matrix cores, then column by column
                // T[0:jj, jj] = -2 T[0:jj, 0:jj] Z[0:jj, jj] and T[jj][jj] = 2
                // (unit reflectors, H = I - 2 v v^T); zero below the diagonal.
                if (fit_status == kSpaiSuccess) {
                    float *wp = s_wy, *tt = s_wy + kNW * 256;
                    const int cl = lane & 15, quad = lane >> 4;
                    const int nchunks = (m + 15) / 16;
                    for (int g = qf / kWyG; g * kWyG < q; ++g) {
                        const int g0 = g * kWyG;
                        wy_f4 acc = {0.f, 0.f, 0.f, 0.f};
                        for (int ch = g + ((wave - g) & (kNW - 1)); ch < nchunks; ch += kNW) {
                            const wy_f4 va = wy_refl4(R, M, vhead, g0 + cl, q, 16 * ch + 4 * quad, m);
#pragma unroll
                            for (int e = 0; e < 4; ++e)
                                acc = wy_mfma(va[e], va[e], acc);
                        }
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            wp[wave * 256 + (4 * quad + e) * 16 + cl] = acc[e];
                        tt[tid] = Tg[(std::int64_t)g * 256 + tid];
                        __syncthreads();
                        if (wave == 0) {
                            const int j1 = min(q, g0 + kWyG) - g0;
                            for (int jj = max(qf, g0) - g0; jj < j1; ++jj) {
                                float t = 0.f;
                                if (lane < jj) {
                                    for (int l = lane; l < jj; ++l) {
                                        const int z = l * 16 + jj;
                                        t += tt[lane * 16 + l] *
                                             (wp[z] + wp[256 + z] + wp[512 + z] + wp[768 + z]);
                                    }
                                    t *= -2.f;
                                } else if (lane == jj) {
                                    t = 2.f;
                                }
                                wave_sync();
                                if (lane < 16)
                                    tt[lane * 16 + jj] = t;
                                wave_sync();
                            }
                            for (int e = lane; e < 256; e += kWave)
                                Tg[(std::int64_t)g * 256 + e] = tt[e];
                        }
                        __syncthreads();
                    }
                }
                if (fit_status == kSpaiSuccess)
                    qf = q;
                min_pivot = piv_min;
                // Back substitution by one wave, column by column (R is stored by
                // columns): y_j = y_j / R_jj, then y_i -= R_ij y_j for i < j.
                if (fit_status == kSpaiSuccess) {
                    if (wave == 0) {
                        T *y = q <= kYLds ? s_y : coef;
                        for (int r = lane; r < q; r += kWave)
                            y[r] = rhs[r];
                        wave_sync();
                        int st = kSpaiSuccess;
                        for (int j = q - 1; j >= 0; --j) {
                            const T *cj = R + (std::int64_t)j * M;
                            const T dg = cj[j];
                            if (!isfinite(dg)) {
                                st = kSpaiNumericalFailure;
                                break;
                            }
                            if (dg == 0) {
                                st = kSpaiRankDeficient;
                                break;
                            }
                            const T yj = y[j] / dg;
                            if (!isfinite(yj)) {
                                st = kSpaiNumericalFailure;
                                break;
                            }
                            for (int i = lane; i < j; i += kWave)
                                y[i] = y[i] - cj[i] * yj;
                            wave_sync();
                            if (lane == 0)
                                y[j] = yj;
                            wave_sync();
                        }
                        if (y != coef)
                            for (int r = lane; r < q; r += kWave)
                                coef[r] = y[r];
                        if (lane == 0)
                            s_fail = st;
                    }
                    __syncthreads();
                    fit_status = s_fail;
                }
                if (fit_status != kSpaiSuccess) {
                    status = fit_status;
                    break;
                }
            }
            const T pivot_ratio = inf_norm == 0 ? T(0) : min_pivot / inf_norm;
            // ---- fitted residual s = B(I,J) m - e_k(I): per row, the stored
            // entries in pattern order (as the lane kernel's scatter)
            // (stored entries of B are nonzero: the scaling step rejects zero
            // scaled entries, so G == 0 marks exactly the absent ones)
            for (int r = tid; r < m; r += kBlk) {
                T acc = 0;
                for (int c0 = 0; c0 < q; c0 += kBatch) {
                    T gb[kBatch];
#pragma unroll
                    for (int u = 0; u < kBatch; ++u)
                        gb[u] = G[(std::int64_t)min(c0 + u, q - 1) * M + r];
#pragma unroll
                    for (int u = 0; u < kBatch; ++u)
                        if (c0 + u < q && gb[u] != T(0))
                            acc += gb[u] * coef[c0 + u];
                }
                if (r == pk)
                    acc -= T(1);
                res[r] = acc;
                if (!kExact && m <= kYLds)
                    s_y[r] = acc; // (FP32) the scores' gathers read it from LDS
            }
            __syncthreads();
            T rnorm;
            if (kExact) {
                if (tid == 0) {
                    bool fin = true;
                    for (int r = 0; r < m; ++r)
                        fin = fin && isfinite(res[r]);
                    const T rn = hypot_chain(res, m);
                    s_scal[2] = rn;
                    s_fail = fin && isfinite(rn) ? 0 : 1;
                }
                __syncthreads();
                rnorm = s_scal[2];
            } else {
                double acc = 0;
                for (int r = tid; r < m; r += kBlk)
                    acc += (double)res[r] * (double)res[r];
                const double ss = block_sum_fixed(acc, s_dred);
                rnorm = (T)sqrt(ss);
                if (tid == 0)
                    s_fail = isfinite(ss) && isfinite(rnorm) ? 0 : 1;
                __syncthreads();
            }
            if (s_fail) {
                status = kSpaiNumericalFailure;
                break;
            }
            if (fits != 0)
                ++growths;
            ++fits;
            {
                const double fm = m, fq = q;
                qr_flops += 2 * fm * fq * fq - 2 * fq * fq * fq / 3 + 4 * fm * fq + fq * fq;
            }
            if constexpr (kExact) {
                for (int c = tid; c < q; c += kBlk) {
                    jout[c] = J[c];
                    cout_[c] = coef[c];
                }
            } else { // (FP32) coef follows Jf; the pattern is written sorted
                for (int c = tid; c < q; c += kBlk) {
                    const int at = lower_bound(J, q, Jf[c]);
                    jout[at] = Jf[c];
                    cout_[at] = coef[c];
                }
            }
            qdone = q;
            last_res = (double)rnorm;
            last_piv = (double)pivot_ratio;
            if (rnorm <= eps) {
                status = kSpaiSuccess;
                break;
            }
            if ((long long)growths == alpha) {
                status = kSpaiIterationLimit;
                break;
            }
            // ---- candidates: union of B rows i in I (and k), minus J
            if (tid == 0) {
                s_wlo = INT_MAX;
                s_whi = -1;
            }
            __syncthreads();
            {
                int lo = INT_MAX, hi = -1;
                for (int r = tid; r < m; r += kBlk) {
                    const int i = I[r];
                    const int e1 = bptr[i + 1];
                    for (int e0 = bptr[i]; e0 < e1; e0 += kBatch) {
                        int cb[kBatch];
#pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            cb[u] = bcol[min(e0 + u, e1 - 1)];
#pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            if (e0 + u < e1)
                                mark(bm, cb[u], lo, hi);
                    }
                }
                for (int e = bptr[k] + tid; e < bptr[k + 1]; e += kBlk)
                    mark(bm, bcol[e], lo, hi);
                publish_range(lo, hi, &s_wlo, &s_whi);
            }
            __syncthreads();
            for (int jj = tid; jj < q; jj += kBlk)
                atomicAnd(bm + (J[jj] >> 5), ~(1u << (J[jj] & 31)));
            __syncthreads();
            const int cw0 = s_wlo, cw1 = s_whi;
            const int c = kExact ? bitmap_compact(bm, cw0, cw1, cand, C, sh)
                                 : bitmap_rank_list(bm, cw0, cw1, cand, C, sh, wpre);
            if (c > C || c == 0) {
                if (!kExact)
                    bitmap_clear(bm, cw0, cw1);
                status = c > C ? kSpaiOverflow : kSpaiEmptyCandidates;
                break;
            }
            // ---- restricted scores (src/spai.cpp:198-215), one thread per
            // candidate; the first failing candidate decides the status
            const T rr = rnorm * rnorm;
            if (tid == 0)
                s_fail = INT_MAX;
            __syncthreads();
            // (FP32) the thread's share of the mean and of the clamps, and its
            // kTop smallest (score, index) pairs, ascending
            double msum = 0;
            int mycl = 0;
            T tops[kTop];
            int topj[kTop];
#pragma unroll
            for (int w = 0; w < kTop; ++w) {
                tops[w] = T(INFINITY);
                topj[w] = INT_MAX;
            }
            if constexpr (kExact) {
                for (int t = tid; t < c; t += kBlk) {
                    const int j = cand[t];
                    const T dj = dscale[j];
                    T dt = 0, den = 0;
                    const int e1 = aptr[j + 1];
                    for (int e0 = aptr[j]; e0 < e1; e0 += kRowB) {
                        int pb[kRowB];
                        T ab[kRowB], rb[kRowB];
    #pragma unroll
                        for (int u = 0; u < kRowB; ++u)
                            pb[u] = acol[min(e0 + u, e1 - 1)];
    #pragma unroll
                        for (int u = 0; u < kRowB; ++u)
                            pb[u] = (inI[pb[u] >> 5] >> (pb[u] & 31)) & 1u ? pos[pb[u]] : -1;
    #pragma unroll
                        for (int u = 0; u < kRowB; ++u) {
                            ab[u] = aval[min(e0 + u, e1 - 1)];
                            rb[u] = res[max(pb[u], 0)];
                        }
    #pragma unroll
                        for (int u = 0; u < kRowB; ++u)
                            if (e0 + u < e1 && pb[u] >= 0) {
                                const T value = ab[u] * dj;
                                dt += rb[u] * value;
                                den += value * value;
                            }
                    }
                    T sv = 0;
                    bool cl = false;
                    const int st = score_of(rr, dt, den, sv, cl);
                    sc[t] = sv;
                    clampf[t] = st == kSpaiSuccess ? (cl ? 1 : 0) : -st;
                    if (st != kSpaiSuccess)
                        atomicMin(&s_fail, t);
                }
            } else {
                // (FP32) Scores from the side of I.  A candidate's dot
                // products run over its row of A restricted to I, and on the
                // power-law systems that row is long (31 entries on average
                // at n = 200,000) while only one or two of its entries lie in
                // I.  So the pairs are found from I instead: the entries of
                // B's rows i in I whose column j is a candidate (j's position
                // in cand[] from the bitmap's rank dictionary).  Pass A stages
                // them as (rank, slot, scaled value) in any order and counts
                // each candidate's pairs; pass B buckets the staged pairs per
                // candidate (counting sort); pass C orders each bucket by slot
                // and sums it, and, over the candidates a thread scores (in a
                // fixed order), accumulates the mean's partial sum, the clamp
                // count and the thread's kTop best (score, index).  Every pass
                // issues its independent loads together (the work area lives
                // in HBM, so each dependent load costs microseconds).
                const std::int64_t kPairCap = caps.pcap;
                for (int t = tid; t <= c; t += kBlk)
                    cnt[t] = 0;
                if (tid == 0)
                    s_np = 0;
                __syncthreads();
                for (int r = tid; r < m; r += kBlk) {
                    const int i = I[r];
                    const int e1 = bptr[i + 1];
                    for (int e0 = bptr[i]; e0 < e1; e0 += kPairB) {
                        int jb[kPairB], rk[kPairB];
                        unsigned wb[kPairB];
                        T vb[kPairB], db[kPairB];
#pragma unroll
                        for (int u = 0; u < kPairB; ++u)
                            jb[u] = bcol[min(e0 + u, e1 - 1)];
#pragma unroll
                        for (int u = 0; u < kPairB; ++u) {
                            wb[u] = bm[jb[u] >> 5];
                            vb[u] = bval[min(e0 + u, e1 - 1)];
                            db[u] = dscale[jb[u]];
                        }
                        int nk = 0;
#pragma unroll
                        for (int u = 0; u < kPairB; ++u) {
                            const bool ok = e0 + u < e1 && ((wb[u] >> (jb[u] & 31)) & 1u);
                            rk[u] = ok ? (jb[u] >> 5) - cw0 : -1;
                            nk += ok ? 1 : 0;
                        }
#pragma unroll
                        for (int u = 0; u < kPairB; ++u)
                            if (rk[u] >= 0)
                                rk[u] = wpre[rk[u]] +
                                        __popc(wb[u] & ((1u << (jb[u] & 31)) - 1u));
                        if (nk == 0)
                            continue;
                        int at = atomicAdd(&s_np, nk);
#pragma unroll
                        for (int u = 0; u < kPairB; ++u)
                            if (rk[u] >= 0) {
                                atomicAdd(cnt + rk[u], 1);
                                if (at < kPairCap) {
                                    sk[at] = rk[u];
                                    sr[at] = r;
                                    spv[at] = vb[u] * db[u];
                                }
                                ++at;
                            }
                    }
                }
                __syncthreads();
                const int np = s_np;
                if (np > kPairCap) {
                    bitmap_clear(bm, cw0, cw1);
                    status = kSpaiOverflow;
                    break;
                }
                // exclusive scan of the counts (each thread a run of candidates)
                {
                    const int per = (c + kBlk - 1) / kBlk;
                    const int a = tid * per, b = min(c, a + per);
                    int sum = 0;
                    for (int t = a; t < b; ++t)
                        sum += cnt[t];
                    int total;
                    int at = block_scan_excl(sum, sh, total);
                    __syncthreads();
                    for (int t = a; t < b; ++t) {
                        const int x = cnt[t];
                        cnt[t] = at;
                        clampf[t] = at; // fill cursor
                        at += x;
                    }
                    if (tid == 0)
                        cnt[c] = total;
                    __syncthreads();
                }
                // pass B: the staged pairs into their candidates' buckets
                for (int a0 = tid; a0 < np; a0 += kBlk * kPairB) {
                    int kb[kPairB], rb[kPairB], pb[kPairB];
                    T vb[kPairB];
#pragma unroll
                    for (int u = 0; u < kPairB; ++u) {
                        const int x = min(a0 + u * kBlk, np - 1);
                        kb[u] = sk[x];
                        rb[u] = sr[x];
                        vb[u] = spv[x];
                    }
#pragma unroll
                    for (int u = 0; u < kPairB; ++u)
                        pb[u] = a0 + u * kBlk < np ? atomicAdd(clampf + kb[u], 1) : -1;
#pragma unroll
                    for (int u = 0; u < kPairB; ++u)
                        if (pb[u] >= 0) {
                            pr[pb[u]] = rb[u];
                            pv[pb[u]] = vb[u];
                        }
                }
                __syncthreads();
                bitmap_clear(bm, cw0, cw1);
                // pass C, kCandB candidates of the thread at a time; buckets of
                // up to kPairS pairs are put in slot order in registers
                const T *rs = m <= kYLds ? s_y : res;
                for (int t0 = tid; t0 < c; t0 += kBlk * kCandB) {
                    int p0[kCandB], pn[kCandB], jc[kCandB];
#pragma unroll
                    for (int u = 0; u < kCandB; ++u) {
                        const int t = min(t0 + u * kBlk, c - 1);
                        p0[u] = cnt[t];
                        pn[u] = cnt[t + 1];
                        jc[u] = cand[t];
                    }
                    int rw[kCandB][kPairS];
                    T vw[kCandB][kPairS], xw[kCandB][kPairS];
#pragma unroll
                    for (int u = 0; u < kCandB; ++u) {
                        pn[u] -= p0[u];
#pragma unroll
                        for (int w = 0; w < kPairS; ++w) {
                            const bool ok = w < pn[u];
                            rw[u][w] = ok ? pr[p0[u] + w] : INT_MAX;
                            vw[u][w] = ok ? pv[p0[u] + w] : T(0);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kCandB; ++u) {
                        // sorting network on four (slot, value) pairs
                        auto cswap = [&](int x, int y) {
                            const bool sw = rw[u][x] > rw[u][y];
                            const int ra = rw[u][x], rb2 = rw[u][y];
                            const T va = vw[u][x], vb2 = vw[u][y];
                            rw[u][x] = sw ? rb2 : ra;
                            rw[u][y] = sw ? ra : rb2;
                            vw[u][x] = sw ? vb2 : va;
                            vw[u][y] = sw ? va : vb2;
                        };
                        static_assert(kPairS == 4, "four-input sorting network");
                        cswap(0, 1);
                        cswap(2, 3);
                        cswap(0, 2);
                        cswap(1, 3);
                        cswap(1, 2);
#pragma unroll
                        for (int w = 0; w < kPairS; ++w)
                            xw[u][w] = rw[u][w] != INT_MAX ? rs[rw[u][w]] : T(0);
                    }
#pragma unroll
                    for (int u = 0; u < kCandB; ++u) {
                        const int t = t0 + u * kBlk;
                        if (t >= c)
                            continue;
                        T dt = 0, den = 0;
                        if (pn[u] <= kPairS) {
#pragma unroll
                            for (int w = 0; w < kPairS; ++w)
                                if (w < pn[u]) {
                                    dt += xw[u][w] * vw[u][w];
                                    den += vw[u][w] * vw[u][w];
                                }
                        } else {
                            const int q0 = p0[u], q1 = p0[u] + pn[u];
                            for (int a = q0 + 1; a < q1; ++a) { // by slot
                                const int rr0 = pr[a];
                                const T vv0 = pv[a];
                                int b = a - 1;
                                while (b >= q0 && pr[b] > rr0) {
                                    pr[b + 1] = pr[b];
                                    pv[b + 1] = pv[b];
                                    --b;
                                }
                                pr[b + 1] = rr0;
                                pv[b + 1] = vv0;
                            }
                            for (int a = q0; a < q1; ++a) {
                                const T value = pv[a];
                                dt += rs[pr[a]] * value;
                                den += value * value;
                            }
                        }
                        T sv_ = 0;
                        bool cl = false;
                        const int st = score_of(rr, dt, den, sv_, cl);
                        sc[t] = sv_;
                        clampf[t] = st == kSpaiSuccess ? (cl ? 1 : 0) : -st;
                        if (st != kSpaiSuccess) {
                            atomicMin(&s_fail, t);
                            continue;
                        }
                        msum += (double)sv_;
                        mycl += cl ? 1 : 0;
                        // keep the thread's kTop smallest (score, index)
                        const int j = jc[u];
                        if (sv_ < tops[kTop - 1] || (sv_ == tops[kTop - 1] && j < topj[kTop - 1])) {
                            T ins = sv_;
                            int insj = j;
#pragma unroll
                            for (int w = 0; w < kTop; ++w) {
                                const bool lt = ins < tops[w] || (ins == tops[w] && insj < topj[w]);
                                const T os = tops[w];
                                const int oj = topj[w];
                                tops[w] = lt ? ins : os;
                                topj[w] = lt ? insj : oj;
                                ins = lt ? os : ins;
                                insj = lt ? oj : insj;
                            }
                        }
                    }
                }
            }
            __syncthreads();
            const int first_bad = s_fail;
            if (!kExact && first_bad == INT_MAX) {
                clamps += (int)block_sum_fixed((double)mycl, s_dred);
            } else {
                // clamps are counted up to the first failing candidate, as the
                // reference's sequential loop does
                const int lim = first_bad != INT_MAX ? first_bad : c;
                int cl = 0;
                for (int t = tid; t < lim; t += kBlk)
                    cl += clampf[t];
                sh[tid] = cl;
                __syncthreads();
                int total = 0;
                for (int t = 0; t < kBlk; ++t)
                    total += sh[t];
                clamps += total;
                __syncthreads();
            }
            if (first_bad != INT_MAX) {
                status = -clampf[first_bad];
                break;
            }
            T mean;
            if (kExact) { // std::accumulate order (src/spai.cpp:216-217)
                if (tid == 0) {
                    T acc = 0;
                    for (int t0 = 0; t0 < c; t0 += kBatch) {
                        T xb[kBatch];
#pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            xb[u] = sc[min(t0 + u, c - 1)];
#pragma unroll
                        for (int u = 0; u < kBatch; ++u)
                            if (t0 + u < c)
                                acc += xb[u];
                    }
                    s_scal[0] = acc / T(c);
                }
                __syncthreads();
                mean = s_scal[0];
            } else {
                mean = (T)(block_sum_fixed(msum, s_dred) / (double)c);
            }
            if (!isfinite(mean)) {
                status = kSpaiNumericalFailure;
                break;
            }
            // ---- select <= beta smallest (score, index) with score <= mean
            int nsel = 0;
            T prev_s = T(-INFINITY);
            int prev_j = -1;
            if (!kExact && beta <= kTop) {
                // (FP32) merge of the threads' sorted lists: each round takes
                // the smallest head (a wave minimum, then the four waves'), and
                // its thread drops it.  Rounds alternate between two halves of
                // s_red / s_redi, so one barrier per round suffices.
                for (; nsel < beta; ++nsel) {
                    T bs = tops[0] <= mean ? tops[0] : T(INFINITY);
                    int bj = tops[0] <= mean ? topj[0] : INT_MAX;
#pragma unroll
                    for (int off = kWave / 2; off > 0; off >>= 1) {
                        const T os = __shfl_xor(bs, off, kWave);
                        const int oj = __shfl_xor(bj, off, kWave);
                        const bool lt = os < bs || (os == bs && oj < bj);
                        bs = lt ? os : bs;
                        bj = lt ? oj : bj;
                    }
                    const int half = (nsel & 1) * kNW;
                    if (lane == 0) {
                        s_red[half + wave] = bs;
                        s_redi[half + wave] = bj;
                    }
                    __syncthreads();
                    bs = s_red[half];
                    bj = s_redi[half];
#pragma unroll
                    for (int w = 1; w < kNW; ++w) {
                        const T os = s_red[half + w];
                        const int oj = s_redi[half + w];
                        const bool lt = os < bs || (os == bs && oj < bj);
                        bs = lt ? os : bs;
                        bj = lt ? oj : bj;
                    }
                    if (bj == INT_MAX)
                        break;
                    if (q + nsel + 1 > Q) { // the grown pattern would not fit
                        nsel = -1;
                        break;
                    }
                    if (topj[0] == bj) {
#pragma unroll
                        for (int w = 0; w + 1 < kTop; ++w) {
                            tops[w] = tops[w + 1];
                            topj[w] = topj[w + 1];
                        }
                        tops[kTop - 1] = T(INFINITY);
                        topj[kTop - 1] = INT_MAX;
                    }
                    if (tid == 0)
                        sel[nsel] = bj;
                }
                __syncthreads();
            } else
            for (; nsel < beta; ++nsel) {
                T best_s = T(INFINITY);
                int best_j = -1;
                for (int t0 = tid; t0 < c; t0 += kBlk * kRowB) {
                    T sb[kRowB];
                    int jb[kRowB];
#pragma unroll
                    for (int u = 0; u < kRowB; ++u) {
                        const int t = min(t0 + u * kBlk, c - 1);
                        sb[u] = sc[t];
                        jb[u] = cand[t];
                    }
#pragma unroll
                    for (int u = 0; u < kRowB; ++u) {
                        const T sv = sb[u];
                        const int j = jb[u];
                        if (t0 + u * kBlk >= c || sv > mean)
                            continue;
                        const bool after = sv > prev_s || (sv == prev_s && j > prev_j);
                        const bool before =
                            best_j < 0 || sv < best_s || (sv == best_s && j < best_j);
                        if (after && before) {
                            best_s = sv;
                            best_j = j;
                        }
                    }
                }
                s_red[tid] = best_s;
                s_redi[tid] = best_j;
                __syncthreads();
                for (int off = kBlk / 2; off > 0; off >>= 1) {
                    if (tid < off) {
                        const T os = s_red[tid + off];
                        const int oj = s_redi[tid + off];
                        const T ms = s_red[tid];
                        const int mj = s_redi[tid];
                        if (oj >= 0 && (mj < 0 || os < ms || (os == ms && oj < mj))) {
                            s_red[tid] = os;
                            s_redi[tid] = oj;
                        }
                    }
                    __syncthreads();
                }
                best_s = s_red[0];
                best_j = s_redi[0];
                __syncthreads();
                if (best_j < 0)
                    break;
                if (q + nsel + 1 > Q) { // the grown pattern would not fit
                    nsel = -1;
                    break;
                }
                if (tid == 0)
                    sel[nsel] = best_j;
                prev_s = best_s;
                prev_j = best_j;
            }
            if (nsel < 0) {
                status = kSpaiOverflow;
                break;
            }
            if (nsel == 0) {
                status = kSpaiNoGrowth;
                break;
            }
            // sort the selection by index, then merge into J
            if (tid == 0) {
                for (int a = 1; a < nsel; ++a) {
                    const int x = sel[a];
                    int b = a - 1;
                    while (b >= 0 && sel[b] > x) {
                        sel[b + 1] = sel[b];
                        --b;
                    }
                    sel[b + 1] = x;
                }
                int a = 0, b = 0, t = 0;
                while (a < q || b < nsel) {
                    if (b >= nsel || (a < q && J[a] < sel[b]))
                        Jt[t++] = J[a++];
                    else
                        Jt[t++] = sel[b++];
                }
                for (int r = 0; r < t; ++r)
                    J[r] = Jt[r];
                if (!kExact)
                    for (int r = 0; r < nsel; ++r)
                        Jf[q + r] = sel[r];
            }
            q += nsel;
            __syncthreads();
        }
        __syncthreads();
        for (int r = tid; r < mapped; r += kBlk) {
            pos[I[r]] = -1;
            atomicAnd(inI + (I[r] >> 5), ~(1u << (I[r] & 31)));
        }
        mapped = 0;
        __syncthreads();
        if (tid == 0) {
            out.status[k] = status;
            out.fits[k] = fits;
            out.growths[k] = growths;
            out.clamps[k] = clamps;
            out.q[k] = status == kSpaiOverflow ? 0 : qdone;
            out.residual[k] = last_res;
            out.pivot[k] = last_piv;
            out.flops[k] = qr_flops;
            out.off[k] = (std::int64_t)slot * out_stride;
            out.pass[k] = out.pass_id;
        }
        __syncthreads();
    }
}

template <typename T>
void launch_spai_block(hipStream_t s, int blocks, int n, const int *aptr, const int *acol,
                       const T *aval, const T *dscale, const int *bptr, const int *bcol,
                       const T *bval, T eps, long long alpha, int beta, const int *worklist,
                       int nwork, int *cursor, int mcap, int qcap, int ccap, T *dwork,
                       int *iwork, SpaiColumnOut out, int *out_j, T *out_c,
                       std::int64_t out_stride, int occupancy, std::int64_t pcap) {
    const int words = (n + 31) / 32;
    BlockCaps caps{mcap, qcap, ccap, words, pcap};
    auto kernel = occupancy >= 4 ? spai_block_kernel<T, 4> : spai_block_kernel<T, 3>;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlk), 0, s, n, aptr, acol, aval,
                       dscale, bptr, bcol, bval, eps, alpha, beta, worklist, nwork, cursor, caps,
                       dwork,
                       iwork, spai_block_doubles(mcap, qcap, ccap, pcap),
                       spai_block_ints(mcap, qcap, ccap, pcap, words, n), out, out_j, out_c,
                       out_stride);
    HIP_LAUNCH_CHECK();
}

__global__ void fill_int_kernel(std::int64_t stride, std::int64_t offset, int count, int *p,
                                int value) {
    int *q = p + (std::int64_t)blockIdx.y * stride + offset;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        q[i] = value;
}
void launch_fill_int(hipStream_t s, std::int64_t blocks, std::int64_t stride, std::int64_t offset,
                     int count, int *p, int value) {
    if (blocks == 0 || count == 0)
        return;
    hipLaunchKernelGGL(fill_int_kernel, dim3(std::min<unsigned>(ceil_div(count, 256), 1024u),
                                             (unsigned)blocks),
                       dim3(256), 0, s, stride, offset, count, p, value);
    HIP_LAUNCH_CHECK();
}

#define BSPAI_BLOCK_INST(T)                                                                       \
    template void launch_spai_block<T>(hipStream_t, int, int, const int *, const int *,          \
                                       const T *, const T *, const int *, const int *, const T *, \
                                       T,                                                        \
                                       long long, int, const int *, int, int *, int, int, int,    \
                                       T *, int *, SpaiColumnOut, int *, T *, std::int64_t, int,  \
                                       std::int64_t);
BSPAI_BLOCK_INST(double)
BSPAI_BLOCK_INST(float)
#undef BSPAI_BLOCK_INST

// Single score on the device (unit-test hook for the restricted score).
__global__ void score_kernel(double rr, double dot, double den, int *out_i, double *out_d) {
    double sc = 0;
    bool cl = false;
    out_i[0] = score_of<double>(rr, dot, den, sc, cl);
    out_i[1] = cl;
    out_d[0] = sc;
}

void spai_score_on_device(hipStream_t s, double rr, double dot, double den, int *status,
                          int *clamped, double *score) {
    int *di = nullptr;
    double *dd = nullptr;
    HIP_CHECK(hipMalloc(&di, 2 * sizeof(int)));
    HIP_CHECK(hipMalloc(&dd, sizeof(double)));
    hipLaunchKernelGGL(score_kernel, dim3(1), dim3(1), 0, s, rr, dot, den, di, dd);
    HIP_LAUNCH_CHECK();
    int hi[2];
    HIP_CHECK(hipMemcpyAsync(hi, di, sizeof hi, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(score, dd, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(di);
    (void)hipFree(dd);
    *status = hi[0];
    *clamped = hi[1];
}

// ------------------------------------------------------------------ scaling
// Row scaling d_i = 1/max_j |a_ij| with the reference's validity checks
// (src/spai.cpp:322-348).  row_err[i]: 0 ok, 1 zero diagonal, 2 scaling out of
// range, 3 scaled entry zero/nonfinite.
template <typename T>
__global__ void scale_rows_kernel(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                                  const T *__restrict__ val, T *__restrict__ d,
                                  int *__restrict__ first_bad) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n)
        return;
    bool diag = false;
    T mx = 0;
    for (int e = ptr[row]; e < ptr[row + 1]; ++e) {
        diag = diag || col[e] == row;
        const T a = fabs(val[e]);
        mx = a > mx ? a : mx;
    }
    int err = 0;
    T s = 0;
    if (!diag)
        err = 1;
    else {
        s = T(1) / mx;
        if (!isfinite(s) || s == 0)
            err = 2;
    }
    if (!err)
        for (int e = ptr[row]; e < ptr[row + 1]; ++e) {
            const T v = val[e] * s;
            if (!isfinite(v) || v == 0)
                err = 3;
        }
    d[row] = s;
    // first_bad[0] = min(4*row + kind) over kinds 1/2 (checked row by row in
    // the reference); first_bad[1] = min row of kind 3 (checked after all rows).
    if (err == 3)
        atomicMin(first_bad + 1, row);
    else if (err)
        atomicMin(first_bad, row * 4 + err);
}

template <typename T>
void launch_scale_rows(hipStream_t s, int n, const int *ptr, const int *col, const T *val, T *d,
                       int *first_bad) {
    hipLaunchKernelGGL(scale_rows_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, ptr,
                       col, val, d, first_bad);
    HIP_LAUNCH_CHECK();
}
template void launch_scale_rows<double>(hipStream_t, int, const int *, const int *,
                                        const double *, double *, int *);
template void launch_scale_rows<float>(hipStream_t, int, const int *, const int *, const float *,
                                       float *, int *);

// ------------------------------------------------------------------ transpose structure
__global__ void count_cols_kernel(std::int64_t nnz, const int *__restrict__ col,
                                  int *__restrict__ counts) {
    const std::int64_t e = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nnz)
        atomicAdd(counts + col[e] + 1, 1);
}

template <typename T>
__global__ void transpose_fill_kernel(int n, const int *__restrict__ ptr,
                                      const int *__restrict__ col, const T *__restrict__ val,
                                      int *__restrict__ cursor, int *__restrict__ tcol,
                                      T *__restrict__ tval) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n)
        return;
    for (int e = ptr[row]; e < ptr[row + 1]; ++e) {
        const int pos = atomicAdd(cursor + col[e], 1);
        tcol[pos] = row;
        if (tval)
            tval[pos] = val[e];
    }
}

// Insertion-sort each row segment by column (rows of a transpose are short).
template <typename T>
__global__ void sort_rows_kernel(int n, const int *__restrict__ ptr, int *__restrict__ col,
                                 T *__restrict__ val) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n)
        return;
    const int b = ptr[row], e = ptr[row + 1];
    for (int a = b + 1; a < e; ++a) {
        const int x = col[a];
        const T xv = val ? val[a] : T(0);
        int c = a - 1;
        while (c >= b && col[c] > x) {
            col[c + 1] = col[c];
            if (val)
                val[c + 1] = val[c];
            --c;
        }
        col[c + 1] = x;
        if (val)
            val[c + 1] = xv;
    }
}

void launch_count_cols(hipStream_t s, std::int64_t nnz, const int *col, int *counts) {
    if (nnz == 0)
        return;
    hipLaunchKernelGGL(count_cols_kernel, dim3(ceil_div(nnz, 256)), dim3(256), 0, s, nnz, col,
                       counts);
    HIP_LAUNCH_CHECK();
}
// cursor must hold a copy of tptr[0..n-1] on entry.
template <typename T>
void launch_transpose_fill(hipStream_t s, int n, const int *ptr, const int *col, const T *val,
                           const int *tptr, int *cursor, int *tcol, T *tval) {
    hipLaunchKernelGGL(transpose_fill_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, ptr,
                       col, val, cursor, tcol, tval);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(sort_rows_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, tptr,
                       tcol, tval);
    HIP_LAUNCH_CHECK();
}
template void launch_transpose_fill<double>(hipStream_t, int, const int *, const int *,
                                            const double *, const int *, int *, int *, double *);
template void launch_transpose_fill<float>(hipStream_t, int, const int *, const int *,
                                           const float *, const int *, int *, int *, float *);

} // namespace bspai_hip
