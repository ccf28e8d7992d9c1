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

constexpr int kFuse = 2; // trailing columns reflected together (register budget)

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
template <typename T>
__device__ T householder_vector(const Lane<T> &L, int ov, int len, T length) {
    const T diagonal = -copysign(length, L.D(ov));
    T vnorm = 0;
    for (int r = 0; r < len; ++r) {
        T x = L.D(ov + r) / length;
        if (r == 0)
            x += copysign(T(1), x);
        L.D(ov + r) = x;
        vnorm = hypot_ref(vnorm, x);
    }
    for (int r = 0; r < len; ++r)
        L.D(ov + r) = L.D(ov + r) / vnorm;
    return diagonal;
}

} // namespace

// Elements per lane for the given caps.
std::int64_t spai_lane_doubles(int mcap, int qcap, int ccap) {
    return (std::int64_t)mcap * qcap + 3LL * mcap + ccap + qcap;
}
std::int64_t spai_lane_ints(int mcap, int qcap, int ccap) {
    return 2LL * mcap + 2LL * qcap + 2LL * ccap + qcap;
}

template <typename T>
__global__ void __launch_bounds__(256)
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
            T length = 0;
            for (int r = 0; r < m; ++r) {
                const T x = L.D(oR + r);
                L.D(oVc + r) = x;
                length = hypot_ref(length, x);
            }
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
                    for (int r = 1; r < len; ++r) {
                        T &x = L.D(base + r);
                        x -= T(2) * L.D(oVc + r) * dt;
                        ok = ok && isfinite(x);
                        L.D(oVn + r - 1) = x;
                        nlength = hypot_ref(nlength, x);
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
            T rnorm = 0;
            bool res_finite = true;
            for (int r = 0; r < m; ++r) {
                const T x = L.D(oRes + r);
                res_finite = res_finite && isfinite(x);
                rnorm = hypot_ref(rnorm, x);
            }
            if (!res_finite || !isfinite(rnorm)) {
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
