// Left-preconditioned GMRES and iterative-refinement kernels for gfx950.
//
// Algorithm (reference src/gmres.cpp:108-262, src/ir.cpp:83-204): unrestarted
// GMRES on P A d = P r with one-pass modified Gram-Schmidt, Givens rotations,
// and a candidate x0 + V y that is formed and re-verified (true residual and
// recomputed preconditioned residual) exactly on the iterations whose
// acceptance test can pass: estimate <= tau, Arnoldi breakdown, zero Givens
// pivot, or the last iteration.
//
// Layout: all vectors are concatenations of independent systems; each kernel
// runs one 256-thread block per row tile (<= 2048 rows, inside one system) and
// skips tiles of systems that are not active, so one launch advances every
// system of a batch.  Per-system reductions are deterministic: each tile
// writes its partial, the last tile of the system to finish (atomic ticket)
// sums the partials in tile order and runs the per-system scalar step (norm,
// Hessenberg entry, Givens update, back substitution, accept decision) on one
// thread in FP64.  No host round trip happens inside an iteration.
//
// Precision: T is the Krylov vector/matrix type (double for the plain path,
// float for the mixed path).  Hessenberg, rotations, y and the candidate are
// FP64 in both.
#include "common.hpp"
#include "krylov.hpp"

namespace bspai_hip {

// --------------------------------------------------------------- reductions
// Robust 2-norm partials: per tile (e, s) with s = sum (x * 2^-e)^2, e the
// exponent of the tile's max |x|; combined relative to the largest e.
struct NormPart {
    double e, s;
};

__device__ inline double combine_norm(const double *part, int t0, int t1) {
    double emax = -1e300;
    for (int t = t0; t < t1; ++t)
        if (part[2 * t + 1] > 0)
            emax = fmax(emax, part[2 * t]);
    if (emax == -1e300)
        return 0.0;
    double s = 0;
    for (int t = t0; t < t1; ++t)
        if (part[2 * t + 1] > 0)
            s += ldexp(part[2 * t + 1], 2 * (int)(part[2 * t] - emax));
    return ldexp(sqrt(s), (int)emax);
}

// Tile-level robust sum of squares of the kTilePer values v[] held by each
// thread; writes the tile's (e, s) to part[2t..2t+1].
__device__ inline void tile_norm_partial(const double *v, int cnt, double *part, int t,
                                         double *lds) {
    double mx = 0;
    for (int i = 0; i < cnt; ++i)
        mx = fmax(mx, fabs(v[i]));
    mx = block_max(mx, lds);
    double s = 0;
    int e = 0;
    if (mx > 0 && isfinite(mx)) {
        frexp(mx, &e);
        for (int i = 0; i < cnt; ++i) {
            const double y = ldexp(v[i], -e);
            s += y * y;
        }
    } else if (!isfinite(mx)) {
        s = mx; // propagate inf/nan
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        part[2 * t] = (double)e;
        part[2 * t + 1] = s;
    }
}

// Returns true on the one thread (thread 0 of the system's last tile) that
// must finalize system s.  Call after the tile's partial is stored.
__device__ inline bool last_tile(int *cnt, int s, int ntiles_s) {
    __shared__ int is_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const int ticket = atomicAdd(cnt + s, 1);
        is_last = ticket == ntiles_s - 1;
        if (is_last) {
            cnt[s] = 0; // reset for the next reduction
            __threadfence();
        }
    }
    __syncthreads();
    return is_last && threadIdx.x == 0;
}

// ------------------------------------------------------------------ GMRES
// Per-system scalar step after h_{k+1,k} is known (src/gmres.cpp:183-254).
__device__ void givens_step(SysState &st, double *H, double *cs, double *sn, double *g, double *y,
                            int k, int cap, double hnorm) {
    double *col = H + hcol_offset(k);
    col[k + 1] = hnorm;
    st.hnorm = hnorm;
    const bool breakdown = hnorm == 0;
    bool bad = !isfinite(hnorm);
    for (int r = 0; r < k; ++r) {
        const double up = col[r], lo = col[r + 1];
        col[r] = cs[r] * up + sn[r] * lo;
        col[r + 1] = -sn[r] * up + cs[r] * lo;
        bad = bad || !isfinite(col[r]) || !isfinite(col[r + 1]);
    }
    const double upper = col[k];
    const double radius = hypot_ref(upper, col[k + 1]);
    bad = bad || !isfinite(radius);
    st.iterations = k + 1;
    st.breakdown = breakdown;
    int count;
    if (radius != 0) {
        cs[k] = upper / radius;
        sn[k] = col[k + 1] / radius;
        col[k] = radius;
        col[k + 1] = 0;
        const double old = g[k];
        g[k] = cs[k] * old;
        g[k + 1] = -sn[k] * old;
        st.est = fabs(g[k + 1]) / st.den;
        bad = bad || !isfinite(cs[k]) || !isfinite(sn[k]) || !isfinite(g[k]) ||
              !isfinite(g[k + 1]) || !isfinite(st.est);
        st.zero_pivot = 0;
        st.triggered = breakdown || st.est <= st.tol || k + 1 == cap;
        count = k + 1;
    } else {
        // Zero pivot: re-verify the previous candidate (or x0), never a new y.
        st.zero_pivot = 1;
        st.triggered = 1;
        count = k;
    }
    if (bad) {
        st.status = kGmNumericalFailure;
        st.active = 0;
        st.triggered = 0;
        return;
    }
    st.ccount = count;
    if (st.triggered) {
        for (int row = count - 1; row >= 0; --row) {
            double value = g[row];
            for (int c = row + 1; c < count; ++c)
                value = value - H[hcol_offset(c) + row] * y[c];
            y[row] = value / H[hcol_offset(row) + row];
            if (!isfinite(y[row])) {
                st.status = kGmNumericalFailure;
                st.active = 0;
                st.triggered = 0;
                return;
            }
        }
    }
}

// One MGS round of iteration k (j = 0..k+1):
//   j > 0:    w -= h_{j-1,k} V_{j-1}
//   j <= k:   h_{j,k} = <V_j, w>
//   j == k+1: h_{k+1,k} = ||w||, then the Givens step
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
mgs_round_kernel(Tiles tiles, KrylovArgs<T> a, int j, int k) {
    __shared__ double lds[kTileThreads / kWave];
    const int t = blockIdx.x, s = tiles.sys[t];
    SysState &st = a.st[s];
    if (!st.active)
        return;
    const int r0 = tiles.row0[t], r1 = tiles.row1[t];
    T *w = a.w;
    double *H = a.H + (std::int64_t)s * a.hstride;
    const T *vprev = j > 0 ? a.V[j - 1] : nullptr;
    const T *vcur = j <= k ? a.V[j] : nullptr;
    const T hprev = j > 0 ? (T)H[hcol_offset(k) + j - 1] : T(0);
    double vals[kTilePer];
    int cnt = 0;
    double dot = 0;
    for (int r = r0 + threadIdx.x; r < r1; r += kTileThreads) {
        T x = w[r];
        if (j > 0) {
            x = x - hprev * vprev[r];
            w[r] = x;
        }
        if (j <= k)
            dot += (double)vcur[r] * (double)x;
        else
            vals[cnt++] = (double)x;
    }
    if (j <= k) {
        dot = block_sum(dot, lds);
        if (threadIdx.x == 0)
            a.part[2 * t + 1] = dot;
    } else {
        tile_norm_partial(vals, cnt, a.part, t, lds);
    }
    if (!last_tile(a.cnt, s, tiles.sys_tile1[s] - tiles.sys_tile0[s]))
        return;
    __threadfence();
    const int t0 = tiles.sys_tile0[s], t1 = tiles.sys_tile1[s];
    if (j <= k) {
        double h = 0;
        for (int u = t0; u < t1; ++u)
            h += a.part[2 * u + 1];
        H[hcol_offset(k) + j] = h;
        if (!isfinite(h)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
        }
    } else {
        const double hn = combine_norm(a.part, t0, t1);
        givens_step(st, H, a.cs + (std::int64_t)s * a.cap, a.sn + (std::int64_t)s * a.cap,
                    a.g + (std::int64_t)s * (a.cap + 1), a.y + (std::int64_t)s * a.cap, k, a.cap,
                    hn);
    }
}

// V_{k+1} = w / h_{k+1,k} (skipped on breakdown).
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
normalize_kernel(Tiles tiles, KrylovArgs<T> a, int k) {
    const int t = blockIdx.x, s = tiles.sys[t];
    const SysState &st = a.st[s];
    if (!st.active || st.breakdown || st.iterations != k + 1)
        return;
    const double *H = a.H + (std::int64_t)s * a.hstride;
    // The unrotated h_{k+1,k} was overwritten by the rotation; it is kept in hnorm.
    const T h = (T)st.hnorm;
    T *v = a.V[k + 1];
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads)
        v[r] = a.w[r] / h;
    (void)H;
}

// Candidate x = x0 + sum_{j<count} y_j V_j (FP64, j ascending, src/gmres.cpp:60-63).
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
candidate_kernel(Tiles tiles, KrylovArgs<T> a) {
    const int t = blockIdx.x, s = tiles.sys[t];
    const SysState &st = a.st[s];
    if (!st.active || !st.triggered)
        return;
    const double *y = a.y + (std::int64_t)s * a.cap;
    const int count = st.ccount;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads) {
        double x = a.x0 ? a.x0[r] : 0.0;
        for (int j = 0; j < count; ++j)
            x = x + (double)a.V[j][r] * y[j];
        a.x[r] = x;
    }
}

// Norm of (u - v) over active systems (flag chosen by `gate`), result used by
// the per-system step `what` (see krylov.hpp: NormWhat).
template <typename TU, typename TV, typename T>
__global__ void __launch_bounds__(kTileThreads)
diff_norm_kernel(Tiles tiles, KrylovArgs<T> a, const TU *u, const TV *v, double *out, int what) {
    __shared__ double lds[kTileThreads / kWave];
    const int t = blockIdx.x, s = tiles.sys[t];
    SysState &st = a.st[s];
    const bool gate = what == kNormCheckOrig || what == kNormCheckRel ? (st.active && st.triggered)
                                                                       : st.active;
    if (!gate)
        return;
    double vals[kTilePer];
    int cnt = 0;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads) {
        const double d = (double)u[r] - (v ? (double)v[r] : 0.0);
        vals[cnt++] = d;
        if (out)
            out[r] = d;
    }
    tile_norm_partial(vals, cnt, a.part, t, lds);
    if (!last_tile(a.cnt, s, tiles.sys_tile1[s] - tiles.sys_tile0[s]))
        return;
    __threadfence();
    const double nrm = combine_norm(a.part, tiles.sys_tile0[s], tiles.sys_tile1[s]);
    if (!isfinite(nrm)) {
        st.status = kGmNumericalFailure;
        st.active = 0;
        st.triggered = 0;
        return;
    }
    switch (what) {
    case kNormInitOrig: // ||rhs - A x0||
        st.init_orig = st.final_orig = nrm;
        if (nrm == 0) {
            st.status = st.rhs_zero ? kGmZeroRhs : kGmExactInitial;
            st.converged = 1;
            st.init_rel = st.final_rel = 0;
            st.active = 0;
        }
        break;
    case kNormRhs:
        st.rhs_zero = nrm == 0;
        break;
    case kNormDen: // ||P rhs||
        st.den = nrm;
        if (nrm == 0) {
            st.status = kGmZeroPreconditionedRhs;
            st.active = 0;
        }
        break;
    case kNormBeta: { // ||P rhs - P A x0||
        st.beta = nrm;
        st.hnorm = nrm;
        st.init_rel = st.final_rel = nrm / st.den;
        a.g[(std::int64_t)s * (a.cap + 1)] = nrm;
        if (!isfinite(st.init_rel)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
        } else if (st.init_rel <= st.tol) {
            st.status = kGmSuccess;
            st.converged = 1;
            st.active = 0;
            st.x0_is_solution = 1;
        }
        break;
    }
    case kNormCheckOrig:
        st.final_orig = nrm;
        break;
    case kNormCheckRel: { // accept decision (src/gmres.cpp:218-253)
        const double rel = nrm / st.den;
        st.final_rel = rel;
        st.triggered = 0;
        if (!isfinite(rel)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
            break;
        }
        const bool ok = rel <= st.tol;
        if (st.zero_pivot) {
            st.status = ok ? kGmSuccess : kGmBreakdown;
            st.converged = ok;
            st.active = 0;
        } else if (ok) {
            st.status = kGmSuccess;
            st.converged = 1;
            st.active = 0;
        } else if (st.breakdown) {
            st.status = kGmBreakdown;
            st.active = 0;
        } else if (st.iterations == a.cap) {
            st.status = a.cap < st.nrows ? kGmIterationLimit : kGmNonconvergence;
            st.active = 0;
        }
        st.has_solution = 1;
        break;
    }
    default:
        break;
    }
}

// Normalize the start vector: V_0 = r0 / beta.
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
start_vector_kernel(Tiles tiles, KrylovArgs<T> a, const T *r0) {
    const int t = blockIdx.x, s = tiles.sys[t];
    const SysState &st = a.st[s];
    if (!st.active)
        return;
    const T beta = (T)st.beta;
    T *v = a.V[0];
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads)
        v[r] = r0[r] / beta;
}

// Systems whose GMRES finished at x0 (init_rel <= tau, zero rhs) or never
// formed a candidate return x0.
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
finish_kernel(Tiles tiles, KrylovArgs<T> a) {
    const int t = blockIdx.x, s = tiles.sys[t];
    const SysState &st = a.st[s];
    if (st.has_solution || !st.participating)
        return;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads)
        a.x[r] = a.x0 ? a.x0[r] : 0.0;
}

// ------------------------------------------------------------------ launchers
template <typename T> void gmres_mgs_round(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int j, int k) {
    hipLaunchKernelGGL(mgs_round_kernel<T>, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, a, j, k);
    HIP_LAUNCH_CHECK();
}
template <typename T> void gmres_normalize(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k) {
    hipLaunchKernelGGL(normalize_kernel<T>, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, a, k);
    HIP_LAUNCH_CHECK();
}
template <typename T> void gmres_candidate(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a) {
    hipLaunchKernelGGL(candidate_kernel<T>, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, a);
    HIP_LAUNCH_CHECK();
}
template <typename T> void gmres_start_vector(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, const T *r0) {
    hipLaunchKernelGGL(start_vector_kernel<T>, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, a, r0);
    HIP_LAUNCH_CHECK();
}
template <typename T> void gmres_finish(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a) {
    hipLaunchKernelGGL(finish_kernel<T>, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, a);
    HIP_LAUNCH_CHECK();
}
template <typename TU, typename TV, typename T>
void gmres_diff_norm(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, const TU *u,
                     const TV *v, double *out, int what) {
    hipLaunchKernelGGL((diff_norm_kernel<TU, TV, T>), dim3(tl.ntiles), dim3(kTileThreads), 0, s,
                       tl, a, u, v, out, what);
    HIP_LAUNCH_CHECK();
}

#define BSPAI_KRYLOV_INST(T)                                                                      \
    template void gmres_mgs_round<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int, int); \
    template void gmres_normalize<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int);     \
    template void gmres_candidate<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &);          \
    template void gmres_start_vector<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &,        \
                                        const T *);                                               \
    template void gmres_finish<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &);             \
    template void gmres_diff_norm<T, T, T>(hipStream_t, const Tiles &, const KrylovArgs<T> &,     \
                                           const T *, const T *, double *, int);
BSPAI_KRYLOV_INST(double)
BSPAI_KRYLOV_INST(float)
#undef BSPAI_KRYLOV_INST
template void gmres_diff_norm<float, double, float>(hipStream_t, const Tiles &,
                                                    const KrylovArgs<float> &, const float *,
                                                    const double *, double *, int);

// ================================================================== IR (FP64)
// ||A||_inf per system: max row sum of |a_ij| (src/ir.cpp:118-125).
__global__ void __launch_bounds__(kTileThreads)
ainf_kernel(Tiles tiles, const int *__restrict__ ptr, const double *__restrict__ val,
            IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    const int t = blockIdx.x, s = tiles.sys[t];
    double mx = 0;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads) {
        double rs = 0;
        for (int e = ptr[r]; e < ptr[r + 1]; ++e)
            rs = rs + fabs(val[e]);
        mx = fmax(mx, rs);
    }
    mx = block_max(mx, lds);
    if (threadIdx.x == 0)
        a.part[4 * t] = mx;
    if (!last_tile(a.cnt, s, tiles.sys_tile1[s] - tiles.sys_tile0[s]))
        return;
    __threadfence();
    double m = 0;
    for (int u = tiles.sys_tile0[s]; u < tiles.sys_tile1[s]; ++u)
        m = fmax(m, a.part[4 * u]);
    a.st[s].ainf = m;
}

// r = b - A x with the reference's sequential row sums (bit-identical), and
// the diagnostics of src/ir.cpp:83-133: ||r||_inf, ||r||_2, normwise and
// componentwise backward errors.  Finalizes the outer-loop decision.
__global__ void __launch_bounds__(kTileThreads)
diagnostics_kernel(Tiles tiles, const int *__restrict__ ptr, const int *__restrict__ col,
                   const double *__restrict__ val, const double *__restrict__ b,
                   const double *__restrict__ x, double *__restrict__ r, IrArgs a, int step) {
    __shared__ double lds[kTileThreads / kWave];
    const int t = blockIdx.x, s = tiles.sys[t];
    IrSysState &st = a.st[s];
    if (!st.active)
        return;
    double rinf = 0, xinf = 0, binf = 0, cbe = 0;
    double vals[kTilePer];
    int cnt = 0;
    bool bad = false;
    for (int row = tiles.row0[t] + threadIdx.x; row < tiles.row1[t]; row += kTileThreads) {
        double ax = 0, cd = fabs(b[row]);
        for (int e = ptr[row]; e < ptr[row + 1]; ++e) {
            const double xv = x[col[e]];
            ax += val[e] * xv;
            cd = cd + fabs(val[e]) * fabs(xv);
        }
        const double res = b[row] - ax;
        bad = bad || !isfinite(ax) || !isfinite(res) || !isfinite(cd);
        r[row] = res;
        vals[cnt++] = res;
        rinf = fmax(rinf, fabs(res));
        xinf = fmax(xinf, fabs(x[row]));
        binf = fmax(binf, fabs(b[row]));
        const double ratio = cd == 0 ? (res == 0 ? 0.0 : INFINITY) : fabs(res) / cd;
        cbe = fmax(cbe, ratio);
    }
    rinf = block_max(rinf, lds);
    xinf = block_max(xinf, lds);
    binf = block_max(binf, lds);
    cbe = block_max(cbe, lds);
    const double badv = block_max(bad ? 1.0 : 0.0, lds);
    double *p4 = a.part + 4 * (std::int64_t)t;
    if (threadIdx.x == 0) {
        p4[0] = rinf;
        p4[1] = xinf;
        p4[2] = binf;
        p4[3] = cbe;
        a.part2[2 * t + 0] = badv;
    }
    tile_norm_partial(vals, cnt, a.part3, t, lds);
    if (!last_tile(a.cnt, s, tiles.sys_tile1[s] - tiles.sys_tile0[s]))
        return;
    __threadfence();
    double R = 0, X = 0, B = 0, C = 0, BAD = 0;
    for (int u = tiles.sys_tile0[s]; u < tiles.sys_tile1[s]; ++u) {
        R = fmax(R, a.part[4 * u]);
        X = fmax(X, a.part[4 * u + 1]);
        B = fmax(B, a.part[4 * u + 2]);
        C = fmax(C, a.part[4 * u + 3]);
        BAD = fmax(BAD, a.part2[2 * u]);
    }
    const double r2 = combine_norm(a.part3, tiles.sys_tile0[s], tiles.sys_tile1[s]);
    const double den = st.ainf * X + B;
    const double nbe = den == 0 ? (R == 0 ? 0.0 : INFINITY) : R / den;
    st.rinf = R;
    st.r2 = r2;
    st.nbe = nbe;
    st.cbe = C;
    st.xinf = X;
    if (BAD != 0 || !isfinite(den) || !isfinite(r2)) {
        st.status = kIrNumericalFailure;
        st.active = 0;
        st.done = 1;
        return;
    }
    if (step < kMaxOuterHistory) {
        double *h = a.history + ((std::int64_t)s * kMaxOuterHistory + step) * 4;
        h[0] = R;
        h[1] = r2;
        h[2] = nbe;
        h[3] = C;
    }
    st.history_len = step + 1;
    const bool conv = nbe <= a.be_tol && (a.relres_target <= 0 || r2 <= a.relres_target * st.bnorm);
    if (conv) {
        st.status = kIrSuccess;
        st.converged = 1;
        st.active = 0;
        st.done = 1;
    } else if (st.unchanged) {
        st.status = kIrStagnation;
        st.active = 0;
        st.done = 1;
    }
}

// ||b||_2 per system (robust), for the relative-residual stop.
__global__ void __launch_bounds__(kTileThreads)
bnorm_kernel(Tiles tiles, const double *__restrict__ b, IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    const int t = blockIdx.x, s = tiles.sys[t];
    double vals[kTilePer];
    int cnt = 0;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads)
        vals[cnt++] = b[r];
    tile_norm_partial(vals, cnt, a.part3, t, lds);
    if (!last_tile(a.cnt, s, tiles.sys_tile1[s] - tiles.sys_tile0[s]))
        return;
    __threadfence();
    a.st[s].bnorm = combine_norm(a.part3, tiles.sys_tile0[s], tiles.sys_tile1[s]);
}

// x += d for systems whose inner solve converged; records whether x changed.
__global__ void __launch_bounds__(kTileThreads)
update_kernel(Tiles tiles, double *__restrict__ x, const double *__restrict__ d, IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    const int t = blockIdx.x, s = tiles.sys[t];
    IrSysState &st = a.st[s];
    if (!st.active)
        return;
    double changed = 0;
    bool bad = false;
    const double scale = st.dscale;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads) {
        const double dv = scale == 1.0 ? d[r] : d[r] * scale;
        const double nx = x[r] + dv;
        bad = bad || !isfinite(nx);
        changed = nx != x[r] ? 1.0 : changed;
        x[r] = nx;
    }
    changed = block_max(changed, lds);
    const double badv = block_max(bad ? 1.0 : 0.0, lds);
    if (threadIdx.x == 0) {
        a.part[4 * t] = changed;
        a.part[4 * t + 1] = badv;
    }
    if (!last_tile(a.cnt, s, tiles.sys_tile1[s] - tiles.sys_tile0[s]))
        return;
    __threadfence();
    double ch = 0, bd = 0;
    for (int u = tiles.sys_tile0[s]; u < tiles.sys_tile1[s]; ++u) {
        ch = fmax(ch, a.part[4 * u]);
        bd = fmax(bd, a.part[4 * u + 1]);
    }
    st.unchanged = ch == 0;
    ++st.outer_steps;
    if (bd != 0) {
        st.status = kIrNumericalFailure;
        st.active = 0;
        st.done = 1;
    }
}

// Mixed path: the FP32 right side is r * 2^-e (e from ||r||_inf), so the
// inner solve sees O(1) data; the correction is scaled back exactly by 2^e.
__global__ void __launch_bounds__(kTileThreads)
scale_rhs_kernel(Tiles tiles, const double *__restrict__ r, float *__restrict__ rf, IrArgs a) {
    const int t = blockIdx.x, s = tiles.sys[t];
    IrSysState &st = a.st[s];
    if (!st.active)
        return;
    int e = 0;
    if (st.rinf > 0)
        frexp(st.rinf, &e);
    const double sc = ldexp(1.0, -e);
    for (int row = tiles.row0[t] + threadIdx.x; row < tiles.row1[t]; row += kTileThreads)
        rf[row] = (float)(r[row] * sc);
    if (t == tiles.sys_tile0[s] && threadIdx.x == 0)
        st.pending_dscale = ldexp(1.0, e);
}

// Copy per-system inner-solve outcome into the IR state (one thread per system).
template <typename T>
__global__ void absorb_inner_kernel(int nsys, KrylovArgs<T> k, IrArgs a) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys)
        return;
    IrSysState &st = a.st[s];
    if (!st.active)
        return;
    const SysState &g = k.st[s];
    if (st.inner_attempts < kMaxOuterHistory) {
        int *rec = a.inner + ((std::int64_t)s * kMaxOuterHistory + st.inner_attempts) * 3;
        rec[0] = g.status;
        rec[1] = g.converged;
        rec[2] = g.iterations;
    }
    ++st.inner_attempts;
    st.dscale = st.pending_dscale;
    if (!g.converged) {
        st.status = kIrInnerFailure;
        st.inner_status = g.status;
        st.active = 0;
        st.done = 1;
    }
}

// Reset GMRES per-system state for a new solve; systems take part iff their
// IR state is active (or all, when ir == nullptr).
template <typename T>
__global__ void gmres_reset_kernel(int nsys, KrylovArgs<T> k, const IrSysState *ir,
                                   const int *nrows, double tol) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys)
        return;
    SysState st = {};
    st.tol = tol;
    st.nrows = nrows[s];
    st.participating = ir ? ir[s].active : 1;
    st.active = st.participating;
    st.init_orig = st.init_rel = st.final_orig = st.final_rel = __builtin_nan("");
    k.st[s] = st;
}

void ir_ainf(hipStream_t s, const Tiles &tl, const int *ptr, const double *val, const IrArgs &a) {
    hipLaunchKernelGGL(ainf_kernel, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, ptr, val, a);
    HIP_LAUNCH_CHECK();
}
void ir_bnorm(hipStream_t s, const Tiles &tl, const double *b, const IrArgs &a) {
    hipLaunchKernelGGL(bnorm_kernel, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, b, a);
    HIP_LAUNCH_CHECK();
}
void ir_diagnostics(hipStream_t s, const Tiles &tl, const int *ptr, const int *col,
                    const double *val, const double *b, const double *x, double *r,
                    const IrArgs &a, int step) {
    hipLaunchKernelGGL(diagnostics_kernel, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, ptr,
                       col, val, b, x, r, a, step);
    HIP_LAUNCH_CHECK();
}
void ir_update(hipStream_t s, const Tiles &tl, double *x, const double *d, const IrArgs &a) {
    hipLaunchKernelGGL(update_kernel, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, x, d, a);
    HIP_LAUNCH_CHECK();
}
void ir_scale_rhs(hipStream_t s, const Tiles &tl, const double *r, float *rf, const IrArgs &a) {
    hipLaunchKernelGGL(scale_rhs_kernel, dim3(tl.ntiles), dim3(kTileThreads), 0, s, tl, r, rf, a);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void ir_absorb_inner(hipStream_t s, int nsys, const KrylovArgs<T> &k, const IrArgs &a) {
    hipLaunchKernelGGL(absorb_inner_kernel<T>, dim3(ceil_div(nsys, 64)), dim3(64), 0, s, nsys, k,
                       a);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_reset(hipStream_t s, int nsys, const KrylovArgs<T> &k, const IrSysState *ir,
                 const int *nrows, double tol) {
    hipLaunchKernelGGL(gmres_reset_kernel<T>, dim3(ceil_div(nsys, 64)), dim3(64), 0, s, nsys, k,
                       ir, nrows, tol);
    HIP_LAUNCH_CHECK();
}
template void ir_absorb_inner<double>(hipStream_t, int, const KrylovArgs<double> &, const IrArgs &);
template void ir_absorb_inner<float>(hipStream_t, int, const KrylovArgs<float> &, const IrArgs &);
template void gmres_reset<double>(hipStream_t, int, const KrylovArgs<double> &, const IrSysState *,
                                  const int *, double);
template void gmres_reset<float>(hipStream_t, int, const KrylovArgs<float> &, const IrSysState *,
                                 const int *, double);

} // namespace bspai_hip
