// Left-/right-preconditioned GMRES and iterative-refinement kernels (gfx950).
//
// Algorithm (reference src/gmres.cpp:108-262, src/ir.cpp:83-204): unrestarted
// GMRES with modified Gram-Schmidt and Givens rotations; the candidate
// x0 + V y is formed and re-verified (true and recomputed preconditioned
// residuals) exactly on the iterations whose acceptance test can pass --
// estimate <= tau, Arnoldi breakdown, zero Givens pivot, or the last
// iteration -- which leaves every decision of the reference unchanged.
//
// Gram-Schmidt runs in compact form.  The MGS coefficients h of w against
// V_0..V_k (h_j = <w - sum_{i<j} h_i V_i, V_j>, src/gmres.cpp:176-181) satisfy
//     (I + L) h = V^T w,   L = strictly lower triangle of V^T V,
// an identity of exact arithmetic that holds for any V (the inverse compact
// WY form of MGS: Swirydowicz, Langou, Ananthan, Yang and Thomas, "Low
// synchronization Gram-Schmidt and GMRES algorithms", Numer. Linear Algebra
// Appl. 28, 2021).  One streaming pass forms V^T w and the new Gram row
// V_k^T V_{0..k-1}, a 64-lane forward substitution gives h, and a second pass
// forms w - V h: two passes over the basis and four launches per iteration
// instead of k+2 dependent rounds, with the same coefficients up to rounding.
//
// Layout: vectors concatenate independent systems; every kernel runs one
// block per row tile (<= 2048 rows, inside one system) and skips tiles of
// inactive systems, so one launch advances a whole batch.
//
// Reductions never synchronize blocks inside a kernel (a device-scope release
// on a multi-XCD part writes back the L2, which is far costlier than a launch).
// Instead each tile writes its partial into a named slot, and the NEXT kernel
// reduces the slot: every block of the system sums the partials in the same
// fixed order, so all blocks obtain the same value bit for bit.  Per-system
// state (Hessenberg entries, Givens data, decisions) has one writer, thread 0
// of the system's first tile, and is read by later kernels only.
//
// Precision: T is the Krylov vector/matrix type (double for the plain path,
// float for the mixed path); Hessenberg, rotations, y and candidates are FP64.
#include <cstdlib>
#include <string>
#include <type_traits>
#include "common.hpp"
#include "krylov.hpp"

namespace bspai_hip {

// ------------------------------------------------------------- partial slots
// Dot partial of tile t: part[2t+1] (part[2t] = 0).  Norm partial: (e, s) with
// s = sum (x 2^-e)^2 and e the exponent of the tile's max |x| (robust 2-norm).
__device__ inline double *slot_ptr(double *part, int ntiles, int slot) {
    return part + (std::int64_t)slot * 2 * ntiles;
}

__device__ inline void store_dot_partial(double v, double *p, int t, double *lds) {
    v = block_sum(v, lds);
    if (threadIdx.x == 0) {
        p[2 * t] = 0;
        p[2 * t + 1] = v;
    }
}

__device__ inline void store_norm_partial(const double *v, int cnt, double *p, int t,
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
        s = INFINITY;
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) {
        p[2 * t] = (double)e;
        p[2 * t + 1] = s;
    }
}

// Deterministic block-cooperative reductions over tiles [t0, t1): the thread
// -> tile assignment and the combining tree are fixed, so every block that
// evaluates them gets identical bits.
__device__ inline double reduce_dot(const double *p, int t0, int t1, double *lds) {
    double s = 0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x)
        s += p[2 * t + 1];
    return block_sum(s, lds);
}

__device__ inline double reduce_norm(const double *p, int t0, int t1, double *lds) {
    double emax = -1e300;
#pragma unroll 4
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x)
        if (p[2 * t + 1] > 0)
            emax = fmax(emax, p[2 * t]);
    emax = block_max(emax, lds);
    if (emax == -1e300)
        return 0.0;
    double s = 0;
#pragma unroll 4
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x)
        if (p[2 * t + 1] > 0)
            s += ldexp(p[2 * t + 1], 2 * (int)(p[2 * t] - emax));
    s = block_sum(s, lds);
    return ldexp(sqrt(s), (int)emax);
}

__device__ inline double reduce_max(const double *p, int t0, int t1, double *lds, int stride,
                                    int off) {
    double m = 0;
    for (int t = t0 + (int)threadIdx.x; t < t1; t += blockDim.x)
        m = fmax(m, p[(std::int64_t)stride * t + off]);
    return block_max(m, lds);
}

#define BSPAI_SYS_TILES                                                                           \
    const int t = blockIdx.x, s = tiles.sys[t];                                                  \
    const int t0 = tiles.sys_tile0[s], t1 = tiles.sys_tile1[s];                                  \
    const bool writer = t == t0 && threadIdx.x == 0;                                             \
    (void)t1;                                                                                    \
    (void)writer;

// ------------------------------------------------------------------ GMRES
// Givens update of column k once h_{k+1,k} is known (src/gmres.cpp:190-217),
// the older rotations having been applied already.  Sets st.triggered when
// the candidate must be formed and verified; y is then solved by
// back_substitute.
__device__ void givens_step(SysState &st, double *H, double *cs, double *sn, double *g, int k,
                            int cap, double hnorm, int defer) {
    double *col = H + hcol_offset(k);
    col[k + 1] = hnorm;
    st.hnorm = hnorm;
    const bool breakdown = hnorm == 0;
    bool bad = !isfinite(hnorm) || st.rot_bad;
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
        st.triggered = breakdown || st.est <= st.tol || k + 1 == st.scap;
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
        st.run[(k + 1) & 1] = 0;
        return;
    }
    st.ccount = count;
    st.run[(k + 1) & 1] = !(st.triggered && defer);
}

// y = R^{-1} g for the leading count x count block (src/gmres.cpp:47-59), by
// one wave: row by row from the bottom, the sum over c > row is split by
// c mod 64 across the lanes (each lane ascending) and combined by a fixed
// butterfly; lane row mod 64 owns y[row], so every y entry is written and
// read by one lane only.  Returns false on a non-finite y.
__device__ bool back_substitute(const double *__restrict__ H, const double *__restrict__ g,
                                double *__restrict__ y, int count) {
    const int lane = threadIdx.x % kWave;
    bool bad = false;
    for (int row = count - 1; row >= 0; --row) {
        double part = 0;
        for (int c = row + 1 + ((lane - row - 1) % kWave + kWave) % kWave; c < count; c += kWave)
            part += H[hcol_offset(c) + row] * y[c];
        part = wave_sum(part);
        if (lane == row % kWave) {
            const double v = (g[row] - part) / H[hcol_offset(row) + row];
            y[row] = v;
            bad = bad || !isfinite(v);
        }
    }
    return !__any(bad);
}

// ------------------------------------------------------- compact MGS passes
constexpr int kProjWaves = 8;                  // waves per projection block
constexpr int kProjPerWave = 4;                // basis vectors streamed together by a wave
constexpr int kProjVecs = kProjPerWave * kProjWaves; // basis vectors per block
constexpr int kReduceWaves = 16;               // one value per wave
constexpr int kUpdateThreads = 256;             // 8 rows of a tile per thread
// Basis vectors per load batch of the update: 8 16-byte loads per thread.
template <typename T> constexpr int kUpdBatch = sizeof(T) == 8 ? 2 : 4;

// 16-byte accesses: E rows of T per access.
template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };
template <typename T> __device__ inline T lane_of(const typename Vec16<T>::type &v, int e);
template <> __device__ inline float lane_of<float>(const float4 &v, int e) {
    return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
}
template <> __device__ inline double lane_of<double>(const double2 &v, int e) {
    return e == 0 ? v.x : v.y;
}
// 16 bytes at p: one access when p is 16-byte aligned (A), else element by
// element -- the same values either way, so a tile's result does not depend
// on its alignment.  Checked (C): only the first lim elements exist, the
// others read as zero and are not stored (the last tile of a system).
template <typename T, bool A, bool C = false>
__device__ inline typename Vec16<T>::type ld16(const T *p, int lim = 16) {
    using V16 = typename Vec16<T>::type;
    if constexpr (A && !C) {
        return gld_vec(reinterpret_cast<const V16 *>(p));
    } else {
        V16 v;
        T *q = reinterpret_cast<T *>(&v);
#pragma unroll
        for (int e = 0; e < (int)(16 / sizeof(T)); ++e)
            q[e] = !C || e < lim ? gld(p + e) : T(0);
        return v;
    }
}
template <typename T, bool A, bool C = false>
__device__ inline void st16(T *p, const typename Vec16<T>::type &v, int lim = 16) {
    using V16 = typename Vec16<T>::type;
    if constexpr (A && !C) {
        gst_vec(reinterpret_cast<V16 *>(p), v);
    } else {
        const T *q = reinterpret_cast<const T *>(&v);
#pragma unroll
        for (int e = 0; e < (int)(16 / sizeof(T)); ++e)
            if (!C || e < lim)
                gst(p + e, q[e]);
    }
}

// Pass 1, grid (tile, group of kProjVecs vectors): w and V_k of the tile are
// staged in LDS; each wave streams kProjPerWave basis vectors of the group
// together (one 16-byte access of each per step, so that several are in
// flight) and forms <V_i, w> and <V_i, V_k> (FP64 accumulation, each lane
// over its rows, then a fixed butterfly), written to gp[i][tile] and
// gp[cap+1+i][tile].  Full tiles whose start is 16-byte aligned (every tile
// of a single system) use 16-byte accesses; the others read the same rows
// per lane in the same order, one element at a time.
template <typename T, bool kUniform>
__global__ void __launch_bounds__(kProjWaves *kWave)
project_kernel(Tiles tiles, KrylovArgs<T> a, int k) {
    k = step_index(a, k);
    if (k >= a.cap) // a replayed graph's steps beyond the cap do nothing
        return;
    using V16 = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
    constexpr int G = kTileRows / (kWave * E); // 16-byte accesses per lane per vector
    constexpr int P = kProjPerWave;
    __shared__ V16 ws4[kTileRows / E], vs4[kTileRows / E];
    T *ws = reinterpret_cast<T *>(ws4), *vs = reinterpret_cast<T *>(vs4);
    const int t = blockIdx.x, s = tiles.sys[t];
    const int i0 = blockIdx.y * kProjVecs;
    if (!a.st[s].active || !a.st[s].run[k & 1] || i0 > k)
        return;
    const int r0 = tiles.row0[t], nr = tiles.row1[t] - r0;
    const T *__restrict__ w = a.w + r0;
    const T *__restrict__ vk = a.V[k] + r0;
    for (int r = threadIdx.x; r < kTileRows; r += kProjWaves * kWave) {
        ws[r] = r < nr ? gld(w + r) : T(0);
        vs[r] = r < nr ? gld(vk + r) : T(0);
    }
    __syncthreads();
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int ib = i0 + P * wave; // this wave's vectors ib .. ib+P-1 (those <= k)
    if (ib > k)
        return;
    const T *__restrict__ vp[P];
#pragma unroll
    for (int u = 0; u < P; ++u)
        vp[u] = a.V[min(ib + u, k)] + r0;
    double dw[P], dv[P];
#pragma unroll
    for (int u = 0; u < P; ++u)
        dw[u] = dv[u] = 0;
    // 16-byte groups: one access each (aligned full tiles), element by
    // element (full tiles that start mid-group, as in batched systems), or
    // element by element with the rows past the tile's end reading as zero
    // (a system's last tile).  Same rows per lane and order in all three.
    auto groups = [&](auto aligned, auto checked) {
        constexpr bool A = decltype(aligned)::value, C = decltype(checked)::value;
        auto step = [&](int g) {
            const int q = g * kWave + lane;
            V16 x[P];
#pragma unroll
            for (int u = 0; u < P; ++u)
                x[u] = ld16<T, A, C>(vp[u] + (std::int64_t)q * E, nr - q * E);
            const V16 wq = ws4[q], vq = vs4[q];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const double we = (double)lane_of<T>(wq, e), ve = (double)lane_of<T>(vq, e);
#pragma unroll
                for (int u = 0; u < P; ++u) {
                    const double xe = (double)lane_of<T>(x[u], e);
                    dw[u] += xe * we;
                    dv[u] += xe * ve;
                }
            }
        };
        if constexpr (C) { // a system's last tile: keep its registers small
#pragma unroll 1
            for (int g = 0; g < G; ++g)
                step(g);
        } else {
            // 4 steps in flight (4P loads per lane): measured at 128^3 with
            // k ~ 75, FP64 280 us per launch vs 301 at 2 steps; 8 and 16
            // are slower for FP32 (136 vs 131 us).
#pragma unroll 4
            for (int g = 0; g < G; ++g)
                step(g);
        }
    };
    // (kUniform: the host found every tile full and aligned, and this
    // instantiation holds that path only, with its smaller register budget)
    if (kUniform || (nr == kTileRows && r0 % E == 0))
        groups(std::true_type{}, std::false_type{});
    else if constexpr (!kUniform) {
        if (nr == kTileRows)
            groups(std::false_type{}, std::false_type{});
        else
            groups(std::false_type{}, std::true_type{});
    }
#pragma unroll
    for (int u = 0; u < P; ++u) {
        dw[u] = wave_sum(dw[u]);
        dv[u] = wave_sum(dv[u]);
    }
    if (lane == 0) {
        const std::int64_t nt = tiles.ntiles;
#pragma unroll
        for (int u = 0; u < P; ++u) {
            const int i = ib + u;
            if (i <= k)
                a.gp[(2 * (std::int64_t)i) * nt + t] = dw[u];
            if (i < k)
                a.gp[(2 * (std::int64_t)i + 1) * nt + t] = dv[u];
        }
    }
}

// Pass 1 reduction, grid (system, group of kReduceWaves values): value
// v <= k is q_v = <V_v, w>, value k < v <= 2k is the Gram entry
// <V_k, V_{v-k-1}>.  Each wave sums the system's tiles (lane l: tiles
// t0+l, t0+l+64, ... in order, then a fixed butterfly).
template <typename T>
__global__ void __launch_bounds__(kReduceWaves *kWave)
gram_reduce_kernel(Tiles tiles, KrylovArgs<T> a, int k) {
    k = step_index(a, k);
    if (k >= a.cap) // a replayed graph's steps beyond the cap do nothing
        return;
    const int s = blockIdx.x;
    if (!a.st[s].active || !a.st[s].run[k & 1])
        return;
    const int v = blockIdx.y * kReduceWaves + (int)(threadIdx.x / kWave);
    const int lane = threadIdx.x % kWave;
    if (v > 2 * k)
        return;
    const std::int64_t nt = tiles.ntiles;
    const double *row = a.gp + (v <= k ? 2 * (std::int64_t)v : 2 * (std::int64_t)(v - k - 1) + 1) * nt;
    double sum = 0;
#pragma unroll 8
    for (int t = tiles.sys_tile0[s] + lane; t < tiles.sys_tile1[s]; t += kWave)
        sum += row[t];
    sum = wave_sum(sum);
    if (lane == 0) {
        if (v <= k)
            a.q[(std::int64_t)s * (a.cap + 1) + v] = sum;
        else
            a.gram[(std::int64_t)s * a.gstride + gram_index(k, v - k - 1, a.gsteps)] = sum;
    }
}

// (I + L) h = q by forward substitution, one wave per system: lane j owns
// unknown j of the current block of 64, first subtracts the finished blocks
// (i ascending), then the block is solved with the pivots broadcast lane by
// lane -- every h_j sees its terms in the order i = 0, 1, ..., j-1.  h is the
// MGS column H[0..k][k].  Lane 0 then applies the older Givens rotations to
// it, staged through LDS in chunks so the serial chain reads no global memory.
constexpr int kRotChunk = 512;
constexpr int kLoadBatch = 16; // divides kWave
template <typename T>
__global__ void __launch_bounds__(kWave) gram_solve_kernel(Tiles tiles, KrylovArgs<T> a, int k) {
    k = step_index(a, k);
    if (k >= a.cap) // a replayed graph's steps beyond the cap do nothing
        return;
    __shared__ double lcs[kRotChunk], lsn[kRotChunk], lrow[kRotChunk];
    __shared__ double tile[kWave][kWave + 1]; // the block's own triangle, [i][lane]
    const int s = blockIdx.x;
    if (!a.st[s].active || !a.st[s].run[k & 1])
        return;
    const int lane = threadIdx.x, cap = a.cap;
    const double *q = a.q + (std::int64_t)s * (cap + 1);
    const double *L = a.gram + (std::int64_t)s * a.gstride;
    double *h = a.h + (std::int64_t)s * (cap + 1);
    double *col = a.H + (std::int64_t)s * a.hstride + hcol_offset(k);
    for (int b0 = 0; b0 <= k; b0 += kWave) {
        const int j = b0 + lane;
        const bool mine = j <= k;
        const int last = min(k, b0 + kWave - 1);
        // Stage L(j, b0..j-1) first: these loads do not depend on the chain,
        // and are issued in batches so their latencies overlap.
        for (int i0 = b0; i0 <= last; i0 += kLoadBatch) {
            double v[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) {
                const int i = i0 + u;
                v[u] = i <= last && mine && j > i ? L[gram_index(j, i, a.gsteps)] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                if (i0 + u <= last)
                    tile[i0 + u - b0][lane] = v[u];
        }
        double acc = mine ? q[j] : 0.0;
        for (int i0 = 0; i0 < b0; i0 += kLoadBatch) { // b0 is a multiple of the batch
            double lv[kLoadBatch], hv[kLoadBatch];
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u) {
                lv[u] = mine ? L[gram_index(j, i0 + u, a.gsteps)] : 0.0;
                hv[u] = h[i0 + u];
            }
#pragma unroll
            for (int u = 0; u < kLoadBatch; ++u)
                acc = acc - lv[u] * hv[u];
        }
        __syncthreads();
        for (int i = b0; i <= last; ++i) {
            const double hi = __shfl(acc, i - b0, kWave);
            if (mine && j > i)
                acc = acc - tile[i - b0][lane] * hi;
        }
        if (mine) {
            h[j] = acc;
            col[j] = acc;
        }
        __syncthreads();
    }
    // Rotations 0..k-1 on rows 0..k (src/gmres.cpp:183-189); the carried row
    // stays in a register of lane 0.
    const double *cs = a.cs + (std::int64_t)s * cap, *sn = a.sn + (std::int64_t)s * cap;
    double up = col[0];
    bool bad = false;
    for (int c0 = 0; c0 < k; c0 += kRotChunk) {
        const int cnt = min(kRotChunk, k - c0);
#pragma unroll
        for (int u = 0; u < kRotChunk / kWave; ++u) {
            const int r = lane + u * kWave;
            if (r < cnt) {
                lcs[r] = cs[c0 + r];
                lsn[r] = sn[c0 + r];
                lrow[r] = col[c0 + 1 + r];
            }
        }
        __syncthreads();
        if (lane == 0) {
            for (int r = 0; r < cnt; ++r) {
                const double lo = lrow[r], c = lcs[r], z = lsn[r];
                const double top = c * up + z * lo;
                up = -z * up + c * lo;
                lrow[r] = top; // now row c0 + r
                bad = bad || !isfinite(top) || !isfinite(up);
            }
        }
        __syncthreads();
        for (int r = lane; r < cnt; r += kWave)
            col[c0 + r] = lrow[r];
        __syncthreads();
    }
    if (lane == 0) {
        col[k] = up;
        a.st[s].rot_bad = bad;
    }
}

// The same solve for k < 512, fused into the update pass (gram_update_kernel
// with S > 0): every update block solves its system's (I + L) h = q itself
// before it streams the basis, so no launch separates the two passes.  All
// waves stage column chunks of L in LDS (every load of a chunk in flight at
// once), then wave 0 runs the substitution from LDS, unknown j in lane
// j % 64, slot j / 64, each pivot broadcast by readlane.  Every h_j again
// sees its terms in the order i = 0, 1, ..., j-1 (bit-identical to the
// kernel above), so every block of a system holds the same h.  The system's
// first block then applies the older rotations from LDS and writes the
// column of H.
// S unknowns per lane (k < 64 S): 4 for k < 256 with 24 KiB chunks of L, 8
// up to k < 512 with 18 KiB chunks: with h and the pads, 30 KiB of LDS, so
// that five update blocks (their register budget) still share a CU.
template <int S> constexpr int kSolveChunk = S == 4 ? 3072 : 2304;
static_assert(kGramSolveLdsMaxK == 8 * kWave - 1, "LDS Gram solve capacity");

__device__ inline double readlane_f64(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Steps i in [i, min(iend, 64 (B+1))) of the substitution, whose pivot h_i
// sits in slot B (lane i - 64 B).  Column i of L starts at Lc + lo; row j's
// entry is at Lc + lo + (j - i - 1).  Rows of slot B below or at i keep their
// value; every row of a later slot lies below i, so no mask is needed there
// (rows beyond k compute values that are never stored).  Lc is padded by
// kSolvePad entries on both sides, so these reads stay in bounds.
template <int S, int B>
__device__ inline void gram_steps(double (&acc)[S], const double *Lc, int &i, int &lo, int iend,
                                  int k, int lane) {
    constexpr int kSolvePad = S * kWave;
    // Steps are taken in batches of kNB: the batch's L entries are read from
    // LDS together (one round of LDS latency per batch, not per step), then
    // the batch's steps run from registers.  Reads past `stop` repeat the
    // last step's column, so every read stays inside the staged chunk.
    constexpr int kNB = 32 / S; // 32 doubles of L in registers per batch
    const int stop = min(iend, (B + 1) * kWave);
    const double *base = Lc + kSolvePad + lane;
    while (i < stop) {
        int col0[kNB], piv[kNB];
        {
            int l = lo;
#pragma unroll
            for (int t = 0; t < kNB; ++t) {
                const bool ok = i + t < stop;
                piv[t] = ok ? i + t : piv[t > 0 ? t - 1 : 0];
                col0[t] = ok ? l : col0[t > 0 ? t - 1 : 0];
                if (ok)
                    l += k - (i + t);
            }
            lo = l;
        }
        double lv[kNB][S];
#pragma unroll
        for (int t = 0; t < kNB; ++t) {
            const double *colp = base + col0[t] - piv[t] - 1;
#pragma unroll
            for (int u = B; u < S; ++u)
                lv[t][u] = colp[u * kWave];
        }
#pragma unroll
        for (int t = 0; t < kNB; ++t) {
            if (i + t < stop) {
                const int ii = i + t;
                const double hi = readlane_f64(acc[B], ii - B * kWave);
                const double y = acc[B] - lv[t][B] * hi;
                acc[B] = lane + B * kWave > ii ? y : acc[B];
#pragma unroll
                for (int u = B + 1; u < S; ++u)
                    acc[u] = acc[u] - lv[t][u] * hi;
            }
        }
        i = min(i + kNB, stop);
    }
}

// (I + L) h = q for system s at step k < 64 S by a block of NT threads (the
// chunk loop and substitution described above); h_0..h_k end in hs, visible
// to the whole block.
template <typename T, int S, int NT>
__device__ void gram_substitute(const KrylovArgs<T> &a, int s, int k, double *Lc, double *hs) {
    constexpr int kSolvePad = S * kWave, kChunk = kSolveChunk<S>;
    static_assert(kChunk % NT == 0, "chunk entries per thread");
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave, cap = a.cap;
    const double *q = a.q + (std::int64_t)s * (cap + 1);
    const double *L = a.gram + (std::int64_t)s * a.gstride;
    double acc[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        const int j = lane + u * kWave;
        acc[u] = wave == 0 && j <= k ? q[j] : 0.0;
    }
    // Steps i = 0..k-1 (column i of L holds rows j = i+1..k), in chunks of
    // whole columns; column i of a chunk starting at c0 begins at entry
    // sum_{c0 <= i' < i} (k - i').
    for (int c0 = 0; c0 < k;) {
        int c1 = c0, used = 0;
        while (c1 < k && used + (k - c1) <= kChunk) {
            used += k - c1;
            ++c1;
        }
        // flat entry e of the chunk -> column i, row j = i + 1 + e - off(i).
        // Absolute offsets from column 0, F(i) = i (2k + 1 - i) / 2, are
        // inverted in closed form (then corrected by one step either way);
        // all of a thread's loads are issued together.  The root's argument
        // (2k+1)^2 - 8e < 2^24 is exact in FP32, whose rounded root is within
        // the one-step correction.
        {
            constexpr int R = kChunk / NT;
            const long long base = (long long)c0 * (2 * k + 1 - c0) / 2;
            const float b2 = 2.0f * (float)k + 1.0f;
            int src[R];
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const long long e = base + min(tid + u * NT, used - 1);
                int i = (int)floorf((b2 - sqrtf(b2 * b2 - 8.0f * (float)e)) * 0.5f);
                i = max(c0, min(i, c1 - 1));
                if ((long long)i * (2 * k + 1 - i) / 2 > e)
                    --i;
                else if ((long long)(i + 1) * (2 * k - i) / 2 <= e)
                    ++i;
                const long long fi = (long long)i * (2 * k + 1 - i) / 2;
                src[u] = gram_index(i + 1 + (int)(e - fi), i, a.gsteps);
            }
            double v[R];
#pragma unroll
            for (int u = 0; u < R; ++u)
                v[u] = L[src[u]];
#pragma unroll
            for (int u = 0; u < R; ++u)
                if (tid + u * NT < used)
                    Lc[kSolvePad + tid + u * NT] = v[u];
        }
        __syncthreads();
        if (wave == 0) {
            int i = c0, lo = 0;
            gram_steps<S, 0>(acc, Lc, i, lo, c1, k, lane);
            gram_steps<S, 1>(acc, Lc, i, lo, c1, k, lane);
            gram_steps<S, 2>(acc, Lc, i, lo, c1, k, lane);
            gram_steps<S, 3>(acc, Lc, i, lo, c1, k, lane);
            if constexpr (S > 4) {
                gram_steps<S, 4>(acc, Lc, i, lo, c1, k, lane);
                gram_steps<S, 5>(acc, Lc, i, lo, c1, k, lane);
                gram_steps<S, 6>(acc, Lc, i, lo, c1, k, lane);
                gram_steps<S, 7>(acc, Lc, i, lo, c1, k, lane);
            }
        }
        __syncthreads();
        c0 = c1;
    }
    if (wave == 0) {
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int j = lane + u * kWave;
            if (j <= k)
                hs[j] = acc[u];
        }
    }
    __syncthreads();
}

// Rotations 0..k-1 on rows 0..k of h (src/gmres.cpp:183-189) by lane 0 of
// wave 0, cs and sn staged in buf (2 x 512 entries) by the block; the
// rotated column goes to H and its finiteness to rot_bad.  A non-finite
// intermediate carried row stays non-finite to the end, so finiteness is
// checked once over the rotated column afterwards.  hs is overwritten.
template <typename T, int NT>
__device__ void rotate_column(const KrylovArgs<T> &a, int s, int k, double *hs, double *buf) {
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave, cap = a.cap;
    const double *cs = a.cs + (std::int64_t)s * cap, *sn = a.sn + (std::int64_t)s * cap;
    __syncthreads(); // every read of hs and buf by the block is done
    for (int r = tid; r < k; r += NT) {
        buf[r] = cs[r];
        buf[kGramSolveLdsMaxK + 1 + r] = sn[r];
    }
    __syncthreads();
    if (wave != 0)
        return;
    const double *lcs = buf, *lsn = buf + kGramSolveLdsMaxK + 1;
    if (lane == 0) {
        double up = hs[0];
        constexpr int B = 8; // LDS reads batched ahead of the rotation chain
        for (int r0 = 0; r0 < k; r0 += B) {
            double lo[B], c[B], z[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const int r = min(r0 + u, k - 1);
                lo[u] = hs[r + 1];
                c[u] = lcs[r];
                z[u] = lsn[r];
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                if (r0 + u < k) {
                    const double top = c[u] * up + z[u] * lo[u];
                    up = -z[u] * up + c[u] * lo[u];
                    hs[r0 + u] = top;
                }
            }
        }
        hs[k] = up;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double *col = a.H + (std::int64_t)s * a.hstride + hcol_offset(k);
    bool bad = false;
    for (int r = lane; r <= k; r += kWave) {
        const double v = hs[r];
        col[r] = v;
        bad = bad || !isfinite(v);
    }
    bad = __any(bad) && k > 0;
    if (lane == 0)
        a.st[s].rot_bad = bad;
}

// Pass 2: w <- w - sum_i h_i V_i (i ascending, in T as the reference's
// w -= h V_j), and the norm partial of the new w into kSlotNorm.  Each thread
// owns G 16-byte groups of the tile's rows; the basis is streamed in batches
// of kUpdBatch<T> vectors, two batches in registers, so that one batch's loads
// are in flight while the other is applied.  Tiles that are not full and
// 16-byte aligned take the same rows per thread, element by element.
template <typename T, int G, bool A, bool C>
__device__ inline void upd_load(typename Vec16<T>::type (&buf)[kUpdBatch<T>][G],
                                const KrylovArgs<T> &a, int base, int k, int r0, int tid,
                                int nr) {
    constexpr int E = 16 / sizeof(T);
#pragma unroll
    for (int u = 0; u < kUpdBatch<T>; ++u) {
        const T *__restrict__ v = a.V[min(base + u, k)] + r0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int r = (g * kUpdateThreads + tid) * E;
            buf[u][g] = ld16<T, A, C>(v + r, nr - r);
        }
    }
}
template <typename T, int G, int E>
__device__ inline void upd_apply(T (&x)[G][E], const typename Vec16<T>::type (&buf)[kUpdBatch<T>][G],
                                 const double *h, int base, int k) {
#pragma unroll
    for (int u = 0; u < kUpdBatch<T>; ++u) {
        if (base + u <= k) {
            const T hi = (T)h[base + u];
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int e = 0; e < E; ++e)
                    x[g][e] = x[g][e] - hi * lane_of<T>(buf[u][g], e);
        }
    }
}

// S = 0: h from the Gram solve kernel; S = 4 or 8: the block solves for h
// first (gram_substitute, k < 64 S), and the system's first block applies
// the rotations after its update.
template <typename T, int S>
__global__ void __launch_bounds__(kUpdateThreads)
gram_update_kernel(Tiles tiles, KrylovArgs<T> a, int k) {
    k = step_index(a, k);
    if (k >= a.cap) // a replayed graph's steps beyond the cap do nothing
        return;
    using V16 = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
    constexpr int G = kTileRows / (kUpdateThreads * E);
    __shared__ double lds[kUpdateThreads / kWave];
    BSPAI_SYS_TILES
    if (!a.st[s].active || !a.st[s].run[k & 1])
        return;
    const double *h;
    [[maybe_unused]] double *hsolved = nullptr, *lbuf = nullptr;
    if constexpr (S > 0) {
        __shared__ double Lc[2 * S * kWave + kSolveChunk<S>];
        __shared__ double hs[S * kWave];
        gram_substitute<T, S, kUpdateThreads>(a, s, k, Lc, hs);
        h = hs;
        hsolved = hs;
        lbuf = Lc;
    } else {
        h = a.h + (std::int64_t)s * (a.cap + 1);
    }
    const int r0 = tiles.row0[t], nr = tiles.row1[t] - r0;
    const int tid = threadIdx.x;
    T *__restrict__ w = a.w + r0;
    T x[G][E];
    // 16-byte groups: one access each (aligned full tiles), element by
    // element (full tiles that start mid-group, as in batched systems), or
    // element by element with the rows past the tile's end reading as zero
    // and left unwritten (a system's last tile).  Same rows per thread and
    // order in all three.
    auto groups = [&](auto aligned, auto checked) {
        constexpr bool A = decltype(aligned)::value, C = decltype(checked)::value;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int r = (g * kUpdateThreads + tid) * E;
            const V16 y = ld16<T, A, C>(w + r, nr - r);
#pragma unroll
            for (int e = 0; e < E; ++e)
                x[g][e] = lane_of<T>(y, e);
        }
        constexpr int B = kUpdBatch<T>;
        V16 b0[B][G], b1[B][G];
        upd_load<T, G, A, C>(b0, a, 0, k, r0, tid, nr);
        for (int base = 0; base <= k; base += 2 * B) {
            upd_load<T, G, A, C>(b1, a, base + B, k, r0, tid, nr);
            upd_apply<T, G, E>(x, b0, h, base, k);
            if (base + B > k)
                break;
            upd_load<T, G, A, C>(b0, a, base + 2 * B, k, r0, tid, nr);
            upd_apply<T, G, E>(x, b1, h, base + B, k);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            V16 y;
            T *yp = reinterpret_cast<T *>(&y);
#pragma unroll
            for (int e = 0; e < E; ++e)
                yp[e] = x[g][e];
            const int r = (g * kUpdateThreads + tid) * E;
            st16<T, A, C>(w + r, y, nr - r);
        }
    };
    if (nr == kTileRows && r0 % E == 0)
        groups(std::true_type{}, std::false_type{});
    else if (nr == kTileRows)
        groups(std::false_type{}, std::false_type{});
    else
        groups(std::false_type{}, std::true_type{});
    double vals[G * E];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int e = 0; e < E; ++e)
            vals[g * E + e] = (double)x[g][e];
    store_norm_partial(vals, G * E, slot_ptr(a.part, tiles.ntiles, kSlotNorm), t, lds);
    if constexpr (S > 0) {
        if (t == t0)
            rotate_column<T, kUpdateThreads>(a, s, k, hsolved, lbuf);
    }
}

// h_{k+1,k} = ||w||; V_{k+1} = w / h (unless breakdown); the writer runs the
// Givens step for the system.
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
normalize_givens_kernel(Tiles tiles, KrylovArgs<T> a, int k) {
    k = step_index(a, k);
    if (k >= a.cap) // a replayed graph's steps beyond the cap do nothing
        return;
    __shared__ double lds[kTileThreads / kWave];
    BSPAI_SYS_TILES
    SysState &st = a.st[s];
    if (!st.active)
        return;
    if (!st.run[k & 1]) { // held: the hold carries to the next step
        if (writer)
            st.run[(k + 1) & 1] = 0;
        return;
    }
    const double hn = reduce_norm(slot_ptr(a.part, tiles.ntiles, kSlotNorm), t0, t1, lds);
    if (hn != 0 && isfinite(hn) && k + 1 <= a.cap) {
        // V_{k+1} = w / h, a full aligned tile in 16-byte accesses (all of a
        // thread's loads issued together), other tiles element by element
        using V16 = typename Vec16<T>::type;
        constexpr int E = 16 / sizeof(T);
        constexpr int G = kTileRows / (kTileThreads * E);
        const T h = (T)hn;
        const int r0 = tiles.row0[t], nr = tiles.row1[t] - r0;
        T *v = a.V[k + 1] + r0;
        const T *w = a.w + r0;
        auto full = [&](auto aligned) {
            constexpr bool A = decltype(aligned)::value;
            V16 x[G];
#pragma unroll
            for (int g = 0; g < G; ++g)
                x[g] = ld16<T, A>(w + (std::int64_t)(g * kTileThreads + threadIdx.x) * E);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                T *xp = reinterpret_cast<T *>(&x[g]);
#pragma unroll
                for (int e = 0; e < E; ++e)
                    xp[e] = xp[e] / h;
                st16<T, A>(v + (std::int64_t)(g * kTileThreads + threadIdx.x) * E, x[g]);
            }
        };
        if (nr == kTileRows && r0 % E == 0) {
            full(std::true_type{});
        } else if (nr == kTileRows) {
            full(std::false_type{});
        } else {
#pragma unroll 8
            for (int r = threadIdx.x; r < nr; r += kTileThreads)
                v[r] = w[r] / h;
        }
    }
    if (t != t0)
        return;
    double *H = a.H + (std::int64_t)s * a.hstride, *g = a.g + (std::int64_t)s * (a.cap + 1);
    if (writer)
        givens_step(st, H, a.cs + (std::int64_t)s * a.cap, a.sn + (std::int64_t)s * a.cap, g, k,
                    a.cap, hn, a.defer);
    __syncthreads();
    if (threadIdx.x < kWave && st.triggered) {
        const bool ok = back_substitute(H, g, a.y + (std::int64_t)s * a.cap, st.ccount);
        if (!ok && threadIdx.x == 0) {
            st.status = kGmNumericalFailure;
            st.active = 0;
            st.triggered = 0;
        }
    }
}

// Candidate x = x0 + sum_{j<count} y_j V_j (FP64, j ascending, src/gmres.cpp:60-63).
// Each thread owns the update pass's rows (G 16-byte groups) and streams the
// basis the same way, two batches of loads in flight; every row still sees
// x0 + y_0 V_0 + y_1 V_1 + ... in that order.
template <typename T>
__global__ void __launch_bounds__(kUpdateThreads)
candidate_kernel(Tiles tiles, KrylovArgs<T> a) {
    const int t = blockIdx.x, s = tiles.sys[t];
    const SysState &st = a.st[s];
    if (!st.active || !st.triggered)
        return;
    using V16 = typename Vec16<T>::type;
    constexpr int E = 16 / sizeof(T);
    constexpr int G = kTileRows / (kUpdateThreads * E);
    constexpr int B = kUpdBatch<T>;
    const double *y = a.y + (std::int64_t)s * a.cap;
    const int count = st.ccount;
    const int r0 = tiles.row0[t], nr = tiles.row1[t] - r0;
    const int tid = threadIdx.x;
    double x[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int r = (g * kUpdateThreads + tid) * E + e;
            x[g][e] = a.x0 && r < nr ? a.x0[r0 + r] : 0.0;
        }
    auto apply = [&](const V16(&buf)[B][G], int base) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (base + u < count) {
                const double yu = y[base + u];
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int e = 0; e < E; ++e)
                        x[g][e] = x[g][e] + (double)lane_of<T>(buf[u][g], e) * yu;
            }
        }
    };
    auto groups = [&](auto aligned, auto checked) {
        constexpr bool A = decltype(aligned)::value, C = decltype(checked)::value;
        V16 b0[B][G], b1[B][G];
        upd_load<T, G, A, C>(b0, a, 0, count - 1, r0, tid, nr);
        for (int base = 0; base < count; base += 2 * B) {
            upd_load<T, G, A, C>(b1, a, base + B, count - 1, r0, tid, nr);
            apply(b0, base);
            if (base + B >= count)
                break;
            upd_load<T, G, A, C>(b0, a, base + 2 * B, count - 1, r0, tid, nr);
            apply(b1, base + B);
        }
    };
    if (count > 0) {
        if (nr == kTileRows && r0 % E == 0)
            groups(std::true_type{}, std::false_type{});
        else if (nr == kTileRows)
            groups(std::false_type{}, std::false_type{});
        else
            groups(std::false_type{}, std::true_type{});
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int r = (g * kUpdateThreads + tid) * E + e;
            if (r < nr)
                a.x[r0 + r] = x[g][e];
        }
}

// y = M x on the tile's rows (sequential row order), fused with the norm
// partial of (u - y) into `slot`; optional store of y.  gate: 0 active,
// 1 active && triggered.
template <typename TV, typename TU>
__global__ void __launch_bounds__(kTileThreads)
spmv_resid_kernel(Tiles tiles, SysState *st, double *part, const int *__restrict__ ptr,
                  const int *__restrict__ col, const TV *__restrict__ val,
                  const double *__restrict__ x, const TU *__restrict__ u, double *__restrict__ y,
                  int slot, int gate) {
    __shared__ double lds[kTileThreads / kWave];
    BSPAI_SYS_TILES
    const SysState &S = st[s];
    if (!S.active || (gate == 1 && !S.triggered))
        return;
    double vals[kTilePer];
    int cnt = 0;
    for (int row = tiles.row0[t] + threadIdx.x; row < tiles.row1[t]; row += kTileThreads) {
        double acc = 0;
        for (int e = ptr[row]; e < ptr[row + 1]; ++e)
            acc += (double)val[e] * x[col[e]];
        if (y)
            y[row] = acc;
        vals[cnt++] = (u ? (double)u[row] : 0.0) - acc;
    }
    store_norm_partial(vals, cnt, slot_ptr(part, tiles.ntiles, slot), t, lds);
}

// Norm partial of (u - v) into `slot` (v optional); optional store of the difference.
template <typename TU, typename TV, typename TO>
__global__ void __launch_bounds__(kTileThreads)
diff_partial_kernel(Tiles tiles, SysState *st, double *part, const TU *__restrict__ u,
                    const TV *__restrict__ v, TO *__restrict__ out, int slot) {
    __shared__ double lds[kTileThreads / kWave];
    BSPAI_SYS_TILES
    if (!st[s].active)
        return;
    double vals[kTilePer];
    int cnt = 0;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads) {
        const double d = (double)u[r] - (v ? (double)v[r] : 0.0);
        vals[cnt++] = d;
        if (out)
            out[r] = (TO)d;
    }
    store_norm_partial(vals, cnt, slot_ptr(part, tiles.ntiles, slot), t, lds);
}

// Per-system decisions; one block per system.
//   phase kPhaseInit: slots kSlotRhs, kSlotOrig -> zero_rhs / exact_initial
//   phase kPhaseDen:  slot kSlotDen -> zero_preconditioned_rhs
//   phase kPhaseBeta: slot kSlotBeta -> g_0, init_rel <= tau
//   phase kPhaseCheck: slots kSlotChkOrig, kSlotChkRel -> accept / fail / go on
template <typename T>
__global__ void __launch_bounds__(kTileThreads)
decide_kernel(Tiles tiles, KrylovArgs<T> a, int phase, int right) {
    __shared__ double lds[kTileThreads / kWave];
    const int s = blockIdx.x;
    SysState &st = a.st[s];
    if (!st.active || (phase == kPhaseCheck && !st.triggered))
        return;
    const int t0 = tiles.sys_tile0[s], t1 = tiles.sys_tile1[s];
    const int nt = tiles.ntiles;
    auto norm_of = [&](int slot) { return reduce_norm(slot_ptr(a.part, nt, slot), t0, t1, lds); };
    if (phase == kPhaseInit) {
        const double rn = norm_of(kSlotRhs), on = norm_of(kSlotOrig);
        if (threadIdx.x)
            return;
        if (!isfinite(on) || !isfinite(rn)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
            return;
        }
        st.init_orig = st.final_orig = on;
        if (on == 0) {
            st.status = rn == 0 ? kGmZeroRhs : kGmExactInitial;
            st.converged = 1;
            st.init_rel = st.final_rel = 0;
            st.active = 0;
        }
    } else if (phase == kPhaseDen) {
        const double dn = norm_of(kSlotDen);
        if (threadIdx.x)
            return;
        st.den = dn;
        if (!isfinite(dn)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
        } else if (dn == 0) {
            st.status = kGmZeroPreconditionedRhs;
            st.active = 0;
        }
    } else if (phase == kPhaseBeta) {
        const double bn = norm_of(kSlotBeta);
        if (threadIdx.x)
            return;
        st.beta = bn;
        st.hnorm = bn;
        a.g[(std::int64_t)s * (a.cap + 1)] = bn;
        st.init_rel = st.final_rel = bn / st.den;
        if (!isfinite(st.init_rel)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
        } else if (st.init_rel <= st.tol) {
            st.status = kGmSuccess;
            st.converged = 1;
            st.active = 0;
        }
    } else { // kPhaseCheck (src/gmres.cpp:218-253)
        const double on = norm_of(kSlotChkOrig);
        const double rn = right ? on : norm_of(kSlotChkRel);
        if (threadIdx.x)
            return;
        const double rel = rn / st.den;
        st.final_orig = on;
        st.final_rel = rel;
        st.triggered = 0;
        st.has_solution = 1;
        if (!isfinite(rel) || !isfinite(on)) {
            st.status = kGmNumericalFailure;
            st.active = 0;
            return;
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
        } else if (st.iterations == st.scap) {
            st.status = st.scap < st.nrows ? kGmIterationLimit : kGmNonconvergence;
            st.active = 0;
        }
        if (st.active) // not accepted: the next step (a held one resumes) runs
            st.run[st.iterations & 1] = 1;
    }
}

// V_0 = r0 / beta.
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

// Systems that ended without forming a candidate return x0 (or 0).
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

template <typename T>
__global__ void gmres_reset_kernel(int nsys, KrylovArgs<T> k, const IrSysState *ir,
                                   const int *nrows, double tol, InnerTolRule rule) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys)
        return;
    SysState st = {};
    st.tol = tol;
    if (ir && rule.theta > 0) {
        // Adaptive inner tolerance: the contraction this refinement step
        // still needs to meet the outer stop (backward error, and the
        // relative residual when asked for), times a safety factor; never
        // tighter than the caller's tolerance, never looser than 1/2.
        const IrSysState &q = ir[s];
        double need = q.nbe > 0 ? rule.be_tol / q.nbe : 1.0;
        if (rule.relres_target > 0 && q.r2 > 0 && q.bnorm > 0)
            need = fmin(need, rule.relres_target / (q.r2 / q.bnorm));
        const double t = rule.theta * need;
        if (t == t)
            st.tol = fmin(0.5, fmax(tol, t));
    }
    st.nrows = nrows[s];
    // Each system stops at its own cap: the workspace's cap (the caller's, or
    // the largest system) bounded by its own dimension, as a single solve of
    // that system would be (src/gmres.cpp:128-130).
    st.scap = k.cap < st.nrows ? k.cap : st.nrows;
    st.participating = ir ? ir[s].active : 1;
    st.active = st.participating;
    st.run[0] = st.run[1] = 1;
    st.init_orig = st.init_rel = st.final_orig = st.final_rel = __builtin_nan("");
    k.st[s] = st;
}

// Gram triangle re-laid from K0 to K steps (column i moves from
// gcol_offset(i, K0) to gcol_offset(i, K); its rows i+1..K0-1 are kept).
// Grid (system, column group): block y copies columns y, y + gridDim.y, ...
__global__ void gram_relayout_kernel(const double *__restrict__ src, std::int64_t sstride, int K0,
                                     double *__restrict__ dst, std::int64_t dstride, int K) {
    const int s = blockIdx.x;
    for (int i = blockIdx.y; i < K0 - 1; i += gridDim.y) {
        const double *a = src + s * sstride + gcol_offset(i, K0);
        double *b = dst + s * dstride + gcol_offset(i, K);
        for (int r = threadIdx.x; r < K0 - 1 - i; r += blockDim.x)
            b[r] = a[r];
    }
}
void gmres_gram_relayout(hipStream_t s, int nsys, const double *src, std::int64_t sstride, int K0,
                         double *dst, std::int64_t dstride, int K) {
    if (K0 < 2 || nsys == 0)
        return;
    hipLaunchKernelGGL(gram_relayout_kernel, dim3(nsys, std::min(K0 - 1, 1024)), dim3(256), 0, s,
                       src, sstride, K0, dst, dstride, K);
    HIP_LAUNCH_CHECK();
}

__global__ void set_step_kernel(int *kdev, int k) {
    *kdev = k;
}
__global__ void advance_step_kernel(int *kdev, int d) {
    *kdev += d;
}

// ------------------------------------------------------------------ launchers
#define BSPAI_GRID dim3(tl.ntiles), dim3(kTileThreads), 0, s

void gmres_set_step(hipStream_t s, int *kdev, int k) {
    hipLaunchKernelGGL(set_step_kernel, dim3(1), dim3(1), 0, s, kdev, k);
    HIP_LAUNCH_CHECK();
}
void gmres_advance_step(hipStream_t s, int *kdev, int d) {
    hipLaunchKernelGGL(advance_step_kernel, dim3(1), dim3(1), 0, s, kdev, d);
    HIP_LAUNCH_CHECK();
}

template <typename T>
void gmres_project(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax) {
    const dim3 grid(tl.ntiles, ceil_div(kmax + 1, kProjVecs)), block(kProjWaves * kWave);
    if (tl.uniform)
        hipLaunchKernelGGL((project_kernel<T, true>), grid, block, 0, s, tl, a, k);
    else
        hipLaunchKernelGGL((project_kernel<T, false>), grid, block, 0, s, tl, a, k);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_gram_reduce(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax) {
    hipLaunchKernelGGL(gram_reduce_kernel<T>, dim3(tl.nsys, ceil_div(2 * kmax + 1, kReduceWaves)),
                       dim3(kReduceWaves * kWave), 0, s, tl, a, k);
    HIP_LAUNCH_CHECK();
}
bool gmres_gram_solve_fused(int kmax) {
    // BSPAI_GRAM_FUSED=0 keeps the separate one-wave solve kernel at every
    // step (the same bits; for comparisons)
    const char *env = std::getenv("BSPAI_GRAM_FUSED");
    return kmax <= kGramSolveLdsMaxK && !(env && std::string(env) == "0");
}
template <typename T>
void gmres_gram_solve(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax) {
    if (gmres_gram_solve_fused(kmax)) // the update pass solves
        return;
    hipLaunchKernelGGL(gram_solve_kernel<T>, dim3(tl.nsys), dim3(kWave), 0, s, tl, a, k);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_gram_update(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax) {
    // the variant that holds the largest step the launch may see
    if (kmax < 4 * kWave)
        hipLaunchKernelGGL((gram_update_kernel<T, 4>), dim3(tl.ntiles), dim3(kUpdateThreads), 0,
                           s, tl, a, k);
    else if (gmres_gram_solve_fused(kmax))
        hipLaunchKernelGGL((gram_update_kernel<T, 8>), dim3(tl.ntiles), dim3(kUpdateThreads), 0,
                           s, tl, a, k);
    else
        hipLaunchKernelGGL((gram_update_kernel<T, 0>), dim3(tl.ntiles), dim3(kUpdateThreads), 0,
                           s, tl, a, k);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_normalize_givens(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k) {
    hipLaunchKernelGGL(normalize_givens_kernel<T>, BSPAI_GRID, tl, a, k);
    HIP_LAUNCH_CHECK();
}
template <typename T> void gmres_candidate(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a) {
    hipLaunchKernelGGL(candidate_kernel<T>, dim3(tl.ntiles), dim3(kUpdateThreads), 0, s, tl, a);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_start_vector(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, const T *r0) {
    hipLaunchKernelGGL(start_vector_kernel<T>, BSPAI_GRID, tl, a, r0);
    HIP_LAUNCH_CHECK();
}
template <typename T> void gmres_finish(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a) {
    hipLaunchKernelGGL(finish_kernel<T>, BSPAI_GRID, tl, a);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_decide(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int phase, int right) {
    hipLaunchKernelGGL(decide_kernel<T>, dim3(tl.nsys), dim3(kTileThreads), 0, s, tl, a, phase,
                       right);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void gmres_reset(hipStream_t s, int nsys, const KrylovArgs<T> &k, const IrSysState *ir,
                 const int *nrows, double tol, const InnerTolRule &rule) {
    hipLaunchKernelGGL(gmres_reset_kernel<T>, dim3(ceil_div(nsys, 64)), dim3(64), 0, s, nsys, k,
                       ir, nrows, tol, rule);
    HIP_LAUNCH_CHECK();
}
template <typename TV, typename TU>
void gmres_spmv_resid(hipStream_t s, const Tiles &tl, SysState *st, double *part, const int *ptr,
                      const int *col, const TV *val, const double *x, const TU *u, double *y,
                      int slot, int gate) {
    hipLaunchKernelGGL((spmv_resid_kernel<TV, TU>), BSPAI_GRID, tl, st, part, ptr, col, val, x, u,
                       y, slot, gate);
    HIP_LAUNCH_CHECK();
}
template <typename TU, typename TV, typename TO>
void gmres_diff_partial(hipStream_t s, const Tiles &tl, SysState *st, double *part, const TU *u,
                        const TV *v, TO *out, int slot) {
    hipLaunchKernelGGL((diff_partial_kernel<TU, TV, TO>), BSPAI_GRID, tl, st, part, u, v, out,
                       slot);
    HIP_LAUNCH_CHECK();
}

#define BSPAI_KRYLOV_INST(T)                                                                      \
    template void gmres_project<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int, int);  \
    template void gmres_gram_reduce<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int,    \
                                       int);                                                      \
    template void gmres_gram_solve<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int,     \
                                      int);                                                       \
    template void gmres_gram_update<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int,    \
                                       int);                                                      \
    template void gmres_normalize_givens<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &,    \
                                            int);                                                 \
    template void gmres_candidate<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &);          \
    template void gmres_start_vector<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &,        \
                                        const T *);                                               \
    template void gmres_finish<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &);             \
    template void gmres_decide<T>(hipStream_t, const Tiles &, const KrylovArgs<T> &, int, int);   \
    template void gmres_reset<T>(hipStream_t, int, const KrylovArgs<T> &, const IrSysState *,     \
                                 const int *, double, const InnerTolRule &);                      \
    template void gmres_spmv_resid<T, T>(hipStream_t, const Tiles &, SysState *, double *,        \
                                         const int *, const int *, const T *, const double *,     \
                                         const T *, double *, int, int);                          \
    template void gmres_diff_partial<T, T, T>(hipStream_t, const Tiles &, SysState *, double *,   \
                                              const T *, const T *, T *, int);
BSPAI_KRYLOV_INST(double)
BSPAI_KRYLOV_INST(float)
#undef BSPAI_KRYLOV_INST
template void gmres_diff_partial<float, double, float>(hipStream_t, const Tiles &, SysState *,
                                                       double *, const float *, const double *,
                                                       float *, int);
template void gmres_spmv_resid<float, double>(hipStream_t, const Tiles &, SysState *, double *,
                                              const int *, const int *, const float *,
                                              const double *, const double *, double *, int, int);

// ================================================================== IR (FP64)
// Per-tile statistics of r = b - A x (sequential row sums, bit-identical to
// the reference) and of the diagnostics of src/ir.cpp:83-133:
//   part4[4t..4t+3] = max |r|, max |x|, max |b|, max componentwise ratio
//   norm partial of r in slot kIrSlotR2, nonfinite flag in flags[2t].
__global__ void __launch_bounds__(kTileThreads)
diagnostics_kernel(Tiles tiles, const int *__restrict__ ptr, const int *__restrict__ col,
                   const double *__restrict__ val, const double *__restrict__ b,
                   const double *__restrict__ x, double *__restrict__ r, IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    BSPAI_SYS_TILES
    if (!a.st[s].active)
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
    if (threadIdx.x == 0) {
        double *p4 = a.part4 + 4 * (std::int64_t)t;
        p4[0] = rinf;
        p4[1] = xinf;
        p4[2] = binf;
        p4[3] = cbe;
        a.flags[2 * t] = badv;
    }
    store_norm_partial(vals, cnt, slot_ptr(a.part, tiles.ntiles, kIrSlotR2), t, lds);
}

// ||A||_inf partials (row sums of |a_ij| in CSR order, src/ir.cpp:118-125)
// and ||b||_2 partials.
__global__ void __launch_bounds__(kTileThreads)
setup_partials_kernel(Tiles tiles, const int *__restrict__ ptr, const double *__restrict__ val,
                      const double *__restrict__ b, IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    BSPAI_SYS_TILES
    double mx = 0;
    double vals[kTilePer];
    int cnt = 0;
    for (int r = tiles.row0[t] + threadIdx.x; r < tiles.row1[t]; r += kTileThreads) {
        double rs = 0;
        for (int e = ptr[r]; e < ptr[r + 1]; ++e)
            rs = rs + fabs(val[e]);
        mx = fmax(mx, rs);
        vals[cnt++] = b[r];
    }
    mx = block_max(mx, lds);
    if (threadIdx.x == 0)
        a.part4[4 * (std::int64_t)t] = mx;
    store_norm_partial(vals, cnt, slot_ptr(a.part, tiles.ntiles, kIrSlotB), t, lds);
}

__global__ void __launch_bounds__(kTileThreads)
setup_decide_kernel(Tiles tiles, IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    const int s = blockIdx.x;
    const int t0 = tiles.sys_tile0[s], t1 = tiles.sys_tile1[s];
    const double ainf = reduce_max(a.part4, t0, t1, lds, 4, 0);
    const double bn = reduce_norm(slot_ptr(a.part, tiles.ntiles, kIrSlotB), t0, t1, lds);
    if (threadIdx.x == 0) {
        a.st[s].ainf = ainf;
        a.st[s].bnorm = bn;
    }
}

// Outer decision after the diagnostics of iterate `step` (src/ir.cpp:150-191);
// one block per system.
__global__ void __launch_bounds__(kTileThreads)
ir_decide_kernel(Tiles tiles, IrArgs a, int step) {
    __shared__ double lds[kTileThreads / kWave];
    const int s = blockIdx.x;
    IrSysState &st = a.st[s];
    if (!st.active)
        return;
    const int t0 = tiles.sys_tile0[s], t1 = tiles.sys_tile1[s];
    const double R = reduce_max(a.part4, t0, t1, lds, 4, 0);
    const double X = reduce_max(a.part4, t0, t1, lds, 4, 1);
    const double B = reduce_max(a.part4, t0, t1, lds, 4, 2);
    const double C = reduce_max(a.part4, t0, t1, lds, 4, 3);
    const double BAD = reduce_max(a.flags, t0, t1, lds, 2, 0);
    const double CH = step > 0 ? reduce_max(a.flags, t0, t1, lds, 2, 1) : 1.0;
    const double BADU = step > 0 ? reduce_max(a.upd_bad, t0, t1, lds, 1, 0) : 0.0;
    const double r2 = reduce_norm(slot_ptr(a.part, tiles.ntiles, kIrSlotR2), t0, t1, lds);
    if (threadIdx.x)
        return;
    if (step > 0)
        ++st.outer_steps;
    if (BADU != 0) {
        st.status = kIrNumericalFailure;
        st.active = 0;
        st.done = 1;
        return;
    }
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
    const bool conv =
        nbe <= a.be_tol && (a.relres_target <= 0 || r2 <= a.relres_target * st.bnorm);
    if (conv) {
        st.status = kIrSuccess;
        st.converged = 1;
        st.active = 0;
        st.done = 1;
    } else if (CH == 0) {
        st.status = kIrStagnation;
        st.active = 0;
        st.done = 1;
    }
}

// x += d for systems whose inner solve converged; per-tile "x changed" flag.
__global__ void __launch_bounds__(kTileThreads)
update_kernel(Tiles tiles, double *__restrict__ x, const double *__restrict__ d, IrArgs a) {
    __shared__ double lds[kTileThreads / kWave];
    BSPAI_SYS_TILES
    const IrSysState &st = a.st[s];
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
        a.flags[2 * t + 1] = changed;
        a.upd_bad[t] = badv;
    }
}

// Mixed path: the FP32 right side is r * 2^-e (e from ||r||_inf), so the
// inner solve sees O(1) data; the correction is scaled back exactly by 2^e.
__global__ void __launch_bounds__(kTileThreads)
scale_rhs_kernel(Tiles tiles, const double *__restrict__ r, float *__restrict__ rf, IrArgs a) {
    const int t = blockIdx.x, s = tiles.sys[t];
    const IrSysState &st = a.st[s];
    if (!st.active)
        return;
    int e = 0;
    if (st.rinf > 0)
        frexp(st.rinf, &e);
    const double sc = ldexp(1.0, -e);
    for (int row = tiles.row0[t] + threadIdx.x; row < tiles.row1[t]; row += kTileThreads)
        rf[row] = (float)(r[row] * sc);
}

// Absorb the inner outcome into the IR state (one thread per system).
template <typename T>
__global__ void absorb_inner_kernel(int nsys, KrylovArgs<T> k, IrArgs a, int scaled) {
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
    int e = 0;
    if (scaled && st.rinf > 0)
        frexp(st.rinf, &e);
    st.dscale = scaled ? ldexp(1.0, e) : 1.0;
    if (!g.converged) {
        st.status = kIrInnerFailure;
        st.inner_status = g.status;
        st.active = 0;
        st.done = 1;
    }
}

__global__ void ir_reset_kernel(int nsys, IrSysState *st) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys)
        return;
    IrSysState z = {};
    z.active = 1;
    z.dscale = 1.0;
    z.status = kIrIterationLimit;
    st[s] = z;
}

__global__ void ir_cap_kernel(int nsys, IrSysState *st) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys)
        return;
    if (st[s].active) {
        st[s].status = kIrIterationLimit;
        st[s].active = 0;
        st[s].done = 1;
    }
}

void ir_setup(hipStream_t s, const Tiles &tl, const int *ptr, const double *val, const double *b,
              const IrArgs &a) {
    hipLaunchKernelGGL(ir_reset_kernel, dim3(ceil_div(tl.nsys, 64)), dim3(64), 0, s, tl.nsys,
                       a.st);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(setup_partials_kernel, BSPAI_GRID, tl, ptr, val, b, a);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(setup_decide_kernel, dim3(tl.nsys), dim3(kTileThreads), 0, s, tl, a);
    HIP_LAUNCH_CHECK();
}
void ir_diagnostics(hipStream_t s, const Tiles &tl, const int *ptr, const int *col,
                    const double *val, const double *b, const double *x, double *r,
                    const IrArgs &a, int step) {
    hipLaunchKernelGGL(diagnostics_kernel, BSPAI_GRID, tl, ptr, col, val, b, x, r, a);
    HIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(ir_decide_kernel, dim3(tl.nsys), dim3(kTileThreads), 0, s, tl, a, step);
    HIP_LAUNCH_CHECK();
}
void ir_update(hipStream_t s, const Tiles &tl, double *x, const double *d, const IrArgs &a) {
    hipLaunchKernelGGL(update_kernel, BSPAI_GRID, tl, x, d, a);
    HIP_LAUNCH_CHECK();
}
void ir_scale_rhs(hipStream_t s, const Tiles &tl, const double *r, float *rf, const IrArgs &a) {
    hipLaunchKernelGGL(scale_rhs_kernel, BSPAI_GRID, tl, r, rf, a);
    HIP_LAUNCH_CHECK();
}
void ir_cap(hipStream_t s, int nsys, IrSysState *st) {
    hipLaunchKernelGGL(ir_cap_kernel, dim3(ceil_div(nsys, 64)), dim3(64), 0, s, nsys, st);
    HIP_LAUNCH_CHECK();
}
template <typename T>
void ir_absorb_inner(hipStream_t s, int nsys, const KrylovArgs<T> &k, const IrArgs &a, int scaled) {
    hipLaunchKernelGGL(absorb_inner_kernel<T>, dim3(ceil_div(nsys, 64)), dim3(64), 0, s, nsys, k,
                       a, scaled);
    HIP_LAUNCH_CHECK();
}
template void ir_absorb_inner<double>(hipStream_t, int, const KrylovArgs<double> &, const IrArgs &,
                                      int);
template void ir_absorb_inner<float>(hipStream_t, int, const KrylovArgs<float> &, const IrArgs &,
                                     int);

} // namespace bspai_hip
