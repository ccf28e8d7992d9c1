// Device state and launchers of the GMRES / IR kernels (internal interface).
#pragma once

#include "kernels.hpp"

namespace bspai_hip {

// GMRES statuses mirror bspai::GmresStatus (reference include/bspai/gmres.hpp:10-20).
enum : int {
    kGmSuccess = 0,
    kGmZeroRhs = 1,
    kGmExactInitial = 2,
    kGmInvalidInput = 3,
    kGmZeroPreconditionedRhs = 4,
    kGmBreakdown = 5,
    kGmIterationLimit = 6,
    kGmNonconvergence = 7,
    kGmNumericalFailure = 8,
};
// IR statuses mirror bspai::IrStatus (reference include/bspai/ir.hpp:9-17).
enum : int {
    kIrSuccess = 0,
    kIrInvalidInput = 1,
    kIrConstructionFailure = 2,
    kIrInnerFailure = 3,
    kIrStagnation = 4,
    kIrIterationLimit = 5,
    kIrNumericalFailure = 6,
};
// Partial-sum slots of the GMRES reductions (each 2 doubles per tile).
enum : int {
    kSlotNorm = 0, // ||w|| after orthogonalization
    kSlotRhs = 1,
    kSlotOrig = 2,
    kSlotDen = 3,
    kSlotBeta = 4,
    kSlotChkOrig = 5,
    kSlotChkRel = 6,
    kNumSlots = 7,
};
enum : int { kPhaseInit = 0, kPhaseDen = 1, kPhaseBeta = 2, kPhaseCheck = 3 };
enum : int { kIrSlotR2 = 0, kIrSlotB = 1, kIrNumSlots = 2 };

constexpr int kMaxOuterHistory = 24; // >= outer cap 20 + initial

// Column k of the rotated Hessenberg holds rows 0..k+1, packed.
__host__ __device__ inline std::int64_t hcol_offset(int k) {
    return (std::int64_t)k * (k + 3) / 2;
}
inline std::int64_t hessenberg_size(int cap) {
    return hcol_offset(cap) + 2;
}
// Strictly lower Gram triangle L(j, i) = <V_j, V_i>, i < j < K, packed by
// columns: column i holds rows i+1..K-1, so the lanes of a wave that own
// consecutive rows j read consecutive words.  K is the number of Krylov steps
// the workspace currently holds (it grows; a larger K re-lays the columns).
__host__ __device__ inline std::int64_t gcol_offset(int i, int K) {
    return (std::int64_t)i * K - (std::int64_t)i * (i + 1) / 2;
}
__host__ __device__ inline std::int64_t gram_index(int j, int i, int K) {
    return gcol_offset(i, K) + (j - i - 1);
}
inline std::int64_t gram_size(int K) {
    return (std::int64_t)K * (K - 1) / 2 + 1;
}

struct SysState {
    double tol;
    double den, beta, hnorm, est;
    double init_orig, init_rel, final_orig, final_rel;
    int nrows;
    int scap; // this system's Krylov cap: min(workspace cap, nrows)
    int participating, active, status, converged, iterations;
    int triggered, ccount, zero_pivot, breakdown;
    int rot_bad; // a previous rotation made column k non-finite
    int rhs_zero, x0_is_solution, has_solution;
    // run[k & 1]: Arnoldi step k does work for this system (active and not
    // held for a deferred verification).  Step k's kernels read run[k & 1];
    // the step's normalize writer and the check of step k write run[(k+1) & 1],
    // so no launch reads a flag that it also writes.
    int run[2];
};

// Largest k solved by the LDS-staged Gram solve (beyond it, the one-wave kernel).
constexpr int kGramSolveLdsMaxK = 511;

template <typename T> struct KrylovArgs {
    SysState *st;
    T *const *V; // device array of basis-vector pointers (cap+1)
    T *w;
    double *H, *cs, *sn, *g, *y;
    std::int64_t hstride;
    int cap;
    double *part; // kNumSlots slots x 2 doubles per tile
    double *x;    // candidate / solution (FP64)
    const double *x0;
    // Compact Gram-Schmidt workspace: gp[v][tile] per-tile partials of
    // <V_i, w> (v = 2i) and <V_i, V_k> (v = 2i+1); per system q = V^T w,
    // h = MGS coefficients ([cap+1] each) and the packed Gram triangle.  H,
    // the Gram triangle and gp are sized for the Krylov steps reached so far
    // (hstride, gstride per system) and grow with them.
    double *gp, *q, *h, *gram;
    std::int64_t gstride;
    int gsteps; // K of the Gram layout
    // 1: a triggered step holds the system (run cleared) until the host
    // launches the verification; 0: the verification follows every step.
    int defer;
    // Replayed graphs: the step kernels take their step as *kdev + the
    // offset they were launched with (nullptr: the argument is the step).
    const int *kdev;
};

// The Arnoldi step a step kernel works on (see KrylovArgs::kdev).
template <typename T> __device__ inline int step_index(const KrylovArgs<T> &a, int k) {
    return a.kdev ? *a.kdev + k : k;
}

struct IrSysState {
    double ainf, bnorm, rinf, r2, nbe, cbe, xinf;
    double dscale;
    int active, done, status, converged, outer_steps, inner_attempts, inner_status;
    int history_len;
};

// Inner tolerance of a refinement step (theta > 0): the larger of the
// caller's tolerance and theta times the contraction the step still needs
// (be_tol / backward error, and relres_target / relative residual when set).
struct InnerTolRule {
    double theta = 0, be_tol = 0, relres_target = 0;
};

struct IrArgs {
    IrSysState *st;
    double *part;    // kIrNumSlots slots x 2 doubles per tile (norm partials)
    double *part4;   // 4 per tile: max |r|, max |x|, max |b|, max cbe (or ||A||_inf)
    double *flags;   // 2 per tile: nonfinite residual, x changed
    double *upd_bad; // 1 per tile: nonfinite update
    double *history; // [nsys][kMaxOuterHistory][4]: rinf, r2, nbe, cbe
    int *inner;      // [nsys][kMaxOuterHistory][3]: status, converged, iterations
    double be_tol, relres_target;
};

// Compact MGS of iteration k: project (V^T w and the Gram row), reduce over
// tiles, triangular solve for h (+ older rotations), update w -= V h.
// Copies the Gram triangles of nsys systems from a K0-step to a K-step layout.
void gmres_gram_relayout(hipStream_t s, int nsys, const double *src, std::int64_t sstride, int K0,
                         double *dst, std::int64_t dstride, int K);
// k: the step (a.kdev == nullptr) or the offset to *a.kdev; kmax: the
// largest step the launch may work on, which sizes the grids (k itself for
// a direct launch, the Krylov cap - 1 for a replayed graph).
template <typename T>
void gmres_project(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax);
template <typename T>
void gmres_gram_reduce(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax);
template <typename T>
void gmres_gram_solve(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax);
// Gram solves for launches that may see steps up to kmax <= kGramSolveLdsMaxK
// run inside the update pass (gmres_gram_solve then launches nothing).
bool gmres_gram_solve_fused(int kmax);
template <typename T>
void gmres_gram_update(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k, int kmax);
template <typename T>
void gmres_normalize_givens(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int k);
// *kdev = k (one thread): the step a replayed graph starts at.
void gmres_set_step(hipStream_t s, int *kdev, int k);
// *kdev += d (one thread): the last node of a replayed graph.
void gmres_advance_step(hipStream_t s, int *kdev, int d);
template <typename T> void gmres_candidate(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a);
template <typename T>
void gmres_start_vector(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, const T *r0);
template <typename T> void gmres_finish(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a);
template <typename T>
void gmres_decide(hipStream_t s, const Tiles &tl, const KrylovArgs<T> &a, int phase, int right);
template <typename T>
void gmres_reset(hipStream_t s, int nsys, const KrylovArgs<T> &k, const IrSysState *ir,
                 const int *nrows, double tol, const InnerTolRule &rule);
template <typename TV, typename TU>
void gmres_spmv_resid(hipStream_t s, const Tiles &tl, SysState *st, double *part, const int *ptr,
                      const int *col, const TV *val, const double *x, const TU *u, double *y,
                      int slot, int gate);
template <typename TU, typename TV, typename TO>
void gmres_diff_partial(hipStream_t s, const Tiles &tl, SysState *st, double *part, const TU *u,
                        const TV *v, TO *out, int slot);

void ir_setup(hipStream_t s, const Tiles &tl, const int *ptr, const double *val, const double *b,
              const IrArgs &a);
void ir_diagnostics(hipStream_t s, const Tiles &tl, const int *ptr, const int *col,
                    const double *val, const double *b, const double *x, double *r,
                    const IrArgs &a, int step);
void ir_update(hipStream_t s, const Tiles &tl, double *x, const double *d, const IrArgs &a);
void ir_scale_rhs(hipStream_t s, const Tiles &tl, const double *r, float *rf, const IrArgs &a);
void ir_cap(hipStream_t s, int nsys, IrSysState *st);
template <typename T>
void ir_absorb_inner(hipStream_t s, int nsys, const KrylovArgs<T> &k, const IrArgs &a, int scaled);

} // namespace bspai_hip
