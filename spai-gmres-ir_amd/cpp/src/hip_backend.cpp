// HIP back end of the bspai C++ API, implemented over the C ABI of
// libbspai_hip.so.  Argument validation and result conventions follow the
// reference's CUDA entry points (src/cuda_ir.cpp:9-23, 118-169): invalid
// options yield status invalid_input with a message; device/runtime failures
// throw std::runtime_error.
#include "bspai/hip_ir.hpp"

#include "bspai_hip.h"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>

namespace bspai {
namespace {

struct Ctx {
    bspai_hip_ctx *h = nullptr;
    explicit Ctx(int device) {
        const int rc = bspai_hip_ctx_create(device, nullptr, &h);
        if (rc != BSPAI_HIP_OK)
            throw std::runtime_error("bspai_hip_ctx_create failed (code " + std::to_string(rc) +
                                     ")");
    }
    ~Ctx() { bspai_hip_ctx_destroy(h); }
    void check(int rc) const {
        if (rc != BSPAI_HIP_OK)
            throw std::runtime_error(std::string("HIP backend: ") + bspai_hip_last_error(h));
    }
};

struct Mat {
    const Ctx &c;
    bspai_hip_matrix *h = nullptr;
    Mat(const Ctx &ctx, const CsrMatrix &a) : c(ctx) {
        if (a.rows() >= (std::size_t)INT_MAX || a.values().size() >= (std::size_t)INT_MAX)
            throw std::runtime_error("HIP backend: matrix exceeds int32 indexing");
        std::vector<std::int32_t> ptr(a.row_offsets().begin(), a.row_offsets().end());
        std::vector<std::int32_t> col(a.column_indices().begin(), a.column_indices().end());
        c.check(bspai_hip_matrix_create(c.h, (std::int64_t)a.rows(), (std::int64_t)a.values().size(),
                                        ptr.data(), col.data(), a.values().data(), BSPAI_HIP_HOST, 1,
                                        nullptr, &h));
    }
    Mat(const Ctx &ctx, bspai_hip_matrix *m) : c(ctx), h(m) {}
    ~Mat() { bspai_hip_matrix_destroy(h); }
};

struct Build {
    bspai_hip_spai *h = nullptr;
    ~Build() { bspai_hip_spai_destroy(h); }
};

bool valid(const CsrMatrix &a, const std::vector<double> &rhs, const IrOptions &o) {
    if (!a.rows() || a.rows() != a.cols() || rhs.size() != a.rows() || o.max_outer_steps > 20 ||
        !std::isfinite(o.gmres.tolerance) || o.gmres.tolerance <= 0 ||
        o.gmres.max_iterations > a.rows())
        return false;
    for (double v : rhs)
        if (!std::isfinite(v))
            return false;
    return true;
}

int prec(HipPrecision p) {
    return p == HipPrecision::mixed ? BSPAI_HIP_MIXED : BSPAI_HIP_FP64;
}

bspai_hip_ir_options ir_opts(const IrOptions &o, HipPrecision p, double relres_target) {
    bspai_hip_ir_options r{};
    r.spai_epsilon = o.spai.epsilon;
    r.spai_alpha = (std::int64_t)o.spai.alpha;
    r.spai_beta = (std::int64_t)o.spai.beta;
    r.gmres_tolerance = o.gmres.tolerance;
    r.gmres_max_iterations = (std::int64_t)o.gmres.max_iterations;
    r.max_outer_steps = (std::int64_t)o.max_outer_steps;
    r.precision = prec(p);
    // The reference's stop (normwise backward error <= 10u, src/ir.cpp:182)
    // alone, unless the caller also asks for ||b-Ax||_2 <= target ||b||_2.
    r.relres_target = relres_target;
    return r;
}

void fill_construction(const Ctx &c, const Build &b, std::size_t n, HipSpaiResult &out) {
    bspai_hip_spai_info info{};
    c.check(bspai_hip_spai_get_info(b.h, &info));
    SpaiResult &r = out.construction;
    r.status = static_cast<SpaiStatus>(info.status);
    r.message = info.message;
    out.telemetry.spai_ms = info.build_ms;
    out.telemetry.spai_passes = (std::size_t)info.passes;
    out.telemetry.spai_fits = (std::size_t)info.fits;
    out.telemetry.spai_growths = (std::size_t)info.growths;
    out.telemetry.p_nnz = (std::size_t)info.p_nnz;
    if (info.passes == 0) // rejected before any column was built
        return;
    r.scaling.resize(n);
    c.check(bspai_hip_spai_scaling(c.h, b.h, r.scaling.data()));
    std::vector<std::int32_t> st(n), fi(n), gr(n), cl(n);
    std::vector<double> res(n), piv(n);
    c.check(bspai_hip_spai_columns(c.h, b.h, st.data(), fi.data(), gr.data(), cl.data(),
                                   res.data(), piv.data()));
    std::vector<std::int64_t> pptr(n + 1);
    std::vector<std::int32_t> pidx((std::size_t)std::max<std::int64_t>(info.pattern_nnz, 1));
    std::vector<double> pcoef(pidx.size());
    c.check(bspai_hip_spai_patterns(c.h, b.h, pptr.data(), pidx.data(), pcoef.data()));
    r.columns.resize(n);
    std::vector<Entry> pe;
    pe.reserve((std::size_t)info.p_nnz);
    for (std::size_t k = 0; k < n; ++k) {
        SpaiColumn &col = r.columns[k];
        col.status = static_cast<SpaiStatus>(st[k]);
        col.fits = (std::size_t)fi[k];
        col.growths = (std::size_t)gr[k];
        col.score_clamps = (std::size_t)cl[k];
        col.reduced_residual = col.full_residual = res[k];
        col.qr_pivot_ratio = piv[k];
        for (std::int64_t q = pptr[k]; q < pptr[k + 1]; ++q) {
            col.pattern.push_back((std::size_t)pidx[q]);
            col.coefficients.push_back(pcoef[q]);
        }
    }
    if (r.status == SpaiStatus::success) {
        std::vector<std::int32_t> ptr(n + 1), idx((std::size_t)std::max<std::int64_t>(info.p_nnz, 1));
        std::vector<double> val(idx.size());
        c.check(bspai_hip_spai_preconditioner(c.h, b.h, ptr.data(), idx.data(), val.data()));
        for (std::size_t k = 0; k < n; ++k)
            for (std::int32_t q = ptr[k]; q < ptr[k + 1]; ++q)
                pe.push_back(Entry{k, (std::size_t)idx[q], val[q]});
        r.preconditioner = CsrMatrix::from_entries(n, n, std::move(pe));
    }
}

void fill_ir(const bspai_hip_ir_info &i, std::vector<double> x, IrResult &r) {
    r.status = static_cast<IrStatus>(i.status);
    r.converged = i.converged != 0;
    r.message = i.message;
    r.outer_steps = (std::size_t)i.outer_steps;
    r.inner_attempts = (std::size_t)i.inner_attempts;
    r.solution = std::move(x);
    r.history.clear();
    for (int h = 0; h < i.history_len; ++h) {
        IrIteration it;
        it.step = (std::size_t)h;
        it.has_inner = h > 0;
        it.diagnostics.residual_inf = i.history[h][0];
        it.diagnostics.residual_2 = i.history[h][1];
        it.diagnostics.normwise_backward_error = i.history[h][2];
        it.diagnostics.componentwise_backward_error = i.history[h][3];
        r.history.push_back(it);
    }
    r.inner_history.clear();
    for (std::int64_t k = 0; k < i.inner_attempts && k < BSPAI_HIP_MAX_OUTER; ++k) {
        GmresResult g;
        g.status = static_cast<GmresStatus>(i.inner_status[k]);
        g.converged = g.status == GmresStatus::success || g.status == GmresStatus::zero_rhs ||
                      g.status == GmresStatus::exact_initial_solution;
        g.iterations = (std::size_t)i.inner_iterations[k];
        r.inner_history.push_back(g);
    }
}

} // namespace

HipSpaiResult hip_construct_spai(const CsrMatrix &a, const SpaiOptions &o, HipPrecision p,
                                 int device) {
    HipSpaiResult out(a.rows());
    if (!a.rows() || a.rows() != a.cols() || !std::isfinite(o.epsilon) || o.epsilon <= 0) {
        out.construction.status = SpaiStatus::invalid_input;
        out.construction.message = "Require positive square dimensions and positive finite epsilon";
        return out;
    }
    Ctx c(device);
    Mat m(c, a);
    bspai_hip_spai_options so{o.epsilon, (std::int64_t)o.alpha, (std::int64_t)o.beta, prec(p)};
    Build b;
    c.check(bspai_hip_spai_build(c.h, m.h, &so, &b.h));
    fill_construction(c, b, a.rows(), out);
    return out;
}

GmresResult hip_left_preconditioned_gmres(const CsrMatrix &a, const CsrMatrix &p,
                                          const std::vector<double> &rhs,
                                          const GmresOptions &o,
                                          const std::vector<double> &x0, int device) {
    GmresResult r;
    const std::size_t n = a.rows();
    if (!n || a.cols() != n || p.rows() != n || p.cols() != n || rhs.size() != n ||
        (!x0.empty() && x0.size() != n) || !std::isfinite(o.tolerance) || o.tolerance <= 0 ||
        o.max_iterations > n) {
        r.message = "Invalid dimensions, initial guess, tolerance, or iteration cap";
        return r;
    }
    for (double v : rhs)
        if (!std::isfinite(v)) {
            r.message = "RHS must be finite";
            return r;
        }
    for (double v : x0)
        if (!std::isfinite(v)) {
            r.message = "Initial guess must be finite";
            return r;
        }
    Ctx c(device);
    Mat ma(c, a), mp(c, p);
    bspai_hip_gmres_options go{o.tolerance, (std::int64_t)o.max_iterations};
    bspai_hip_gmres_info info{};
    r.solution.assign(n, 0.0);
    c.check(bspai_hip_gmres(c.h, ma.h, mp.h, rhs.data(), x0.empty() ? nullptr : x0.data(), &go,
                            r.solution.data(), BSPAI_HIP_HOST, &info));
    r.status = static_cast<GmresStatus>(info.status);
    r.converged = info.converged != 0;
    r.iterations = (std::size_t)info.iterations;
    r.initial_original_residual = info.initial_original;
    r.initial_preconditioned_relative = info.initial_relative;
    r.final_original_residual = info.final_original;
    r.final_preconditioned_relative = info.final_relative;
    return r;
}

HipIrResult hip_refine_with_preconditioner(const CsrMatrix &a, const CsrMatrix &p,
                                           const std::vector<double> &rhs, const IrOptions &o,
                                           HipPrecision precision, int device,
                                           double relres_target) {
    HipIrResult out(a.rows());
    if (!valid(a, rhs, o) || p.rows() != a.rows() || p.cols() != a.cols()) {
        out.solver.message = "Invalid dimensions, finite inputs, inner options, or outer cap";
        return out;
    }
    Ctx c(device);
    Mat ma(c, a), mp(c, p);
    const bspai_hip_ir_options io = ir_opts(o, precision, relres_target);
    bspai_hip_ir_info info{};
    std::vector<double> x(a.rows());
    c.check(bspai_hip_ir_solve(c.h, ma.h, mp.h, rhs.data(), &io, x.data(), BSPAI_HIP_HOST,
                               &info));
    fill_ir(info, std::move(x), out.solver);
    out.telemetry.solve_ms = info.solve_ms;
    return out;
}

HipIrResult hip_gpu_spai_gmres_ir(const CsrMatrix &a, const std::vector<double> &rhs,
                                  const IrOptions &o, HipPrecision precision, int device,
                                  double relres_target) {
    HipIrResult out(a.rows());
    if (!valid(a, rhs, o)) {
        out.solver.message = "Invalid dimensions, finite inputs, inner options, or outer cap";
        return out;
    }
    Ctx c(device);
    Mat ma(c, a);
    bspai_hip_spai_options so{o.spai.epsilon, (std::int64_t)o.spai.alpha, (std::int64_t)o.spai.beta,
                              prec(precision)};
    Build b;
    c.check(bspai_hip_spai_build(c.h, ma.h, &so, &b.h));
    HipSpaiResult built(a.rows());
    fill_construction(c, b, a.rows(), built);
    out.telemetry = built.telemetry;
    out.solver.construction = std::move(built.construction);
    out.solver.construction_performed = true;
    if (out.solver.construction.status != SpaiStatus::success) {
        out.solver.status = IrStatus::construction_failure;
        out.solver.message = "SPAI failed: " +
                             std::string(spai_status_name(out.solver.construction.status)) +
                             ": " + out.solver.construction.message;
        return out;
    }
    bspai_hip_matrix *pm = nullptr;
    c.check(bspai_hip_spai_matrix(c.h, b.h, 0, &pm));
    Mat mp(c, pm);
    const bspai_hip_ir_options io = ir_opts(o, precision, relres_target);
    bspai_hip_ir_info info{};
    std::vector<double> x(a.rows());
    c.check(bspai_hip_ir_solve(c.h, ma.h, mp.h, rhs.data(), &io, x.data(), BSPAI_HIP_HOST,
                               &info));
    fill_ir(info, std::move(x), out.solver);
    out.telemetry.solve_ms = info.solve_ms;
    return out;
}

} // namespace bspai
