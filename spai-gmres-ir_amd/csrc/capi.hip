// extern "C" boundary of libbspai_hip.so (declared in include/bspai_hip.h).
// No exception crosses it: every entry point returns a status code and keeps
// the message for bspai_hip_last_error().
#include "../../include/bspai_hip.h"
#include "engine.hpp"
#include "mmgpu.hpp"
#include "mmio_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

using namespace bspai_hip;

struct bspai_hip_ctx {
    Context c;
};
struct bspai_hip_matrix {
    DevCsr a;
};
struct bspai_hip_mm_gpu {
    MmGpu m;
};
struct bspai_hip_spai {
    std::unique_ptr<SpaiBuild> b;
};

namespace {

// Every entry point runs on the context's device and gives the caller's
// current device back on return, as the reference's DeviceGuard does
// (src/cuda_gmres.cu:78-87).
struct DeviceRestore {
    int prev = -1;
    ~DeviceRestore() {
        if (prev >= 0)
            (void)hipSetDevice(prev);
    }
};

template <typename F> int guard(bspai_hip_ctx *ctx, F &&f) {
    DeviceRestore restore;
    try {
        if (ctx) {
            int cur = -1;
            HIP_CHECK(hipGetDevice(&cur));
            restore.prev = cur;
            if (cur != ctx->c.device)
                HIP_CHECK(hipSetDevice(ctx->c.device));
        }
        f();
        return BSPAI_HIP_OK;
    } catch (const std::invalid_argument &e) {
        if (ctx)
            ctx->c.last_error = e.what();
        return BSPAI_HIP_EINVAL;
    } catch (const std::bad_alloc &e) {
        if (ctx)
            ctx->c.last_error = e.what();
        return BSPAI_HIP_ENOMEM;
    } catch (const std::exception &e) {
        if (ctx)
            ctx->c.last_error = e.what();
        const std::string m = e.what();
        return m.find("hipMalloc") != std::string::npos ? BSPAI_HIP_ENOMEM : BSPAI_HIP_EDEVICE;
    }
}

void require(bool ok, const char *msg) {
    if (!ok)
        throw std::invalid_argument(msg);
}

// Copy-in helper: returns a device pointer to n elements (uploading host data).
template <typename T> struct In {
    DeviceArray<T> own;
    const T *p = nullptr;
    In(const Context &c, const T *src, std::size_t n, int where) {
        if (where == BSPAI_HIP_DEVICE || n == 0) {
            p = src;
            return;
        }
        own.resize(n);
        HIP_CHECK(hipMemcpyAsync(own.data(), src, n * sizeof(T), hipMemcpyHostToDevice, c.stream));
        p = own.data();
    }
};
template <typename T> struct Out {
    DeviceArray<T> own;
    T *p = nullptr, *dst;
    std::size_t n;
    int where;
    const Context &c;
    Out(const Context &c_, T *d, std::size_t n_, int w) : dst(d), n(n_), where(w), c(c_) {
        if (where == BSPAI_HIP_DEVICE) {
            p = d;
            return;
        }
        own.resize(std::max<std::size_t>(n, 1));
        p = own.data();
    }
    // Host outputs are copied back and the call is synchronous; device
    // outputs are stream-ordered on the context stream (the caller syncs).
    void finish() {
        if (where == BSPAI_HIP_DEVICE)
            return;
        if (n)
            HIP_CHECK(hipMemcpyAsync(dst, p, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
        HIP_CHECK(hipStreamSynchronize(c.stream));
    }
};

void set_msg(char *dst, const std::string &m) {
    std::strncpy(dst, m.c_str(), 255);
    dst[255] = 0;
}

const char *kGmNames[] = {"success",
                          "zero_rhs",
                          "exact_initial_solution",
                          "invalid_input",
                          "zero_preconditioned_rhs",
                          "breakdown",
                          "iteration_limit",
                          "nonconvergence",
                          "numerical_failure"};
const char *kSpaiNames[] = {"success",           "not_attempted",    "invalid_input",
                            "rank_deficient",    "numerical_failure", "empty_candidates",
                            "zero_denominator",  "invalid_score",     "no_growth",
                            "iteration_limit"};

const char *gm_message(int status) {
    switch (status) {
    case kGmZeroPreconditionedRhs:
        return "Zero preconditioned RHS with nonzero original initial residual";
    case kGmBreakdown:
        return "Arnoldi breakdown candidate failed recomputed residual criterion";
    case kGmIterationLimit:
    case kGmNonconvergence:
        return "Iteration cap reached without verified convergence";
    case kGmNumericalFailure:
        return "Nonfinite GMRES arithmetic";
    default:
        return "";
    }
}

void fill_ir_info(bspai_hip_ir_info &info, const IrResultHost &r) {
    const IrSysState &s = r.st;
    info.status = s.status;
    info.converged = s.converged;
    info.outer_steps = s.outer_steps;
    info.inner_attempts = s.inner_attempts;
    for (int i = 0; i < BSPAI_HIP_MAX_OUTER; ++i) {
        info.inner_status[i] = i < s.inner_attempts ? r.inner[i][0] : 0;
        info.inner_iterations[i] = i < s.inner_attempts ? r.inner[i][2] : 0;
    }
    info.history_len = s.history_len;
    for (int h = 0; h < BSPAI_HIP_MAX_OUTER; ++h)
        for (int q = 0; q < 4; ++q)
            info.history[h][q] = h < s.history_len ? r.history[h][q] : 0.0;
    info.residual_inf = s.rinf;
    info.residual_2 = s.r2;
    info.normwise_backward_error = s.nbe;
    info.componentwise_backward_error = s.cbe;
    std::string m;
    switch (s.status) {
    case kIrInnerFailure:
        m = std::string("Correction GMRES failed: ") + kGmNames[s.inner_status] + ": " +
            gm_message(s.inner_status);
        break;
    case kIrStagnation:
        m = "Unchanged outer iterate above original backward-error tolerance";
        break;
    case kIrIterationLimit:
        m = "Outer correction cap reached above original backward-error tolerance";
        break;
    case kIrNumericalFailure:
        m = "Nonfinite outer residual/backward-error arithmetic";
        break;
    default:
        break;
    }
    set_msg(info.message, m);
}

void validate_ir(const bspai_hip_matrix *a, const bspai_hip_ir_options *o) {
    require(a && o, "null matrix or options");
    require(o->max_outer_steps >= 0 && o->max_outer_steps <= 20,
            "Invalid dimensions, finite inputs, inner options, or outer cap");
    require(std::isfinite(o->gmres_tolerance) && o->gmres_tolerance > 0,
            "Invalid dimensions, finite inputs, inner options, or outer cap");
    std::int64_t minrows = a->a.n;
    const Layout &L = *a->a.layout;
    for (int s = 0; s < L.nsys; ++s)
        minrows = std::min(minrows, L.offsets[s + 1] - L.offsets[s]);
    require(o->gmres_max_iterations >= 0 && o->gmres_max_iterations <= minrows,
            "Invalid dimensions, finite inputs, inner options, or outer cap");
    require(o->precision == BSPAI_HIP_FP64 || o->precision == BSPAI_HIP_MIXED,
            "precision must be BSPAI_HIP_FP64 or BSPAI_HIP_MIXED");
}

IrOptionsC ir_opts(const bspai_hip_ir_options *o) {
    IrOptionsC r;
    r.tol = o->gmres_tolerance;
    r.cap = (int)o->gmres_max_iterations;
    r.outer = (int)o->max_outer_steps;
    r.precision = o->precision;
    r.relres_target = o->relres_target;
    return r;
}

// FP64 copy of an FP32-valued CSR (a mixed-precision P).
std::unique_ptr<DevCsr> widen(Context &c, const DevCsr &src) {
    auto p = std::make_unique<DevCsr>();
    p->n = src.n;
    p->nnz = src.nnz;
    p->ptr.resize(p->n + 1);
    p->col.resize(std::max<std::int64_t>(p->nnz, 1));
    p->val.resize(std::max<std::int64_t>(p->nnz, 1));
    HIP_CHECK(hipMemcpyAsync(p->ptr.data(), src.ptr.data(), (p->n + 1) * sizeof(int),
                             hipMemcpyDeviceToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(p->col.data(), src.col.data(), p->nnz * sizeof(int),
                             hipMemcpyDeviceToDevice, c.stream));
    launch_convert<float, double>(c.stream, p->nnz, src.val32.data(), p->val.data());
    p->has64 = true;
    p->lanes = src.lanes;
    p->layout = src.layout;
    return p;
}

// IR of A x = b with the preconditioner P (values in FP64 and/or FP32).
void run_ir(bspai_hip_ctx *ctx, const DevCsr &A0, const DevCsr &P0, const double *b,
            const bspai_hip_ir_options *o, double *x, int where, bspai_hip_ir_info *info) {
    Context &c = ctx->c;
    require(P0.n == A0.n, "preconditioner dimension mismatch");
    const int prec = o->precision;
    DevCsr &A = const_cast<DevCsr &>(A0);
    DevCsr *P = const_cast<DevCsr *>(&P0);
    std::unique_ptr<DevCsr> tmp;
    if (prec == BSPAI_HIP_FP64 && !P->has64) {
        tmp = widen(c, *P);
        P = tmp.get();
    }
    if (prec == BSPAI_HIP_MIXED) {
        A.ensure32(c);
        P->ensure32(c);
    }
    In<double> bin(c, b, A.n, where);
    Out<double> xo(c, x, A.n, where);
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    auto res = ir_run(c, A, A, *P, kLeft, bin.p, ir_opts(o), xo.p);
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    xo.finish();
    HIP_CHECK(hipStreamSynchronize(c.stream));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    for (std::size_t s = 0; s < res.size(); ++s) {
        fill_ir_info(info[s], res[s]);
        info[s].solve_ms = ms;
    }
}

} // namespace

extern "C" {

int bspai_hip_abi_version(void) {
    return BSPAI_HIP_ABI_VERSION;
}

int bspai_hip_ctx_create(int device, void *stream, bspai_hip_ctx **out) {
    if (!out)
        return BSPAI_HIP_EINVAL;
    *out = nullptr;
    try {
        auto *ctx = new bspai_hip_ctx;
        ctx->c.device = device;
        const int rc = guard(ctx, [&] {
            if (stream) {
                ctx->c.stream = static_cast<hipStream_t>(stream);
            } else {
                HIP_CHECK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
                ctx->c.own_stream = true;
            }
            HIP_CHECK(hipEventCreate(&ctx->c.ev0));
            HIP_CHECK(hipEventCreate(&ctx->c.ev1));
            // the allocation-free entry point's scratch address on this device
            noalloc_setup(ctx->c);
        });
        if (rc != BSPAI_HIP_OK) {
            delete ctx;
            return rc;
        }
        *out = ctx;
        return BSPAI_HIP_OK;
    } catch (...) {
        return BSPAI_HIP_EDEVICE;
    }
}

void bspai_hip_ctx_destroy(bspai_hip_ctx *ctx) {
    if (!ctx)
        return;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.ev0)
        (void)hipEventDestroy(ctx->c.ev0);
    if (ctx->c.ev1)
        (void)hipEventDestroy(ctx->c.ev1);
    if (ctx->c.own_stream && ctx->c.stream)
        (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

const char *bspai_hip_last_error(const bspai_hip_ctx *ctx) {
    return ctx ? ctx->c.last_error.c_str() : "null context";
}

int bspai_hip_synchronize(bspai_hip_ctx *ctx) {
    return guard(ctx, [&] { HIP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}

int bspai_hip_ctx_set_timing(bspai_hip_ctx *ctx, int enable) {
    if (!ctx)
        return BSPAI_HIP_EINVAL;
    return guard(ctx, [&] {
        HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
        ctx->c.timer.clear();
        ctx->c.timer.on = enable != 0;
        ctx->c.timer.spai_only = enable == 2;
    });
}

int bspai_hip_ctx_graph_stats(const bspai_hip_ctx *ctx, int64_t *captured, int64_t *replayed) {
    if (!ctx)
        return BSPAI_HIP_EINVAL;
    if (captured)
        *captured = ctx->c.graphs_captured;
    if (replayed)
        *replayed = ctx->c.graphs_replayed;
    return BSPAI_HIP_OK;
}

int bspai_hip_ctx_kernel_times(bspai_hip_ctx *ctx, bspai_hip_kernel_time *out, int capacity,
                               int *count) {
    if (!ctx || !count || capacity < 0 || (capacity > 0 && !out))
        return BSPAI_HIP_EINVAL;
    return guard(ctx, [&] {
        const auto totals = ctx->c.timer.collect();
        *count = (int)totals.size();
        for (int i = 0; i < capacity && i < (int)totals.size(); ++i) {
            const auto &t = totals[i];
            std::memset(out[i].name, 0, sizeof(out[i].name));
            std::strncpy(out[i].name, t.name.c_str(), sizeof(out[i].name) - 1);
            out[i].launches = t.launches;
            out[i].total_ms = t.ms;
            out[i].bytes = t.bytes;
            out[i].flops = t.flops;
        }
    });
}

int bspai_hip_matrix_create(bspai_hip_ctx *ctx, int64_t n, int64_t nnz, const int32_t *ptr,
                            const int32_t *col, const double *val, int where, int nsys,
                            const int64_t *sys_offsets, bspai_hip_matrix **out) {
    return guard(ctx, [&] {
        require(out && ctx, "null argument");
        require(n > 0 && n < INT_MAX && nnz >= 0 && nnz < INT_MAX, "matrix size out of range");
        require(nsys >= 1 && (nsys == 1 || sys_offsets), "invalid system layout");
        Context &c = ctx->c;
        auto m = std::make_unique<bspai_hip_matrix>();
        DevCsr &a = m->a;
        a.n = n;
        a.nnz = nnz;
        a.ptr.resize(n + 1);
        a.col.resize(std::max<int64_t>(nnz, 1));
        a.val.resize(std::max<int64_t>(nnz, 1));
        const hipMemcpyKind kind =
            where == BSPAI_HIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        std::vector<std::int64_t> offs;
        if (nsys == 1)
            offs = {0, n};
        else {
            offs.assign(sys_offsets, sys_offsets + nsys + 1);
            require(offs.front() == 0 && offs.back() == n, "system offsets must span [0, n]");
            for (int s = 0; s < nsys; ++s)
                require(offs[s] < offs[s + 1], "systems must be nonempty");
        }
        if (where == BSPAI_HIP_HOST) {
            // The canonical form every kernel relies on (the reference's
            // CsrMatrix is canonical by construction): offsets 0..nnz
            // nondecreasing, columns strictly increasing within a row and
            // inside the row's system, finite values.
            // (all offsets first, so that no row walk leaves the arrays)
            require(ptr[0] == 0 && ptr[n] == nnz,
                    "row pointers must span [0, nnz] and be nondecreasing");
            for (int64_t r = 0; r < n; ++r)
                require(ptr[r] <= ptr[r + 1],
                        "row pointers must span [0, nnz] and be nondecreasing");
            int sys = 0;
            for (int64_t r = 0; r < n; ++r) {
                while (r >= offs[sys + 1])
                    ++sys;
                for (int64_t q = ptr[r]; q < ptr[r + 1]; ++q) {
                    require(col[q] >= 0 && col[q] < n, "column index out of range");
                    require(q == ptr[r] || col[q] > col[q - 1],
                            "column indices must be strictly increasing within each row");
                    require(col[q] >= offs[sys] && col[q] < offs[sys + 1],
                            "entry outside its row's system (a batch must be block-diagonal)");
                    require(std::isfinite(val[q]), "matrix values must be finite");
                }
            }
        }
        HIP_CHECK(hipMemcpyAsync(a.ptr.data(), ptr, (n + 1) * sizeof(int), kind, c.stream));
        HIP_CHECK(hipMemcpyAsync(a.col.data(), col, nnz * sizeof(int), kind, c.stream));
        HIP_CHECK(hipMemcpyAsync(a.val.data(), val, nnz * sizeof(double), kind, c.stream));
        if (where == BSPAI_HIP_DEVICE)
            check_device_csr(c, n, nnz, a.ptr.data(), a.col.data(), a.val.data(), offs);
        a.layout = std::make_shared<Layout>();
        a.layout->build(c, n, offs);
        a.lanes = choose_lanes(n, nnz);
        HIP_CHECK(hipStreamSynchronize(c.stream));
        *out = m.release();
    });
}

void bspai_hip_matrix_destroy(bspai_hip_matrix *m) {
    delete m;
}

int bspai_hip_matrix_info(const bspai_hip_matrix *m, int64_t *n, int64_t *nnz, int *nsys) {
    if (!m)
        return BSPAI_HIP_EINVAL;
    if (n)
        *n = m->a.n;
    if (nnz)
        *nnz = m->a.nnz;
    if (nsys)
        *nsys = m->a.layout->nsys;
    return BSPAI_HIP_OK;
}

int bspai_hip_matrix_set_values(bspai_hip_ctx *ctx, bspai_hip_matrix *m, const double *val,
                                int where) {
    return guard(ctx, [&] {
        require(m && val, "null argument");
        const hipMemcpyKind kind =
            where == BSPAI_HIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        HIP_CHECK(hipMemcpyAsync(m->a.val.data(), val, m->a.nnz * sizeof(double), kind,
                                 ctx->c.stream));
        m->a.values_changed();
        HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    });
}

int bspai_hip_matrix_download(bspai_hip_ctx *ctx, const bspai_hip_matrix *m, int32_t *ptr,
                              int32_t *col, double *val) {
    return guard(ctx, [&] {
        require(m, "null matrix");
        const Context &c = ctx->c;
        if (ptr)
            HIP_CHECK(hipMemcpyAsync(ptr, m->a.ptr.data(), (m->a.n + 1) * sizeof(int),
                                     hipMemcpyDeviceToHost, c.stream));
        if (col)
            HIP_CHECK(hipMemcpyAsync(col, m->a.col.data(), m->a.nnz * sizeof(int),
                                     hipMemcpyDeviceToHost, c.stream));
        if (val)
            HIP_CHECK(hipMemcpyAsync(val, m->a.val.data(), m->a.nnz * sizeof(double),
                                     hipMemcpyDeviceToHost, c.stream));
        HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int bspai_hip_csr_from_entries(bspai_hip_ctx *ctx, int64_t rows, int64_t cols, int64_t count,
                               const int64_t *row, const int64_t *col, const double *val,
                               int where, int32_t *ptr, int32_t *col_out, double *val_out,
                               int out_where, int64_t *nnz) {
    return guard(ctx, [&] {
        require(ctx && ptr && nnz, "null argument");
        require(count == 0 || (row && col && val && col_out && val_out), "null argument");
        require(rows >= 0 && rows < INT_MAX && cols >= 0 && cols < INT_MAX && count >= 0 &&
                    count < INT_MAX,
                "matrix size out of range");
        Context &c = ctx->c;
        In<int64_t> r(c, row, count, where), cl(c, col, count, where);
        In<double> v(c, val, count, where);
        if (out_where == BSPAI_HIP_DEVICE) {
            *nnz = csr_from_entries(c, rows, cols, count, r.p, cl.p, v.p, ptr, col_out, val_out);
            return;
        }
        CanonWork &w = c.canon;
        w.out_ptr.resize(rows + 1);
        w.out_col.resize(std::max<int64_t>(count, 1));
        w.out_val.resize(std::max<int64_t>(count, 1));
        const int64_t k = csr_from_entries(c, rows, cols, count, r.p, cl.p, v.p, w.out_ptr.data(),
                                           w.out_col.data(), w.out_val.data());
        HIP_CHECK(hipMemcpyAsync(ptr, w.out_ptr.data(), (rows + 1) * sizeof(int),
                                 hipMemcpyDeviceToHost, c.stream));
        if (k > 0) {
            HIP_CHECK(hipMemcpyAsync(col_out, w.out_col.data(), k * sizeof(int),
                                     hipMemcpyDeviceToHost, c.stream));
            HIP_CHECK(hipMemcpyAsync(val_out, w.out_val.data(), k * sizeof(double),
                                     hipMemcpyDeviceToHost, c.stream));
        }
        HIP_CHECK(hipStreamSynchronize(c.stream));
        *nnz = k;
    });
}

int bspai_hip_csr_from_entries_noalloc(bspai_hip_ctx *ctx, int64_t rows, int64_t cols,
                                       int64_t count, const int64_t *row, const int64_t *col,
                                       const double *val, int32_t *ptr, int32_t *col_out,
                                       double *val_out, int64_t *nnz) {
    return guard(ctx, [&] {
        require(ctx && ptr && nnz, "null argument");
        require(count == 0 || (row && col && val && col_out && val_out), "null argument");
        require(rows >= 0 && rows < INT_MAX && cols >= 0 && cols < INT_MAX && count >= 0 &&
                    count < INT_MAX,
                "matrix size out of range");
        // every array must be device memory: nothing is staged through the host
        auto on_device = [](const void *p) {
            if (!p)
                return true;
            hipPointerAttribute_t at{};
            const hipError_t e = hipPointerGetAttributes(&at, p);
            (void)hipGetLastError(); // an unknown (host) pointer is not an error to keep
            return e == hipSuccess && (at.type == hipMemoryTypeDevice || at.isManaged);
        };
        require(on_device(row) && on_device(col) && on_device(val) && on_device(ptr) &&
                    on_device(col_out) && on_device(val_out),
                "allocation-free canonicalization takes device arrays only");
        *nnz = csr_from_entries_noalloc(ctx->c, rows, cols, count, row, col, val, ptr, col_out,
                                        val_out);
    });
}

int bspai_hip_matrix_transpose(bspai_hip_ctx *ctx, const bspai_hip_matrix *m,
                               bspai_hip_matrix **out) {
    return guard(ctx, [&] {
        require(m && out, "null argument");
        Context &c = ctx->c;
        const DevCsr &a = m->a;
        auto t = std::make_unique<bspai_hip_matrix>();
        DevCsr &b = t->a;
        const int n = (int)a.n;
        b.n = a.n;
        b.nnz = a.nnz;
        b.ptr.resize(n + 1);
        b.col.resize(std::max<int64_t>(a.nnz, 1));
        b.val.resize(std::max<int64_t>(a.nnz, 1));
        DeviceArray<int> cursor(n + 1), tmp(scan_tmp_ints(n + 1));
        HIP_CHECK(hipMemsetAsync(b.ptr.data(), 0, (n + 1) * sizeof(int), c.stream));
        launch_count_cols(c.stream, a.nnz, a.col.data(), b.ptr.data());
        scan_inclusive_int(c.stream, b.ptr.data(), n + 1, tmp.data());
        HIP_CHECK(hipMemcpyAsync(cursor.data(), b.ptr.data(), (n + 1) * sizeof(int),
                                 hipMemcpyDeviceToDevice, c.stream));
        launch_transpose_fill<double>(c.stream, n, a.ptr.data(), a.col.data(), a.val.data(),
                                      b.ptr.data(), cursor.data(), b.col.data(), b.val.data());
        b.layout = a.layout; // A^T keeps the block-diagonal system layout
        b.lanes = choose_lanes(b.n, b.nnz);
        HIP_CHECK(hipStreamSynchronize(c.stream));
        *out = t.release();
    });
}

namespace {
void check_spmv_lanes(int lanes) {
    require(lanes == 0 || lanes == -1 || lanes == BSPAI_HIP_SPMV_SLICED ||
                lanes == BSPAI_HIP_SPMV_SLICED16 ||
                (lanes > 0 && lanes <= 64 && (lanes & (lanes - 1)) == 0),
            "lanes: 0, -1, 2^k <= 64, BSPAI_HIP_SPMV_SLICED or BSPAI_HIP_SPMV_SLICED16");
}
// The sliced layout a lanes argument asks for (0: none), built on first use.
int sliced_kind(const Context &c, const DevCsr &a, int lanes) {
    if (lanes != BSPAI_HIP_SPMV_SLICED && lanes != BSPAI_HIP_SPMV_SLICED16)
        return 0;
    require(a.ensure_sliced(c),
            "sliced SpMV: a row has more than 255 entries, or the slices would hold more than "
            "3x the entries");
    if (lanes == BSPAI_HIP_SPMV_SLICED16) {
        require(a.s16 > 0, "sliced SpMV: a column offset col - row does not fit 16 bits");
        return 2;
    }
    return 1;
}
} // namespace

int bspai_hip_spmv(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const double *x, double *y,
                   int where, int lanes) {
    return guard(ctx, [&] {
        require(a && x && y, "null argument");
        check_spmv_lanes(lanes);
        Context &c = ctx->c;
        const int kind = sliced_kind(c, a->a, lanes);
        In<double> xi(c, x, a->a.n, where);
        Out<double> yo(c, y, a->a.n, where);
        if (kind)
            launch_spmv_sliced<double, double, double, double>(
                c.stream, a->a.layout->tiles(), nullptr, 0, a->a.slice_view(kind),
                a->a.sliced_values<double>(c), xi.p, yo.p);
        else
            spmv_any(c, a->a, xi.p, yo.p, lanes);
        yo.finish();
    });
}

int bspai_hip_spmv_f32(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const float *x, float *y,
                       int where, int lanes) {
    return guard(ctx, [&] {
        require(a && x && y, "null argument");
        check_spmv_lanes(lanes);
        Context &c = ctx->c;
        DevCsr &A = const_cast<DevCsr &>(a->a);
        A.ensure32(c);
        const int kind = sliced_kind(c, A, lanes);
        In<float> xi(c, x, A.n, where);
        Out<float> yo(c, y, A.n, where);
        if (kind)
            launch_spmv_sliced<float, float, float, float>(c.stream, A.layout->tiles(), nullptr, 0,
                                                           A.slice_view(kind),
                                                           A.sliced_values<float>(c), xi.p, yo.p);
        else
            launch_spmv<float, float, float, float>(c.stream, lanes ? lanes : A.lanes,
                                                    A.layout->tiles(), nullptr, 0, A.ptr.data(),
                                                    A.col.data(), A.val32.data(), xi.p, yo.p);
        yo.finish();
    });
}

int bspai_hip_spai_build(bspai_hip_ctx *ctx, const bspai_hip_matrix *a,
                         const bspai_hip_spai_options *opt, bspai_hip_spai **out) {
    return guard(ctx, [&] {
        require(a && opt && out, "null argument");
        require(opt->precision == BSPAI_HIP_FP64 || opt->precision == BSPAI_HIP_MIXED,
                "precision must be BSPAI_HIP_FP64 or BSPAI_HIP_MIXED");
        require(opt->alpha >= 0 && opt->beta >= 0 && opt->beta < INT_MAX, "alpha/beta range");
        auto p = std::make_unique<bspai_hip_spai>();
        p->b = spai_build(ctx->c, a->a, opt->epsilon, (long long)opt->alpha, (int)opt->beta,
                          opt->precision);
        *out = p.release();
    });
}

void bspai_hip_spai_destroy(bspai_hip_spai *p) {
    delete p;
}

int bspai_hip_spai_get_info(const bspai_hip_spai *p, bspai_hip_spai_info *info) {
    if (!p || !info)
        return BSPAI_HIP_EINVAL;
    const SpaiBuild &b = *p->b;
    info->status = b.status;
    info->failing_column = b.failing_column;
    info->n = b.n;
    info->p_nnz = b.p_nnz;
    info->pattern_nnz = b.pattern_nnz;
    info->fits = b.fits;
    info->growths = b.growths;
    info->clamps = b.clamps;
    info->passes = b.passes;
    info->build_ms = b.build_ms;
    info->qr_flops = b.qr_flops;
    set_msg(info->message, b.message);
    return BSPAI_HIP_OK;
}

int bspai_hip_spai_preconditioner(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int32_t *ptr,
                                  int32_t *col, double *val) {
    return guard(ctx, [&] {
        require(p, "null argument");
        const SpaiBuild &b = *p->b;
        const Context &c = ctx->c;
        require(b.p.ptr.size() > 0, "no preconditioner (construction failed before assembly)");
        HIP_CHECK(hipMemcpyAsync(ptr, b.p.ptr.data(), (b.n + 1) * sizeof(int),
                                 hipMemcpyDeviceToHost, c.stream));
        HIP_CHECK(hipMemcpyAsync(col, b.p.col.data(), b.p_nnz * sizeof(int), hipMemcpyDeviceToHost,
                                 c.stream));
        if (b.precision == BSPAI_HIP_FP64) {
            HIP_CHECK(hipMemcpyAsync(val, b.p.val.data(), b.p_nnz * sizeof(double),
                                     hipMemcpyDeviceToHost, c.stream));
        } else {
            std::vector<float> v(b.p_nnz);
            HIP_CHECK(hipMemcpyAsync(v.data(), b.p.val32.data(), b.p_nnz * sizeof(float),
                                     hipMemcpyDeviceToHost, c.stream));
            HIP_CHECK(hipStreamSynchronize(c.stream));
            for (std::int64_t i = 0; i < b.p_nnz; ++i)
                val[i] = v[i];
        }
        HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int bspai_hip_spai_columns(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int32_t *status,
                           int32_t *fits, int32_t *growths, int32_t *clamps, double *residual,
                           double *pivot_ratio) {
    return guard(ctx, [&] {
        require(p, "null argument");
        const SpaiBuild &b = *p->b;
        const Context &c = ctx->c;
        require(b.c_status.size() == (std::size_t)b.n, "no column data (invalid input)");
        auto d = [&](auto *dst, const auto &src) {
            if (dst)
                HIP_CHECK(hipMemcpyAsync(dst, src.data(), b.n * sizeof(*dst),
                                         hipMemcpyDeviceToHost, c.stream));
        };
        d(status, b.c_status);
        d(fits, b.c_fits);
        d(growths, b.c_growths);
        d(clamps, b.c_clamps);
        d(residual, b.c_res);
        d(pivot_ratio, b.c_piv);
        HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int bspai_hip_spai_patterns(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int64_t *ptr,
                            int32_t *idx, double *coef) {
    return guard(ctx, [&] {
        require(p && ptr && idx && coef, "null argument");
        require(p->b->c_q.size() == (std::size_t)p->b->n, "no column data (invalid input)");
        spai_patterns(ctx->c, *p->b, ptr, idx, coef);
    });
}

int bspai_hip_spai_scaling(bspai_hip_ctx *ctx, const bspai_hip_spai *p, double *d) {
    return guard(ctx, [&] {
        require(p && d, "null argument");
        const SpaiBuild &b = *p->b;
        const Context &c = ctx->c;
        if (b.precision == BSPAI_HIP_FP64) {
            HIP_CHECK(hipMemcpyAsync(d, b.d64.data(), b.n * sizeof(double), hipMemcpyDeviceToHost,
                                     c.stream));
        } else {
            std::vector<float> v(b.n);
            HIP_CHECK(hipMemcpyAsync(v.data(), b.d32.data(), b.n * sizeof(float),
                                     hipMemcpyDeviceToHost, c.stream));
            HIP_CHECK(hipStreamSynchronize(c.stream));
            for (std::int64_t i = 0; i < b.n; ++i)
                d[i] = v[i];
        }
        HIP_CHECK(hipStreamSynchronize(c.stream));
    });
}

int bspai_hip_gmres(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const bspai_hip_matrix *p,
                    const double *rhs, const double *x0, const bspai_hip_gmres_options *opt,
                    double *x, int where, bspai_hip_gmres_info *info) {
    return guard(ctx, [&] {
        require(a && p && rhs && opt && x && info, "null argument");
        require(a->a.n == p->a.n, "Invalid dimensions, initial guess, tolerance, or iteration cap");
        Context &c = ctx->c;
        const Layout &L = *a->a.layout;
        // The cap applies to every system: it may not exceed the smallest
        // one (src/gmres.cpp:115 for a single system).
        std::int64_t minrows = a->a.n;
        for (int s = 0; s < L.nsys; ++s)
            minrows = std::min(minrows, L.offsets[s + 1] - L.offsets[s]);
        require(std::isfinite(opt->tolerance) && opt->tolerance > 0 &&
                    opt->max_iterations >= 0 && opt->max_iterations <= minrows,
                "Invalid dimensions, initial guess, tolerance, or iteration cap");
        const int cap = opt->max_iterations ? (int)opt->max_iterations : (int)a->a.n;
        std::int64_t maxrows = 0;
        for (int s = 0; s < L.nsys; ++s)
            maxrows = std::max(maxrows, L.offsets[s + 1] - L.offsets[s]);
        GmresEngine<double> eng(c, L, (int)std::min<std::int64_t>(cap, maxrows));
        In<double> bi(c, rhs, a->a.n, where);
        In<double> x0i(c, x0, x0 ? a->a.n : 0, where);
        Out<double> xo(c, x, a->a.n, where);
        eng.solve(a->a, p->a, kLeft, bi.p, x0 ? x0i.p : nullptr, opt->tolerance, xo.p, nullptr);
        xo.finish();
        const auto oc = eng.outcomes();
        for (std::size_t s = 0; s < oc.size(); ++s) {
            info[s].status = oc[s].status;
            info[s].converged = oc[s].converged;
            info[s].iterations = oc[s].iterations;
            info[s].initial_original = oc[s].init_orig;
            info[s].initial_relative = oc[s].init_rel;
            info[s].final_original = oc[s].final_orig;
            info[s].final_relative = oc[s].final_rel;
        }
    });
}

int bspai_hip_ir_solve(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const bspai_hip_matrix *p,
                       const double *b, const bspai_hip_ir_options *o, double *x, int where,
                       bspai_hip_ir_info *info) {
    return guard(ctx, [&] {
        validate_ir(a, o);
        require(p && b && x && info, "null argument");
        const int nsys = a->a.layout->nsys;
        run_ir(ctx, a->a, p->a, b, o, x, where, info);
        for (int s = 0; s < nsys; ++s) {
            info[s].construction_performed = 0;
            info[s].spai_status = 0;
            info[s].spai_ms = 0;
        }
    });
}

int bspai_hip_spai_gmres_ir(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, const double *b,
                            const bspai_hip_ir_options *o, double *x, int where,
                            bspai_hip_ir_info *info) {
    return guard(ctx, [&] {
        validate_ir(a, o);
        require(b && x && info, "null argument");
        const int nsys = a->a.layout->nsys;
        const long long alpha = o->spai_alpha > 0 ? o->spai_alpha : (a->a.n + 7) / 8;
        bspai_hip_spai p;
        p.b = spai_build(ctx->c, a->a, o->spai_epsilon, alpha, (int)o->spai_beta, o->precision);
        for (int s = 0; s < nsys; ++s) {
            std::memset(&info[s], 0, sizeof(info[s]));
            info[s].construction_performed = 1;
            info[s].spai_status = p.b->status;
            info[s].spai_ms = p.b->build_ms;
        }
        if (p.b->status != kSpaiSuccess) {
            for (int s = 0; s < nsys; ++s) {
                info[s].status = kIrConstructionFailure;
                set_msg(info[s].message, std::string("SPAI failed: ") + kSpaiNames[p.b->status] +
                                             ": " + p.b->message);
            }
            return;
        }
        run_ir(ctx, a->a, p.b->p, b, o, x, where, info);
        for (int s = 0; s < nsys; ++s) {
            info[s].construction_performed = 1;
            info[s].spai_status = p.b->status;
            info[s].spai_ms = p.b->build_ms;
        }
    });
}

int bspai_hip_spai_matrix(bspai_hip_ctx *ctx, const bspai_hip_spai *p, int transpose,
                          bspai_hip_matrix **out) {
    return guard(ctx, [&] {
        require(p && out, "null argument");
        Context &c = ctx->c;
        const SpaiBuild &b = *p->b;
        require(b.status == kSpaiSuccess, "construction did not succeed");
        auto m = std::make_unique<bspai_hip_matrix>();
        if (b.precision == BSPAI_HIP_FP64) {
            DevCsr &d = m->a;
            d.n = b.p.n;
            d.nnz = b.p.nnz;
            d.ptr.resize(d.n + 1);
            d.col.resize(std::max<std::int64_t>(d.nnz, 1));
            d.val.resize(std::max<std::int64_t>(d.nnz, 1));
            HIP_CHECK(hipMemcpyAsync(d.ptr.data(), b.p.ptr.data(), (d.n + 1) * sizeof(int),
                                     hipMemcpyDeviceToDevice, c.stream));
            HIP_CHECK(hipMemcpyAsync(d.col.data(), b.p.col.data(), d.nnz * sizeof(int),
                                     hipMemcpyDeviceToDevice, c.stream));
            HIP_CHECK(hipMemcpyAsync(d.val.data(), b.p.val.data(), d.nnz * sizeof(double),
                                     hipMemcpyDeviceToDevice, c.stream));
            d.layout = b.p.layout;
            d.lanes = b.p.lanes;
        } else {
            auto w = widen(c, b.p);
            m->a = std::move(*w);
        }
        HIP_CHECK(hipStreamSynchronize(c.stream));
        if (transpose) {
            bspai_hip_matrix *t = nullptr;
            const int rc = bspai_hip_matrix_transpose(ctx, m.get(), &t);
            if (rc != BSPAI_HIP_OK)
                throw std::runtime_error(ctx->c.last_error);
            *out = t;
        } else {
            *out = m.release();
        }
    });
}

int bspai_hip_adjoint_solve(bspai_hip_ctx *ctx, const bspai_hip_matrix *at,
                            const bspai_hip_matrix *pt, const double *g,
                            const bspai_hip_ir_options *o, double *lambda, int where,
                            bspai_hip_ir_info *info) {
    return guard(ctx, [&] {
        validate_ir(at, o);
        require(pt && g && lambda && info, "null argument");
        require(at->a.n == pt->a.n, "dimension mismatch");
        Context &c = ctx->c;
        DevCsr &AT = const_cast<DevCsr &>(at->a);
        DevCsr &PT = const_cast<DevCsr &>(pt->a);
        if (o->precision == BSPAI_HIP_MIXED) {
            AT.ensure32(c);
            PT.ensure32(c);
        }
        In<double> gi(c, g, AT.n, where);
        Out<double> lo(c, lambda, AT.n, where);
        HIP_CHECK(hipEventRecord(c.ev0, c.stream));
        // Right preconditioning: A^T (P^T u) = g, lambda = P^T u.
        auto res = ir_run(c, AT, PT, AT, kRight, gi.p, ir_opts(o), lo.p);
        HIP_CHECK(hipEventRecord(c.ev1, c.stream));
        lo.finish();
        HIP_CHECK(hipEventSynchronize(c.ev1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        for (std::size_t s = 0; s < res.size(); ++s) {
            fill_ir_info(info[s], res[s]);
            info[s].solve_ms = ms;
            info[s].construction_performed = 0;
        }
    });
}

int bspai_hip_sddmm(bspai_hip_ctx *ctx, const bspai_hip_matrix *a, double alpha, const double *u,
                    const double *v, double *out, int where) {
    return guard(ctx, [&] {
        require(a && u && v && out, "null argument");
        Context &c = ctx->c;
        In<double> ui(c, u, a->a.n, where), vi(c, v, a->a.n, where);
        Out<double> oo(c, out, a->a.nnz, where);
        launch_sddmm(c.stream, a->a.layout->tiles(), a->a.ptr.data(), a->a.col.data(), alpha,
                     ui.p, vi.p, oo.p);
        oo.finish();
    });
}

int bspai_hip_score(bspai_hip_ctx *ctx, double residual_squared, double dot,
                    double denominator_squared, int *status, int *clamped, double *score) {
    return guard(ctx, [&] {
        require(status && clamped && score, "null argument");
        spai_score_on_device(ctx->c.stream, residual_squared, dot, denominator_squared, status,
                             clamped, score);
    });
}


int bspai_hip_mm_parse_gpu(bspai_hip_ctx *ctx, const char *text, int64_t len,
                           bspai_hip_mm_gpu **out, char *message, int message_len) {
    auto put = [&](const char *m) {
        if (message && message_len > 0) {
            std::strncpy(message, m, (std::size_t)message_len - 1);
            message[message_len - 1] = 0;
        }
    };
    put("");
    std::string format_error;
    const int rc = guard(ctx, [&] {
        require(ctx && out && len >= 0 && (len == 0 || text), "null argument");
        *out = nullptr;
        auto m = std::make_unique<bspai_hip_mm_gpu>();
        try {
            mm_parse_gpu(ctx->c, text, (std::size_t)len, m->m);
        } catch (const bspai_mmio::FormatError &e) {
            format_error = e.what();
            return;
        }
        *out = m.release();
    });
    if (!format_error.empty()) {
        ctx->c.last_error = format_error;
        put(format_error.c_str());
        return BSPAI_HIP_EFORMAT;
    }
    if (rc != BSPAI_HIP_OK && ctx)
        put(ctx->c.last_error.c_str());
    return rc;
}

int bspai_hip_mm_gpu_info(const bspai_hip_mm_gpu *mm, int64_t *rows, int64_t *cols,
                          int64_t *count, int *integer_field, int *symmetric) {
    if (!mm)
        return BSPAI_HIP_EINVAL;
    if (rows)
        *rows = mm->m.rows;
    if (cols)
        *cols = mm->m.cols;
    if (count)
        *count = mm->m.count;
    if (integer_field)
        *integer_field = mm->m.integer;
    if (symmetric)
        *symmetric = mm->m.symmetric;
    return BSPAI_HIP_OK;
}

int bspai_hip_mm_gpu_entries(const bspai_hip_mm_gpu *mm, const int64_t **row, const int64_t **col,
                             const double **val) {
    if (!mm)
        return BSPAI_HIP_EINVAL;
    if (row)
        *row = mm->m.row.data();
    if (col)
        *col = mm->m.col.data();
    if (val)
        *val = mm->m.val.data();
    return BSPAI_HIP_OK;
}

int bspai_hip_mm_gpu_copy(bspai_hip_ctx *ctx, const bspai_hip_mm_gpu *mm, int64_t *row,
                          int64_t *col, double *val, int where) {
    return guard(ctx, [&] {
        require(ctx && mm, "null argument");
        require(where == BSPAI_HIP_HOST || where == BSPAI_HIP_DEVICE, "where: HOST or DEVICE");
        const std::size_t n = (std::size_t)mm->m.count;
        if (n == 0)
            return;
        require(row && col && val, "null argument");
        const hipMemcpyKind kind =
            where == BSPAI_HIP_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        hipStream_t s = ctx->c.stream;
        HIP_CHECK(hipMemcpyAsync(row, mm->m.row.data(), n * sizeof(int64_t), kind, s));
        HIP_CHECK(hipMemcpyAsync(col, mm->m.col.data(), n * sizeof(int64_t), kind, s));
        HIP_CHECK(hipMemcpyAsync(val, mm->m.val.data(), n * sizeof(double), kind, s));
        HIP_CHECK(hipStreamSynchronize(s));
    });
}

void bspai_hip_mm_gpu_destroy(bspai_hip_mm_gpu *mm) {
    delete mm;
}

} // extern "C"
