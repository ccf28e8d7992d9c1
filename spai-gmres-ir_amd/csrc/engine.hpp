// Host-side engine: device buffers, matrices, SPAI / GMRES / IR drivers.
#pragma once

#include "common.hpp"
#include "kernels.hpp"
#include "krylov.hpp"

#include <map>
#include <memory>
#include <tuple>
#include <string>
#include <vector>

namespace bspai_hip {

// ---------------------------------------------------------------- buffers
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t bytes) { resize(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    void resize(std::size_t bytes) {
        if (bytes <= n_)
            return;
        release();
        if (bytes) {
            const hipError_t e = hipMalloc(&p_, bytes);
            if (e != hipSuccess) {
                p_ = nullptr;
                throw HipError("hipMalloc of " + std::to_string(bytes) +
                               " bytes failed: " + hipGetErrorString(e));
            }
            n_ = bytes;
        }
    }
    void release() {
        if (p_)
            (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p_); }
    std::size_t bytes() const { return n_; }

  private:
    void *p_ = nullptr;
    std::size_t n_ = 0;
};

template <typename T> class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) { resize(n); }
    void resize(std::size_t n) {
        buf_.resize(n * sizeof(T));
        n_ = n;
    }
    T *data() const { return buf_.as<T>(); }
    std::size_t size() const { return n_; }

  private:
    DeviceBuffer buf_;
    std::size_t n_ = 0;
};

// Per-kernel timing with HIP events on the context's stream (telemetry, off
// unless enabled): every timed launch is bracketed by two events, and
// collect() synchronizes and aggregates the durations by kernel name.
class KernelTimer {
  public:
    struct Total {
        std::string name;
        long long launches = 0;
        double ms = 0, bytes = 0, flops = 0;
    };
    KernelTimer() = default;
    KernelTimer(const KernelTimer &) = delete;
    KernelTimer &operator=(const KernelTimer &) = delete;
    ~KernelTimer();
    bool on = false;
    bool spai_only = false; // time only the SPAI construction kernels
    int begin(hipStream_t s);
    void end(hipStream_t s, int first, const char *name, double bytes, double flops);
    // Adds flops known only after the launch (device-side counts) to the
    // most recent record of `name`.
    void add_flops(const char *name, double flops);
    // Folds the pending records into the running totals (synchronizing on
    // their events) and returns the totals since the last clear().
    const std::vector<Total> &collect();
    void clear() { recs_.clear(), totals_.clear(), used_ = 0; }

  private:
    struct Rec {
        std::string name;
        int first;
        double bytes, flops;
    };
    hipEvent_t event(int i);
    std::vector<hipEvent_t> pool_;
    std::vector<Rec> recs_;
    std::vector<Total> totals_;
    int used_ = 0;
};

// SpMV launch configuration (lanes per row; entries per thread of the
// streaming kernel).
struct SpmvChoice {
    int lanes = 1, per = 8;
    bool c16 = false; // streaming kernel on 16-bit column offsets
    int sliced = 0;   // sliced kernel: 1 int32 columns, 2 int16 offsets
};

// Scratch of the coordinate-entry canonicalization, kept by the context and
// grown on demand (entries of one call at a time).
struct CanonWork {
    DeviceArray<unsigned long long> keys, keys_sorted, flags64;
    DeviceArray<double> vals_sorted, sums;
    DeviceArray<int> keep, slot, flags, scan_tmp, row_tmp;
    DeviceArray<unsigned char> sort_tmp;
    DeviceArray<int> out_ptr, out_col;
    DeviceArray<double> out_val;
};

// Scratch of the SPAI construction, kept by the context and grown on demand
// (one construction at a time): B's structure and values, the column work
// lists, the column kernels' work areas and the assembly counters.
struct SpaiWork {
    DeviceArray<int> bad, bptr, bcol, cursor, scan_tmp, work, next, counter, ff;
    DeviceArray<unsigned long long> tot;
    DeviceArray<double> ftot;
    DeviceBuffer bval, dwork, iwork;
    int cus = 0; // compute units of the device
};

template <typename T> class GmresEngine;
struct Layout;

// Scratch of the refinement loop (per-system state, partials, history, the
// residual and correction vectors), grown on demand.
struct IrWork {
    DeviceBuffer st;
    DeviceArray<double> part, part4, flags, upd, hist, r, d;
    DeviceArray<int> inner;
    DeviceArray<float> rf;
};

// The GMRES workspace of the last refinement per precision (basis, H, Gram
// triangle, partials), kept for the next solve of the same layout and Krylov
// capacity, as the reference's CudaGmresContext keeps its workspace.
template <typename T> struct EngineSlot {
    std::shared_ptr<const Layout> layout;
    int cap = 0;
    std::shared_ptr<GmresEngine<T>> engine;
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string last_error;
    // timing events
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    KernelTimer timer;
    // SpMV autotuning results by matrix shape (n, nnz, value bytes)
    std::map<std::tuple<std::int64_t, std::int64_t, int>, SpmvChoice> spmv_tuned;
    // SPAI capacity-bucket launch shapes by (n, nnz, value bytes, bucket):
    // waves per SIMD the workgroup kernel is compiled for, workgroups per CU
    std::map<std::tuple<std::int64_t, std::int64_t, int, int>, std::pair<int, int>> spai_tuned;
    CanonWork canon;
    SpaiWork spai;
    EngineSlot<double> engine64;
    EngineSlot<float> engine32;
    IrWork ir;
    // Arnoldi-step graphs captured and replayed by this context's solves
    long long graphs_captured = 0, graphs_replayed = 0;
};

// Runs `launch` (one kernel launch on c.stream), timed when the context's
// timer is on.
template <class F>
inline void timed(Context &c, const char *name, double bytes, double flops, F &&launch,
                  bool spai = false) {
    if (!c.timer.on || (c.timer.spai_only && !spai)) {
        launch();
        return;
    }
    const int first = c.timer.begin(c.stream);
    launch();
    c.timer.end(c.stream, first, name, bytes, flops);
}

// Block-diagonal system layout and its row tiles.
struct Layout {
    int nsys = 1;
    std::int64_t n = 0;
    std::vector<std::int64_t> offsets; // nsys+1
    DeviceArray<int> t_sys, t_row0, t_row1, s_tile0, s_tile1, s_rows;
    int ntiles = 0;
    int uniform = 0; // every tile full and 16-byte aligned for FP32 and FP64
    Tiles tiles() const {
        return Tiles{ntiles,         nsys,           t_sys.data(), t_row0.data(),
                     t_row1.data(),  s_tile0.data(), s_tile1.data(), uniform, (int)n};
    }
    void build(const Context &c, std::int64_t n_, const std::vector<std::int64_t> &offs);
};

// Device CSR with FP64 values (and an FP32 copy made on demand).
// Generation numbers of matrices: a new matrix (or a matrix given a new
// structure) never shares one with an older one, even where the allocator
// hands it the older one's addresses.
std::uint64_t next_matrix_generation();

struct DevCsr {
    std::int64_t n = 0, nnz = 0;
    std::uint64_t gen = next_matrix_generation();
    DeviceArray<int> ptr, col;
    DeviceArray<double> val;
    DeviceArray<float> val32;
    bool has32 = false, has64 = true;
    int lanes = 1; // SpMV lanes per row
    std::shared_ptr<Layout> layout;
    // 16-bit column offsets for the streaming SpMV (made on first use):
    // c16 0 = not tried, 1 = available, -1 = an offset does not fit.
    mutable DeviceArray<short> col16;
    mutable int c16 = 0;
    // Sliced layout for the wavefront-per-slice SpMV (structure made on
    // first use; values filled per type on demand and again after new
    // values).  sliced: 0 = not tried, 1 = available, -1 = a row is longer
    // than 255 entries or the slices would hold more than 3x the entries.
    // s16: the int16 column offsets fit (1) or not (-1).
    mutable DeviceArray<int> soff, scol;
    mutable DeviceArray<short> scol16;
    mutable DeviceArray<unsigned char> slen;
    mutable DeviceArray<double> sval;
    mutable DeviceArray<float> sval32;
    mutable std::int64_t spadded = 0;
    mutable int sliced = 0, s16 = 0;
    mutable bool sval_ok = false, sval32_ok = false;
    void ensure32(const Context &c);
    const short *ensure_col16(const Context &c) const;
    // kind 1: int32 columns, 2: int16 offsets; false if not available.
    bool ensure_sliced(const Context &c) const;
    SliceView slice_view(int kind) const;
    template <typename T> const T *sliced_values(const Context &c) const;
    void values_changed() {
        has32 = false;
        sval_ok = sval32_ok = false;
    }
    template <typename T> const T *values() const;
};
template <> inline const double *DevCsr::values<double>() const { return val.data(); }
template <> inline const float *DevCsr::values<float>() const { return val32.data(); }

int choose_lanes(std::int64_t n, std::int64_t nnz);

struct SpaiBuild {
    std::int64_t n = 0;
    int precision = 0;
    int status = 0;
    std::int64_t failing_column = -1;
    std::string message;
    DevCsr p; // values in FP64 (plain) or FP32 (mixed, val32)
    DeviceArray<double> d64;
    DeviceArray<float> d32;
    DeviceArray<int> c_status, c_fits, c_growths, c_clamps, c_q;
    DeviceArray<double> c_res, c_piv, c_flops;
    DeviceArray<std::int64_t> c_off;   // per column: offset into its pass pool
    DeviceArray<int> c_pass;           // per column: pass id
    std::vector<DeviceArray<int>> pool_j;
    std::vector<DeviceBuffer> pool_c;
    std::int64_t p_nnz = 0, pattern_nnz = 0, fits = 0, growths = 0, clamps = 0;
    int passes = 0;
    double build_ms = 0;
    double qr_flops = 0; // algorithmic QR flops of the accepted fits
};

std::unique_ptr<SpaiBuild> spai_build(Context &c, const DevCsr &a, double eps, long long alpha,
                                      int beta, int precision);

struct GmresOutcome {
    int status, converged, iterations;
    double init_orig, init_rel, final_orig, final_rel;
};

// GMRES modes: left-preconditioned P(A x) = P r (forward) or right-
// preconditioned A^T (P^T u) = g, x = P^T u (adjoint).
enum GmresMode { kLeft = 0, kRight = 1 };

template <typename T> class GmresEngine {
  public:
    GmresEngine(Context &c, const Layout &L, int cap);
    ~GmresEngine();
    GmresEngine(const GmresEngine &) = delete;
    GmresEngine &operator=(const GmresEngine &) = delete;
    // m1 then m2 applied per Arnoldi step: left (A, P), right (P^T, A^T).
    // rhs in T; x (FP64) receives the solution; ir (optional) masks systems.
    // rule (with ir): per-system adaptive tolerance, see InnerTolRule.
    void solve(const DevCsr &m1, const DevCsr &m2, GmresMode mode, const T *rhs, const double *x0,
               double tol, double *x, const IrSysState *ir, const InnerTolRule &rule = {});
    std::vector<GmresOutcome> outcomes();
    KrylovArgs<T> args() const { return args_; }
    int cap() const { return cap_; }
    // Host check points every sync_every steps: two replayed graphs queued
    // per check, so the card is not idle while the host reads the state.
    // Growth of the workspace synchronizes before it moves buffers, and
    // steps queued behind a triggered or finished one do nothing.
    int sync_every = 8;
    // Arnoldi steps per replayed graph (BSPAI_GMRES_GRAPH=0 launches every
    // kernel on the stream instead); graphs_replayed counts the replays of
    // the last solve.
    static constexpr int kGraphSteps = 4;
    long long graphs_replayed = 0, graphs_captured = 0;

  private:
    // A captured group of kGraphSteps Arnoldi steps (with or without the
    // per-step verification), valid while its key -- every pointer and
    // launch choice baked into its nodes -- is unchanged.
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        std::vector<const void *> key;
    };
    StepGraph graphs_[2];
    DeviceArray<int> kdev_; // the step the next replayed graph starts at
    int kdev_val_ = -1;     // its value as the host last set or advanced it
    void ensure_basis(int count);
    // Capacity of H, the Gram triangle and the projection partials for
    // Krylov steps 0..k (grown by doubling, contents kept).
    void ensure_steps(int k);
    Context &c_;
    const Layout &L_;
    int cap_;
    std::vector<DeviceArray<T>> chunks_;
    std::vector<T *> vptrs_;
    DeviceArray<T *> vdev_;
    int vcount_ = 0;
    int ksteps_ = 0; // Krylov steps H, gram_ and gp_ hold
    DeviceArray<T> w_, t_, prhs_, r0_;
    DeviceArray<double> xz_, ax_, pax_, ax0_, tmpd_;
    DeviceArray<double> H_, cs_, sn_, g_, y_, part_, gp_, q_, h_, gram_;
    DeviceArray<SysState> st_;
    KrylovArgs<T> args_{};
};

struct IrResultHost {
    IrSysState st;
    double history[kMaxOuterHistory][4];
    int inner[kMaxOuterHistory][3];
};

struct IrOptionsC {
    double tol = 1e-8;
    int cap = 0;
    int outer = 20;
    int precision = 0;
    double relres_target = 0;
};

// Iterative refinement on a (batched) system; a64 is the FP64 matrix for the
// residual, m1/m2 the inner operators (FP64 or FP32 values per precision).
std::vector<IrResultHost> ir_run(Context &c, const DevCsr &a64, const DevCsr &m1, const DevCsr &m2,
                                 GmresMode mode, const double *b, const IrOptionsC &o, double *x);

void spmv_any(Context &c, const DevCsr &a, const double *x, double *y, int lanes);
void spai_patterns(Context &c, const SpaiBuild &b, std::int64_t *hptr, int *hidx, double *hcoef);
void launch_sddmm(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col, double alpha,
                  const double *u, const double *v, double *out);
void spai_score_on_device(hipStream_t s, double rr, double dot, double den, int *status,
                          int *clamped, double *score);

// Canonical rows x cols CSR from `count` coordinate entries held on the
// device (reference CsrMatrix::from_entries, src/cpu.cpp:48-76), written to
// device arrays ptr (rows + 1), out_col and out_val (room for count).
// Returns nnz.  Throws std::invalid_argument with the reference's messages
// for an index out of range, a non-finite value or a non-finite duplicate
// sum; the outputs are then unspecified.
// Allocation-free canonicalization (bspai_hip_csr_from_entries_noalloc): its
// scratch is a static device array of kNoAllocScratchBytes per GPU, set up by
// noalloc_setup (once per device, at context creation).  Limits: count <=
// kNoAllocMaxEntries, rows + 1 <= kNoAllocMaxRows.  Calls that overlap on one
// GPU share the array and race.
constexpr std::int64_t kNoAllocMaxEntries = 1 << 19;
constexpr std::int64_t kNoAllocMaxRows = 1 << 22;
constexpr std::size_t kNoAllocScratchBytes = (std::size_t)48 << 20;
void noalloc_setup(Context &c);
std::int64_t csr_from_entries_noalloc(Context &c, std::int64_t rows, std::int64_t cols,
                                      std::int64_t count, const std::int64_t *row,
                                      const std::int64_t *col, const double *val, int *ptr,
                                      int *out_col, double *out_val);

// Checks a device CSR before use: offsets 0..nnz nondecreasing, then every
// row's columns in range, strictly increasing and inside the row's system,
// and finite values.  Throws std::invalid_argument naming the first problem
// class found.  ptr/col/val are device arrays; sys_offsets: nsys + 1 offsets.
void check_device_csr(Context &c, std::int64_t n, std::int64_t nnz, const int *ptr,
                      const int *col, const double *val,
                      const std::vector<std::int64_t> &sys_offsets);

std::int64_t csr_from_entries(Context &c, std::int64_t rows, std::int64_t cols,
                              std::int64_t count, const std::int64_t *row,
                              const std::int64_t *col, const double *val, int *ptr, int *out_col,
                              double *out_val);
} // namespace bspai_hip
