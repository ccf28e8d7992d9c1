// Host orchestration of the SPAI / GMRES / IR kernels.
#include "engine.hpp"

#include <atomic>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <type_traits>

namespace bspai_hip {

namespace {
template <typename T> void upload(const Context &c, T *dst, const T *src, std::size_t n) {
    if (n)
        HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c.stream));
}
template <typename T> void download(const Context &c, T *dst, const T *src, std::size_t n) {
    if (n)
        HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
}
void sync(const Context &c) {
    HIP_CHECK(hipStreamSynchronize(c.stream));
}
} // namespace

// ---------------------------------------------------------------- timing
// Timer labels: the kernel names rocprofv3 reports, plus the operand role.
template <typename T> constexpr const char *kSpaiKernelName = "spai_columns_kernel<double>";
template <> constexpr const char *kSpaiKernelName<float> = "spai_columns_kernel<float>";
template <typename T> constexpr const char *kSpaiBlockKernelName = "spai_block_kernel<double>";
template <> constexpr const char *kSpaiBlockKernelName<float> = "spai_block_kernel<float>";

KernelTimer::~KernelTimer() {
    for (hipEvent_t e : pool_)
        (void)hipEventDestroy(e);
}
hipEvent_t KernelTimer::event(int i) {
    while ((int)pool_.size() <= i) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        pool_.push_back(e);
    }
    return pool_[i];
}
int KernelTimer::begin(hipStream_t s) {
    const int first = used_;
    HIP_CHECK(hipEventRecord(event(first), s));
    used_ += 2;
    return first;
}
void KernelTimer::end(hipStream_t s, int first, const char *name, double bytes, double flops) {
    HIP_CHECK(hipEventRecord(event(first + 1), s));
    recs_.push_back(Rec{name, first, bytes, flops});
}
void KernelTimer::add_flops(const char *name, double flops) {
    for (auto it = recs_.rbegin(); it != recs_.rend(); ++it)
        if (it->name == name) {
            it->flops += flops;
            return;
        }
}
const std::vector<KernelTimer::Total> &KernelTimer::collect() {
    std::vector<Total> &out = totals_;
    for (const Rec &r : recs_) {
        HIP_CHECK(hipEventSynchronize(pool_[r.first + 1]));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, pool_[r.first], pool_[r.first + 1]));
        auto it = std::find_if(out.begin(), out.end(),
                               [&](const Total &t) { return t.name == r.name; });
        if (it == out.end()) {
            out.push_back(Total{r.name});
            it = out.end() - 1;
        }
        it->launches += 1;
        it->ms += ms;
        it->bytes += r.bytes;
        it->flops += r.flops;
    }
    recs_.clear();
    used_ = 0;
    return out;
}

std::string spmv_label(int lanes, const char *tname, const char *role, int sliced = 0) {
    std::string k = sliced      ? std::string("spmv_slice_kernel")
                    : lanes == 1 ? std::string("spmv_stream_kernel")
                                 : "spmv_group_kernel<" + std::to_string(lanes < 0 ? 1 : lanes) + ">";
    return k + "<" + tname + "> " + role;
}

// ---------------------------------------------------------------- layout
void Layout::build(const Context &c, std::int64_t n_, const std::vector<std::int64_t> &offs) {
    n = n_;
    offsets = offs;
    nsys = (int)offs.size() - 1;
    std::vector<int> ts, r0, r1, st0(nsys), st1(nsys), rows(nsys);
    for (int s = 0; s < nsys; ++s) {
        st0[s] = (int)ts.size();
        rows[s] = (int)(offs[s + 1] - offs[s]);
        for (std::int64_t r = offs[s]; r < offs[s + 1]; r += kTileRows) {
            ts.push_back(s);
            r0.push_back((int)r);
            r1.push_back((int)std::min<std::int64_t>(r + kTileRows, offs[s + 1]));
        }
        st1[s] = (int)ts.size();
    }
    ntiles = (int)ts.size();
    uniform = 1;
    for (int t = 0; t < ntiles; ++t)
        if (r1[t] - r0[t] != kTileRows || r0[t] % 4 != 0)
            uniform = 0;
    t_sys.resize(ntiles);
    t_row0.resize(ntiles);
    t_row1.resize(ntiles);
    s_tile0.resize(nsys);
    s_tile1.resize(nsys);
    s_rows.resize(nsys);
    upload(c, t_sys.data(), ts.data(), ntiles);
    upload(c, t_row0.data(), r0.data(), ntiles);
    upload(c, t_row1.data(), r1.data(), ntiles);
    upload(c, s_tile0.data(), st0.data(), nsys);
    upload(c, s_tile1.data(), st1.data(), nsys);
    upload(c, s_rows.data(), rows.data(), nsys);
    sync(c);
}

// Streaming (lanes = 1, sequential row order) for short rows; lane groups
// for long rows, where one thread per row would serialize.
int choose_lanes(std::int64_t n, std::int64_t nnz) {
    const double avg = n ? (double)nnz / (double)n : 0;
    if (avg <= 48)
        return 1;
    if (avg <= 96)
        return 16;
    return 32;
}

void DevCsr::ensure32(const Context &c) {
    if (has32)
        return;
    val32.resize(nnz);
    launch_convert<double, float>(c.stream, nnz, val.data(), val32.data());
    has32 = true;
}

// SpMV autotuning.  For a matrix whose rows are short enough for the
// sequential row order (lanes = 1), the candidates are the streaming kernel
// with 8 or 16 entries per thread and one thread per row (lanes = -1): all
// three add each row's products in CSR order, so the choice changes timing
// only, never a result bit.  Each candidate runs once untimed and then three
// times between HIP events on the context's stream; the fastest is kept per
// matrix shape (n, nnz, value type), so a rebuilt preconditioner of the same
// shape reuses it.  Long-row matrices keep their lane-group choice.
// BSPAI_SPMV_AUTOTUNE=0 turns tuning off (the row-length rule decides).
const short *DevCsr::ensure_col16(const Context &c) const {
    if (c16 == 0) {
        c16 = -1;
        if (nnz > 0 && layout) {
            col16.resize((std::size_t)nnz);
            DeviceArray<int> bad(1);
            HIP_CHECK(hipMemsetAsync(bad.data(), 0, sizeof(int), c.stream));
            launch_encode_col16(c.stream, layout->tiles(), ptr.data(), col.data(), col16.data(),
                                bad.data());
            int hb = 1;
            download(c, &hb, bad.data(), 1);
            sync(c);
            c16 = hb ? -1 : 1;
        }
    }
    return c16 > 0 ? col16.data() : nullptr;
}

std::uint64_t next_matrix_generation() {
    static std::atomic<std::uint64_t> counter{0};
    return ++counter;
}

bool DevCsr::ensure_sliced(const Context &c) const {
    if (sliced == 0) {
        sliced = -1;
        s16 = -1;
        if (nnz > 0 && layout && layout->ntiles > 0) {
            const Tiles tl = layout->tiles();
            const std::size_t ns = (std::size_t)tl.ntiles * (kTileRows / kWave);
            slen.resize((std::size_t)n);
            DeviceArray<int> width(ns), bad(1);
            HIP_CHECK(hipMemsetAsync(bad.data(), 0, sizeof(int), c.stream));
            launch_slice_width(c.stream, tl, ptr.data(), col.data(), slen.data(), width.data(),
                               bad.data());
            std::vector<int> w(ns);
            int hb = 0;
            download(c, w.data(), width.data(), ns);
            download(c, &hb, bad.data(), 1);
            sync(c);
            std::vector<int> off(ns + 1);
            std::int64_t acc = 0;
            for (std::size_t i = 0; i < ns; ++i) {
                off[i] = (int)std::min<std::int64_t>(acc, INT_MAX);
                acc += (std::int64_t)w[i] * kWave;
            }
            off[ns] = (int)std::min<std::int64_t>(acc, INT_MAX);
            // (no layout for rows longer than 255, or when the slices would
            // hold more than 3x the entries: ragged rows gain nothing)
            if (!(hb & 1) && acc < INT_MAX && acc <= 3 * nnz + (std::int64_t)ns * kWave) {
                spadded = acc;
                soff.resize(ns + 1);
                upload(c, soff.data(), off.data(), ns + 1);
                scol.resize((std::size_t)acc);
                const bool fits16 = !(hb & 2);
                if (fits16)
                    scol16.resize((std::size_t)acc);
                launch_slice_fill<double>(c.stream, tl, ptr.data(), col.data(), nullptr,
                                          slice_view(1), scol.data(),
                                          fits16 ? scol16.data() : nullptr, nullptr);
                sync(c);
                sliced = 1;
                s16 = fits16 ? 1 : -1;
            }
        }
    }
    return sliced > 0;
}

SliceView DevCsr::slice_view(int kind) const {
    SliceView v;
    v.soff = soff.data();
    v.slen = slen.data();
    if (kind == 2)
        v.scol16 = scol16.data();
    else
        v.scol = scol.data();
    return v;
}

template <> const double *DevCsr::sliced_values<double>(const Context &c) const {
    if (sliced <= 0)
        throw std::logic_error("sliced SpMV values of a matrix without the sliced layout");
    if (!sval_ok) {
        sval.resize((std::size_t)spadded);
        launch_slice_fill<double>(c.stream, layout->tiles(), ptr.data(), col.data(), val.data(),
                                  slice_view(1), nullptr, nullptr, sval.data());
        sval_ok = true;
    }
    return sval.data();
}
template <> const float *DevCsr::sliced_values<float>(const Context &c) const {
    if (sliced <= 0)
        throw std::logic_error("sliced SpMV values of a matrix without the sliced layout");
    if (!sval32_ok) {
        if (!has32)
            throw std::logic_error("sliced FP32 values before the FP32 copy");
        sval32.resize((std::size_t)spadded);
        launch_slice_fill<float>(c.stream, layout->tiles(), ptr.data(), col.data(), val32.data(),
                                 slice_view(1), nullptr, nullptr, sval32.data());
        sval32_ok = true;
    }
    return sval32.data();
}

// A choice is tuned per matrix shape, and another matrix of that shape (A and
// its transpose, say) can lack what it asks for: its sliced layout is built
// here, and a form it cannot take falls back (int16 -> int32 offsets -> the
// streaming kernel).
SpmvChoice usable_choice(const Context &c, const DevCsr &m, SpmvChoice k) {
    if (k.sliced && !m.ensure_sliced(c))
        k.sliced = 0;
    if (k.sliced == 2 && m.s16 <= 0)
        k.sliced = 1;
    return k;
}

// comp: the product is formed with compensated FP32 sums (the FP32 solve's
// product with A), which the candidates are timed with.
template <typename T> SpmvChoice spmv_choice(Context &c, const DevCsr &m, bool comp = false) {
    SpmvChoice rule{m.lanes, stream_per(m.n, m.nnz)};
    if (m.lanes != 1 || m.n == 0)
        return rule;
    if (const char *env = std::getenv("BSPAI_SPMV_AUTOTUNE"))
        if (std::string(env) == "0")
            return rule;
    // BSPAI_SPMV_C16=1 takes the streaming kernel on 16-bit column offsets
    // whenever the matrix allows, =0 the row-length rule on 32-bit columns;
    // either bypasses the tuning (for comparisons).
    if (const char *env = std::getenv("BSPAI_SPMV_C16")) {
        if (std::string(env) == "1" && m.ensure_col16(c))
            return SpmvChoice{1, rule.per, true};
        return rule;
    }
    // BSPAI_SPMV_SLICED=1 takes the sliced kernel whenever the matrix allows
    // (int16 offsets when they fit), =0 leaves it out of the tuning.
    const char *senv = std::getenv("BSPAI_SPMV_SLICED");
    const bool no_sliced = senv && std::string(senv) == "0";
    if (senv && std::string(senv) == "1") {
        if (m.ensure_sliced(c))
            return SpmvChoice{1, rule.per, false, m.s16 > 0 ? 2 : 1};
        return rule;
    }
    comp = comp && sizeof(T) == 4;
    const auto key =
        std::make_tuple(m.n, m.nnz, (int)sizeof(T) + (no_sliced ? 16 : 0) + (comp ? 32 : 0));
    const auto it = c.spmv_tuned.find(key);
    if (it != c.spmv_tuned.end())
        return usable_choice(c, m, it->second);
    const short *c16 = m.ensure_col16(c);
    std::vector<SpmvChoice> cand = {{1, 8}, {1, 16}, {-1, 8}};
    if (c16) {
        cand.push_back({1, 8, true});
        cand.push_back({1, 16, true});
    }
    if (!no_sliced && m.ensure_sliced(c)) {
        cand.push_back({1, 8, false, 1});
        if (m.s16 > 0)
            cand.push_back({1, 8, false, 2});
    }
    const int ncand = (int)cand.size();
    DeviceArray<T> x((std::size_t)m.n), y((std::size_t)m.n);
    HIP_CHECK(hipMemsetAsync(x.data(), 0, (std::size_t)m.n * sizeof(T), c.stream));
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    const Tiles tl = m.layout->tiles();
    SpmvChoice best = rule;
    float best_ms = 0;
    for (int ci = 0; ci < ncand; ++ci) {
        const SpmvChoice &k = cand[ci];
        const T *sv = k.sliced ? m.sliced_values<T>(c) : nullptr;
        auto run = [&] {
            if constexpr (sizeof(T) == 4) {
                if (comp) {
                    if (k.sliced)
                        launch_spmv_sliced<T, T, Comp32, T>(c.stream, tl, nullptr, 0,
                                                            m.slice_view(k.sliced), sv, x.data(),
                                                            y.data());
                    else
                        launch_spmv<T, T, Comp32, T>(c.stream, k.lanes, tl, nullptr, 0,
                                                     m.ptr.data(), m.col.data(), m.values<T>(),
                                                     x.data(), y.data(), k.per,
                                                     k.c16 ? c16 : nullptr);
                    return;
                }
            }
            if (k.sliced)
                launch_spmv_sliced<T, T, T, T>(c.stream, tl, nullptr, 0, m.slice_view(k.sliced),
                                               sv, x.data(), y.data());
            else
                launch_spmv<T, T, T, T>(c.stream, k.lanes, tl, nullptr, 0, m.ptr.data(),
                                        m.col.data(), m.values<T>(), x.data(), y.data(), k.per,
                                        k.c16 ? c16 : nullptr);
        };
        run();
        HIP_CHECK(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 3; ++r)
            run();
        HIP_CHECK(hipEventRecord(e1, c.stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ci == 0 || ms < best_ms) {
            best_ms = ms;
            best = k;
        }
    }
    HIP_CHECK(hipEventDestroy(e0));
    HIP_CHECK(hipEventDestroy(e1));
    c.spmv_tuned[key] = best;
    return best;
}

void spmv_any(Context &c, const DevCsr &a, const double *x, double *y, int lanes) {
    launch_spmv<double, double, double, double>(c.stream, lanes ? lanes : a.lanes,
                                                a.layout->tiles(), nullptr, 0, a.ptr.data(),
                                                a.col.data(), a.val.data(), x, y);
}

// ---------------------------------------------------------------- SPAI
namespace {

__global__ void iota_kernel(int n, int *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = i;
}
__global__ void fill_int_kernel(std::int64_t n, int *out, int v) {
    const std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = v;
}
// Append columns with status == kSpaiOverflow to list (order irrelevant).
__global__ void collect_overflow_kernel(int nwork, const int *work, const int *status, int *list,
                                        int *count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nwork)
        return;
    const int k = work[i];
    if (status[k] == kSpaiOverflow)
        list[atomicAdd(count, 1)] = k;
}
// Reduce per-column outcomes: first failing column, totals.
__global__ void column_summary_kernel(int n, const int *status, const int *fits, const int *growths,
                                      const int *clamps, const int *q, const double *flops,
                                      const int *pass, int npass, int *first_fail,
                                      unsigned long long *tot, double *ftot) {
    // Each thread sums a strided share of the columns, then block sums
    // (exact: the counts stay far below 2^53), one atomic per block and
    // counter.
    __shared__ double lds[256 / kWave];
    double cnt[4] = {0, 0, 0, 0}, pf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int fail = INT_MAX;
    for (std::int64_t k = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (std::int64_t)gridDim.x * blockDim.x) {
        if (status[k] != kSpaiSuccess)
            fail = min(fail, (int)k);
        cnt[0] += fits[k];
        cnt[1] += growths[k];
        cnt[2] += clamps[k];
        cnt[3] += q[k];
        const int p = pass[k];
        for (int u = 0; u < 8; ++u)
            pf[u] += u == p ? flops[k] : 0.0;
    }
    if (fail != INT_MAX)
        atomicMin(first_fail, fail);
    for (int c = 0; c < 4; ++c) {
        const double v = block_sum(cnt[c], lds);
        if (threadIdx.x == 0)
            atomicAdd(tot + c, (unsigned long long)v);
    }
    // QR flops per capacity pass (telemetry only: summation order does not matter)
    for (int p = 0; p < npass; ++p) {
        const double f = block_sum(pf[p], lds);
        if (threadIdx.x == 0)
            atomicAdd(ftot + p, f);
    }
}
// Column pool pointers: per column (pass, offset) -> the pools' base pointers.
struct PoolTable {
    const int *j[8];
    const void *c[8];
};
// P = N^T D from the column pools: row k of P is column k of N, entries
// t = 0..q_k-1 in pattern order, value c_t d_{j_t}, zeros dropped.  One group
// of 16 lanes per column takes 16 entries at a time, so that a wavefront's
// loads cover four columns' consecutive pool slots (the first pass stores
// column k at slot k) and its stores four consecutive rows of P; an entry's
// place in its row is the count of kept entries before it (a ballot within
// the group), so the row keeps pattern order.
constexpr int kPoolGroup = 16;
template <typename T, bool kFill>
__device__ inline void pool_row(int k, const int *q, const int *pass, const std::int64_t *off,
                                const PoolTable &pt, const T *d, int *counts, int *asm_fail,
                                const int *pptr, int *pcol, T *pval) {
    const int g = (int)(threadIdx.x % kWave) / kPoolGroup; // group within the wave
    const int l = (int)(threadIdx.x % kPoolGroup);
    const bool has = k >= 0;
    const int qk = has ? q[k] : 0;
    const int *pj = has ? pt.j[pass[k]] + off[k] : nullptr;
    const T *pc = has ? static_cast<const T *>(pt.c[pass[k]]) + off[k] : nullptr;
    int w = kFill && has ? pptr[k] : 0;
    int cnt = 0;
    const unsigned long long below = (1ull << (g * kPoolGroup + l)) - 1ull;
    const unsigned long long gmask = 0xffffull << (g * kPoolGroup);
    // every lane runs the largest group's rounds of its wave (ballots)
    int rounds = (qk + kPoolGroup - 1) / kPoolGroup;
    for (int o = 1; o < kWave; o <<= 1)
        rounds = max(rounds, __shfl_xor(rounds, o, kWave));
    for (int i = 0; i < rounds; ++i) {
        const int t = i * kPoolGroup + l;
        const bool in = t < qk;
        T v = 0;
        int j = 0;
        if (in) {
            j = pj[t];
            const T cf = pc[t];
            v = cf * d[j];
            if (!kFill && (!isfinite(v) || (cf != 0 && v == 0)))
                atomicMin(asm_fail, k);
        }
        const bool keep = in && v != 0;
        const unsigned long long m = __ballot(keep) & gmask;
        if (kFill && keep) {
            const int at = w + __popcll(m & below);
            pcol[at] = j;
            pval[at] = v;
        }
        const int kept = __popcll(m);
        w += kept;
        cnt += kept;
    }
    if (!kFill && has && l == 0)
        counts[k + 1] = cnt;
}
template <typename T>
__global__ void __launch_bounds__(256)
p_count_pools_kernel(int n, const int *q, const int *pass, const std::int64_t *off, PoolTable pt,
                     const T *d, int *counts, int *asm_fail) {
    const std::int64_t k = ((std::int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kPoolGroup;
    pool_row<T, false>(k < n ? (int)k : -1, q, pass, off, pt, d, counts, asm_fail, nullptr,
                       nullptr, nullptr);
}
template <typename T>
__global__ void __launch_bounds__(256)
p_fill_pools_kernel(int n, const int *q, const int *pass, const std::int64_t *off, PoolTable pt,
                    const T *d, const int *pptr, int *pcol, T *pval) {
    const std::int64_t k = ((std::int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kPoolGroup;
    pool_row<T, true>(k < n ? (int)k : -1, q, pass, off, pt, d, nullptr, nullptr, pptr, pcol,
                      pval);
}
template <typename T>
__global__ void patterns_kernel(int n, const int *q, const int *pass, const std::int64_t *off,
                                PoolTable pt, const std::int64_t *optr, int *oidx, double *ocoef) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const int *pj = pt.j[pass[k]] + off[k];
    const T *pc = static_cast<const T *>(pt.c[pass[k]]) + off[k];
    const std::int64_t o = optr[k];
    for (int t = 0; t < q[k]; ++t) {
        oidx[o + t] = pj[t];
        ocoef[o + t] = (double)pc[t];
    }
}

const char *spai_column_message(int status) {
    switch (status) {
    case kSpaiRankDeficient:
    case kSpaiNumericalFailure:
        return "Householder QR/triangular fit failed";
    case kSpaiEmptyCandidates:
        return "No candidate remains before convergence";
    case kSpaiNoGrowth:
        return "No acceptable growth permitted before convergence";
    case kSpaiIterationLimit:
        return "Requested accuracy not reached within growth/refit budget";
    case kSpaiZeroDenominator:
    case kSpaiInvalidScore:
        return "Candidate scoring failed";
    default:
        return "Column construction failed";
    }
}

template <typename T>
void spai_build_typed(Context &c, const DevCsr &a, const T *aval, T *d, double eps,
                      long long alpha, int beta, SpaiBuild &out) {
    const int n = (int)a.n;
    hipStream_t s = c.stream;
    // ---- scaling and validity (src/spai.cpp:322-348)
    SpaiWork &w = c.spai;
    DeviceArray<int> &bad = w.bad;
    bad.resize(2);
    {
        const int init[2] = {INT_MAX, INT_MAX};
        upload(c, bad.data(), init, 2);
    }
    launch_scale_rows<T>(s, n, a.ptr.data(), a.col.data(), aval, d, bad.data());
    int hb[2];
    download(c, hb, bad.data(), 2);
    sync(c);
    if (hb[0] != INT_MAX) {
        const int row = hb[0] / 4, kind = hb[0] % 4;
        out.status = kSpaiInvalidInput;
        out.message = kind == 1 ? "Zero diagonal at row " + std::to_string(row)
                                : std::string("Reciprocal row scaling outside binary64 range");
        return;
    }
    if (hb[1] != INT_MAX) {
        out.status = kSpaiInvalidInput;
        out.message = "Scaling changes nonzero structure/range";
        return;
    }
    // ---- B structure = transpose of A's pattern (row i lists rows j with a_ji != 0)
    // FP32 first bucket: one wavefront per column (measured faster than one
    // lane per column from 4k to 2.1M columns: 58 against 103 ms at 128^3),
    // up to the column count its 32-bit sort keys hold; the lane kernel
    // beyond.  BSPAI_SPAI_FP32_KERNEL=lane|wave forces one.
    if (w.cus == 0)
        HIP_CHECK(hipDeviceGetAttribute(&w.cus, hipDeviceAttributeMultiprocessorCount, c.device));
    const int cus = w.cus;
    bool wave_fp32 = std::is_same<T, float>::value;
    if (const char *env = std::getenv("BSPAI_SPAI_FP32_KERNEL"))
        wave_fp32 = std::is_same<T, float>::value && std::string(env) == "wave";
    wave_fp32 = wave_fp32 && n <= spai_wave_max_n();
    DeviceArray<int> &bptr = w.bptr, &bcol = w.bcol, &cursor = w.cursor, &scan_tmp = w.scan_tmp;
    bptr.resize(n + 1);
    bcol.resize(a.nnz > 0 ? a.nnz : 1);
    cursor.resize(n + 1);
    scan_tmp.resize(scan_tmp_ints(n + 1));
    // B's values: the FP32 kernels score candidates from them
    const bool need_bval = std::is_same<T, float>::value;
    w.bval.resize((need_bval && a.nnz > 0 ? a.nnz : 1) * sizeof(T));
    T *const bval = w.bval.as<T>();
    HIP_CHECK(hipMemsetAsync(bptr.data(), 0, (n + 1) * sizeof(int), s));
    launch_count_cols(s, a.nnz, a.col.data(), bptr.data());
    scan_inclusive_int(s, bptr.data(), n + 1, scan_tmp.data());
    HIP_CHECK(hipMemcpyAsync(cursor.data(), bptr.data(), (n + 1) * sizeof(int),
                             hipMemcpyDeviceToDevice, s));
    launch_transpose_fill<T>(s, n, a.ptr.data(), a.col.data(), need_bval ? aval : nullptr,
                             bptr.data(), cursor.data(), bcol.data(),
                             need_bval ? bval : nullptr);
    // ---- columns, in capacity passes
    out.c_status.resize(n);
    out.c_fits.resize(n);
    out.c_growths.resize(n);
    out.c_clamps.resize(n);
    out.c_q.resize(n);
    out.c_res.resize(n);
    out.c_piv.resize(n);
    out.c_flops.resize(n);
    out.c_off.resize(n);
    out.c_pass.resize(n);
    DeviceArray<int> &work = w.work, &next = w.next, &counter = w.counter;
    work.resize(n);
    next.resize(n);
    counter.resize(2);
    hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, work.data());
    int nwork = n;
    int mcap = 64, qcap = 16, ccap = 128;
    // Columns in flight: the column kernel is latency-bound (dependent loads
    // of the lane's work area), so occupancy pays.  Overridable for tuning.
    int waves_per_cu = 32;
    if (const char *env = std::getenv("BSPAI_SPAI_WAVES_PER_CU"))
        waves_per_cu = std::max(1, std::atoi(env));
    const int max_waves = cus * waves_per_cu;
    std::size_t free_bytes = 0, total_bytes = 0;
    HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
    const std::int64_t area_cap =
        std::min<std::int64_t>(8LL << 30, std::max<std::int64_t>(1LL << 28, free_bytes / 4));
    DeviceBuffer &dwork = w.dwork, &iwork = w.iwork;
    // Launch shape of the workgroup kernel per capacity bucket: the waves per
    // SIMD its registers are held to (3 or 4) and the workgroups per CU
    // launched (persistent: each takes columns from the bucket's worklist
    // until it is empty).  The rule: 4 waves in the first workgroup bucket,
    // 3 beyond, and 2 workgroups per CU (measured on the power-law buckets,
    // whose work areas of tens of MB thrash the caches at 4 per CU: the
    // tuning below chose 2 per CU in every large bucket).  A bucket of at
    // least tune_min columns is tuned instead: each candidate shape fits the
    // same slice of the worklist (one round of the largest grid) between HIP
    // events, the
    // fastest is kept per (n, nnz, value type, bucket) in the context, and
    // the bucket goes on from the slice's end.  A column's fit does not
    // depend on the workgroup that runs it, so the shape changes timing
    // only, never a result bit.  BSPAI_SPAI_BLOCK_OCC=3 or 4 sets the waves
    // of every bucket (no tuning); BSPAI_SPAI_AUTOTUNE=0 keeps the rule.
    int occ_env = 0;
    if (const char *env = std::getenv("BSPAI_SPAI_BLOCK_OCC"))
        occ_env = std::atoi(env);
    if (occ_env != 3 && occ_env != 4)
        occ_env = 0;
    bool tune = occ_env == 0;
    if (const char *env = std::getenv("BSPAI_SPAI_AUTOTUNE"))
        tune = tune && std::string(env) != "0";
    int tune_min = 16 * cus * 4;
    if (const char *env = std::getenv("BSPAI_SPAI_TUNE_MIN"))
        tune_min = std::max(1, std::atoi(env));
    const std::pair<int, int> shapes[4] = {{4, 4}, {3, 3}, {4, 2}, {3, 2}};
    std::vector<std::string> pass_labels; // timer label of each capacity pass
    while (nwork > 0) {
        if (out.passes >= 8)
            throw std::runtime_error("SPAI column exceeds the largest capacity bucket");
        const int pass = out.passes++;
        if (pass > 0) {
            // Columns that outgrew a bucket: one workgroup per column, caps
            // growing 4x per pass (never beyond n).
            if (pass == 1) {
                mcap = 2048;
                qcap = 128;
                ccap = 16384;
            } else {
                mcap *= 4;
                qcap *= 4;
                ccap *= 4;
            }
            // slots in multiples of 16: 16-byte aligned columns for the FP32
            // build's matrix-core sweeps (a cap above n is never reached)
            mcap = (int)((std::min<std::int64_t>(mcap, n) + 15) / 16 * 16);
            qcap = (int)std::min<std::int64_t>(qcap, n);
            ccap = (int)std::min<std::int64_t>(ccap, n);
            const int words = (int)((n + 31) / 32);
            // (the pair arrays serve the FP32 scoring only)
            const std::int64_t pcap =
                sizeof(T) == sizeof(double) ? 1 : spai_pair_cap(ccap, n, a.nnz);
            const std::int64_t bd = spai_block_doubles(mcap, qcap, ccap, pcap);
            const std::int64_t bi = spai_block_ints(mcap, qcap, ccap, pcap, words, n);
            const std::int64_t per_block = bd * (std::int64_t)sizeof(T) + bi * 4;
            // the bucket's launch shape: forced, cached, tuned on a slice, or the rule
            std::pair<int, int> shape = std::make_pair(occ_env ? occ_env : (pass == 1 ? 4 : 3), 2);
            const auto key = std::make_tuple((std::int64_t)n, a.nnz, (int)sizeof(T), pass);
            const bool cached = tune && c.spai_tuned.count(key) > 0;
            if (cached)
                shape = c.spai_tuned[key];
            const bool tune_now = tune && !cached && nwork >= tune_min;
            // work areas for the largest grid the bucket launches
            std::int64_t blocks = std::min<std::int64_t>(nwork, (std::int64_t)cus * (tune_now ? 4 : shape.second));
            // Large buckets hold tens of MB per workgroup: allow up to a quarter
            // of free memory (48 GiB at most) so that a full wave of
            // workgroups stays resident.
            const std::int64_t block_cap =
                std::min<std::int64_t>(48LL << 30, std::max<std::int64_t>(area_cap, free_bytes / 4));
            blocks = std::max<std::int64_t>(1, std::min<std::int64_t>(blocks, block_cap / per_block));
            dwork.resize((std::size_t)(blocks * bd * (std::int64_t)sizeof(T)));
            iwork.resize((std::size_t)(blocks * bi * 4));
            // bitmaps clear, position maps -1 (the kernel keeps them so)
            HIP_CHECK(hipMemsetAsync(iwork.as<int>(), 0, (std::size_t)(blocks * bi * 4), s));
            launch_fill_int(s, blocks, bi, bi - n, n, iwork.as<int>(), -1);
            out.pool_j.emplace_back((std::size_t)nwork * qcap);
            out.pool_c.emplace_back((std::size_t)nwork * qcap * sizeof(T));
            HIP_CHECK(hipMemsetAsync(counter.data(), 0, 2 * sizeof(int), s));
            SpaiColumnOut o{out.c_status.data(), out.c_fits.data(),  out.c_growths.data(),
                            out.c_clamps.data(), out.c_q.data(),     out.c_res.data(),
                            out.c_piv.data(),    out.c_flops.data(), out.c_off.data(),
                            out.c_pass.data(),   pass};
            auto run = [&](int occ, std::int64_t nb, int count) {
                launch_spai_block<T>(s, (int)nb, n, a.ptr.data(), a.col.data(), aval, d,
                                     bptr.data(), bcol.data(), need_bval ? bval : nullptr,
                                     (T)eps, alpha, beta, work.data(), count, counter.data(),
                                     mcap, qcap, ccap, dwork.as<T>(), iwork.as<int>(), o,
                                     out.pool_j.back().data(), out.pool_c.back().as<T>(), qcap,
                                     occ, pcap);
            };
            int start = 0; // worklist slots already fitted (by the tuning)
            if (tune_now) {
                const int slice = (int)std::min<std::int64_t>(nwork, blocks);
                hipEvent_t t0, t1;
                HIP_CHECK(hipEventCreate(&t0));
                HIP_CHECK(hipEventCreate(&t1));
                float best_ms = 0;
                bool first = true;
                for (const auto &cand : shapes) {
                    const std::int64_t nb =
                        std::min<std::int64_t>(blocks, (std::int64_t)cus * cand.second);
                    HIP_CHECK(hipMemsetAsync(counter.data(), 0, sizeof(int), s));
                    HIP_CHECK(hipEventRecord(t0, s));
                    run(cand.first, nb, slice);
                    HIP_CHECK(hipEventRecord(t1, s));
                    HIP_CHECK(hipEventSynchronize(t1));
                    float ms = 0;
                    HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
                    if (first || ms < best_ms) {
                        best_ms = ms;
                        shape = cand;
                        first = false;
                    }
                }
                HIP_CHECK(hipEventDestroy(t0));
                HIP_CHECK(hipEventDestroy(t1));
                c.spai_tuned[key] = shape;
                start = slice;
            }
            const std::int64_t nb = std::min<std::int64_t>(blocks, (std::int64_t)cus * shape.second);
            if (start > 0) {
                static thread_local int start_host = 0;
                start_host = start;
                upload(c, counter.data(), &start_host, 1);
                sync(c);
            }
            const std::string label = std::string(kSpaiBlockKernelName<T>) + " pass " +
                                      std::to_string(pass) + " (" + std::to_string(nwork) +
                                      " columns, occ " + std::to_string(shape.first) + ", " +
                                      std::to_string(shape.second) + "/CU" +
                                      (start > 0 ? ", tuned" : "") + ")";
            pass_labels.push_back(label);
            timed(c, label.c_str(), 0.0, 0.0, [&] { run(shape.first, nb, nwork); }, true);
            hipLaunchKernelGGL(collect_overflow_kernel, dim3(ceil_div(nwork, 256)), dim3(256), 0,
                               s, nwork, work.data(), out.c_status.data(), next.data(),
                               counter.data() + 1);
            int nover = 0;
            download(c, &nover, counter.data() + 1, 1);
            sync(c);
            std::swap(work, next);
            nwork = nover;
            continue;
        }
        if (wave_fp32) {
            // FP32: one wavefront per column (a one-wave workgroup).
            const int wq = spai_wave_qcap();
            const int blocks = (int)std::max<std::int64_t>(
                1, std::min<std::int64_t>(nwork, cus * 16LL));
            out.pool_j.emplace_back((std::size_t)nwork * wq);
            out.pool_c.emplace_back((std::size_t)nwork * wq * sizeof(T));
            HIP_CHECK(hipMemsetAsync(counter.data(), 0, 2 * sizeof(int), s));
            SpaiColumnOut o{out.c_status.data(), out.c_fits.data(), out.c_growths.data(),
                            out.c_clamps.data(), out.c_q.data(), out.c_res.data(),
                            out.c_piv.data(), out.c_flops.data(), out.c_off.data(),
                            out.c_pass.data(), pass};
            pass_labels.push_back("spai_wave_kernel<float>");
            timed(c, "spai_wave_kernel<float>", 0.0, 0.0, [&] {
                launch_spai_wave(s, blocks, a.ptr.data(), a.col.data(), (const float *)aval,
                                 (const float *)d, bptr.data(), bcol.data(),
                                 (const float *)bval, (float)eps, alpha, beta,
                                 work.data(), nwork, counter.data(), o,
                                 out.pool_j.back().data(), out.pool_c.back().as<float>(), wq);
            }, true);
            hipLaunchKernelGGL(collect_overflow_kernel, dim3(ceil_div(nwork, 256)), dim3(256), 0,
                               s, nwork, work.data(), out.c_status.data(), next.data(),
                               counter.data() + 1);
            int nover = 0;
            download(c, &nover, counter.data() + 1, 1);
            sync(c);
            std::swap(work, next);
            nwork = nover;
            continue;
        }
        const std::int64_t ld = spai_lane_doubles(mcap, qcap, ccap);
        const std::int64_t li = spai_lane_ints(mcap, qcap, ccap);
        // Lanes in flight: bounded by the device and by the work-area cap
        // (8 GiB, or a quarter of free memory).
        const int threads = 256, waves_per_block = threads / kWave;
        std::int64_t waves = std::min<std::int64_t>(max_waves, (nwork + kWave - 1) / kWave);
        const std::int64_t per_wave = (ld * (std::int64_t)sizeof(T) + li * 4) * kWave;
        waves = std::max<std::int64_t>(1, std::min<std::int64_t>(waves, area_cap / per_wave));
        // Every launched lane may take a column, so the work area covers whole blocks.
        const int blocks = (int)((waves + waves_per_block - 1) / waves_per_block);
        waves = (std::int64_t)blocks * waves_per_block;
        dwork.resize((std::size_t)(waves * ld * kWave * sizeof(T)));
        iwork.resize((std::size_t)(waves * li * kWave * 4));
        out.pool_j.emplace_back((std::size_t)nwork * qcap);
        out.pool_c.emplace_back((std::size_t)nwork * qcap * sizeof(T));
        HIP_CHECK(hipMemsetAsync(counter.data(), 0, 2 * sizeof(int), s));
        SpaiColumnOut o{out.c_status.data(), out.c_fits.data(), out.c_growths.data(),
                        out.c_clamps.data(), out.c_q.data(), out.c_res.data(),
                        out.c_piv.data(), out.c_flops.data(), out.c_off.data(),
                        out.c_pass.data(), pass};
        pass_labels.push_back(kSpaiKernelName<T>);
        timed(c, kSpaiKernelName<T>, 0.0, 0.0, [&] {
            launch_spai_columns<T>(s, blocks, threads, n, a.ptr.data(), a.col.data(), aval, d,
                                   bptr.data(), bcol.data(), (T)eps, alpha, beta, work.data(),
                                   nwork, counter.data(), mcap, qcap, ccap, dwork.as<T>(),
                                   iwork.as<int>(), o, out.pool_j.back().data(),
                                   out.pool_c.back().as<T>(), qcap);
        }, true);
        hipLaunchKernelGGL(collect_overflow_kernel, dim3(ceil_div(nwork, 256)), dim3(256), 0, s,
                           nwork, work.data(), out.c_status.data(), next.data(),
                           counter.data() + 1);
        int nover = 0;
        download(c, &nover, counter.data() + 1, 1);
        sync(c);
        std::swap(work, next);
        nwork = nover;
        mcap *= 4;
        qcap *= 3;
        ccap *= 4;
    }
    // ---- summary and P assembly (src/spai.cpp:352-372)
    DeviceArray<int> &ff = w.ff;
    DeviceArray<unsigned long long> &tot = w.tot;
    DeviceArray<double> &ftot = w.ftot;
    ff.resize(2);
    tot.resize(4);
    ftot.resize(8);
    {
        const int init[2] = {INT_MAX, INT_MAX};
        upload(c, ff.data(), init, 2);
        HIP_CHECK(hipMemsetAsync(tot.data(), 0, 4 * sizeof(unsigned long long), s));
        HIP_CHECK(hipMemsetAsync(ftot.data(), 0, 8 * sizeof(double), s));
    }
    hipLaunchKernelGGL(column_summary_kernel, dim3(std::min<unsigned>(ceil_div(n, 256), 1024)),
                       dim3(256), 0, s, n,
                       out.c_status.data(), out.c_fits.data(), out.c_growths.data(),
                       out.c_clamps.data(), out.c_q.data(), out.c_flops.data(),
                       out.c_pass.data(), out.passes, ff.data(), tot.data(), ftot.data());
    PoolTable pt{};
    for (int p = 0; p < out.passes; ++p) {
        pt.j[p] = out.pool_j[p].data();
        pt.c[p] = out.pool_c[p].as<T>();
    }
    out.p.n = n;
    out.p.ptr.resize(n + 1);
    HIP_CHECK(hipMemsetAsync(out.p.ptr.data(), 0, sizeof(int), s));
    const unsigned pool_blocks = (unsigned)ceil_div((std::int64_t)n * kPoolGroup, 256);
    hipLaunchKernelGGL(p_count_pools_kernel<T>, dim3(pool_blocks), dim3(256), 0, s, n,
                       out.c_q.data(), out.c_pass.data(), out.c_off.data(), pt, (const T *)d,
                       out.p.ptr.data(), ff.data() + 1);
    scan_inclusive_int(s, out.p.ptr.data(), n + 1, scan_tmp.data());
    int hff[2];
    unsigned long long htot[4];
    int pnnz = 0;
    download(c, hff, ff.data(), 2);
    download(c, htot, tot.data(), 4);
    download(c, &pnnz, out.p.ptr.data() + n, 1);
    double pass_flops[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    download(c, pass_flops, ftot.data(), 8);
    sync(c);
    out.qr_flops = 0;
    for (int p = 0; p < out.passes; ++p) {
        out.qr_flops += pass_flops[p];
        c.timer.add_flops(pass_labels[p].c_str(), pass_flops[p]);
    }
    out.fits = (std::int64_t)htot[0];
    out.growths = (std::int64_t)htot[1];
    out.clamps = (std::int64_t)htot[2];
    out.pattern_nnz = (std::int64_t)htot[3];
    out.p_nnz = pnnz;
    out.p.nnz = pnnz;
    out.p.col.resize(pnnz > 0 ? pnnz : 1);
    if (sizeof(T) == 8) {
        out.p.val.resize(pnnz > 0 ? pnnz : 1);
        out.p.has64 = true;
        out.p.has32 = false;
    } else {
        out.p.val32.resize(pnnz > 0 ? pnnz : 1);
        out.p.has32 = true;
        out.p.has64 = false;
    }
    T *pval = sizeof(T) == 8 ? (T *)(void *)out.p.val.data() : (T *)(void *)out.p.val32.data();
    hipLaunchKernelGGL(p_fill_pools_kernel<T>, dim3(pool_blocks), dim3(256), 0, s, n,
                       out.c_q.data(), out.c_pass.data(), out.c_off.data(), pt, (const T *)d,
                       out.p.ptr.data(), out.p.col.data(), pval);
    out.p.lanes = choose_lanes(n, pnnz);
    out.p.layout = a.layout;
    if (hff[1] != INT_MAX) {
        out.status = kSpaiNumericalFailure;
        out.failing_column = hff[1];
        out.message = "Preconditioner assembly outside binary64 range";
    } else if (hff[0] != INT_MAX) {
        int st = 0;
        download(c, &st, out.c_status.data() + hff[0], 1);
        sync(c);
        out.status = st;
        out.failing_column = hff[0];
        out.message = "Column " + std::to_string(hff[0]) + ": " + spai_column_message(st);
    } else {
        out.status = kSpaiSuccess;
    }
    sync(c);
}

} // namespace

std::unique_ptr<SpaiBuild> spai_build(Context &c, const DevCsr &a, double eps, long long alpha,
                                      int beta, int precision) {
    auto out = std::make_unique<SpaiBuild>();
    out->n = a.n;
    out->precision = precision;
    out->status = kSpaiInvalidInput;
    if (a.n == 0 || !std::isfinite(eps) || eps <= 0) {
        out->message = "Require positive square dimensions and positive finite epsilon";
        return out;
    }
    HIP_CHECK(hipEventRecord(c.ev0, c.stream));
    if (precision == 0) {
        out->d64.resize(a.n);
        spai_build_typed<double>(c, a, a.val.data(), out->d64.data(), eps, alpha, beta, *out);
    } else {
        const_cast<DevCsr &>(a).ensure32(c);
        out->d32.resize(a.n);
        spai_build_typed<float>(c, a, a.val32.data(), out->d32.data(), (float)eps, alpha, beta,
                                *out);
    }
    HIP_CHECK(hipEventRecord(c.ev1, c.stream));
    HIP_CHECK(hipEventSynchronize(c.ev1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    out->build_ms = ms;
    return out;
}

// Gather per-column patterns into compact arrays (host ptr computed from q).
void spai_patterns(Context &c, const SpaiBuild &b, std::int64_t *hptr, int *hidx,
                   double *hcoef) {
    const int n = (int)b.n;
    std::vector<int> q(n);
    download(c, q.data(), b.c_q.data(), n);
    sync(c);
    hptr[0] = 0;
    for (int k = 0; k < n; ++k)
        hptr[k + 1] = hptr[k] + q[k];
    const std::int64_t tot = hptr[n];
    DeviceArray<std::int64_t> optr(n + 1);
    DeviceArray<int> oidx(tot > 0 ? tot : 1);
    DeviceArray<double> ocoef(tot > 0 ? tot : 1);
    upload(c, optr.data(), hptr, n + 1);
    PoolTable pt{};
    for (int p = 0; p < b.passes; ++p) {
        pt.j[p] = b.pool_j[p].data();
        pt.c[p] = b.pool_c[p].as<void>();
    }
    if (b.precision == 0)
        hipLaunchKernelGGL(patterns_kernel<double>, dim3(ceil_div(n, 256)), dim3(256), 0,
                           c.stream, n, b.c_q.data(), b.c_pass.data(), b.c_off.data(), pt,
                           optr.data(), oidx.data(), ocoef.data());
    else
        hipLaunchKernelGGL(patterns_kernel<float>, dim3(ceil_div(n, 256)), dim3(256), 0,
                           c.stream, n, b.c_q.data(), b.c_pass.data(), b.c_off.data(), pt,
                           optr.data(), oidx.data(), ocoef.data());
    HIP_LAUNCH_CHECK();
    download(c, hidx, oidx.data(), tot);
    download(c, hcoef, ocoef.data(), tot);
    sync(c);
}

// ---------------------------------------------------------------- GMRES
template <typename T>
GmresEngine<T>::GmresEngine(Context &c, const Layout &L, int cap) : c_(c), L_(L), cap_(cap) {
    const std::int64_t n = L.n;
    w_.resize(n);
    t_.resize(n);
    prhs_.resize(n);
    r0_.resize(n);
    xz_.resize(n);
    ax_.resize(n);
    pax_.resize(n);
    cs_.resize((std::size_t)cap * L.nsys);
    sn_.resize((std::size_t)cap * L.nsys);
    g_.resize((std::size_t)(cap + 1) * L.nsys);
    y_.resize((std::size_t)cap * L.nsys);
    part_.resize((std::size_t)kNumSlots * 2 * std::max(1, L.ntiles));
    q_.resize((std::size_t)(cap + 1) * L.nsys);
    h_.resize((std::size_t)(cap + 1) * L.nsys);
    st_.resize(L.nsys);
    vdev_.resize(cap + 1);
    vptrs_.reserve(cap + 1);
    args_.st = st_.data();
    args_.V = vdev_.data();
    args_.w = w_.data();
    args_.cs = cs_.data();
    args_.sn = sn_.data();
    args_.g = g_.data();
    args_.y = y_.data();
    args_.cap = cap;
    args_.part = part_.data();
    args_.q = q_.data();
    args_.h = h_.data();
    args_.defer = 0;
    args_.kdev = nullptr;
    ensure_steps(0);
}

template <typename T> GmresEngine<T>::~GmresEngine() {
    for (StepGraph &g : graphs_)
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
}

template <typename T> void GmresEngine<T>::ensure_steps(int k) {
    // H and the Gram triangle are O(steps^2) per system: a cap of n (the
    // reference CLI's innercap 0) must not allocate n^2 up front, so they
    // start at 64 steps and double as the Krylov space grows.
    if (k < ksteps_)
        return;
    const int K = std::min(cap_, std::max({k + 1, 2 * ksteps_, 64}));
    const std::size_t nsys = (std::size_t)L_.nsys;
    const std::int64_t hs = hessenberg_size(K), gs = gram_size(K);
    DeviceArray<double> H2((std::size_t)hs * nsys), G2((std::size_t)gs * nsys);
    if (ksteps_ > 0) {
        const std::int64_t hs0 = hessenberg_size(ksteps_), gs0 = gram_size(ksteps_);
        HIP_CHECK(hipMemcpy2DAsync(H2.data(), hs * sizeof(double), H_.data(),
                                   hs0 * sizeof(double), hs0 * sizeof(double), nsys,
                                   hipMemcpyDeviceToDevice, c_.stream));
        gmres_gram_relayout(c_.stream, (int)nsys, gram_.data(), gs0, ksteps_, G2.data(), gs, K);
        sync(c_); // the old buffers are released below
    }
    H_ = std::move(H2);
    gram_ = std::move(G2);
    // per-tile partials of one step (2k+2 values): no contents to keep
    gp_.resize((std::size_t)2 * K * std::max(1, L_.ntiles));
    args_.H = H_.data();
    args_.hstride = hs;
    args_.gram = gram_.data();
    args_.gstride = gs;
    args_.gsteps = K;
    args_.gp = gp_.data();
    ksteps_ = K;
}

template <typename T> void GmresEngine<T>::ensure_basis(int count) {
    // Basis vectors are allocated in chunks as the Krylov space grows.
    const int chunk = 16;
    if (count <= vcount_)
        return;
    // Vector stride: n rounded up to 64 elements (every basis vector starts
    // 256-byte aligned, which the 16-byte accesses of the Gram-Schmidt passes
    // rely on), plus a skew of 576 elements.  The passes read several vectors
    // at the same rows at once; at a stride that is a multiple of a large
    // power of two (n = 2^21 at 128^3) those reads fall on the same HBM
    // channels.  Measured at 128^3 (mixed): projection 127 -> 117 us,
    // update 116 -> 109 us per launch; FP64 projection 280 -> 247 us.
    constexpr std::size_t kBasisSkew = 576;
    const std::size_t ld = (std::size_t)((L_.n + 63) / 64 * 64) + kBasisSkew;
    // (the table's earlier upload may still be queued behind a running
    // graph: let it finish before the host vector can move)
    sync(c_);
    while ((int)vptrs_.size() < count) {
        const int take = std::min(chunk, cap_ + 1 - (int)vptrs_.size());
        chunks_.emplace_back((std::size_t)take * ld);
        for (int i = 0; i < take; ++i)
            vptrs_.push_back(chunks_.back().data() + (std::size_t)i * ld);
    }
    upload(c_, vdev_.data(), vptrs_.data(), vptrs_.size());
    vcount_ = (int)vptrs_.size();
}

template <typename T>
void GmresEngine<T>::solve(const DevCsr &m1, const DevCsr &m2, GmresMode mode, const T *rhs,
                           const double *x0, double tol, double *x, const IrSysState *ir,
                           const InnerTolRule &rule) {
    hipStream_t s = c_.stream;
    const Tiles tl = L_.tiles();
    const bool left = mode == kLeft;
    args_.x = left ? x : xz_.data();
    args_.x0 = left ? x0 : nullptr;
    ensure_basis(std::min(cap_ + 1, 2));
    gmres_reset<T>(s, L_.nsys, args_, ir, L_.s_rows.data(), tol, rule);
    SysState *st = st_.data();
    double *part = part_.data();
    const int *active = reinterpret_cast<const int *>(reinterpret_cast<const char *>(st) +
                                                      offsetof(SysState, active));
    const int *trig = reinterpret_cast<const int *>(reinterpret_cast<const char *>(st) +
                                                    offsetof(SysState, triggered));
    const int stride = (int)(sizeof(SysState) / sizeof(int));
    const T *m1v = m1.values<T>(), *m2v = m2.values<T>();
    const SpmvChoice k1 = spmv_choice<T>(c_, m1, left), k2 = spmv_choice<T>(c_, m2, !left);
    // The FP32 solve forms its products with A (left) or A^T (adjoint) as
    // compensated FP32 row sums (Acc<Comp32>): rows of A cancel (a
    // stencil's interior rows sum to zero), and plain FP32 sums there lose
    // digits in proportion to the cancellation.  The compensated sums keep
    // each product at one FP32 rounding, which lowers the inner iteration
    // counts at no extra bytes.
    const T *none = nullptr;
    // ---- start (src/gmres.cpp:134-167)
    gmres_diff_partial<T, T, T>(s, tl, st, part, rhs, none, (T *)nullptr, kSlotRhs);
    if (left && x0)
        gmres_spmv_resid<T, T>(s, tl, st, part, m1.ptr.data(), m1.col.data(), m1v, x0, rhs,
                               ax_.data(), kSlotOrig, 0);
    else
        gmres_diff_partial<T, T, T>(s, tl, st, part, rhs, none, (T *)nullptr, kSlotOrig);
    gmres_decide<T>(s, tl, args_, kPhaseInit, !left);
    const T *prhs = rhs;
    if (left) {
        launch_spmv<T, T, T, T>(s, k2.lanes, tl, active, stride, m2.ptr.data(), m2.col.data(),
                                m2v, rhs, prhs_.data(), k2.per);
        prhs = prhs_.data();
    }
    gmres_diff_partial<T, T, T>(s, tl, st, part, prhs, none, (T *)nullptr, kSlotDen);
    gmres_decide<T>(s, tl, args_, kPhaseDen, !left);
    const T *r0 = prhs;
    if (left && x0) {
        launch_spmv<T, double, double, double>(s, m2.lanes, tl, active, stride, m2.ptr.data(),
                                               m2.col.data(), m2v, ax_.data(), pax_.data());
        gmres_diff_partial<T, double, T>(s, tl, st, part, prhs, pax_.data(), r0_.data(),
                                         kSlotBeta);
        r0 = r0_.data();
    } else {
        gmres_diff_partial<T, T, T>(s, tl, st, part, prhs, none, (T *)nullptr, kSlotBeta);
    }
    gmres_decide<T>(s, tl, args_, kPhaseBeta, !left);
    gmres_start_vector<T>(s, tl, args_, r0);
    // ---- Arnoldi loop (src/gmres.cpp:172-254)
    std::vector<SysState> host(L_.nsys);
    // Timer labels and algorithmic bytes: SpMV reads the matrix (value +
    // int32 column per entry, int32 row offsets) and x once, writes y; the
    // Gram-Schmidt passes stream V_0..V_k plus w and V_k (project) or w in
    // and out (update).
    const char *tname = sizeof(T) == 8 ? "double" : "float";
    const char *role1 = left ? "A" : "P^T", *role2 = left ? "P" : "A^T";
    const std::string l1 = spmv_label(k1.lanes, tname, role1, k1.sliced),
                      l2 = spmv_label(k2.lanes, tname, role2, k2.sliced);
    const std::string lproj = std::string("project_kernel<") + tname + ">";
    const std::string lred = std::string("gram_reduce_kernel<") + tname + ">";
    const std::string lsol = std::string("gram_solve_kernel<") + tname + ">";
    const std::string lupd = std::string("gram_update_kernel<") + tname + ">";
    const std::string lnrm = std::string("normalize_givens_kernel<") + tname + ">";
    const double sz = sizeof(T), nn = (double)L_.n;
    // (the algorithmic bytes count a 4-byte column index per entry, also
    // where the kernel reads 16-bit offsets: value + int32 column per entry,
    // int32 row offsets, x once, y once)
    auto spmv_bytes = [&](const DevCsr &m, const SpmvChoice &) {
        return (double)m.nnz * (sz + 4) + 4.0 * (double)(m.n + 1) + 2.0 * (double)m.n * sz;
    };
    const short *c16a = k1.c16 ? m1.ensure_col16(c_) : nullptr;
    const short *c16b = k2.c16 ? m2.ensure_col16(c_) : nullptr;
    const T *m1sv = k1.sliced ? m1.sliced_values<T>(c_) : nullptr;
    const T *m2sv = k2.sliced ? m2.sliced_values<T>(c_) : nullptr;
    // (every pointer a replayed graph bakes in is part of its key below)
    const SliceView sv1 = k1.sliced ? m1.slice_view(k1.sliced) : SliceView{};
    const SliceView sv2 = k2.sliced ? m2.slice_view(k2.sliced) : SliceView{};
    // Step k's kernels are gated by run[k & 1] (active and not held).
    auto run_flag = [&](int k) {
        return reinterpret_cast<const int *>(reinterpret_cast<const char *>(st) +
                                             offsetof(SysState, run) + (k & 1) * sizeof(int));
    };
    // Verification of a triggered step (src/gmres.cpp:215-253): candidate,
    // recomputed residual(s), accept / fail / go on.  Every kernel is gated on
    // active && triggered.
    auto verify = [&]() {
        gmres_candidate<T>(s, tl, args_);
        if (left) {
            gmres_spmv_resid<T, T>(s, tl, st, part, m1.ptr.data(), m1.col.data(), m1v, x, rhs,
                                   ax_.data(), kSlotChkOrig, 1);
            gmres_spmv_resid<T, T>(s, tl, st, part, m2.ptr.data(), m2.col.data(), m2v, ax_.data(),
                                   prhs, pax_.data(), kSlotChkRel, 1);
        } else {
            // x = P^T z, then ||g - A^T x|| (absolute; relative to ||g||).
            launch_spmv<T, double, double, double>(s, m1.lanes, tl, trig, stride, m1.ptr.data(),
                                                   m1.col.data(), m1v, xz_.data(), x);
            gmres_spmv_resid<T, T>(s, tl, st, part, m2.ptr.data(), m2.col.data(), m2v, x, rhs,
                                   ax_.data(), kSlotChkOrig, 1);
        }
        gmres_decide<T>(s, tl, args_, kPhaseCheck, !left);
    };
    // A single system defers verification: until its first trigger, the
    // verification kernels are not launched after every step (they would
    // exit at once).  A triggered step holds the system -- later launched
    // steps do nothing -- and the host, at its next check point, runs the
    // verification of the held step and resumes after it, so every accept
    // decision is the one the step would have made at once.  From then on
    // every step is verified.  Batched systems step in lockstep and verify
    // every step.
    // (BSPAI_GMRES_VERIFY_EVERY_STEP=1 keeps per-step verification, so the
    // two schedules can be compared.)
    const char *every = std::getenv("BSPAI_GMRES_VERIFY_EVERY_STEP");
    bool defer = L_.nsys == 1 && !(every && std::string(every) == "1");
    // One Arnoldi step (src/gmres.cpp:172-254): the two products, the compact
    // Gram-Schmidt passes and the Givens step, then (unless deferred) the
    // verification.  a.kdev == nullptr: k is the step; else k is the offset
    // to the device's step counter, and xs picks V_k on the device.
    auto step = [&](const KrylovArgs<T> &a, int k, int kmax, bool verify_now, XSel xs) {
        const int *run = run_flag(k);
        const T *vk = a.kdev ? nullptr : vptrs_[k];
        // (TA: the row-sum accumulator; Comp32 for the product with A)
        auto product = [&](const DevCsr &mm, const SpmvChoice &kc, const T *mv, const T *msv,
                           const T *in, T *out, const short *c16, XSel sel, bool with_a) {
            if (kc.sliced) {
                const SliceView &view = &mm == &m1 ? sv1 : sv2;
                if constexpr (sizeof(T) == 4) {
                    if (with_a) {
                        launch_spmv_sliced<T, T, Comp32, T>(s, tl, run, stride, view, msv, in,
                                                            out, sel);
                        return;
                    }
                }
                launch_spmv_sliced<T, T, T, T>(s, tl, run, stride, view, msv, in, out, sel);
                return;
            }
            if constexpr (sizeof(T) == 4) {
                if (with_a) {
                    launch_spmv<T, T, Comp32, T>(s, kc.lanes, tl, run, stride, mm.ptr.data(),
                                                 mm.col.data(), mv, in, out, kc.per, c16, sel);
                    return;
                }
            }
            launch_spmv<T, T, T, T>(s, kc.lanes, tl, run, stride, mm.ptr.data(), mm.col.data(),
                                    mv, in, out, kc.per, c16, sel);
        };
        timed(c_, l1.c_str(), spmv_bytes(m1, k1), 2.0 * (double)m1.nnz,
              [&] { product(m1, k1, m1v, m1sv, vk, t_.data(), c16a, xs, left); });
        timed(c_, l2.c_str(), spmv_bytes(m2, k2), 2.0 * (double)m2.nnz,
              [&] { product(m2, k2, m2v, m2sv, t_.data(), w_.data(), c16b, XSel{}, !left); });
        timed(c_, lproj.c_str(), (k + 3) * nn * sz, 2.0 * (2 * k + 1) * nn,
              [&] { gmres_project<T>(s, tl, a, k, kmax); });
        timed(c_, lred.c_str(), 2.0 * (2 * k + 1) * tl.ntiles * 8.0, 0.0,
              [&] { gmres_gram_reduce<T>(s, tl, a, k, kmax); });
        // (for kmax <= kGramSolveLdsMaxK the update pass also solves for h)
        if (!gmres_gram_solve_fused(kmax))
            timed(c_, lsol.c_str(), 0.0, 0.0, [&] { gmres_gram_solve<T>(s, tl, a, k, kmax); });
        timed(c_, lupd.c_str(), (k + 3) * nn * sz, 2.0 * (k + 1) * nn,
              [&] { gmres_gram_update<T>(s, tl, a, k, kmax); });
        timed(c_, lnrm.c_str(), 2.0 * nn * sz, 0.0,
              [&] { gmres_normalize_givens<T>(s, tl, a, k); });
        if (verify_now)
            verify();
    };
    // Replayed graphs: kGraphSteps steps captured once per (operators,
    // workspace, schedule) and replayed between the host's check points.
    // Their launch shapes hold any step the workspace holds (blocks beyond
    // the step exit at once); the workspace still grows on demand (doubling
    // from 64 steps), and a growth moves H, the Gram triangle and the
    // partials, so the graph is captured again for the larger capacity.
    const char *genv = std::getenv("BSPAI_GMRES_GRAPH");
    // With the step kernels timed (not only the SPAI build's), every kernel
    // is launched on the stream with its events, and the host checks after
    // every step, so that no held step (which exits at once) is launched
    // and counted with a full step's bytes.
    const bool timed_steps = c_.timer.on && !c_.timer.spai_only;
    const int check_every = timed_steps ? 1 : sync_every;
    const bool graphs = !(genv && std::string(genv) == "0") && !timed_steps &&
                        cap_ >= kGraphSteps && check_every % kGraphSteps == 0;
    graphs_replayed = 0;
    if (graphs && kdev_.size() == 0)
        kdev_.resize(1);
    auto graph_for = [&](bool verify_each, int kmax) -> hipGraphExec_t {
        KrylovArgs<T> ga = args_;
        ga.defer = verify_each ? 0 : 1;
        ga.kdev = kdev_.data();
        std::vector<const void *> key = {
            m1.ptr.data(), m1.col.data(), m1v, c16a, m2.ptr.data(), m2.col.data(), m2v, c16b,
            rhs, prhs, x, vdev_.data(), H_.data(), gram_.data(), gp_.data(), w_.data(),
            t_.data(), ax_.data(), pax_.data(), xz_.data(),
            reinterpret_cast<const void *>((std::intptr_t)(k1.lanes * 64 + k1.per)),
            reinterpret_cast<const void *>((std::intptr_t)(k2.lanes * 64 + k2.per)),
            reinterpret_cast<const void *>((std::intptr_t)(k1.sliced * 4 + k2.sliced)), m1sv,
            m2sv, sv1.soff, sv1.slen, sv1.scol, sv1.scol16, sv2.soff, sv2.slen, sv2.scol,
            sv2.scol16, reinterpret_cast<const void *>((std::uintptr_t)m1.gen),
            reinterpret_cast<const void *>((std::uintptr_t)m2.gen),
            reinterpret_cast<const void *>((std::intptr_t)mode),
            reinterpret_cast<const void *>((std::intptr_t)kmax)};
        StepGraph &g = graphs_[verify_each ? 1 : 0];
        if (g.exec && g.key == key)
            return g.exec;
        if (g.exec) {
            HIP_CHECK(hipGraphExecDestroy(g.exec));
            g.exec = nullptr;
        }
        // the graph's arguments are read once here; the captured kernels see
        // ga.defer and the device step counter
        const KrylovArgs<T> saved = args_;
        args_ = ga; // verify() launches with args_
        hipGraph_t graph = nullptr;
        HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            for (int o = 0; o < kGraphSteps; ++o)
                step(ga, o, kmax, verify_each,
                     XSel{reinterpret_cast<const void *const *>(vdev_.data()), kdev_.data(), o,
                          cap_});
            gmres_advance_step(s, kdev_.data(), kGraphSteps);
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph)
                (void)hipGraphDestroy(graph);
            args_ = saved;
            throw;
        }
        HIP_CHECK(hipStreamEndCapture(s, &graph));
        args_ = saved;
        const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_CHECK(e);
        g.key = std::move(key);
        ++graphs_captured;
        ++c_.graphs_captured;
        return g.exec;
    };
    int k = 0;
    while (k < cap_) {
        args_.defer = defer ? 1 : 0;
        const int k0 = k;
        if (graphs && k % kGraphSteps == 0) {
            // room for the graph's steps k .. k + kGraphSteps - 1
            ensure_basis(std::min(cap_ + 1, k + kGraphSteps + 1));
            ensure_steps(std::min(cap_ - 1, k + kGraphSteps - 1));
            hipGraphExec_t g = graph_for(!defer, ksteps_ - 1);
            if (kdev_val_ != k)
                gmres_set_step(s, kdev_.data(), k);
            HIP_CHECK(hipGraphLaunch(g, s));
            kdev_val_ = k + kGraphSteps;
            ++graphs_replayed;
            ++c_.graphs_replayed;
            k = std::min(cap_, k + kGraphSteps);
        } else {
            ensure_basis(std::min(cap_ + 1, k + 2));
            ensure_steps(k);
            step(args_, k, k, !defer, XSel{});
            ++k;
        }
        (void)k0;
        if (k % check_every == 0 || k == cap_) {
            download(c_, host.data(), st_.data(), L_.nsys);
            sync(c_);
            if (defer && host[0].active && host[0].triggered) {
                // the held step's verification; the loop resumes after it
                args_.defer = 1;
                verify();
                download(c_, host.data(), st_.data(), L_.nsys);
                sync(c_);
                defer = false;
                k = host[0].iterations;
            }
            bool any = false;
            for (const SysState &h : host)
                any = any || h.active;
            if (!any)
                break;
        }
    }
    args_.x = x;
    gmres_finish<T>(s, tl, args_);
    args_.x = left ? x : xz_.data();
}

template <typename T> std::vector<GmresOutcome> GmresEngine<T>::outcomes() {
    std::vector<SysState> host(L_.nsys);
    download(c_, host.data(), st_.data(), L_.nsys);
    sync(c_);
    std::vector<GmresOutcome> out(L_.nsys);
    for (int i = 0; i < L_.nsys; ++i) {
        const SysState &h = host[i];
        out[i] = GmresOutcome{h.status,     h.converged,  h.iterations, h.init_orig,
                              h.init_rel,   h.final_orig, h.final_rel};
        if (h.active) { // cap reached without an accept (loop ended)
            out[i].status = h.scap < h.nrows ? kGmIterationLimit : kGmNonconvergence;
        }
    }
    return out;
}

template class GmresEngine<double>;
template class GmresEngine<float>;

// ---------------------------------------------------------------- IR
// The context's engine when it was made for this layout and capacity, else a
// new one in its place (the old workspace is freed first).
template <typename T>
GmresEngine<T> &cached_engine(Context &c, EngineSlot<T> &slot,
                              const std::shared_ptr<Layout> &layout, int cap) {
    if (!slot.engine || slot.layout != layout || slot.cap != cap) {
        slot.engine.reset();
        slot.engine = std::make_shared<GmresEngine<T>>(c, *layout, cap);
        slot.layout = layout;
        slot.cap = cap;
    }
    return *slot.engine;
}

// Safety factor of the mixed path's adaptive inner tolerance (InnerTolRule):
// the inner solve of a refinement step targets 0.3 x the contraction the
// outer stop still needs, or the caller's tolerance if that is looser.
constexpr double kMixedInnerTheta = 0.3;

std::vector<IrResultHost> ir_run(Context &c, const DevCsr &a64, const DevCsr &m1, const DevCsr &m2,
                                 GmresMode mode, const double *b, const IrOptionsC &o, double *x) {
    hipStream_t s = c.stream;
    const Layout &L = *a64.layout;
    const Tiles tl = L.tiles();
    const int nsys = L.nsys;
    const std::size_t nt = (std::size_t)std::max(1, L.ntiles);
    IrWork &w = c.ir;
    w.st.resize((std::size_t)nsys * sizeof(IrSysState));
    IrSysState *const st = w.st.as<IrSysState>();
    DeviceArray<double> &part = w.part, &part4 = w.part4, &flags = w.flags, &upd = w.upd;
    DeviceArray<double> &hist = w.hist, &r = w.r, &d = w.d;
    DeviceArray<int> &inner = w.inner;
    part.resize(kIrNumSlots * 2 * nt);
    part4.resize(4 * nt);
    flags.resize(2 * nt);
    upd.resize(nt);
    const std::size_t nhist = (std::size_t)nsys * kMaxOuterHistory * 4;
    const std::size_t ninner = (std::size_t)nsys * kMaxOuterHistory * 3;
    hist.resize(nhist);
    inner.resize(ninner);
    r.resize(L.n);
    d.resize(L.n);
    HIP_CHECK(hipMemsetAsync(hist.data(), 0, nhist * sizeof(double), s));
    HIP_CHECK(hipMemsetAsync(inner.data(), 0, ninner * sizeof(int), s));
    IrArgs a{st, part.data(), part4.data(), flags.data(), upd.data(), hist.data(),
             inner.data(), 10.0 * std::ldexp(1.0, -53), o.relres_target};
    ir_setup(s, tl, a64.ptr.data(), a64.val.data(), b, a);
    // x0 = M b with M the preconditioner (src/ir.cpp:149), in the reference's
    // sequential row order.
    const DevCsr &pre = mode == kLeft ? m2 : m1;
    if (o.precision == 0)
        launch_spmv<double, double, double, double>(s, 1, tl, nullptr, 0, pre.ptr.data(),
                                                    pre.col.data(), pre.val.data(), b, x);
    else
        launch_spmv<float, double, double, double>(s, 1, tl, nullptr, 0, pre.ptr.data(),
                                                   pre.col.data(), pre.val32.data(), b, x);
    ir_diagnostics(s, tl, a64.ptr.data(), a64.col.data(), a64.val.data(), b, x, r.data(), a, 0);
    const int cap = o.cap > 0 ? o.cap : (int)std::min<std::int64_t>(L.n, INT_MAX - 2);
    // The largest system bounds the Krylov dimension of every system.
    std::int64_t maxrows = 0;
    for (int i = 0; i < nsys; ++i)
        maxrows = std::max(maxrows, L.offsets[i + 1] - L.offsets[i]);
    const int kcap = (int)std::min<std::int64_t>(cap, maxrows);
    GmresEngine<double> *g64 = nullptr;
    GmresEngine<float> *g32 = nullptr;
    DeviceArray<float> &rf = w.rf;
    if (o.precision == 0)
        g64 = &cached_engine(c, c.engine64, a64.layout, kcap);
    else {
        g32 = &cached_engine(c, c.engine32, a64.layout, kcap);
        rf.resize(L.n);
    }
    std::vector<IrSysState> host(nsys);
    auto any_active = [&]() {
        download(c, host.data(), st, nsys);
        sync(c);
        for (const auto &h : host)
            if (h.active)
                return true;
        return false;
    };
    bool running = any_active();
    for (int step = 0; step < o.outer && running; ++step) {
        if (o.precision == 0) {
            g64->solve(m1, m2, mode, r.data(), nullptr, o.tol, d.data(), st);
            ir_absorb_inner<double>(s, nsys, g64->args(), a, 0);
        } else {
            ir_scale_rhs(s, tl, r.data(), rf.data(), a);
            // The caller's tolerance is the FP32 solve's floor; a step that
            // needs less contraction to meet the outer stop asks for less.
            InnerTolRule rule;
            rule.theta = kMixedInnerTheta;
            rule.be_tol = a.be_tol;
            rule.relres_target = a.relres_target;
            g32->solve(m1, m2, mode, rf.data(), nullptr, o.tol, d.data(), st, rule);
            ir_absorb_inner<float>(s, nsys, g32->args(), a, 1);
        }
        ir_update(s, tl, x, d.data(), a);
        ir_diagnostics(s, tl, a64.ptr.data(), a64.col.data(), a64.val.data(), b, x, r.data(), a,
                       step + 1);
        running = any_active();
    }
    ir_cap(s, nsys, st);
    std::vector<IrResultHost> out(nsys);
    std::vector<double> hh(nhist);
    std::vector<int> hi(ninner);
    download(c, host.data(), st, nsys);
    download(c, hh.data(), hist.data(), hh.size());
    download(c, hi.data(), inner.data(), hi.size());
    sync(c);
    for (int i = 0; i < nsys; ++i) {
        out[i].st = host[i];
        std::copy(hh.begin() + (std::size_t)i * kMaxOuterHistory * 4,
                  hh.begin() + (std::size_t)(i + 1) * kMaxOuterHistory * 4, &out[i].history[0][0]);
        std::copy(hi.begin() + (std::size_t)i * kMaxOuterHistory * 3,
                  hi.begin() + (std::size_t)(i + 1) * kMaxOuterHistory * 3, &out[i].inner[0][0]);
    }
    return out;
}

} // namespace bspai_hip
