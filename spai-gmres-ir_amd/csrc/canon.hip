// Canonical CSR from coordinate entries on the device: the reference's
// CsrMatrix::from_entries (src/cpu.cpp:48-76).
//
//   1. entries_check_kernel: every entry's index range and finiteness; the
//      first offending entry in input order (atomicMin over its position)
//      decides the error, as the reference's validation loop does; each valid
//      entry gets the 64-bit sort key row * cols + col.
//   2. a stable radix sort of (key, value) pairs (hipCUB): least-significant-
//      digit radix sort keeps equal keys in input order, which is the
//      reference's accumulation order for duplicates (its std::stable_sort).
//   3. run_sums_kernel: the first entry of each run of equal keys adds the
//      run's values one by one in that order, flags a non-finite partial sum
//      and keeps the run when its sum is not zero (+0 and -0 both drop).
//   4. an inclusive scan of the kept flags numbers the kept runs;
//      scatter_kernel writes each one's column and value and counts its row;
//      a scan of the row counts gives the row offsets.
// The whole pass streams the entries a few times (HBM-bound); the sort is
// the bulk of it.
#include "engine.hpp"
#include "kernels.hpp"


#include <climits>
#include <map>
#include <mutex>
#include <stdexcept>

namespace bspai_hip {

namespace {

constexpr int kCanonThreads = 256;

__global__ void __launch_bounds__(kCanonThreads)
entries_check_kernel(std::int64_t count, std::int64_t rows, std::int64_t cols,
                     const std::int64_t *__restrict__ row, const std::int64_t *__restrict__ col,
                     const double *__restrict__ val, unsigned long long *__restrict__ keys,
                     unsigned long long *__restrict__ first_bad) {
    const std::int64_t stride = (std::int64_t)gridDim.x * blockDim.x;
    unsigned long long bad = ULLONG_MAX;
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        const std::int64_t r = row[i], c = col[i];
        const bool ok = r >= 0 && r < rows && c >= 0 && c < cols && isfinite(val[i]);
        keys[i] = ok ? (unsigned long long)r * (unsigned long long)cols + (unsigned long long)c
                     : 0ull;
        if (!ok && (unsigned long long)i < bad)
            bad = (unsigned long long)i;
    }
    if (bad != ULLONG_MAX)
        atomicMin(first_bad, bad);
}

__global__ void __launch_bounds__(kCanonThreads)
run_sums_kernel(std::int64_t count, const unsigned long long *__restrict__ keys,
                const double *__restrict__ val, int *__restrict__ keep,
                double *__restrict__ sums, int *__restrict__ nonfinite) {
    const std::int64_t stride = (std::int64_t)gridDim.x * blockDim.x;
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        const unsigned long long k = keys[i];
        int kept = 0;
        if (i == 0 || keys[i - 1] != k) {
            double sum = val[i];
            bool finite = true;
            for (std::int64_t j = i + 1; j < count && keys[j] == k; ++j) {
                sum += val[j];
                finite = finite && isfinite(sum);
            }
            if (!finite)
                *nonfinite = 1;
            kept = sum != 0.0 ? 1 : 0;
            sums[i] = sum;
        }
        keep[i] = kept;
    }
}

__global__ void __launch_bounds__(kCanonThreads)
scatter_kernel(std::int64_t count, std::int64_t cols, const unsigned long long *__restrict__ keys,
               const int *__restrict__ keep, const int *__restrict__ slot,
               const double *__restrict__ sums, int *__restrict__ out_col,
               double *__restrict__ out_val, int *__restrict__ row_count) {
    const std::int64_t stride = (std::int64_t)gridDim.x * blockDim.x;
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += stride) {
        if (!keep[i])
            continue;
        const unsigned long long k = keys[i];
        const unsigned long long r = k / (unsigned long long)cols;
        const int s = slot[i] - 1;
        out_col[s] = (int)(k - r * (unsigned long long)cols);
        out_val[s] = sums[i];
        atomicAdd(row_count + r + 1, 1);
    }
}

int canon_blocks(std::int64_t count) {
    const std::int64_t b = (count + kCanonThreads - 1) / kCanonThreads;
    return (int)std::max<std::int64_t>(1, std::min<std::int64_t>(b, 1 << 16));
}

__global__ void fill_u32_kernel(unsigned *__restrict__ p, std::int64_t n, unsigned v) {
    const std::int64_t stride = (std::int64_t)gridDim.x * blockDim.x;
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

// Clears with a kernel rather than hipMemsetAsync, so the sequence stays
// valid inside a stream that a caller captures into a graph.
void fill_u32(hipStream_t s, void *p, std::int64_t words, unsigned v) {
    if (words <= 0)
        return;
    const int blocks = (int)std::min<std::int64_t>(1024, (words + 255) / 256);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, s, static_cast<unsigned *>(p),
                       words, v);
    HIP_LAUNCH_CHECK();
}

// The scratch of one canonicalization: either the context's own arrays or a
// carve of the static scratch array (bspai_hip_csr_from_entries_noalloc).
struct CanonBuffers {
    unsigned long long *keys, *keys_sorted, *flags64;
    double *vals_sorted, *sums;
    int *keep, *slot, *flags, *scan_tmp, *row_tmp;
    void *sort_tmp;
    std::size_t sort_tmp_bytes;
};

// Radix-sort temporary storage for `count` pairs (radix.hip).
std::size_t sort_tmp_bytes(std::int64_t count) { return radix_sort_tmp_bytes(count); }

int key_bits(std::int64_t rows, std::int64_t cols) {
    const unsigned long long max_key =
        (unsigned long long)rows * (unsigned long long)cols - 1ull;
    return max_key == 0 ? 1 : 64 - __builtin_clzll(max_key);
}

// The canonicalization proper, on scratch w (count > 0).
std::int64_t canon_run(Context &c, std::int64_t rows, std::int64_t cols, std::int64_t count,
                       const std::int64_t *row, const std::int64_t *col, const double *val,
                       int *ptr, int *out_col, double *out_val, const CanonBuffers &w) {
    const hipStream_t s = c.stream;
    const int blocks = canon_blocks(count);
    fill_u32(s, w.flags64, 2, 0xffffffffu);
    fill_u32(s, w.flags, 1, 0u);
    // One pass to the end, then one synchronization: an invalid entry gets
    // key 0 (in range), so the later stages stay in bounds, and the errors
    // are reported in the reference's order (validation first).
    timed(c, "entries_check_kernel", 32.0 * (double)count, 0.0, [&] {
        hipLaunchKernelGGL(entries_check_kernel, dim3(blocks), dim3(kCanonThreads), 0, s, count,
                           rows, cols, row, col, val, w.keys, w.flags64);
        HIP_LAUNCH_CHECK();
    });
    // Sort on the key bits in use only.
    const int end_bit = key_bits(rows, cols);
    timed(c, "entries_radix_sort", 4.0 * 16.0 * (double)count, 0.0, [&] {
        radix_sort_pairs(s, w.keys, val, w.keys_sorted, w.vals_sorted, count, end_bit,
                         w.sort_tmp);
    });
    timed(c, "run_sums_kernel", 28.0 * (double)count, 0.0, [&] {
        hipLaunchKernelGGL(run_sums_kernel, dim3(blocks), dim3(kCanonThreads), 0, s, count,
                           w.keys_sorted, w.vals_sorted, w.keep, w.sums, w.flags);
        HIP_LAUNCH_CHECK();
    });
    HIP_CHECK(hipMemcpyAsync(w.slot, w.keep, count * sizeof(int), hipMemcpyDeviceToDevice, s));
    scan_inclusive_int(s, w.slot, count, w.scan_tmp);
    timed(c, "scatter_kernel", 20.0 * (double)count, 0.0, [&] {
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(kCanonThreads), 0, s, count, cols,
                           w.keys_sorted, w.keep, w.slot, w.sums, out_col, out_val, ptr);
        HIP_LAUNCH_CHECK();
    });
    scan_inclusive_int(s, ptr, rows + 1, w.row_tmp);
    unsigned long long bad = 0;
    int flag = 0, nnz = 0;
    HIP_CHECK(hipMemcpyAsync(&bad, w.flags64, sizeof bad, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&flag, w.flags, sizeof flag, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&nnz, w.slot + (count - 1), sizeof nnz, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (bad != ULLONG_MAX) {
        // The reference checks the indices of an entry before its value.
        std::int64_t r = 0, cc = 0;
        HIP_CHECK(hipMemcpy(&r, row + bad, sizeof r, hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(&cc, col + bad, sizeof cc, hipMemcpyDeviceToHost));
        if (r < 0 || r >= rows || cc < 0 || cc >= cols)
            throw std::invalid_argument("Entry index outside matrix dimensions");
        throw std::invalid_argument("Matrix entries must be finite");
    }
    if (flag)
        throw std::invalid_argument("Duplicate sum is nonfinite");
    return nnz;
}

} // namespace

std::int64_t csr_from_entries(Context &c, std::int64_t rows, std::int64_t cols,
                              std::int64_t count, const std::int64_t *row,
                              const std::int64_t *col, const double *val, int *ptr, int *out_col,
                              double *out_val) {
    const hipStream_t s = c.stream;
    fill_u32(s, ptr, rows + 1, 0u);
    if (count == 0) {
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    }
    CanonWork &w = c.canon;
    w.keys.resize(count);
    w.keys_sorted.resize(count);
    w.vals_sorted.resize(count);
    w.sums.resize(count);
    w.keep.resize(count);
    w.slot.resize(count);
    w.flags64.resize(1);
    w.flags.resize(1);
    w.scan_tmp.resize(scan_tmp_ints(count));
    w.row_tmp.resize(scan_tmp_ints(rows + 1));
    const std::size_t tb = sort_tmp_bytes(count);
    w.sort_tmp.resize(std::max<std::size_t>(tb, 1));
    const CanonBuffers b{w.keys.data(), w.keys_sorted.data(), w.flags64.data(),
                         w.vals_sorted.data(), w.sums.data(), w.keep.data(),
                         w.slot.data(), w.flags.data(), w.scan_tmp.data(),
                         w.row_tmp.data(), w.sort_tmp.data(), tb};
    return canon_run(c, rows, cols, count, row, col, val, ptr, out_col, out_val, b);
}

// ------------------------------------------------- allocation-free variant
// The scratch of bspai_hip_csr_from_entries_noalloc: one static device array,
// each GPU holding its own copy until the process ends.  Host code never uses
// the name `scratch` as an address: noalloc_setup looks the device address up
// once per device (with that device current) and every launch takes it from
// there.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocDevice {
    void *base = nullptr;      // device address of this GPU's scratch
    std::size_t sort_bytes = 0; // radix-sort temporary for kNoAllocMaxEntries pairs
};
std::mutex g_noalloc_mutex;
std::map<int, NoAllocDevice> g_noalloc;

// Offsets of the carve for `count` entries and `rows` rows (each 256-aligned).
struct NoAllocCarve {
    std::size_t keys, keys_sorted, flags64, vals_sorted, sums, keep, slot, flags, scan_tmp,
        row_tmp, sort_tmp, end;
};
NoAllocCarve carve(std::int64_t count, std::int64_t rows, std::size_t sort_bytes) {
    auto up = [](std::size_t x) { return (x + 255) & ~(std::size_t)255; };
    NoAllocCarve o{};
    std::size_t at = 0;
    auto take = [&](std::size_t &field, std::size_t bytes) {
        field = at;
        at = up(at + bytes);
    };
    take(o.keys, 8 * (std::size_t)count);
    take(o.keys_sorted, 8 * (std::size_t)count);
    take(o.flags64, 8);
    take(o.vals_sorted, 8 * (std::size_t)count);
    take(o.sums, 8 * (std::size_t)count);
    take(o.keep, 4 * (std::size_t)count);
    take(o.slot, 4 * (std::size_t)count);
    take(o.flags, 4);
    take(o.scan_tmp, 4 * (std::size_t)scan_tmp_ints(count));
    take(o.row_tmp, 4 * (std::size_t)scan_tmp_ints(rows + 1));
    take(o.sort_tmp, sort_bytes);
    o.end = at;
    return o;
}

} // namespace

void noalloc_setup(Context &c) {
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (g_noalloc.count(c.device))
        return;
    NoAllocDevice d;
    HIP_CHECK(hipGetSymbolAddress(&d.base, HIP_SYMBOL(scratch)));
    if (!d.base)
        throw HipError("hipGetSymbolAddress returned no address for the scratch array");
    d.sort_bytes = sort_tmp_bytes(kNoAllocMaxEntries);
    g_noalloc[c.device] = d;
}

std::int64_t csr_from_entries_noalloc(Context &c, std::int64_t rows, std::int64_t cols,
                                      std::int64_t count, const std::int64_t *row,
                                      const std::int64_t *col, const double *val, int *ptr,
                                      int *out_col, double *out_val) {
    NoAllocDevice d;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mutex);
        const auto it = g_noalloc.find(c.device);
        if (it == g_noalloc.end())
            throw HipError("allocation-free scratch not set up on this device");
        d = it->second;
    }
    if (count > kNoAllocMaxEntries || rows + 1 > kNoAllocMaxRows)
        throw std::invalid_argument("allocation-free canonicalization: at most " +
                                    std::to_string(kNoAllocMaxEntries) + " entries and " +
                                    std::to_string(kNoAllocMaxRows - 1) + " rows");
    const NoAllocCarve o = carve(kNoAllocMaxEntries, kNoAllocMaxRows - 1, d.sort_bytes);
    if (o.end > kNoAllocScratchBytes)
        throw HipError("allocation-free scratch too small for its own limits");
    const hipStream_t s = c.stream;
    fill_u32(s, ptr, rows + 1, 0u);
    if (count == 0) {
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    }
    unsigned char *b = static_cast<unsigned char *>(d.base);
    const CanonBuffers w{reinterpret_cast<unsigned long long *>(b + o.keys),
                         reinterpret_cast<unsigned long long *>(b + o.keys_sorted),
                         reinterpret_cast<unsigned long long *>(b + o.flags64),
                         reinterpret_cast<double *>(b + o.vals_sorted),
                         reinterpret_cast<double *>(b + o.sums),
                         reinterpret_cast<int *>(b + o.keep),
                         reinterpret_cast<int *>(b + o.slot),
                         reinterpret_cast<int *>(b + o.flags),
                         reinterpret_cast<int *>(b + o.scan_tmp),
                         reinterpret_cast<int *>(b + o.row_tmp),
                         b + o.sort_tmp,
                         d.sort_bytes};
    return canon_run(c, rows, cols, count, row, col, val, ptr, out_col, out_val, w);
}

// ------------------------------------------------------------ CSR checks
// A matrix handed over in device memory is checked on the device before use
// (the SPAI and SpMV kernels index by it): row offsets first, then -- only
// when they are consistent, so that no row walk leaves the arrays -- every
// row's columns and values.

namespace {

enum : int {
    kBadPtr = 1,      // offsets not 0 .. nnz, or decreasing
    kBadRange = 2,    // column outside [0, n)
    kBadOrder = 4,    // columns not strictly increasing within a row
    kBadValue = 8,    // nonfinite value
    kBadSystem = 16,  // column outside the row's system (batched layouts)
};

__global__ void __launch_bounds__(256)
csr_ptr_check_kernel(int n, int nnz, const int *__restrict__ ptr, int *__restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n)
        return;
    const int p = ptr[i];
    bool ok = p >= 0 && p <= nnz;
    if (i == 0)
        ok = ok && p == 0;
    if (i == n)
        ok = ok && p == nnz;
    if (i < n)
        ok = ok && p <= ptr[i + 1];
    if (!ok)
        atomicOr(bad, kBadPtr);
}

// One thread per row; sys (nsys + 1 offsets) bounds each row's columns.
__global__ void __launch_bounds__(256)
csr_rows_check_kernel(int n, const int *__restrict__ ptr, const int *__restrict__ col,
                      const double *__restrict__ val, int nsys, const int *__restrict__ sys,
                      int *__restrict__ bad) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    int lo = 0, hi = n;
    if (nsys > 1) { // the system s with sys[s] <= r < sys[s + 1]
        int a = 0, b = nsys;
        while (b - a > 1) {
            const int m = (a + b) / 2;
            if (sys[m] <= r)
                a = m;
            else
                b = m;
        }
        lo = sys[a];
        hi = sys[a + 1];
    }
    int flags = 0, prev = -1;
    for (int q = ptr[r]; q < ptr[r + 1]; ++q) {
        const int c = col[q];
        if (c < 0 || c >= n)
            flags |= kBadRange;
        else if (c < lo || c >= hi)
            flags |= kBadSystem;
        if (c <= prev)
            flags |= kBadOrder;
        if (!isfinite(val[q]))
            flags |= kBadValue;
        prev = c;
    }
    if (flags)
        atomicOr(bad, flags);
}

const char *csr_problem(int flags) {
    if (flags & kBadPtr)
        return "row pointers must span [0, nnz] and be nondecreasing";
    if (flags & kBadRange)
        return "column index out of range";
    if (flags & kBadOrder)
        return "column indices must be strictly increasing within each row";
    if (flags & kBadSystem)
        return "entry outside its row's system (a batch must be block-diagonal)";
    if (flags & kBadValue)
        return "matrix values must be finite";
    return nullptr;
}

} // namespace

void check_device_csr(Context &c, std::int64_t n, std::int64_t nnz, const int *ptr,
                      const int *col, const double *val,
                      const std::vector<std::int64_t> &sys_offsets) {
    hipStream_t s = c.stream;
    const int nsys = (int)sys_offsets.size() - 1;
    DeviceArray<int> flag(1), sys(nsys + 1);
    std::vector<int> hs(sys_offsets.begin(), sys_offsets.end());
    HIP_CHECK(hipMemsetAsync(flag.data(), 0, sizeof(int), s));
    HIP_CHECK(hipMemcpyAsync(sys.data(), hs.data(), hs.size() * sizeof(int),
                             hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(csr_ptr_check_kernel, dim3(ceil_div(n + 1, 256)), dim3(256), 0, s, (int)n,
                       (int)nnz, ptr, flag.data());
    HIP_LAUNCH_CHECK();
    int h = 0;
    HIP_CHECK(hipMemcpyAsync(&h, flag.data(), sizeof h, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (!h) {
        hipLaunchKernelGGL(csr_rows_check_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n,
                           ptr, col, val, nsys, sys.data(), flag.data());
        HIP_LAUNCH_CHECK();
        HIP_CHECK(hipMemcpyAsync(&h, flag.data(), sizeof h, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    if (const char *m = csr_problem(h))
        throw std::invalid_argument(m);
}

} // namespace bspai_hip
