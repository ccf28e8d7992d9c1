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

#include <hipcub/hipcub.hpp>

#include <climits>
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

} // namespace

std::int64_t csr_from_entries(Context &c, std::int64_t rows, std::int64_t cols,
                              std::int64_t count, const std::int64_t *row,
                              const std::int64_t *col, const double *val, int *ptr, int *out_col,
                              double *out_val) {
    const hipStream_t s = c.stream;
    CanonWork &w = c.canon;
    HIP_CHECK(hipMemsetAsync(ptr, 0, (rows + 1) * sizeof(int), s));
    if (count == 0) {
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    }
    const int blocks = canon_blocks(count);
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
    HIP_CHECK(hipMemsetAsync(w.flags64.data(), 0xff, sizeof(unsigned long long), s));
    HIP_CHECK(hipMemsetAsync(w.flags.data(), 0, sizeof(int), s));
    // One pass to the end, then one synchronization: an invalid entry gets
    // key 0 (in range), so the later stages stay in bounds, and the errors
    // are reported in the reference's order (validation first).
    timed(c, "entries_check_kernel", 32.0 * (double)count, 0.0, [&] {
        hipLaunchKernelGGL(entries_check_kernel, dim3(blocks), dim3(kCanonThreads), 0, s, count,
                           rows, cols, row, col, val, w.keys.data(), w.flags64.data());
        HIP_LAUNCH_CHECK();
    });
    // Sort on the key bits in use only.
    const unsigned long long max_key =
        (unsigned long long)rows * (unsigned long long)cols - 1ull;
    const int end_bit = max_key == 0 ? 1 : 64 - __builtin_clzll(max_key);
    std::size_t tmp_bytes = 0;
    HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, w.keys.data(),
                                                 w.keys_sorted.data(), val, w.vals_sorted.data(),
                                                 (int)count, 0, end_bit, s));
    w.sort_tmp.resize(std::max<std::size_t>(tmp_bytes, 1));
    timed(c, "entries_radix_sort", 4.0 * 16.0 * (double)count, 0.0, [&] {
        HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(w.sort_tmp.data(), tmp_bytes, w.keys.data(),
                                                     w.keys_sorted.data(), val,
                                                     w.vals_sorted.data(), (int)count, 0, end_bit,
                                                     s));
    });
    timed(c, "run_sums_kernel", 28.0 * (double)count, 0.0, [&] {
        hipLaunchKernelGGL(run_sums_kernel, dim3(blocks), dim3(kCanonThreads), 0, s, count,
                           w.keys_sorted.data(), w.vals_sorted.data(), w.keep.data(),
                           w.sums.data(), w.flags.data());
        HIP_LAUNCH_CHECK();
    });
    HIP_CHECK(hipMemcpyAsync(w.slot.data(), w.keep.data(), count * sizeof(int),
                             hipMemcpyDeviceToDevice, s));
    scan_inclusive_int(s, w.slot.data(), count, w.scan_tmp.data());
    timed(c, "scatter_kernel", 20.0 * (double)count, 0.0, [&] {
        hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(kCanonThreads), 0, s, count, cols,
                           w.keys_sorted.data(), w.keep.data(), w.slot.data(), w.sums.data(),
                           out_col, out_val, ptr);
        HIP_LAUNCH_CHECK();
    });
    scan_inclusive_int(s, ptr, rows + 1, w.row_tmp.data());
    unsigned long long bad = 0;
    int flag = 0, nnz = 0;
    HIP_CHECK(hipMemcpyAsync(&bad, w.flags64.data(), sizeof bad, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&flag, w.flags.data(), sizeof flag, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&nnz, w.slot.data() + (count - 1), sizeof nnz,
                             hipMemcpyDeviceToHost, s));
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
