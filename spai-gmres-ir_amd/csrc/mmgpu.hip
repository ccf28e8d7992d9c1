// Matrix Market text parsed on the GPU (the reference's read_matrix_market,
// src/cpu.cpp:105-161, with its rules and messages).  The banner and the
// dimension line are read on the host (mm_read_header, a few bytes); the
// body is copied to the device once and parsed there:
//
//   1. line breaks: each thread counts the '\n' of 16 bytes (one 16-byte
//      load), blocks of 4 KiB are scanned, and every break writes the start
//      of the next line (starts[0] = 0);
//   2. lines: one thread per line classifies it (blank or comment, or data)
//      and parses a data line with the rules of mmtext.hpp, which the host
//      reader shares; a value that needs the exact big-integer conversion
//      (more than 19 significant digits, a subnormal or overflowing result,
//      an Eisel-Lemire tie) is listed and converted by a second kernel;
//   3. order: a scan numbers the data lines; the first failing one among
//      the first `count` (its data-line number and code under one 64-bit
//      atomic minimum) is the error a sequential reader meets first; with
//      none, fewer data lines than `count` is a truncated file and more a
//      trailing one;
//   4. entries: a scan of the entries per line (two for an off-diagonal
//      entry of a symmetric file) places (row, col, value) in file order.
//
// The result stays on the device (int64 zero-based indices, FP64 values),
// ready for bspai_hip_csr_from_entries.
#include "common.hpp"
#include "engine.hpp"
#include "kernels.hpp"
#include "mmgpu.hpp"
#include "mmio_internal.hpp"
#include "mmtext.hpp"

#include <climits>
#include <vector>

namespace bspai_hip {
namespace {

constexpr int kNlThreads = 256, kNlPer = 16, kNlChunk = kNlThreads * kNlPer;

__device__ inline int count_breaks(const uint4 &q) {
    const unsigned w[4] = {q.x, q.y, q.z, q.w};
    int c = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            c += ((w[i] >> (8 * b)) & 0xffu) == '\n';
    return c;
}

// Breaks per 4 KiB chunk (the text is padded with zero bytes to a multiple of
// 16, and to whole chunks, on the device).
__global__ void __launch_bounds__(kNlThreads)
break_count_kernel(const uint4 *__restrict__ text, int *__restrict__ chunk) {
    __shared__ int lds[kNlThreads / kWave];
    const std::int64_t g = (std::int64_t)blockIdx.x * kNlThreads + threadIdx.x;
    const int c = block_sum(count_breaks(text[g]), lds);
    if (threadIdx.x == 0)
        chunk[blockIdx.x] = c;
}

// Line starts: starts[1 + (breaks before p)] = p + 1 for every break p.
// chunk_incl: inclusive scan of the chunk counts.
__global__ void __launch_bounds__(kNlThreads)
line_start_kernel(const uint4 *__restrict__ text, const int *__restrict__ chunk_incl,
                  std::int64_t *__restrict__ starts) {
    __shared__ int wsum[kNlThreads / kWave];
    const std::int64_t g = (std::int64_t)blockIdx.x * kNlThreads + threadIdx.x;
    const uint4 q = text[g];
    const int c = count_breaks(q);
    // exclusive scan of c over the block: within the wave, then wave totals
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    int incl = c;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int o = __shfl_up(incl, off, kWave);
        if (lane >= off)
            incl += o;
    }
    if (lane == kWave - 1)
        wsum[wid] = incl;
    __syncthreads();
    int before = incl - c;
    for (int w = 0; w < wid; ++w)
        before += wsum[w];
    if (blockIdx.x > 0)
        before += chunk_incl[blockIdx.x - 1];
    if (c == 0)
        return;
    const unsigned w4[4] = {q.x, q.y, q.z, q.w};
    const std::int64_t p0 = g * kNlPer;
    int k = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (((w4[i] >> (8 * b)) & 0xffu) == '\n')
                starts[1 + before + k++] = p0 + 4 * i + b + 1;
}

struct LineOut {
    int *data;            // 1 for a data line
    unsigned char *code;  // its LineCode
    std::int64_t *row, *col;
    double *val;
    int *exact_list, *exact_count;
};

// Line i spans [starts[i], starts[i + 1] - 1), the last one [starts[L-1], S).
template <bool kExact>
__device__ inline void parse_line(const char *text, std::int64_t S, const std::int64_t *starts,
                                  std::int64_t L, std::int64_t i, const MmParseShape &sh,
                                  const LineOut &o) {
    const char *b = text + starts[i];
    const char *e = text + (i + 1 < L ? starts[i + 1] - 1 : S);
    const bool data = bspai_mm::is_data_line(b, e);
    o.data[i] = data;
    if (!data) {
        o.code[i] = 0;
        return;
    }
    std::uint64_t r = 0, c = 0;
    double v = 0;
    const int code =
        bspai_mm::parse_entry_line<kExact>(b, e, sh.rows, sh.cols, sh.integer != 0, r, c, v);
    o.code[i] = (unsigned char)code;
    o.row[i] = (std::int64_t)r;
    o.col[i] = (std::int64_t)c;
    o.val[i] = v;
    if (!kExact && code == bspai_mm::kNeedsExact)
        o.exact_list[atomicAdd(o.exact_count, 1)] = (int)i;
}

__global__ void __launch_bounds__(256)
parse_lines_kernel(const char *__restrict__ text, std::int64_t S,
                   const std::int64_t *__restrict__ starts, std::int64_t L, MmParseShape sh,
                   LineOut o) {
    const std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L)
        parse_line<false>(text, S, starts, L, i, sh, o);
}

__global__ void __launch_bounds__(64)
parse_exact_kernel(const char *__restrict__ text, std::int64_t S,
                   const std::int64_t *__restrict__ starts, std::int64_t L, MmParseShape sh,
                   LineOut o, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n)
        parse_line<true>(text, S, starts, L, o.exact_list[j], sh, o);
}

// First failing data line: key = data-line number << 8 | code.
__global__ void first_error_kernel(std::int64_t L, const int *__restrict__ data_incl,
                                   const int *__restrict__ data, const unsigned char *__restrict__ code,
                                   unsigned long long *__restrict__ key) {
    const std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L && data[i] && code[i])
        atomicMin(key, ((unsigned long long)(data_incl[i] - 1) << 8) | code[i]);
}

// Entries per line: 0, 1, or 2 for an off-diagonal entry of a symmetric file.
__global__ void entry_count_kernel(std::int64_t L, const int *__restrict__ data,
                                   const std::int64_t *__restrict__ row,
                                   const std::int64_t *__restrict__ col, int symmetric,
                                   int *__restrict__ cnt) {
    const std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L)
        cnt[i] = data[i] ? (symmetric && row[i] != col[i] ? 2 : 1) : 0;
}

__global__ void entry_write_kernel(std::int64_t L, const int *__restrict__ cnt_incl,
                                   const int *__restrict__ data,
                                   const std::int64_t *__restrict__ row,
                                   const std::int64_t *__restrict__ col,
                                   const double *__restrict__ val, int symmetric,
                                   std::int64_t *__restrict__ orow, std::int64_t *__restrict__ ocol,
                                   double *__restrict__ oval) {
    const std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L || !data[i])
        return;
    const std::int64_t r = row[i], c = col[i];
    const bool two = symmetric && r != c;
    const std::int64_t k = cnt_incl[i] - (two ? 2 : 1);
    orow[k] = r;
    ocol[k] = c;
    oval[k] = val[i];
    if (two) {
        orow[k + 1] = c;
        ocol[k + 1] = r;
        oval[k + 1] = val[i];
    }
}

std::int64_t blocks_for(std::int64_t n, int t) { return (n + t - 1) / t; }

} // namespace

void mm_parse_gpu(Context &c, const char *text, std::size_t len, MmGpu &out) {
    const bspai_mmio::MmHeader h = bspai_mmio::mm_read_header(text, len);
    hipStream_t s = c.stream;
    const std::int64_t S = (std::int64_t)(len - h.body);
    const std::int64_t nchunks = std::max<std::int64_t>(1, blocks_for(S, kNlChunk));
    if (nchunks > INT_MAX / 2)
        throw std::runtime_error("Matrix Market body too large for the GPU parser");
    // the body, zero-padded to whole chunks
    DeviceArray<uint4> dtext((std::size_t)nchunks * kNlThreads);
    HIP_CHECK(hipMemsetAsync(dtext.data(), 0, (std::size_t)nchunks * kNlChunk, s));
    if (S > 0)
        HIP_CHECK(hipMemcpyAsync(dtext.data(), text + h.body, (std::size_t)S,
                                 hipMemcpyHostToDevice, s));
    const char *dt = reinterpret_cast<const char *>(dtext.data());
    // 1. line starts
    DeviceArray<int> chunk((std::size_t)nchunks), scan_tmp((std::size_t)scan_tmp_ints(nchunks));
    hipLaunchKernelGGL(break_count_kernel, dim3((unsigned)nchunks), dim3(kNlThreads), 0, s,
                       dtext.data(), chunk.data());
    HIP_LAUNCH_CHECK();
    scan_inclusive_int(s, chunk.data(), nchunks, scan_tmp.data());
    int breaks = 0;
    HIP_CHECK(hipMemcpyAsync(&breaks, chunk.data() + nchunks - 1, sizeof(int),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const std::int64_t L = (std::int64_t)breaks + 1;
    if (L >= INT_MAX - 1)
        throw std::runtime_error("Matrix Market body has too many lines for the GPU parser");
    DeviceArray<std::int64_t> starts((std::size_t)L);
    HIP_CHECK(hipMemsetAsync(starts.data(), 0, sizeof(std::int64_t), s));
    hipLaunchKernelGGL(line_start_kernel, dim3((unsigned)nchunks), dim3(kNlThreads), 0, s,
                       dtext.data(), chunk.data(), starts.data());
    HIP_LAUNCH_CHECK();
    // 2. lines
    DeviceArray<int> data((std::size_t)L), exact((std::size_t)L), exact_n(1);
    DeviceArray<unsigned char> code((std::size_t)L);
    DeviceArray<std::int64_t> row((std::size_t)L), col((std::size_t)L);
    DeviceArray<double> val((std::size_t)L);
    HIP_CHECK(hipMemsetAsync(exact_n.data(), 0, sizeof(int), s));
    const MmParseShape sh{h.rows, h.cols, h.integer ? 1 : 0};
    const LineOut o{data.data(), code.data(), row.data(), col.data(), val.data(), exact.data(),
                    exact_n.data()};
    hipLaunchKernelGGL(parse_lines_kernel, dim3((unsigned)blocks_for(L, 256)), dim3(256), 0, s,
                       dt, S, starts.data(), L, sh, o);
    HIP_LAUNCH_CHECK();
    int nexact = 0;
    HIP_CHECK(hipMemcpyAsync(&nexact, exact_n.data(), sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (nexact > 0) {
        hipLaunchKernelGGL(parse_exact_kernel, dim3((unsigned)blocks_for(nexact, 64)), dim3(64), 0,
                           s, dt, S, starts.data(), L, sh, o, nexact);
        HIP_LAUNCH_CHECK();
    }
    // 3. data-line numbers, the first failure, the count
    DeviceArray<int> dincl((std::size_t)L), tmp2((std::size_t)scan_tmp_ints(L));
    HIP_CHECK(hipMemcpyAsync(dincl.data(), data.data(), (std::size_t)L * sizeof(int),
                             hipMemcpyDeviceToDevice, s));
    scan_inclusive_int(s, dincl.data(), L, tmp2.data());
    DeviceArray<unsigned long long> key(1);
    const unsigned long long none = ~0ull;
    HIP_CHECK(hipMemcpyAsync(key.data(), &none, sizeof none, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(first_error_kernel, dim3((unsigned)blocks_for(L, 256)), dim3(256), 0, s, L,
                       dincl.data(), data.data(), code.data(), key.data());
    HIP_LAUNCH_CHECK();
    unsigned long long hkey = 0;
    int ndata = 0;
    HIP_CHECK(hipMemcpyAsync(&hkey, key.data(), sizeof hkey, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&ndata, dincl.data() + L - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (hkey != none && (std::size_t)(hkey >> 8) < h.count)
        throw bspai_mmio::FormatError(bspai_mm::line_message((int)(hkey & 0xffu)));
    if ((std::size_t)ndata < h.count)
        throw bspai_mmio::FormatError("Truncated coordinate data");
    if ((std::size_t)ndata > h.count)
        throw bspai_mmio::FormatError("Trailing coordinate data");
    // 4. entries in file order
    DeviceArray<int> cnt((std::size_t)L);
    hipLaunchKernelGGL(entry_count_kernel, dim3((unsigned)blocks_for(L, 256)), dim3(256), 0, s, L,
                       data.data(), row.data(), col.data(), h.symmetric ? 1 : 0, cnt.data());
    HIP_LAUNCH_CHECK();
    scan_inclusive_int(s, cnt.data(), L, tmp2.data());
    int total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, cnt.data() + L - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    out.rows = (std::int64_t)h.rows;
    out.cols = (std::int64_t)h.cols;
    out.count = total;
    out.integer = h.integer;
    out.symmetric = h.symmetric;
    out.row.resize((std::size_t)std::max(total, 1));
    out.col.resize((std::size_t)std::max(total, 1));
    out.val.resize((std::size_t)std::max(total, 1));
    hipLaunchKernelGGL(entry_write_kernel, dim3((unsigned)blocks_for(L, 256)), dim3(256), 0, s, L,
                       cnt.data(), data.data(), row.data(), col.data(), val.data(),
                       h.symmetric ? 1 : 0, out.row.data(), out.col.data(), out.val.data());
    HIP_LAUNCH_CHECK();
    HIP_CHECK(hipStreamSynchronize(s));
}

} // namespace bspai_hip
