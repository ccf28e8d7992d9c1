// CSR sparse matrix-vector products and small vector utilities for gfx950.
//
// spmv: a group of LANES consecutive lanes owns one row; it streams the row's
// (col, val) pairs with unit stride (one coalesced run per group) and reduces
// with a fixed butterfly, so results are deterministic.  LANES = 1 keeps the
// reference's sequential row order (src/cpu.cpp:88-89) and is bit-identical to
// CsrMatrix::multiply; wider groups suit long rows.  The grid walks row tiles,
// and a tile whose system is inactive is skipped, which lets one launch serve
// many independent systems.
#include "common.hpp"
#include "kernels.hpp"

namespace bspai_hip {

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b mod 8),
// and each XCD has its own L2.  Remapping gives every XCD one contiguous run
// of row blocks, so the x entries that neighbouring rows share (stencil
// neighbours, the rows of one system) are reused inside one L2.
constexpr int kXcds = 8;
__device__ inline int xcd_block(int b, int nb) {
    const int q = nb / kXcds, r = nb % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Row-group kernel: LANES lanes per row stream the row with unit stride and
// reduce with a fixed butterfly (for long rows).
template <int LANES, typename TV, typename TX, typename TA, typename TY>
__global__ void __launch_bounds__(kSpmvRows)
spmv_group_kernel(Tiles tiles, const int *__restrict__ flag, int flag_stride,
                  const int *__restrict__ ptr, const int *__restrict__ col,
                  const TV *__restrict__ val, const TX *__restrict__ x, TY *__restrict__ y,
                  XSel xs) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int t = b / kSpmvSub, sb = b % kSpmvSub;
    if (flag && !flag[(std::int64_t)tiles.sys[t] * flag_stride])
        return;
    if (xs.tab)
        x = static_cast<const TX *>(xs.tab[min(*xs.idx + xs.off, xs.last)]);
    const int rbeg = tiles.row0[t] + sb * kSpmvRows;
    const int rend = min(tiles.row1[t], rbeg + kSpmvRows);
    const int groups = kSpmvRows / LANES;
    const int g = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    for (int row = rbeg + g; row < rend; row += groups) {
        TA acc = Acc<TA>::zero();
        const int e1 = ptr[row + 1];
        for (int e = ptr[row] + lane; e < e1; e += LANES)
            Acc<TA>::add(acc, Acc<TA>::mul(val[e], x[col[e]]));
        acc = Acc<TA>::template group<LANES>(acc);
        if (lane == 0)
            y[row] = Acc<TA>::template out<TY>(acc);
    }
}

// Streaming kernel for short rows: the block's 256 rows own one contiguous
// range of (val, col); all threads load it with unit stride, form the products
// in LDS, then each thread adds its row's products in CSR order -- the
// reference's sequential row sum (src/cpu.cpp:88-89), bit for bit.
// PER entries per thread per chunk: 8 for rows of up to ~8 entries, 16 for
// longer ones, so that a block's rows usually fit one chunk.
// With C16 the columns come from 16-bit offsets to the block's first row
// (col16[e] = col[e] - rbeg, see encode_col16_kernel): 2 bytes per entry
// instead of 4, the same column and the same sums.
// A block's path is a chain of dependent loads (entry range, entries, x
// gathers), so nothing else is put in front of it: one system's block
// geometry is arithmetic, the run flag loads beside the entry range, and a
// replayed graph's x is looked up behind the entry loads.
template <typename TV, typename TX, typename TA, typename TY, int PER, bool C16>
__global__ void __launch_bounds__(kSpmvRows)
spmv_stream_kernel(Tiles tiles, const int *__restrict__ flag, int flag_stride,
                   const int *__restrict__ ptr, const int *__restrict__ col,
                   const short *__restrict__ col16, const TV *__restrict__ val,
                   const TX *__restrict__ x, TY *__restrict__ y, XSel xs) {
    constexpr int kStreamChunk = kSpmvRows * PER;
    __shared__ typename Acc<TA>::Prod prod[kStreamChunk];
    const int b = xcd_block(blockIdx.x, gridDim.x);
    int rbeg, rend, e0, e1;
    if (tiles.nsys == 1) {
        // one system: tile t is rows [t kTileRows, ...), so block b starts at
        // row 256 b; the entry range and the run flag load together (both
        // in bounds either way)
        rbeg = b * kSpmvRows;
        rend = min(tiles.nrows, rbeg + kSpmvRows);
        if (rbeg >= rend)
            return;
        e0 = ptr[rbeg];
        e1 = ptr[rend];
        if (flag && !flag[0])
            return;
    } else {
        const int t = b / kSpmvSub, sb = b % kSpmvSub;
        if (flag && !flag[(std::int64_t)tiles.sys[t] * flag_stride])
            return;
        rbeg = tiles.row0[t] + sb * kSpmvRows;
        rend = min(tiles.row1[t], rbeg + kSpmvRows);
        if (rbeg >= rend)
            return;
        e0 = ptr[rbeg];
        e1 = ptr[rend];
    }
    const int nr = rend - rbeg;
    const int tid = threadIdx.x;
    // (each thread's row bounds load alongside; the sums need them)
    const int mb = tid < nr ? ptr[rbeg + tid] : 0, me = tid < nr ? ptr[rbeg + tid + 1] : 0;
    TA acc = Acc<TA>::zero();
    constexpr int per = PER;
    for (int c0 = e0; c0 < e1; c0 += kStreamChunk) {
        const int c1 = min(e1, c0 + kStreamChunk);
        // Unconditional loads (indices clamped into the chunk) so that all of
        // them are in flight at once: columns and values, then the gathers.
        int cj[per];
        TV vv[per];
#pragma unroll
        for (int u = 0; u < per; ++u) {
            const int e = min(c0 + tid + u * kSpmvRows, c1 - 1);
            if constexpr (C16)
                cj[u] = rbeg + (int)col16[e];
            else
                cj[u] = col[e];
            vv[u] = val[e];
        }
        if (xs.tab && c0 == e0)
            x = static_cast<const TX *>(xs.tab[min(*xs.idx + xs.off, xs.last)]);
        TX xv[per];
#pragma unroll
        for (int u = 0; u < per; ++u)
            xv[u] = x[cj[u]];
#pragma unroll
        for (int u = 0; u < per; ++u) {
            const int e = c0 + tid + u * kSpmvRows;
            if (e < c1)
                prod[e - c0] = Acc<TA>::mul(vv[u], xv[u]);
        }
        __syncthreads();
        const int lo = max(mb, c0), hi = min(me, c1);
        for (int e = lo; e < hi; ++e)
            Acc<TA>::add(acc, prod[e - c0]);
        __syncthreads();
    }
    if (tid < nr)
        y[rbeg + tid] = Acc<TA>::template out<TY>(acc);
}

// Row blocks of the SpMV kernels: block b of tile t covers rows
// [row0[t] + b kSpmvRows, ...).  col16[e] = col[e] - (first row of e's block)
// when every such offset fits 16 bits; *bad is set otherwise.
__global__ void __launch_bounds__(kSpmvRows)
encode_col16_kernel(Tiles tiles, const int *__restrict__ ptr, const int *__restrict__ col,
                    short *__restrict__ col16, int *__restrict__ bad) {
    const int t = blockIdx.x / kSpmvSub, sb = blockIdx.x % kSpmvSub;
    const int rbeg = tiles.row0[t] + sb * kSpmvRows;
    const int rend = min(tiles.row1[t], rbeg + kSpmvRows);
    if (rbeg >= rend)
        return;
    const int e1 = ptr[rend];
    bool over = false;
    for (int e = ptr[rbeg] + threadIdx.x; e < e1; e += kSpmvRows) {
        const int d = col[e] - rbeg;
        over = over || d < -32768 || d > 32767;
        col16[e] = (short)d;
    }
    if (over)
        *bad = 1;
}

void launch_encode_col16(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col,
                         short *col16, int *bad) {
    if (tiles.ntiles == 0)
        return;
    hipLaunchKernelGGL(encode_col16_kernel, dim3((unsigned)tiles.ntiles * kSpmvSub),
                       dim3(kSpmvRows), 0, s, tiles, ptr, col, col16, bad);
    HIP_LAUNCH_CHECK();
}

template <typename TV, typename TX, typename TA, typename TY>
void launch_spmv(hipStream_t s, int lanes, const Tiles &tiles, const int *flag, int flag_stride,
                 const int *ptr, const int *col, const TV *val, const TX *x, TY *y,
                 int stream_per, const short *col16, XSel xs) {
    if (tiles.ntiles == 0)
        return;
    const dim3 grid((unsigned)tiles.ntiles * kSpmvSub), block(kSpmvRows);
#define BSPAI_SPMV_CASE(L)                                                                        \
    case L:                                                                                      \
        hipLaunchKernelGGL((spmv_group_kernel<L, TV, TX, TA, TY>), grid, block, 0, s, tiles,     \
                           flag, flag_stride, ptr, col, val, x, y, xs);                          \
        break;
    switch (lanes) {
    case 1:
        if (col16 && stream_per > 8)
            hipLaunchKernelGGL((spmv_stream_kernel<TV, TX, TA, TY, 16, true>), grid, block, 0, s,
                               tiles, flag, flag_stride, ptr, col, col16, val, x, y, xs);
        else if (col16)
            hipLaunchKernelGGL((spmv_stream_kernel<TV, TX, TA, TY, 8, true>), grid, block, 0, s,
                               tiles, flag, flag_stride, ptr, col, col16, val, x, y, xs);
        else if (stream_per > 8)
            hipLaunchKernelGGL((spmv_stream_kernel<TV, TX, TA, TY, 16, false>), grid, block, 0, s,
                               tiles, flag, flag_stride, ptr, col, col16, val, x, y, xs);
        else
            hipLaunchKernelGGL((spmv_stream_kernel<TV, TX, TA, TY, 8, false>), grid, block, 0, s,
                               tiles, flag, flag_stride, ptr, col, col16, val, x, y, xs);
        break;
    case -1:
        hipLaunchKernelGGL((spmv_group_kernel<1, TV, TX, TA, TY>), grid, block, 0, s, tiles, flag,
                           flag_stride, ptr, col, val, x, y, xs);
        break;
        BSPAI_SPMV_CASE(2)
        BSPAI_SPMV_CASE(4)
        BSPAI_SPMV_CASE(8)
        BSPAI_SPMV_CASE(16)
        BSPAI_SPMV_CASE(32)
        BSPAI_SPMV_CASE(64)
    default:
        throw std::invalid_argument("spmv lanes must be -1 or a power of two <= 64");
    }
#undef BSPAI_SPMV_CASE
    HIP_LAUNCH_CHECK();
}

#define BSPAI_SPMV_INST(TV, TX, TA, TY)                                                           \
    template void launch_spmv<TV, TX, TA, TY>(hipStream_t, int, const Tiles &, const int *, int,  \
                                              const int *, const int *, const TV *, const TX *,   \
                                              TY *, int, const short *, XSel);
BSPAI_SPMV_INST(double, double, double, double)
BSPAI_SPMV_INST(float, float, float, float)
BSPAI_SPMV_INST(float, double, double, double)
BSPAI_SPMV_INST(double, float, double, double)
BSPAI_SPMV_INST(float, float, Comp32, float)
#undef BSPAI_SPMV_INST

template <typename TS, typename TD>
__global__ void convert_kernel(std::int64_t n, const TS *__restrict__ src, TD *__restrict__ dst) {
    const std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = (TD)src[i];
}
template <typename TS, typename TD>
void launch_convert(hipStream_t s, std::int64_t n, const TS *src, TD *dst) {
    if (n == 0)
        return;
    hipLaunchKernelGGL((convert_kernel<TS, TD>), dim3(ceil_div(n, 256)), dim3(256), 0, s, n, src,
                       dst);
    HIP_LAUNCH_CHECK();
}
template void launch_convert<double, float>(hipStream_t, std::int64_t, const double *, float *);
template void launch_convert<float, double>(hipStream_t, std::int64_t, const float *, double *);
template void launch_convert<double, double>(hipStream_t, std::int64_t, const double *, double *);

// ------------------------------------------------------------------ SDDMM
// out[q] = alpha * u[row(q)] * v[col(q)] over the stored pattern: the gradient
// of a bilinear loss u^T A v restricted to A's nonzeros.
__global__ void __launch_bounds__(kTileThreads)
sddmm_kernel(Tiles tiles, const int *__restrict__ ptr, const int *__restrict__ col, double alpha,
             const double *__restrict__ u, const double *__restrict__ v, double *__restrict__ out) {
    const int t = blockIdx.x;
    const int r1 = tiles.row1[t];
    for (int row = tiles.row0[t] + threadIdx.x / 8; row < r1; row += kTileThreads / 8) {
        const double ur = alpha * u[row];
        for (int e = ptr[row] + (threadIdx.x & 7); e < ptr[row + 1]; e += 8)
            out[e] = ur * v[col[e]];
    }
}

void launch_sddmm(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col, double alpha,
                  const double *u, const double *v, double *out) {
    if (tiles.ntiles == 0)
        return;
    hipLaunchKernelGGL(sddmm_kernel, dim3(tiles.ntiles), dim3(kTileThreads), 0, s, tiles, ptr, col,
                       alpha, u, v, out);
    HIP_LAUNCH_CHECK();
}

// ------------------------------------------------------------------ scan
// Inclusive scan of int32 data[0..n) in place: each block scans a chunk of
// kScanChunk elements, block totals are scanned by one block, then added back.
constexpr int kScanThreads = 256, kScanPer = 8, kScanChunk = kScanThreads * kScanPer;

__global__ void scan_chunks_kernel(int *__restrict__ data, std::int64_t n, int *__restrict__ sums) {
    __shared__ int lds[kScanThreads];
    const std::int64_t base = (std::int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
    int v[kScanPer];
    int run = 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i) {
        run += (base + i < n) ? data[base + i] : 0;
        v[i] = run;
    }
    lds[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int t = threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    const int prefix = threadIdx.x ? lds[threadIdx.x - 1] : 0;
#pragma unroll
    for (int i = 0; i < kScanPer; ++i)
        if (base + i < n)
            data[base + i] = v[i] + prefix;
    if (threadIdx.x == kScanThreads - 1)
        sums[blockIdx.x] = lds[kScanThreads - 1];
}

__global__ void scan_sums_kernel(int *__restrict__ sums, int nblocks) {
    // single block; sequential carry across passes of kScanThreads
    __shared__ int lds[kScanThreads];
    int carry = 0;
    for (int b0 = 0; b0 < nblocks; b0 += kScanThreads) {
        const int i = b0 + threadIdx.x;
        lds[threadIdx.x] = i < nblocks ? sums[i] : 0;
        __syncthreads();
        for (int off = 1; off < kScanThreads; off <<= 1) {
            const int t = threadIdx.x >= off ? lds[threadIdx.x - off] : 0;
            __syncthreads();
            lds[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nblocks)
            sums[i] = lds[threadIdx.x] + carry;
        const int total = lds[kScanThreads - 1];
        __syncthreads();
        carry += total;
    }
}

__global__ void scan_add_kernel(int *__restrict__ data, std::int64_t n, const int *__restrict__ sums) {
    if (blockIdx.x == 0)
        return;
    const int add = sums[blockIdx.x - 1];
    const std::int64_t base = (std::int64_t)blockIdx.x * kScanChunk;
    for (int i = threadIdx.x; i < kScanChunk; i += kScanThreads)
        if (base + i < n)
            data[base + i] += add;
}

std::int64_t scan_tmp_ints(std::int64_t n) {
    return (n + kScanChunk - 1) / kScanChunk + 1;
}

void scan_inclusive_int(hipStream_t s, int *data, std::int64_t n, int *tmp) {
    if (n == 0)
        return;
    const int nb = (int)((n + kScanChunk - 1) / kScanChunk);
    hipLaunchKernelGGL(scan_chunks_kernel, dim3(nb), dim3(kScanThreads), 0, s, data, n, tmp);
    HIP_LAUNCH_CHECK();
    if (nb > 1) {
        hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, s, tmp, nb);
        HIP_LAUNCH_CHECK();
        hipLaunchKernelGGL(scan_add_kernel, dim3(nb), dim3(kScanThreads), 0, s, data, n, tmp);
        HIP_LAUNCH_CHECK();
    }
}

} // namespace bspai_hip
