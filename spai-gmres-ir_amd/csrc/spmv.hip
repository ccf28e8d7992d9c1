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
            Acc<TA>::add(acc, Acc<TA>::mul(val[e], gld(x + col[e])));
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
            xv[u] = gld(x + cj[u]);
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

// ------------------------------------------------------------ sliced SpMV
// Sliced layout (rows in slices of 64, one wavefront per slice): slice s of
// tile t holds rows row0(t) + 64 (s - 32 t) ..., never crossing a tile, so
// never a system.  A slice's entries are stored column-major -- slot
// (j, lane) at soff[s] + 64 j + lane holds entry j of the lane's row -- up to
// the slice's longest row (padded slots: column = the row, value 0);
// slen[row] is the row's length.  Lane l adds its
// row's products for j = 0, 1, ... in CSR order, the reference's sequential
// row sum (src/cpu.cpp:88-89), bit for bit.  Every load is coalesced: the
// entries of 64 consecutive rows are contiguous per j, and for a stencil the
// x gathers of one j are consecutive rows too.  Columns are int32, or int16
// offsets to the row (scol16[q] = col - row) when every entry fits.
constexpr int kSliceRows = kWave;
constexpr int kSlicesPerTile = kTileRows / kSliceRows;
constexpr int kSlicesPerBlock = kSpmvRows / kSliceRows;

template <typename TV, typename TX, typename TA, typename TY, bool C16>
__global__ void __launch_bounds__(kSpmvRows)
spmv_slice_kernel(Tiles tiles, const int *__restrict__ flag, int flag_stride, SliceView sv,
                  const TV *__restrict__ sval, const TX *__restrict__ x, TY *__restrict__ y,
                  XSel xs) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int sl = b * kSlicesPerBlock + (int)(threadIdx.x / kSliceRows);
    const int lane = threadIdx.x % kSliceRows;
    int row0, row1;
    if (tiles.nsys == 1) {
        row0 = sl * kSliceRows;
        row1 = min(tiles.nrows, row0 + kSliceRows);
        if (flag && !flag[0])
            return;
    } else {
        const int t = b / kSpmvSub;
        if (flag && !flag[(std::int64_t)tiles.sys[t] * flag_stride])
            return;
        row0 = tiles.row0[t] + (sl - t * kSlicesPerTile) * kSliceRows;
        row1 = min(tiles.row1[t], row0 + kSliceRows);
    }
    const int row = row0 + lane;
    if (row >= row1)
        return;
    const int s0 = gld(sv.soff + sl), s1 = gld(sv.soff + sl + 1);
    const int len = gld(sv.slen + row);
    const int width = (s1 - s0) / kSliceRows; // the slice's longest row
    const int base = s0 + lane;
    if (xs.tab)
        x = static_cast<const TX *>(xs.tab[min(*xs.idx + xs.off, xs.last)]);
    TA acc = Acc<TA>::zero();
    // U slots in flight per lane: every slot below the slice's width is
    // loaded (a padded slot holds column = the row and value 0), and only the
    // row's own entries are added, in order.
    constexpr int U = 8;
    for (int j0 = 0; j0 < width; j0 += U) {
        int cj[U];
        TV vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = base + min(j0 + u, width - 1) * kSliceRows;
            if constexpr (C16)
                cj[u] = row + (int)gld(sv.scol16 + q);
            else
                cj[u] = gld(sv.scol + q);
            vv[u] = gld(sval + q);
        }
        TX xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
            xv[u] = gld(x + cj[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (j0 + u < len)
                Acc<TA>::add(acc, Acc<TA>::mul(vv[u], xv[u]));
    }
    y[row] = Acc<TA>::template out<TY>(acc);
}

template <typename TV, typename TX, typename TA, typename TY>
void launch_spmv_sliced(hipStream_t s, const Tiles &tiles, const int *flag, int flag_stride,
                        const SliceView &sv, const TV *sval, const TX *x, TY *y, XSel xs) {
    if (tiles.ntiles == 0)
        return;
    const dim3 grid((unsigned)tiles.ntiles * kSpmvSub), block(kSpmvRows);
    if (sv.scol16)
        hipLaunchKernelGGL((spmv_slice_kernel<TV, TX, TA, TY, true>), grid, block, 0, s, tiles,
                           flag, flag_stride, sv, sval, x, y, xs);
    else
        hipLaunchKernelGGL((spmv_slice_kernel<TV, TX, TA, TY, false>), grid, block, 0, s, tiles,
                           flag, flag_stride, sv, sval, x, y, xs);
    HIP_LAUNCH_CHECK();
}
#define BSPAI_SLICED_INST(TV, TX, TA, TY)                                                         \
    template void launch_spmv_sliced<TV, TX, TA, TY>(hipStream_t, const Tiles &, const int *, int, \
                                                     const SliceView &, const TV *, const TX *,   \
                                                     TY *, XSel);
BSPAI_SLICED_INST(double, double, double, double)
BSPAI_SLICED_INST(float, float, float, float)
BSPAI_SLICED_INST(float, float, Comp32, float)
#undef BSPAI_SLICED_INST

// Slice geometry: one wavefront per slice writes its rows' lengths (0 past
// the tile) and the slice's width, the longest row; *bad is set when a row is
// longer than 255 entries or an entry's column offset does not fit 16 bits
// (bit 0: length, bit 1: offset).
__global__ void __launch_bounds__(kSpmvRows)
slice_width_kernel(Tiles tiles, const int *__restrict__ ptr, const int *__restrict__ col,
                   unsigned char *__restrict__ slen, int *__restrict__ width, int *__restrict__ bad) {
    const int b = blockIdx.x;
    const int t = b / kSpmvSub;
    const int sl = b * kSlicesPerBlock + (int)(threadIdx.x / kSliceRows);
    const int lane = threadIdx.x % kSliceRows;
    const int row0 = tiles.row0[t] + (sl - t * kSlicesPerTile) * kSliceRows;
    const int row1 = min(tiles.row1[t], row0 + kSliceRows);
    const int row = row0 + lane;
    int len = 0, flags = 0;
    if (row < row1) {
        const int e0 = ptr[row], e1 = ptr[row + 1];
        len = e1 - e0;
        if (len > 255)
            flags |= 1;
        for (int e = e0; e < e1; ++e) {
            const int d = col[e] - row;
            if (d < -32768 || d > 32767)
                flags |= 2;
        }
        slen[row] = (unsigned char)min(len, 255);
    }
    len = wave_max(len);
    flags = (__any(flags & 1) ? 1 : 0) | (__any(flags & 2) ? 2 : 0);
    if (lane == 0) {
        width[sl] = len;
        if (flags)
            atomicOr(bad, flags);
    }
}

// Entries into the sliced layout (structure: columns, both forms; values:
// a separate pass, so new values refill only those).
template <typename TV>
__global__ void __launch_bounds__(kSpmvRows)
slice_fill_kernel(Tiles tiles, const int *__restrict__ ptr, const int *__restrict__ col,
                  const TV *__restrict__ val, SliceView sv, int *__restrict__ scol,
                  short *__restrict__ scol16, TV *__restrict__ sval) {
    const int b = blockIdx.x;
    const int t = b / kSpmvSub;
    const int sl = b * kSlicesPerBlock + (int)(threadIdx.x / kSliceRows);
    const int lane = threadIdx.x % kSliceRows;
    const int row0 = tiles.row0[t] + (sl - t * kSlicesPerTile) * kSliceRows;
    const int row1 = min(tiles.row1[t], row0 + kSliceRows);
    const int row = row0 + lane;
    if (row >= row1)
        return;
    const int e0 = ptr[row], len = ptr[row + 1] - e0;
    const int base = sv.soff[sl] + lane;
    const int width = (sv.soff[sl + 1] - sv.soff[sl]) / kSliceRows;
    // slots past the row's end: column = the row (offset 0), value 0, so a
    // load of any slot below the width is a valid gather
    for (int j = 0; j < width; ++j) {
        const int q = base + j * kSliceRows;
        const bool in = j < len;
        if (scol)
            scol[q] = in ? col[e0 + j] : row;
        if (scol16)
            scol16[q] = in ? (short)(col[e0 + j] - row) : (short)0;
        if (sval)
            sval[q] = in ? val[e0 + j] : TV(0);
    }
}

void launch_slice_width(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col,
                        unsigned char *slen, int *width, int *bad) {
    if (tiles.ntiles == 0)
        return;
    hipLaunchKernelGGL(slice_width_kernel, dim3((unsigned)tiles.ntiles * kSpmvSub),
                       dim3(kSpmvRows), 0, s, tiles, ptr, col, slen, width, bad);
    HIP_LAUNCH_CHECK();
}
template <typename TV>
void launch_slice_fill(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col,
                       const TV *val, const SliceView &sv, int *scol, short *scol16, TV *sval) {
    if (tiles.ntiles == 0)
        return;
    hipLaunchKernelGGL((slice_fill_kernel<TV>), dim3((unsigned)tiles.ntiles * kSpmvSub),
                       dim3(kSpmvRows), 0, s, tiles, ptr, col, val, sv, scol, scol16, sval);
    HIP_LAUNCH_CHECK();
}
template void launch_slice_fill<double>(hipStream_t, const Tiles &, const int *, const int *,
                                        const double *, const SliceView &, int *, short *,
                                        double *);
template void launch_slice_fill<float>(hipStream_t, const Tiles &, const int *, const int *,
                                       const float *, const SliceView &, int *, short *, float *);

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
