// Launch wrappers of the HIP kernels (internal interface of the library).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bspai_hip {

// SPAI column statuses; values mirror bspai::SpaiStatus (reference
// include/bspai/spai.hpp:10-21) plus an internal capacity code.
enum : int {
    kSpaiSuccess = 0,
    kSpaiNotAttempted = 1,
    kSpaiInvalidInput = 2,
    kSpaiRankDeficient = 3,
    kSpaiNumericalFailure = 4,
    kSpaiEmptyCandidates = 5,
    kSpaiZeroDenominator = 6,
    kSpaiInvalidScore = 7,
    kSpaiNoGrowth = 8,
    kSpaiIterationLimit = 9,
    kSpaiOverflow = 100, // work arrays too small: re-run with larger caps
};

// Per-column results of a construction pass; column k's final pattern J and
// coefficients sit at pool[pass[k]] + off[k], q[k] entries.
struct SpaiColumnOut {
    int *status, *fits, *growths, *clamps, *q;
    double *residual, *pivot;
    double *flops; // QR flops of the column's fits, 2mq^2 - 2q^3/3 + 4mq + q^2 each
    std::int64_t *off;
    int *pass;
    int pass_id;
};

std::int64_t spai_lane_doubles(int mcap, int qcap, int ccap);
std::int64_t spai_lane_ints(int mcap, int qcap, int ccap);

template <typename T>
void launch_spai_columns(hipStream_t s, int blocks, int threads, int n, const int *aptr,
                         const int *acol, const T *aval, const T *dscale, const int *bptr,
                         const int *bcol, T eps, long long alpha, int beta, const int *worklist,
                         int nwork, int *cursor, int mcap, int qcap, int ccap, T *dwork,
                         int *iwork, SpaiColumnOut out, int *out_j, T *out_c,
                         std::int64_t out_stride);
// FP32 wavefront-per-column kernel (first bucket of the mixed path), n up to
// spai_wave_max_n() (32-bit sort keys).  bval: the values of B = A^T
// (unscaled), aligned with bcol.
void launch_spai_wave(hipStream_t s, int blocks, const int *aptr, const int *acol,
                      const float *aval, const float *dscale, const int *bptr, const int *bcol,
                      const float *bval, float eps, long long alpha, int beta,
                      const int *worklist, int nwork, int *cursor, SpaiColumnOut out,
                      int *out_j, float *out_c, std::int64_t out_stride);
int spai_wave_max_n();
int spai_wave_qcap();
// Workgroup-per-column kernel for the columns that outgrow the lane kernel.
// pcap: (FP32) the (row, candidate) pairs one fit may hold (see spai_pair_cap)
std::int64_t spai_block_doubles(int mcap, int qcap, int ccap, std::int64_t pcap);
std::int64_t spai_block_ints(int mcap, int qcap, int ccap, std::int64_t pcap, int words, int n);
// Pair capacity of a workgroup pass: 8 per candidate of the cap, or every
// stored entry of the matrix (no fit can need more) once the caps reach n.
inline std::int64_t spai_pair_cap(int ccap, int n, std::int64_t nnz) {
    const std::int64_t want = ccap >= n ? nnz : 8LL * ccap;
    return want < nnz ? (want > 0 ? want : 1) : (nnz > 0 ? nnz : 1);
}
// p[b*stride + offset + i] = value for b < blocks, i < count.
void launch_fill_int(hipStream_t s, std::int64_t blocks, std::int64_t stride, std::int64_t offset,
                     int count, int *p, int value);
// bval: B's values aligned with bcol (the FP32 build scores candidates from
// them); may be null for FP64.
template <typename T>
void launch_spai_block(hipStream_t s, int blocks, int n, const int *aptr, const int *acol,
                       const T *aval, const T *dscale, const int *bptr, const int *bcol,
                       const T *bval, T eps, long long alpha, int beta, const int *worklist,
                       int nwork, int *cursor, int mcap, int qcap, int ccap, T *dwork,
                       int *iwork, SpaiColumnOut out, int *out_j, T *out_c,
                       std::int64_t out_stride, int occupancy, std::int64_t pcap);
template <typename T>
void launch_scale_rows(hipStream_t s, int n, const int *ptr, const int *col, const T *val, T *d,
                       int *first_bad);
void launch_count_cols(hipStream_t s, std::int64_t nnz, const int *col, int *counts);
template <typename T>
void launch_transpose_fill(hipStream_t s, int n, const int *ptr, const int *col, const T *val,
                           const int *tptr, int *cursor, int *tcol, T *tval);

// exclusive/inclusive scan helpers (int32), in place on counts[0..n]:
// counts[0] must be 0; after the call counts[i] = sum of counts[<i] inputs.
void scan_inclusive_int(hipStream_t s, int *data, std::int64_t n, int *tmp);
// Stable LSD radix sort of (key, value) pairs on the key bits [0, end_bit)
// (radix.hip): keys (overwritten) and vals (read only) in, keys_out and
// vals_out out; tmp holds radix_sort_tmp_bytes(count) bytes.
std::size_t radix_sort_tmp_bytes(std::int64_t count);
void radix_sort_pairs(hipStream_t s, unsigned long long *keys, const double *vals,
                      unsigned long long *keys_out, double *vals_out, std::int64_t count,
                      int end_bit, void *tmp);
std::int64_t scan_tmp_ints(std::int64_t n);

// ------------------------------------------------------------------ tiles
// Row tiles: every vector kernel runs one block per tile; a tile lies inside
// one system.  sys[t] = system of tile t, rows [row0[t], row1[t]); the tiles
// of system s are [sys_tile0[s], sys_tile1[s]).
struct Tiles {
    int ntiles, nsys;
    const int *sys;
    const int *row0;
    const int *row1;
    const int *sys_tile0;
    const int *sys_tile1;
    int uniform; // every tile full (kTileRows rows) and starting at a multiple of 4 rows
    int nrows;   // rows of all systems; with one system, tile t is rows
                 // [t kTileRows, min((t+1) kTileRows, nrows)) of system 0
};
// A tile is one block of the vector kernels: 256 threads x 8 rows, one
// partial per tile and reduction.
constexpr int kTileRows = 2048;
constexpr int kTileThreads = 256;
constexpr int kTilePer = kTileRows / kTileThreads;
// SpMV blocks cover 256 consecutive rows of a tile (one row per thread).
constexpr int kSpmvRows = 256;
constexpr int kSpmvSub = kTileRows / kSpmvRows;

// ------------------------------------------------------------------ SpMV
// y[r] = sum_q val[q] * x[col[q]] over row r, accumulated in TA.  lanes = the
// number of lanes cooperating on one row (1 = sequential CSR order, which is
// bit-identical to the reference's CsrMatrix::multiply, src/cpu.cpp:80-94).
// flag (optional): a tile runs only if flag[sys * flag_stride] != 0.
// stream_per (lanes = 1): entries per thread per chunk, 8 or 16 (for matrices
// averaging more than 8 entries per row); the result does not depend on it.
// col16 (lanes = 1, optional): the columns as 16-bit offsets to each 256-row
// block's first row (launch_encode_col16); the result does not depend on it.
// xs (optional): x is taken on the device as tab[min(*idx + off, last)],
// e.g. the basis vector of the step a replayed graph is at; x is then unused.
struct XSel {
    const void *const *tab = nullptr;
    const int *idx = nullptr;
    int off = 0, last = 0;
};
template <typename TV, typename TX, typename TA, typename TY>
void launch_spmv(hipStream_t s, int lanes, const Tiles &tiles, const int *flag, int flag_stride,
                 const int *ptr, const int *col, const TV *val, const TX *x, TY *y,
                 int stream_per = 8, const short *col16 = nullptr, XSel xs = XSel{});
// Sliced layout of a matrix for the wavefront-per-slice SpMV (64 rows per
// slice, entries column-major within a slice; see spmv.hip).  soff: entry
// offset of each slice (ntiles * 32 + 1); slen: each row's length (<= 255);
// scol (int32 columns) or scol16 (int16 offsets col - row, when set).
struct SliceView {
    const int *soff = nullptr;
    const unsigned char *slen = nullptr;
    const int *scol = nullptr;
    const short *scol16 = nullptr;
};
template <typename TV, typename TX, typename TA, typename TY>
void launch_spmv_sliced(hipStream_t s, const Tiles &tiles, const int *flag, int flag_stride,
                        const SliceView &sv, const TV *sval, const TX *x, TY *y,
                        XSel xs = XSel{});
// Slice geometry: slen per row, width per slice (longest row); *bad (zeroed
// by the caller) gets bit 0 for a row longer than 255, bit 1 for a column
// offset outside int16.
void launch_slice_width(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col,
                        unsigned char *slen, int *width, int *bad);
// Fills the sliced columns (scol and/or scol16) and/or values (sval).
template <typename TV>
void launch_slice_fill(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col,
                       const TV *val, const SliceView &sv, int *scol, short *scol16, TV *sval);
// col16[e] = col[e] - (first row of the SpMV row block holding entry e);
// *bad (zeroed by the caller) is set when an offset does not fit 16 bits.
void launch_encode_col16(hipStream_t s, const Tiles &tiles, const int *ptr, const int *col,
                         short *col16, int *bad);
// Entries per thread for the streaming SpMV of an n x n matrix with nnz entries.
inline int stream_per(long long n, long long nnz) {
    return nnz > 8 * n ? 16 : 8;
}

template <typename TS, typename TD>
void launch_convert(hipStream_t s, std::int64_t n, const TS *src, TD *dst);

} // namespace bspai_hip
