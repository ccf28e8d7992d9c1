// Stable LSD radix sort of (64-bit key, FP64 value) pairs on the key bits
// [0, end_bit), 8 bits per pass, for the canonical CSR from coordinate
// entries (duplicates must stay in input order, src/cpu.cpp:48-76).
//
// One pass (digit d = (key >> shift) & 255), over tiles of 4,096 items:
//   1. histogram: each tile counts its digits in LDS and writes them
//      digit-major, hist[d * tiles + tile];
//   2. scan: an inclusive scan of hist gives every (digit, tile) its start,
//      hist[d * tiles + tile - 1] (all smaller digits, then this digit in
//      earlier tiles);
//   3. scatter: each tile walks its items in order, 256 at a time (one per
//      thread).  A wavefront finds the lanes that share a lane's digit with
//      eight 64-bit ballots (one per digit bit); the lane's rank is the count
//      of such lanes below it, and each group's lowest lane writes the
//      group's size to LDS.  An item goes to its digit's start in the tile,
//      plus the digit's count in the tile's earlier rounds, plus its earlier
//      waves' counts in this round, plus its rank.
// Items keep their order within a digit at every step, so the sort is stable
// and equals any other stable LSD sort on the same bits.
#include "common.hpp"
#include "kernels.hpp"

namespace bspai_hip {
namespace {

constexpr int kRadixThreads = 256, kRadixItems = 16, kRadixTile = kRadixThreads * kRadixItems;
constexpr int kRadixBins = 256, kRadixWaves = kRadixThreads / kWave;

__global__ void __launch_bounds__(kRadixThreads)
radix_hist_kernel(const unsigned long long *__restrict__ keys, int n, int shift,
                  int *__restrict__ hist, int tiles) {
    __shared__ int h[kRadixBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const std::int64_t base = (std::int64_t)blockIdx.x * kRadixTile;
#pragma unroll 4
    for (int i = 0; i < kRadixItems; ++i) {
        const std::int64_t idx = base + i * kRadixThreads + threadIdx.x;
        if (idx < n)
            atomicAdd(&h[(int)((keys[idx] >> shift) & 0xffu)], 1);
    }
    __syncthreads();
    hist[(std::int64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kRadixThreads)
radix_scatter_kernel(const unsigned long long *__restrict__ kin, const double *__restrict__ vin,
                     unsigned long long *__restrict__ kout, double *__restrict__ vout, int n,
                     int shift, const int *__restrict__ incl, int tiles) {
    __shared__ int start[kRadixBins];   // the digit's first position for this tile
    __shared__ int running[kRadixBins]; // the digit's items in this tile's earlier rounds
    __shared__ int wcnt[kRadixWaves][kRadixBins];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    {
        const std::int64_t j = (std::int64_t)tid * tiles + blockIdx.x;
        start[tid] = j == 0 ? 0 : incl[j - 1];
        running[tid] = 0;
    }
    const unsigned long long below = (1ull << lane) - 1ull;
    const std::int64_t base = (std::int64_t)blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixItems; ++r) {
#pragma unroll
        for (int w = 0; w < kRadixWaves; ++w)
            wcnt[w][tid] = 0;
        const std::int64_t idx = base + r * kRadixThreads + tid;
        const bool valid = idx < n;
        const unsigned long long key = valid ? kin[idx] : 0ull;
        const double val = valid ? vin[idx] : 0.0;
        const int d = (int)((key >> shift) & 0xffu);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        __syncthreads(); // wcnt cleared
        if (valid && rank == 0)
            wcnt[wave][d] = __popcll(peers);
        __syncthreads(); // group sizes written
        if (valid) {
            int pos = start[d] + running[d] + rank;
            for (int w = 0; w < wave; ++w)
                pos += wcnt[w][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads(); // every position read running and wcnt
        int add = 0;
#pragma unroll
        for (int w = 0; w < kRadixWaves; ++w)
            add += wcnt[w][tid];
        running[tid] += add;
        __syncthreads(); // running updated before the next round clears wcnt
    }
}

std::size_t align256(std::size_t x) { return (x + 255) & ~(std::size_t)255; }

} // namespace

std::size_t radix_sort_tmp_bytes(std::int64_t count) {
    const std::int64_t tiles = std::max<std::int64_t>(1, (count + kRadixTile - 1) / kRadixTile);
    const std::int64_t hist = tiles * kRadixBins;
    return align256(8 * (std::size_t)count) + align256(8 * (std::size_t)count) +
           align256(4 * (std::size_t)hist) + align256(4 * (std::size_t)scan_tmp_ints(hist));
}

void radix_sort_pairs(hipStream_t s, unsigned long long *keys, const double *vals,
                      unsigned long long *keys_out, double *vals_out, std::int64_t count,
                      int end_bit, void *tmp) {
    if (count <= 0)
        return;
    if (count >= INT_MAX)
        throw std::invalid_argument("radix sort: at most 2^31 - 2 items");
    const int n = (int)count;
    const int tiles = (int)((count + kRadixTile - 1) / kRadixTile);
    const std::int64_t nhist = (std::int64_t)tiles * kRadixBins;
    char *t = static_cast<char *>(tmp);
    auto *tk = reinterpret_cast<unsigned long long *>(t);
    t += align256(8 * (std::size_t)count);
    auto *tv = reinterpret_cast<double *>(t);
    t += align256(8 * (std::size_t)count);
    int *hist = reinterpret_cast<int *>(t);
    t += align256(4 * (std::size_t)nhist);
    int *scan_tmp = reinterpret_cast<int *>(t);
    const int passes = std::max(1, (end_bit + 7) / 8);
    // ping-pong so that the last pass writes (keys_out, vals_out); the first
    // reads the caller's values, which are never written
    const unsigned long long *sk = keys;
    const double *sv = vals;
    for (int p = 0; p < passes; ++p) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        unsigned long long *dk = to_out ? keys_out : tk;
        double *dv = to_out ? vals_out : tv;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(tiles), dim3(kRadixThreads), 0, s, sk, n, 8 * p,
                           hist, tiles);
        HIP_LAUNCH_CHECK();
        scan_inclusive_int(s, hist, nhist, scan_tmp);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(tiles), dim3(kRadixThreads), 0, s, sk, sv,
                           dk, dv, n, 8 * p, hist, tiles);
        HIP_LAUNCH_CHECK();
        sk = dk;
        sv = dv;
    }
}

} // namespace bspai_hip
