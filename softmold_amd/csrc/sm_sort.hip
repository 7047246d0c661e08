// Stable LSD radix sort of (cell key, particle index) pairs for the cell-list build.
//
// The reference sorts with thrust::sort_by_key (include/cuda/cell.h:254-260) and its CPU
// path keeps unsorted linked lists (include/algorithms/cellOpt.h:572-585).  Here: 8-bit
// digits, ceil(bits/8) passes, three kernels per pass:
//   hist     per-tile digit counts in LDS                              (reads 4 B/key)
//   scan     exclusive scan over the digit-major [256 x tiles] table   (one workgroup)
//   scatter  per-tile stable ranking with wave64 ballots, then scatter (reads/writes 8 B/key)
// A tile is 4096 items = 16 rounds of 256; within a round the 4 waves of the workgroup
// rank their 64 items by matching the 8 digit bits with ballots, and the per-wave counts
// go through LDS so the result is stable (equal keys keep their input order).
#include "sm_internal.hpp"

namespace smdev {

namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 16;
constexpr int kTile = kThreads * kRounds;
constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kThreads) void k_hist(int n, const uint32_t *__restrict__ keys, int shift,
                                                   int tiles, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int tile = blockIdx.x;
    const int base = tile * kTile;
    for (int r = 0; r < kRounds; r++) {
        int i = base + r * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * tiles + tile] = h[threadIdx.x];
}

// Exclusive scan, in place, of m entries with one workgroup: each thread owns a contiguous
// segment, segment totals are scanned in LDS, segments are rewritten.
__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t *__restrict__ a, int m)
{
    __shared__ uint32_t s[kScanThreads];
    const int t = threadIdx.x;
    const int seg = (m + kScanThreads - 1) / kScanThreads;
    const int lo = t * seg;
    const int hi = min(lo + seg, m);
    uint32_t sum = 0;
    for (int i = lo; i < hi; i++) sum += a[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
    for (int i = lo; i < hi; i++) {
        uint32_t x = a[i];
        a[i] = run;
        run += x;
    }
}

__global__ __launch_bounds__(kThreads) void k_scatter(int n, const uint32_t *__restrict__ kin,
                                                      const uint32_t *__restrict__ vin,
                                                      uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                      int shift, int tiles, const uint32_t *__restrict__ hist)
{
    __shared__ uint32_t base[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kThreads / kWave][256];
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = t / kWave;
    const int tile = blockIdx.x;
    base[t] = hist[(size_t)t * tiles + tile];
    run[t] = 0;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int r = 0; r < kRounds; r++) {
        const int i = tile * kTile + r * kThreads + t;
        const bool valid = i < n;
        uint32_t k = 0, v = 0, d = 0;
        if (valid) {
            k = kin[i];
            v = vin[i];
            d = (k >> shift) & 255u;
        }
        for (int w = 0; w < kThreads / kWave; w++) wcnt[w][t] = 0;
        __syncthreads();
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + run[d] + (uint32_t)rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][d];
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kThreads / kWave; w++) add += wcnt[w][t];
        run[t] += add;
        __syncthreads();
    }
}

}  // namespace

size_t sort_hist_entries(int n)
{
    int tiles = (n + kTile - 1) / kTile;
    if (tiles < 1) tiles = 1;
    return (size_t)256 * tiles;
}

int radix_sort(hipStream_t st, SortBuffers &b, int bits)
{
    const int n = b.n;
    if (n <= 1) return 0;
    b.tiles = (n + kTile - 1) / kTile;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    uint32_t *ki = b.k0, *vi = b.v0, *ko = b.k1, *vo = b.v1;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_hist, dim3(b.tiles), dim3(kThreads), 0, st, n, ki, shift, b.tiles, b.hist);
        hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, b.hist, 256 * b.tiles);
        hipLaunchKernelGGL(k_scatter, dim3(b.tiles), dim3(kThreads), 0, st, n, ki, vi, ko, vo, shift, b.tiles,
                           b.hist);
        uint32_t *tk = ki, *tv = vi;
        ki = ko; vi = vo; ko = tk; vo = tv;
    }
    return (passes & 1) ? 1 : 0;
}

}  // namespace smdev
