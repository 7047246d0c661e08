// Device helpers of the cell-key radix sort shared by the sort kernels (sm_sort.hip) and
// the kernel that computes the keys (k_first in sm_kernels.hip), which also builds the
// sort's digit histograms while the keys are in registers.
#pragma once
#include "sm_internal.hpp"

namespace smdev {

// lanes of the wave holding the same 8-bit digit as this lane (invalid lanes excluded)
__device__ inline unsigned long long match_digit(uint32_t d, bool valid)
{
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

__device__ inline unsigned long long lanes_below()
{
    const int lane = threadIdx.x & (kWave - 1);
    return (lane == 0) ? 0ull : (~0ull >> (64 - lane));
}

// Adds the digits of one key per thread (all `passes` 8-bit digits) to a block histogram
// h[4][256] in LDS, one LDS atomic per lane and digit.  Aggregating the lanes of a digit
// with ballots first (one atomic per digit and wave) cost more VALU instructions than the
// LDS conflicts it saved: the fused kick + first-half kernel took 17.5 us against 16.1 us
// at 240k beads.
__device__ inline void hist_add(uint32_t (*h)[256], uint32_t key, bool valid, int passes)
{
    if (valid)
        for (int p = 0; p < passes; p++) atomicAdd(&h[p][(key >> (8 * p)) & 255u], 1u);
}

// The global digit histograms are kept in kHistCopies copies, ghist[copy][4][256], and
// block b adds into copy b % kHistCopies, so the atomics on a busy digit (the high digits
// of a compact system are shared by every block) spread over that many addresses; the
// sort sums the copies.
constexpr int kHistCopies = 16;

// Flushes a block histogram into its copy of the global digit histograms.
// Needs blockDim.x == 256 and a __syncthreads() before the call.
__device__ inline void hist_flush(uint32_t (*h)[256], uint32_t *ghist, int passes)
{
    uint32_t *g = ghist + (size_t)(blockIdx.x % kHistCopies) * 4 * 256;
    for (int p = 0; p < passes; p++)
        if (h[p][threadIdx.x]) atomicAdd(&g[p * 256 + threadIdx.x], h[p][threadIdx.x]);
}

}  // namespace smdev
