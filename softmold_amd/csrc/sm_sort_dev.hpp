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

// ---- the window sort: the key kernel's side ----------------------------------------------
struct WinDesc {
    uint32_t x0, y0, z0, wx, wy, wz;
};
__device__ inline WinDesc load_win(const uint32_t *d) { return WinDesc{d[0], d[1], d[2], d[3], d[4], d[5]}; }

// occupied-range words hold max(c + 1) and max(kBoxEnc - c) per axis, so that zero is empty
constexpr uint32_t kBoxEnc = 1u << 20;

// The end of a key kernel's block in a window sort; every thread of the 256-thread block
// calls it and gets the key to store.  A key in the window is counted by its two window
// digits (bits 0-8 and 9-17); a key outside it sets the fallback flag.  With no window the
// key is stored as it is (the fallback counts the digits).  hs: the block's 1024 LDS
// words, zero on entry.
__device__ inline uint32_t win_key_end(const WinBook &wb, const WinDesc &w, uint32_t *hs, uint32_t k, int cx, int cy,
                                       int cz, bool valid)
{
    const int t = threadIdx.x, lane = t & (kWave - 1);
    if (w.wx == 0u) return k;   // workgroup-uniform
    bool inw = false;
    uint32_t wk = 0u;
    if (valid) {
        const uint32_t dx = (uint32_t)cx - w.x0, dy = (uint32_t)cy - w.y0, dz = (uint32_t)cz - w.z0;
        inw = dx < w.wx && dy < w.wy && dz < w.wz;
        wk = (dz * w.wy + dy) * w.wx + dx;
    }
    if (inw) {
        atomicAdd(&hs[wk & 511u], 1u);
        atomicAdd(&hs[512u + ((wk >> 9) & 511u)], 1u);
    }
    if (__ballot(valid && !inw) != 0ull && lane == 0) atomicOr(wb.flag, 1u);
    __syncthreads();
    uint32_t *g = wb.ghw + (size_t)(blockIdx.x % kHistCopies) * 1024;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t c = hs[t + 256 * j];
        if (c) atomicAdd(&g[t + 256 * j], c);
    }
    return inw ? (kWinTag | wk) : k;
}

}  // namespace smdev
