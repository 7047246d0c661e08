// Stable LSD radix sort of (cell key, particle index) pairs for the cell-list build.
//
// The reference sorts with thrust::sort_by_key (include/cuda/cell.h:254-260) and its CPU
// path keeps unsorted linked lists (include/algorithms/cellOpt.h:572-585).  Here: 8-bit
// digits, ceil(bits/8) passes, one kernel per pass ("onesweep"): a workgroup takes the
// next tile id from a counter (so every lower tile is already running), ranks its items
// with wave64 ballots, publishes its per-digit counts, accumulates the counts of the
// tiles before it by decoupled look-back, and scatters.
//
// No kernel of its own clears or counts: the digit histograms of every pass are built by
// the kernel that computes the keys (k_first), and the sort's scratch (histograms, tile
// counters, look-back states) is double-buffered -- while one sort uses scratch[p], the
// reorder kernel after it clears scratch[1-p] for the next sort.
//
// Counting is wave-aggregated: the lanes holding one digit are found with 8 ballots and
// their lowest lane writes the count once, so a digit shared by a whole wave (the high
// digits of a compact system) costs one LDS write instead of 64 contended atomics.
// The look-back spin is bounded: on timeout the kernel records FLAG_SORT_TIMEOUT in the
// caller's flag word, clamps every write in bounds, and the host reports an error.
#include "sm_sort_dev.hpp"

namespace smdev {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr uint32_t kFlagAgg = 1u << 30, kFlagIncl = 2u << 30, kCountMask = (1u << 30) - 1;
constexpr int kSpinLimit = 1 << 20;
constexpr int kLook = 8;

// A tile is kRounds x 256 keys.  Small inputs want short tiles (more of them to spread over
// the chip), large inputs long ones (fewer tiles to look back over): sort_rounds() picks.
template <int kRounds>
__global__ __launch_bounds__(kThreads) void k_onesweep(int n, const uint32_t *__restrict__ kin,
                                                       const uint32_t *__restrict__ vin,
                                                       uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                       int shift, const uint32_t *__restrict__ ghist,
                                                       uint32_t *__restrict__ status, uint32_t *__restrict__ tilectr,
                                                       int *__restrict__ errflag)
{
    __shared__ uint32_t s_tile;
    __shared__ uint32_t scan[256];
    // per (round, wave, digit): item count, then its exclusive output position
    __shared__ uint32_t wc[kRounds][kWavesPerBlock][256];
    const int t = threadIdx.x;
    const int wave = t / kWave;
    if (t == 0) s_tile = atomicAdd(tilectr, 1u);
    // digit t's count over all keys, from every histogram copy: independent of the tile,
    // so its loads are issued first and summed after the look-back
    uint32_t gh[kHistCopies];
#pragma unroll
    for (int c = 0; c < kHistCopies; c++) gh[c] = ghist[(size_t)c * 4 * 256 + t];
    for (int r = 0; r < kRounds; r++)
        for (int w = 0; w < kWavesPerBlock; w++) wc[r][w][t] = 0u;
    __syncthreads();
    const int tile = (int)s_tile;
    const int base = tile * kThreads * kRounds;
    const unsigned long long lt = lanes_below();

    // every round's key and value loads in flight together (indices clamped in bounds;
    // items past n are masked below)
    uint32_t kr[kRounds], vr[kRounds], dr[kRounds], rk[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const int i = min(base + r * kThreads + t, n - 1);
        kr[r] = kin[i];
        vr[r] = vin[i];
    }
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const bool valid = base + r * kThreads + t < n;
        dr[r] = valid ? (kr[r] >> shift) & 255u : 256u;   // 256: no item
        const unsigned long long peers = match_digit(dr[r] & 255u, valid);
        rk[r] = (uint32_t)__popcll(peers & lt);
        if (valid && rk[r] == 0u) wc[r][wave][dr[r]] = (uint32_t)__popcll(peers);
    }
    __syncthreads();

    // tile count of digit t, published for the tiles after this one
    uint32_t mine = 0;
    for (int r = 0; r < kRounds; r++)
        for (int w = 0; w < kWavesPerBlock; w++) mine += wc[r][w][t];
    uint32_t *st = status + (size_t)tile * 256 + t;
    __hip_atomic_store(st, (tile == 0 ? kFlagIncl : kFlagAgg) | mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // Look-back over a window of kLook predecessor tiles at a time: the window's loads are
    // in flight together, so a walk over many aggregate-only tiles costs one memory round
    // trip per kLook tiles.  Entries below tile 0 read as an inclusive zero.
    uint32_t prefix = 0;
    int spins = 0;
    for (int pt = tile - 1; pt >= 0;) {
        uint32_t v[kLook];
#pragma unroll
        for (int w = 0; w < kLook; w++)
            v[w] = pt - w >= 0 ? __hip_atomic_load(status + (size_t)(pt - w) * 256 + t, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                               : kFlagIncl;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int w = 0; w < kLook; w++) {
            if (done || used < w) continue;    // stopped earlier in the window
            const uint32_t f = v[w] & ~kCountMask;
            if (f == 0u) continue;             // not published yet: resume here
            prefix += v[w] & kCountMask;
            used = w + 1;
            done = f == kFlagIncl;
        }
        if (done) break;
        pt -= used;
        if (used < kLook) {                    // waiting on tile pt
            if (++spins >= kSpinLimit) {
                atomicOr(errflag, FLAG_SORT_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    if (tile > 0)
        __hip_atomic_store(st, kFlagIncl | (prefix + mine), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    // digit base: keys of smaller digits (all tiles) + this digit in earlier tiles
    uint32_t g = 0;
#pragma unroll
    for (int c = 0; c < kHistCopies; c++) g += gh[c];
    scan[t] = g;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = (t >= off) ? scan[t - off] : 0u;
        __syncthreads();
        scan[t] += v;
        __syncthreads();
    }
    uint32_t pos = scan[t] - g + prefix;
    for (int r = 0; r < kRounds; r++)
        for (int w = 0; w < kWavesPerBlock; w++) {
            const uint32_t c = wc[r][w][t];
            wc[r][w][t] = pos;
            pos += c;
        }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        if (dr[r] < 256u) {
            uint32_t p = wc[r][wave][dr[r]] + rk[r];
            p = p < (uint32_t)n ? p : (uint32_t)(n - 1);  // only reachable after a timeout
            kout[p] = kr[r];
            vout[p] = vr[r];
        }
    }
}

}  // namespace

// rounds of 256 keys per tile, by key count (measured on 15k, 240k and 3M keys: 3M keys
// sort in 129 us with 3072-key tiles against 148 us with 2048; 240k keys in 37.7 us with
// 2048 against 38.7 with 3072)
int sort_rounds(int n) { return n < 65536 ? 2 : (n < 1000000 ? 8 : 12); }

int sort_passes(int bits)
{
    int passes = (bits + 7) / 8;
    return passes < 1 ? 1 : (passes > 4 ? 4 : passes);
}

size_t sort_hist_words() { return (size_t)kHistCopies * 4 * 256; }

size_t sort_scratch_words(int n, int max_passes)
{
    const int tile = kThreads * sort_rounds(n);
    const size_t tiles = (size_t)((n + tile - 1) / tile) + 1;
    // ghist [copies][4][256] + tile counters [4] + status [passes][tiles][256]
    return (size_t)kHistCopies * 4 * 256 + 4 + (size_t)max_passes * tiles * 256;
}

int radix_sort(hipStream_t st, SortBuffers &b, int bits, int parity)
{
    const int n = b.n;
    if (n <= 1) return 0;
    const int passes = sort_passes(bits);
    const int rounds = sort_rounds(n);
    const int tiles = (n + kThreads * rounds - 1) / (kThreads * rounds);
    b.tiles = tiles;
    uint32_t *ghist = b.scratch[parity];
    uint32_t *tilectr = ghist + (size_t)kHistCopies * 4 * 256;
    uint32_t *status = tilectr + 4;
    uint32_t *ki = b.k0, *vi = b.v0, *ko = b.k1, *vo = b.v1;
    for (int p = 0; p < passes; p++) {
        uint32_t *sp = status + (size_t)p * tiles * 256;
        if (rounds == 12)
            hipLaunchKernelGGL(k_onesweep<12>, dim3(tiles), dim3(kThreads), 0, st, n, ki, vi, ko, vo, 8 * p,
                               ghist + p * 256, sp, tilectr + p, b.errflag);
        else if (rounds == 2)
            hipLaunchKernelGGL(k_onesweep<2>, dim3(tiles), dim3(kThreads), 0, st, n, ki, vi, ko, vo, 8 * p,
                               ghist + p * 256, sp, tilectr + p, b.errflag);
        else
            hipLaunchKernelGGL(k_onesweep<8>, dim3(tiles), dim3(kThreads), 0, st, n, ki, vi, ko, vo, 8 * p,
                               ghist + p * 256, sp, tilectr + p, b.errflag);
        uint32_t *tk = ki, *tv = vi;
        ki = ko; vi = vo; ko = tk; vo = tv;
    }
    return (passes & 1) ? 1 : 0;
}

}  // namespace smdev
