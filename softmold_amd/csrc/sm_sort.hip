// Stable LSD radix sort of (cell key, particle index) pairs for the cell-list build.
//
// The reference sorts with thrust::sort_by_key (include/cuda/cell.h:254-260) and its CPU
// path keeps unsorted linked lists (include/algorithms/cellOpt.h:572-585).  Here: 8-bit
// digits, ceil(bits/8) passes, one kernel per pass ("onesweep"): a workgroup takes the
// next tile id from a counter (so every lower tile is already running), ranks its items
// with wave64 ballots, publishes its per-digit counts, accumulates the counts of the
// tiles before it by decoupled look-back, and scatters.
//
// No kernel of its own clears or builds histograms: the digit histograms of every pass are
// built by the kernel that computes the keys (k_first), and the sort's scratch (histograms, tile
// counters, look-back states) is double-buffered -- while one sort uses scratch[p], the
// reorder kernel after it clears scratch[1-p] for the next sort.
//
// Counting is wave-aggregated: the lanes holding one digit are found with 8 ballots and
// their lowest lane writes the count once, so a digit shared by a whole wave (the high
// digits of a compact system) costs one LDS write instead of 64 contended atomics.
// The look-back spin is bounded: on timeout the kernel records FLAG_SORT_TIMEOUT in the
// caller's flag word, clamps every write in bounds, and the host reports an error.
//
// Incremental re-sort (sort_moved): between two steps only the beads that change cell
// change key.  The previous sorted order, with those beads taken out, is still sorted by
// (key, index); the movers, compacted in index order and sorted by key with the same
// stable passes, are sorted by (key, index) too; merging the two by (key, index) gives the
// unique order that sorting all keys stably from index order gives.  k_count_moves and
// k_compact list the movers and count, per slot of the previous order, the movers before
// it; k_merge places every stay by a binary search among the sorted movers, and every
// mover by a search among the previous slots (inside its cell's previous run when that
// cell was occupied).
#include <algorithm>

#include "sm_sort_dev.hpp"

namespace smdev {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr uint32_t kFlagAgg = 1u << 30, kFlagIncl = 2u << 30, kCountMask = (1u << 30) - 1;
constexpr int kSpinLimit = 1 << 20;
constexpr int kLook = 8;
constexpr uint32_t kMovedBit = 1u << 31;

// A tile is kRounds x 256 keys.  Small inputs want short tiles (more of them to spread over
// the chip), large inputs long ones (fewer tiles to look back over): sort_rounds() picks.
// ndev != NULL: the key count is read from the device (the movers of an incremental
// re-sort), and each workgroup takes tiles one after another until they run out, so a
// small grid serves any count.  Otherwise each workgroup takes one tile.
template <int kRounds>
__global__ __launch_bounds__(kThreads) void k_onesweep(int n, const int *__restrict__ ndev, const uint32_t *__restrict__ kin,
                                                       const uint32_t *__restrict__ vin,
                                                       uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                       int shift, const uint32_t *__restrict__ ghist,
                                                       uint32_t *__restrict__ status, uint32_t *__restrict__ tilectr,
                                                       int *__restrict__ errflag)
{
    __shared__ uint32_t s_tile;
    __shared__ uint32_t scan[kWavesPerBlock];
    // per (round, wave, digit): item count, then its exclusive output position
    __shared__ uint32_t wc[kRounds][kWavesPerBlock][256];
    const int t = threadIdx.x;
    const int wave = t / kWave;
    const bool many = ndev != nullptr;
    if (many) n = *ndev;
    // the tile id reaches LDS only after the histogram loads below are issued, so the
    // atomic's round trip overlaps theirs
    uint32_t first_tile = 0u;
    {
        // the counter's address held in vector registers: for a uniform address the compiler
        // would aggregate the atomic over the wave and wait for its result at once
        typedef __attribute__((address_space(1))) uint32_t global_u32;
        global_u32 *ctr = (global_u32 *)tilectr;
        asm volatile("" : "+v"(ctr));
        if (t == 0) first_tile = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // digit base: keys of smaller digits over all keys, from every histogram copy; a scan
    // inside each wave by shuffles, then the waves' totals (one barrier instead of the 16 of
    // an LDS scan: 37.2 against 38.3 us per 240k-key sort)
    uint32_t g = 0;
#pragma unroll
    for (int c = 0; c < kHistCopies; c++) g += ghist[(size_t)c * 4 * 256 + t];
    const int lane = t & (kWave - 1);
    uint32_t x = g;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) scan[wave] = x;
    if (t == 0) s_tile = first_tile;
    __syncthreads();
    uint32_t wpre = 0;
    for (int w = 0; w < wave; w++) wpre += scan[w];
    const uint32_t dbase = wpre + x - g;
    const unsigned long long lt = lanes_below();
    for (;;) {
        for (int r = 0; r < kRounds; r++)
            for (int w = 0; w < kWavesPerBlock; w++) wc[r][w][t] = 0u;
        __syncthreads();
        const int tile = (int)s_tile;
        const int base = tile * kThreads * kRounds;
        if (base >= n) return;   // past the keys (workgroup-uniform; no tile after it waits on it)

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

        uint32_t pos = dbase + prefix;
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
        if (!many) return;
        __syncthreads();   // every lane is done with this tile's LDS
        if (t == 0) s_tile = atomicAdd(tilectr, 1u);
    }
}

// Movers of an incremental re-sort, in two sequences cut into tiles of kMoveTile: bead
// index order (the mover list) and the previous sorted order (per slot, the movers before
// it).  k_count_moves takes one tile of each: the index tile's mover count from the key
// kernel's per-block counts, and, per slot of the previous order, whether its bead moved
// (one 64-bit ballot per 64 slots) and the slot tile's count.  kold == NULL: every bead
// moves (no previous order) and the slot side is skipped.
__global__ __launch_bounds__(kThreads) void k_count_moves(int n, int nblocks, const uint32_t *__restrict__ knew,
                                                          const uint32_t *__restrict__ kold,
                                                          const uint32_t *__restrict__ skey_old,
                                                          const uint32_t *__restrict__ sidx_old,
                                                          const uint32_t *__restrict__ cnt_b,
                                                          uint32_t *__restrict__ cnt_i, uint32_t *__restrict__ cnt_s,
                                                          unsigned long long *__restrict__ mflags)
{
    constexpr int kR = kMoveTile / kThreads;
    constexpr int kBlocksPerTile = kMoveTile / 256;
    static_assert(kBlocksPerTile <= kWave, "one wave sums a tile's block counts");
    __shared__ uint32_t s_cnt[kWavesPerBlock];
    const int tile = blockIdx.x, t = threadIdx.x, wave = t / kWave, lane = t & (kWave - 1);
    if (wave == 0) {
        const int b = tile * kBlocksPerTile + lane;
        uint32_t c = (lane < kBlocksPerTile && b < nblocks) ? cnt_b[b] : 0u;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) c += __shfl_xor(c, o);
        if (lane == 0) cnt_i[tile] = c;
    }
    if (kold == nullptr) return;
    const int base = tile * kMoveTile;
    uint32_t sk[kR], si[kR], kn[kR];
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int e = min(base + r * kThreads + t, n - 1);
        sk[r] = skey_old[e];
        si[r] = sidx_old[e];
    }
#pragma unroll
    for (int r = 0; r < kR; r++) kn[r] = knew[si[r]];
    uint32_t mine = 0;
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int e = base + r * kThreads + t;
        const unsigned long long b = __ballot(e < n && kn[r] != sk[r]);
        if (lane == 0) {
            if (e < n) mflags[e >> 6] = b;
            mine += (uint32_t)__popcll(b);
        }
    }
    if (lane == 0) s_cnt[wave] = mine;
    __syncthreads();
    if (t == 0) {
        uint32_t c = 0;
        for (int w = 0; w < kWavesPerBlock; w++) c += s_cnt[w];
        cnt_s[tile] = c;
    }
}

// The mover list (tiles 0 .. ntiles-1: movers of an index tile, in index order, after the
// movers of the tiles before) and the movers-before words of the previous slots (tiles
// ntiles .. 2 ntiles - 1), from the counts and ballots of k_count_moves.
__global__ __launch_bounds__(kThreads) void k_compact(int n, int ntiles, const uint32_t *__restrict__ knew,
                                                      const uint32_t *__restrict__ kold,
                                                      const uint32_t *__restrict__ cnt_i,
                                                      const uint32_t *__restrict__ cnt_s,
                                                      const unsigned long long *__restrict__ mflags,
                                                      uint32_t *__restrict__ mkey, uint32_t *__restrict__ midx,
                                                      uint32_t *__restrict__ mbefore, int *__restrict__ mcount)
{
    constexpr int kR = kMoveTile / kThreads;
    constexpr int kWords = kMoveTile / kWave;
    __shared__ uint32_t s_red[2][kWavesPerBlock];
    __shared__ uint32_t s_word[kWords];
    const bool partS = (int)blockIdx.x >= ntiles;   // workgroup-uniform
    const int tile = partS ? (int)blockIdx.x - ntiles : (int)blockIdx.x;
    if (partS && kold == nullptr) return;
    const int t = threadIdx.x, wave = t / kWave, lane = t & (kWave - 1);
    const int base = tile * kMoveTile;
    const unsigned long long lt = lanes_below();
    // this tile's loads first, then the offset from the tiles before it
    uint32_t kn[kR], ko[kR];
    unsigned long long wbits = 0ull;
    if (partS) {
        if (t < kWords) wbits = (base + t * kWave < n) ? mflags[(base >> 6) + t] : 0ull;
    } else {
#pragma unroll
        for (int r = 0; r < kR; r++) {
            const int e = min(base + r * kThreads + t, n - 1);
            kn[r] = knew[e];
            ko[r] = kold ? kold[e] : ~kn[r];
        }
    }
    const uint32_t *cnt = partS ? cnt_s : cnt_i;
    uint32_t pre = 0, all = 0;
    for (int u = t; u < ntiles; u += kThreads) {
        const uint32_t c = cnt[u];
        if (u < tile) pre += c;
        all += c;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        pre += __shfl_xor(pre, o);
        all += __shfl_xor(all, o);
    }
    if (lane == 0) {
        s_red[0][wave] = pre;
        s_red[1][wave] = all;
    }
    if (partS && t < kWords) s_word[t] = (uint32_t)__popcll(wbits);
    __syncthreads();
    pre = 0;
    all = 0;
    for (int w = 0; w < kWavesPerBlock; w++) {
        pre += s_red[0][w];
        all += s_red[1][w];
    }
    if (!partS && tile == 0 && t == 0) *mcount = (int)all;
    if (partS) {
        // exclusive scan of the tile's per-word mover counts; slot r * kThreads + t of the
        // tile is lane `lane` of word r * kWavesPerBlock + wave
        static_assert(kWords <= kThreads && kThreads == kWavesPerBlock * kWave, "one thread per ballot word");
        __shared__ unsigned long long s_bits[kWords];
        if (t < kWords) s_bits[t] = wbits;
        for (int off = 1; off < kWords; off <<= 1) {
            const uint32_t v = (t >= off && t < kWords) ? s_word[t - off] : 0u;
            __syncthreads();
            if (t < kWords) s_word[t] += v;
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < kR; r++) {
            const int e = base + r * kThreads + t;
            const int w = r * kWavesPerBlock + wave;
            const unsigned long long b = s_bits[w];
            if (e < n) {
                const uint32_t before = pre + (s_word[w] - (uint32_t)__popcll(b)) + (uint32_t)__popcll(b & lt);
                mbefore[e] = before | (((b >> lane) & 1ull) ? kMovedBit : 0u);
            }
        }
        return;
    }
    __shared__ uint32_t s_wave[kR][kWavesPerBlock];
    unsigned long long bm[kR];
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int e = base + r * kThreads + t;
        bm[r] = __ballot(e < n && kn[r] != ko[r]);
        if (lane == 0) s_wave[r][wave] = (uint32_t)__popcll(bm[r]);
    }
    __syncthreads();
    uint32_t off = pre;
#pragma unroll
    for (int r = 0; r < kR; r++) {
        const int e = base + r * kThreads + t;
        uint32_t o = off;
        for (int w = 0; w < wave; w++) o += s_wave[r][w];
        if ((bm[r] >> lane) & 1ull) {
            o += (uint32_t)__popcll(bm[r] & lt);
            if (o < (uint32_t)n) {   // counts that disagree are caught by k_merge
                mkey[o] = kn[r];
                midx[o] = (uint32_t)e;
            }
        }
        for (int w = 0; w < kWavesPerBlock; w++) off += s_wave[r][w];
    }
}

// Number of entries (k[j], v[j]), j < len, below (key, idx) in (key, index) order; the
// entries are sorted in that order.
__device__ inline int rank_below(const uint32_t *__restrict__ k, const uint32_t *__restrict__ v, int len,
                                 uint32_t key, uint32_t idx)
{
    const unsigned long long x = ((unsigned long long)key << 32) | idx;
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const unsigned long long y = ((unsigned long long)k[mid] << 32) | v[mid];
        if (y < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// The merge of the stays (previous sorted order minus the movers) with the sorted movers.
// Slot s of the previous order, a stay, goes to (stays before it) + (movers below it);
// mover j goes to j + (stays below it), the stays below it being the previous slots below
// it minus the movers among them.  full: no previous order, the movers are every bead.
// A mover's rank among the previous slots is found inside its cell's run of the previous
// cell table when that cell was occupied (cstart/cend still hold it: the reorder after the
// merge rewrites them), else by a search over all of them.
__global__ __launch_bounds__(kThreads) void k_merge(int n, const int *__restrict__ mcount,
                                                    const uint32_t *__restrict__ mkey,
                                                    const uint32_t *__restrict__ midx,
                                                    const uint32_t *__restrict__ skey_old,
                                                    const uint32_t *__restrict__ sidx_old,
                                                    const uint32_t *__restrict__ mbefore,
                                                    const int *__restrict__ cstart, const int *__restrict__ cend,
                                                    int full, uint32_t *__restrict__ skey,
                                                    uint32_t *__restrict__ sidx, int *__restrict__ errflag)
{
    const int t = blockIdx.x * kThreads + threadIdx.x;
    const int m = *mcount;
    if (!(m >= 0 && m <= n && (full == 0 || m == n))) {   // counts that cannot be: flag, write nothing
        if (t == 0) atomicOr(errflag, FLAG_SORT_COUNT);
        return;
    }
    if (!full && t < n) {
        const uint32_t w = mbefore[t];
        if (!(w & kMovedBit)) {
            const uint32_t key = skey_old[t], idx = sidx_old[t];
            const int out = t - (int)w + rank_below(mkey, midx, m, key, idx);
            if ((unsigned)out < (unsigned)n) {
                skey[out] = key;
                sidx[out] = idx;
            } else {
                atomicOr(errflag, FLAG_SORT_COUNT);
            }
        }
    }
    if (t < m) {
        const uint32_t key = mkey[t], idx = midx[t];
        int below = 0;
        if (!full) {
            const int c0 = cstart[key];
            int p;
            if ((unsigned)c0 < (unsigned)n && skey_old[c0] == key) {
                const int c1 = min(cend[key], n);
                p = c0 + rank_below(skey_old + c0, sidx_old + c0, max(c1 - c0, 0), key, idx);
            } else {
                p = rank_below(skey_old, sidx_old, n, key, idx);
            }
            below = p - (p < n ? (int)(mbefore[p] & ~kMovedBit) : m);
        }
        const int out = t + below;
        if ((unsigned)out < (unsigned)n) {
            skey[out] = key;
            sidx[out] = idx;
        } else {
            atomicOr(errflag, FLAG_SORT_COUNT);
        }
    }
}

}  // namespace

// rounds of 256 keys per tile, by key count (measured on 15k, 240k and 3M keys: 3M keys
// sort in 129 us with 3072-key tiles against 148 us with 2048; 240k keys in 37.7 us with
// 2048 against 38.7 with 3072)
int sort_rounds(int n) { return n < 65536 ? 2 : (n < 1000000 ? 8 : 12); }

// the incremental re-sort's mover sort: tiles sized for about one bead in 16 moving
static int sort_rounds_moved(int n) { return sort_rounds(n / 16 > 0 ? n / 16 : 1); }

int sort_passes(int bits)
{
    int passes = (bits + 7) / 8;
    return passes < 1 ? 1 : (passes > 4 ? 4 : passes);
}

int move_tiles(int n) { return (n + kMoveTile - 1) / kMoveTile; }

size_t sort_hist_words() { return (size_t)kHistCopies * 4 * 256; }

size_t sort_scratch_words(int n, int max_passes)
{
    const int rounds = sort_rounds(n) < sort_rounds_moved(n) ? sort_rounds(n) : sort_rounds_moved(n);
    const int tile = kThreads * rounds;
    const size_t tiles = (size_t)((n + tile - 1) / tile) + 1;
    // ghist [copies][4][256] + tile counters [4] + status [passes][tiles][256]
    return sort_hist_words() + 4 + (size_t)max_passes * tiles * 256;
}

// grid: one workgroup per tile, or `grid` workgroups taking tiles in turn (ndev != NULL)
static int onesweep_passes(hipStream_t st, SortBuffers &b, int bits, int parity, const int *ndev, int rounds, int grid)
{
    const int n = b.n;
    const int passes = sort_passes(bits);
    const int tiles = (n + kThreads * rounds - 1) / (kThreads * rounds);
    b.tiles = tiles;
    if (ndev == nullptr || grid > tiles) grid = tiles;
    uint32_t *ghist = b.scratch[parity];
    uint32_t *tilectr = ghist + sort_hist_words();
    uint32_t *status = tilectr + 4;
    uint32_t *ki = b.k0, *vi = b.v0, *ko = b.k1, *vo = b.v1;
    for (int p = 0; p < passes; p++) {
        uint32_t *sp = status + (size_t)p * tiles * 256;
        if (rounds == 12)
            hipLaunchKernelGGL(k_onesweep<12>, dim3(grid), dim3(kThreads), 0, st, n, ndev, ki, vi, ko, vo, 8 * p,
                               ghist + p * 256, sp, tilectr + p, b.errflag);
        else if (rounds == 2)
            hipLaunchKernelGGL(k_onesweep<2>, dim3(grid), dim3(kThreads), 0, st, n, ndev, ki, vi, ko, vo, 8 * p,
                               ghist + p * 256, sp, tilectr + p, b.errflag);
        else
            hipLaunchKernelGGL(k_onesweep<8>, dim3(grid), dim3(kThreads), 0, st, n, ndev, ki, vi, ko, vo, 8 * p,
                               ghist + p * 256, sp, tilectr + p, b.errflag);
        uint32_t *tk = ki, *tv = vi;
        ki = ko; vi = vo; ko = tk; vo = tv;
    }
    return (passes & 1) ? 1 : 0;
}

int radix_sort(hipStream_t st, SortBuffers &b, int bits, int parity)
{
    if (b.n <= 1) return 0;
    return onesweep_passes(st, b, bits, parity, nullptr, sort_rounds(b.n), 0);
}

void sort_moved(hipStream_t st, SortBuffers &b, const MoveSort &m, int bits, int parity)
{
    const int n = b.n;
    const int ntiles = move_tiles(n);
    hipLaunchKernelGGL(k_count_moves, dim3(ntiles), dim3(kThreads), 0, st, n, move_blocks(n), m.knew, m.kold,
                       m.skey_old, m.sidx_old, m.cnt_b, m.cnt_i, m.cnt_s, m.mflags);
    hipLaunchKernelGGL(k_compact, dim3(2 * ntiles), dim3(kThreads), 0, st, n, ntiles, m.knew, m.kold, m.cnt_i,
                       m.cnt_s, m.mflags, b.k0, b.v0, m.mbefore, m.mcount);
    // the movers' sort: tiles for about one bead in 16 moving, taken in turn by twice as
    // many workgroups as that many movers fill (each further workgroup costs one more
    // atomic on the tile counter)
    const int rounds = sort_rounds_moved(n);
    const int expect = (n / 16 + kThreads * rounds - 1) / (kThreads * rounds);
    const int grid = std::min(512, std::max(32, 2 * expect));
    const int sel = onesweep_passes(st, b, bits, parity, m.mcount, rounds, grid);
    const uint32_t *mk = sel ? b.k1 : b.k0, *mv = sel ? b.v1 : b.v0;
    hipLaunchKernelGGL(k_merge, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st, n, m.mcount, mk, mv,
                       m.skey_old, m.sidx_old, m.mbefore, m.cstart, m.cend, m.kold == nullptr ? 1 : 0, m.skey,
                       m.sidx, b.errflag);
}

int move_blocks(int n) { return (n + 255) / 256; }

}  // namespace smdev
