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
//
// Window sort (sort_mode 2, one replica): the key kernel stores each key whose cell lies in
// the sort's window -- the occupied cell range of the last build that had to measure it,
// widened by one cell, at most 2^18 cells -- as its index in the window's (z, y, x) order,
// which orders keys as the cell index does, and counts its two window digits.  Two launches
// then sort 9 + 9 (or 9 + 8) bits.  If a key fell outside the window (or there is none),
// the same launches sort the cell indices in three 8-bit passes and measure the occupied
// range for the next window.  A vesicle of 240k beads in a 200^3-cell grid has 23-bit keys
// but a 17-bit window: 31.8 against 36.6 us per sort.
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
#pragma unroll
        for (int r = 0; r < kRounds; r++)
#pragma unroll
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

        // the digit's counts read together, then its output positions written
        uint32_t cnt[kRounds][kWavesPerBlock];
#pragma unroll
        for (int r = 0; r < kRounds; r++)
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; w++) cnt[r][w] = wc[r][w][t];
        uint32_t pos = dbase + prefix;
#pragma unroll
        for (int r = 0; r < kRounds; r++)
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; w++) {
                wc[r][w][t] = pos;
                pos += cnt[r][w];
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


// ---------------------------------------------------------------------------------------
// The window sort (sort_mode 2) in two launches.  The window path runs one 9-bit pass in
// each over the window keys.  The fallback sorts the cell indices in three 8-bit passes:
// the first launch converts any window keys back to cell indices and counts their digits
// (when the sort had a window), then runs two passes with a grid-wide barrier before each
// (a launch has at most kWinMaxTiles workgroups, so all of them are resident together); the
// second launch runs the third.  Each launch reads which path to take from the device.
// ---------------------------------------------------------------------------------------
constexpr int kWDig = 512;

template <int kRounds>
struct WLds {
    // per (round, wave, digit): count, then tile-local offset; a thread owning two adjacent
    // digits reads and writes both as one word
    union {
        uint16_t h[kRounds][kWavesPerBlock][kWDig];
        uint32_t w[kRounds][kWavesPerBlock][kWDig / 2];
    } wc;
    uint32_t dpos[kWDig];                          // per digit: output position of the tile's first item
    uint32_t h[3][256];                            // fallback: digit counts of a tile's cell indices
    uint32_t bb[kWavesPerBlock][6];                // fallback: occupied cell range of each wave's keys
    uint32_t scan[kWavesPerBlock];
    uint32_t tile;
};

template <int kBits>
__device__ inline unsigned long long match_bits(uint32_t d, bool valid)
{
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kBits; b++) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    return peers;
}

// Sum of the counts of digit d over the tiles before `tile` (decoupled look-back over
// windows of kLook tiles, bounded like the every-key sort's).
__device__ inline uint32_t look_back(const uint32_t *status, int tile, int d, int *errflag)
{
    uint32_t prefix = 0;
    int spins = 0;
    for (int pt = tile - 1; pt >= 0;) {
        uint32_t v[kLook];
#pragma unroll
        for (int w = 0; w < kLook; w++)
            v[w] = pt - w >= 0 ? __hip_atomic_load(status + (size_t)(pt - w) * kWDig + d, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                               : kFlagIncl;
        int used = 0;
        bool done = false;
#pragma unroll
        for (int w = 0; w < kLook; w++) {
            if (done || used < w) continue;
            const uint32_t f = v[w] & ~kCountMask;
            if (f == 0u) continue;
            prefix += v[w] & kCountMask;
            used = w + 1;
            done = f == kFlagIncl;
        }
        if (done) break;
        pt -= used;
        if (used < kLook) {
            if (++spins >= kSpinLimit) {
                atomicOr(errflag, FLAG_SORT_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return prefix;
}

// Sum of the counts of digits d0 and d1 over the tiles before `tile`: the two look-backs of
// a thread walk together, so the loads of both are in flight at once.
__device__ inline void look_back2(const uint32_t *status, int tile, int d0, int d1, uint32_t &p0, uint32_t &p1,
                                  int *errflag)
{
    p0 = 0u;
    p1 = 0u;
    int pt0 = tile - 1, pt1 = tile - 1, spins = 0;
    while (pt0 >= 0 || pt1 >= 0) {
        uint32_t v0[kLook], v1[kLook];
#pragma unroll
        for (int w = 0; w < kLook; w++) {
            v0[w] = pt0 - w >= 0 ? __hip_atomic_load(status + (size_t)(pt0 - w) * kWDig + d0, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                                 : kFlagIncl;
            v1[w] = pt1 - w >= 0 ? __hip_atomic_load(status + (size_t)(pt1 - w) * kWDig + d1, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                                 : kFlagIncl;
        }
        bool wait = false;
        if (pt0 >= 0) {
            int used = 0;
            bool done = false;
#pragma unroll
            for (int w = 0; w < kLook; w++) {
                if (done || used < w) continue;
                const uint32_t f = v0[w] & ~kCountMask;
                if (f == 0u) continue;
                p0 += v0[w] & kCountMask;
                used = w + 1;
                done = f == kFlagIncl;
            }
            pt0 = done ? -1 : pt0 - used;
            wait = wait || (!done && used < kLook);
        }
        if (pt1 >= 0) {
            int used = 0;
            bool done = false;
#pragma unroll
            for (int w = 0; w < kLook; w++) {
                if (done || used < w) continue;
                const uint32_t f = v1[w] & ~kCountMask;
                if (f == 0u) continue;
                p1 += v1[w] & kCountMask;
                used = w + 1;
                done = f == kFlagIncl;
            }
            pt1 = done ? -1 : pt1 - used;
            wait = wait || (!done && used < kLook);
        }
        if (wait) {
            if (++spins >= kSpinLimit) {
                atomicOr(errflag, FLAG_SORT_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// One stable pass over the kBits-bit digit at `shift` of tile `tile`, whose items take
// digit bases from the per-thread histogram sums g (the thread's kPer digits, over every
// copy).  kIota: the first pass, whose values are the keys' positions (bead indices), not
// loaded from vin.
template <int kRounds, int kBits, bool kIota = false>
__device__ void wpass(WLds<kRounds> &L, int n, int shift, int tile, const uint32_t (&g)[(1 << kBits) / kThreads],
                      const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout,
                      uint32_t *__restrict__ vout, uint32_t *__restrict__ status, int *__restrict__ errflag)
{
    constexpr int kDig = 1 << kBits, kPer = kDig / kThreads;   // thread t owns digits t*kPer .. t*kPer+kPer-1
    const int t = threadIdx.x, wave = t / kWave, lane = t & (kWave - 1);
    const int base = tile * kThreads * kRounds;
    // every round's loads in flight before the LDS work
    uint32_t kr[kRounds], vr[kRounds], dr[kRounds], rk[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const int i = min(base + r * kThreads + t, n - 1);
        kr[r] = kin[i];
        vr[r] = kIota ? (uint32_t)i : vin[i];
    }
    uint32_t gs = 0u;
#pragma unroll
    for (int j = 0; j < kPer; j++) gs += g[j];
    uint32_t x = gs;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == kWave - 1) L.scan[wave] = x;
#pragma unroll
    for (int r = 0; r < kRounds; r++)
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; w++) {
            if (kPer == 2)
                L.wc.w[r][w][t] = 0u;
            else
                L.wc.h[r][w][t] = 0;
        }
    __syncthreads();
    uint32_t wpre = 0;
    for (int w = 0; w < wave; w++) wpre += L.scan[w];
    uint32_t dbase[kPer];
    dbase[0] = wpre + x - gs;
#pragma unroll
    for (int j = 1; j < kPer; j++) dbase[j] = dbase[j - 1] + g[j - 1];
    const unsigned long long lt = lanes_below();
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        const bool valid = base + r * kThreads + t < n;
        dr[r] = valid ? (kr[r] >> shift) & (uint32_t)(kDig - 1) : (uint32_t)kDig;   // kDig: no item
        const unsigned long long peers = match_bits<kBits>(dr[r] & (uint32_t)(kDig - 1), valid);
        rk[r] = (uint32_t)__popcll(peers & lt);
        if (valid && rk[r] == 0u) L.wc.h[r][wave][dr[r]] = (uint16_t)__popcll(peers);
    }
    __syncthreads();

    // per digit: the counts read together, tile-local offsets written, the tile's count
    // published for later tiles
    uint32_t mine[kPer];
    {
        uint32_t cnt[kRounds][kWavesPerBlock];
#pragma unroll
        for (int r = 0; r < kRounds; r++)
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; w++) cnt[r][w] = kPer == 2 ? L.wc.w[r][w][t] : L.wc.h[r][w][t];
        uint32_t off0 = 0, off1 = 0;
#pragma unroll
        for (int r = 0; r < kRounds; r++)
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; w++) {
                if (kPer == 2) {
                    L.wc.w[r][w][t] = off0 | (off1 << 16);
                    off0 += cnt[r][w] & 0xFFFFu;
                    off1 += cnt[r][w] >> 16;
                } else {
                    L.wc.h[r][w][t] = (uint16_t)off0;
                    off0 += cnt[r][w];
                }
            }
        mine[0] = off0;
        if (kPer == 2) mine[kPer - 1] = off1;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++)
        __hip_atomic_store(status + (size_t)tile * kWDig + t * kPer + j, (tile == 0 ? kFlagIncl : kFlagAgg) | mine[j],
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t prefix[kPer];
    if (kPer == 2)
        look_back2(status, tile, t * kPer, t * kPer + kPer - 1, prefix[0], prefix[kPer - 1], errflag);
    else
        prefix[0] = look_back(status, tile, t, errflag);
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int d = t * kPer + j;
        if (tile > 0)
            __hip_atomic_store(status + (size_t)tile * kWDig + d, kFlagIncl | (prefix[j] + mine[j]),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        L.dpos[d] = dbase[j] + prefix[j];
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < kRounds; r++) {
        if (dr[r] < (uint32_t)kDig) {
            uint32_t p = L.dpos[dr[r]] + L.wc.h[r][wave][dr[r]] + rk[r];
            p = p < (uint32_t)n ? p : (uint32_t)(n - 1);   // only reachable after a timeout
            kout[p] = kr[r];
            vout[p] = vr[r];
        }
    }
}

// A tile id from a counter, issued from vector registers (for a uniform address the
// compiler would aggregate the atomic over the wave and wait for its result at once) and
// shared through LDS at the next barrier.
__device__ inline uint32_t draw_tile(uint32_t *tilectr)
{
    typedef __attribute__((address_space(1))) uint32_t global_u32;
    global_u32 *ctr = (global_u32 *)tilectr;
    asm volatile("" : "+v"(ctr));
    uint32_t id = 0u;
    if (threadIdx.x == 0) id = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return id;
}

template <int kPer>
__device__ inline void hist_sums(const uint32_t *__restrict__ ghist, uint32_t (&g)[kPer])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        g[j] = 0u;
#pragma unroll
        for (int c = 0; c < kHistCopies; c++) g[j] += ghist[(size_t)c * 1024 + t * kPer + j];
    }
}

// Every workgroup of the launch arrives at *ctr before any goes on; its stores before are
// visible to every workgroup's loads after.  Bounded like the look-back.
__device__ inline void grid_sync(uint32_t *ctr, uint32_t count, int *errflag)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int spins = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < count) {
            if (++spins >= kSpinLimit) {
                atomicOr(errflag, FLAG_SORT_TIMEOUT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    __threadfence();
}

struct WArgs {
    uint32_t *kA, *vA, *kB, *vB, *kC, *vC;
    uint32_t *ghist;        // the cell indices' digit histograms [copies][4][256]
    const uint32_t *ghw;    // the window digits' [copies][2][512]
    const uint32_t *desc;   // this sort's window
    const uint32_t *flag;   // set: a key outside the window
    uint32_t *bbox;         // the fallback's occupied cell range, [copies][8]
    uint32_t *ctr;          // 3 tile counters: each launch's first pass, the fallback's second
    uint32_t *bar;          // 2 barrier counters
    uint32_t *status;       // look-back words of the same three passes, tiles x 512 each
    int *errflag;
    const DevRep *reps;     // the cell grid: replica 0's
    int tiles;
};

template <int kRounds>
__global__ __launch_bounds__(kThreads) void k_sortw(int n, int launch, WArgs a)
{
    __shared__ WLds<kRounds> L;
    const int t = threadIdx.x;
    // the launch's first pass draws from counter `launch` and uses look-back words `launch`
    // on either path, so its tile id and the window path's histogram sums are requested
    // before the path is known
    const uint32_t tile0 = draw_tile(a.ctr + launch);
    uint32_t gw[2], gw8[1] = {0u};
    hist_sums<2>(a.ghw + 512 * launch, gw);
    if (launch != 0) hist_sums<1>(a.ghw + 512, gw8);   // the second digit read as 8 bits
    const WinDesc w = load_win(a.desc);
    const bool fallback = w.wx == 0u || *a.flag != 0u;
    uint32_t *st0 = a.status, *st1 = st0 + (size_t)a.tiles * kWDig, *st2 = st1 + (size_t)a.tiles * kWDig;
    if (t == 0) L.tile = tile0;
    __syncthreads();
    const int tile = (int)L.tile;
    if (!fallback) {
        if (launch == 0) {
            wpass<kRounds, 9, true>(L, n, 0, tile, gw, a.kA, a.vA, a.kB, a.vB, st0, a.errflag);
        } else if (w.wx * w.wy * w.wz <= (1u << 17)) {   // window keys of 17 bits: an 8-bit second digit
            wpass<kRounds, 8>(L, n, 9, tile, gw8, a.kB, a.vB, a.kA, a.vA, st1, a.errflag);
        } else {
            wpass<kRounds, 9>(L, n, 9, tile, gw, a.kB, a.vB, a.kA, a.vA, st1, a.errflag);
        }
        return;
    }
    uint32_t g8[1];
    if (launch != 0) {
        hist_sums<1>(a.ghist + 512, g8);
        wpass<kRounds, 8>(L, n, 16, tile, g8, a.kC, a.vC, a.kA, a.vA, st1, a.errflag);
        return;
    }
    // every key back to its cell index (window keys), its three digits counted (the key
    // kernel counted none), and the occupied cell range measured for the next window
    const uint32_t nx = (uint32_t)a.reps[0].nc[0], ny = (uint32_t)a.reps[0].nc[1];
    const int lane = t & (kWave - 1), wave = t / kWave;
    for (int p = 0; p < 3; p++) L.h[p][t] = 0u;
    __syncthreads();
    uint32_t m[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    const int base = blockIdx.x * kThreads * kRounds;
    for (int r = 0; r < kRounds; r++) {
        const int i = base + r * kThreads + t;
        if (i < n) {
            uint32_t k = a.kA[i];
            if (k & kWinTag) {
                const uint32_t wk = k & ~kWinTag;
                const uint32_t q = wk / w.wx, x = wk - q * w.wx, z = q / w.wy, y = q - z * w.wy;
                k = ((z + w.z0) * ny + (y + w.y0)) * nx + (x + w.x0);
                a.kA[i] = k;
            }
            for (int p = 0; p < 3; p++) atomicAdd(&L.h[p][(k >> (8 * p)) & 255u], 1u);
            const uint32_t q = k / nx, cx = k - q * nx, cz = q / ny, cy = q - cz * ny;
            m[0] = max(m[0], cx + 1u);
            m[1] = max(m[1], cy + 1u);
            m[2] = max(m[2], cz + 1u);
            m[3] = max(m[3], kBoxEnc - cx);
            m[4] = max(m[4], kBoxEnc - cy);
            m[5] = max(m[5], kBoxEnc - cz);
        }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1)
#pragma unroll
        for (int j = 0; j < 6; j++) m[j] = max(m[j], (uint32_t)__shfl_xor((int)m[j], o));
    if (lane == 0)
        for (int j = 0; j < 6; j++) L.bb[wave][j] = m[j];
    __syncthreads();
    uint32_t *gh = a.ghist + (size_t)(blockIdx.x % kHistCopies) * 1024;
    for (int p = 0; p < 3; p++)
        if (L.h[p][t]) atomicAdd(&gh[p * 256 + t], L.h[p][t]);
    if (t < 6) {
        uint32_t v = 0u;
        for (int q = 0; q < kWavesPerBlock; q++) v = max(v, L.bb[q][t]);
        if (v) atomicMax(&a.bbox[(blockIdx.x % kHistCopies) * 8 + t], v);
    }
    grid_sync(a.bar, gridDim.x, a.errflag);
    hist_sums<1>(a.ghist, g8);
    wpass<kRounds, 8, true>(L, n, 0, tile, g8, a.kA, a.vA, a.kB, a.vB, st0, a.errflag);
    grid_sync(a.bar + 1, gridDim.x, a.errflag);
    const uint32_t tile1 = draw_tile(a.ctr + 2);
    hist_sums<1>(a.ghist + 256, g8);
    if (t == 0) L.tile = tile1;
    __syncthreads();
    wpass<kRounds, 8>(L, n, 8, (int)L.tile, g8, a.kB, a.vB, a.kC, a.vC, st2, a.errflag);
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

// ---- the window sort: host side ----
static int sortw_tiles(int n)
{
    const int tile = kThreads * sort_rounds(n);
    return (n + tile - 1) / tile;
}

bool sortw_fits(int n) { return n >= 1 && sortw_tiles(n) <= kWinMaxTiles && sort_rounds(n) != 12; }

size_t sortw_words(int n) { return kWinKeyWords + (size_t)3 * sortw_tiles(n) * kWDig; }

uint32_t *win_scratch(uint32_t *sort_scratch, int n) { return sort_scratch + sort_scratch_words(n, 4); }

WinBook win_book(uint32_t *sort_scratch, int n, const uint32_t *desc)
{
    uint32_t *ws = win_scratch(sort_scratch, n);
    return WinBook{desc, ws, ws + 16 * 1024, ws + 16 * 1024 + 16 * 8};
}

void sort_window(hipStream_t st, SortBuffers &b, uint32_t *k2, uint32_t *v2, int parity, const uint32_t *desc,
                 const DevRep *reps)
{
    const int n = b.n;
    const int tiles = sortw_tiles(n);
    b.tiles = tiles;
    const WinBook wb = win_book(b.scratch[parity], n, desc);
    uint32_t *misc = wb.flag;   // flag, 3 tile counters, 2 barrier counters
    WArgs a{b.k0, b.v0, b.k1, b.v1, k2, v2, b.scratch[parity], wb.ghw, desc, wb.flag, wb.bbox, misc + 1, misc + 4,
            win_scratch(b.scratch[parity], n) + kWinKeyWords, b.errflag, reps, tiles};
    for (int launch = 0; launch < 2; launch++) {
        if (sort_rounds(n) == 2)
            hipLaunchKernelGGL(k_sortw<2>, dim3(tiles), dim3(kThreads), 0, st, n, launch, a);
        else
            hipLaunchKernelGGL(k_sortw<8>, dim3(tiles), dim3(kThreads), 0, st, n, launch, a);
    }
}

}  // namespace smdev
