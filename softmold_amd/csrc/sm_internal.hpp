// Shared device data layout and kernel launchers of the MI355X MD-step solver.
//
// A context holds R >= 1 independent replica systems that share one set of interaction
// constants, cutoff and time step (a single system is R = 1).  Replica r owns the bead
// range [first_r, first_r + count_r) in molecule order; because the cell key of replica r
// is offset by cellbase_r, the same range is its slot range in sorted (cell) order.
//
// HBM layout (N = total beads):
//   beads   Bead[N]      FP64 position, type, replica index; molecule order       32 B/bead
//   vel     float4[N]    FP32 velocity                                             16 B/bead
//   acc     float4[N]    FP32 acceleration a (= force, unit mass)                  16 B/bead
//   fpair   float4[N]    FP32 pair force of the current step                        16 B/bead
//   key/idx uint32[N]x2  global cell key and particle index, sorted by key (ping-pong)
//   sbeads  Bead[N]      positions in sorted (cell) order                           32 B/bead
//   cstart/cend int32[C] dense cell table over C = sum_r capacity_r cells, never cleared:
//                        an entry is valid only when skey[cstart[k]] == k, so stale
//                        entries of emptied cells are rejected without a reset.
//   chain   int2[N]      per bead: first bead of its CHAIN (or -1), length | const id << 16
//   reps    DevRep[R]    box, cell grid, thermostat and MC state of each replica
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace smdev {

constexpr int kMaxTypes = 8;           // LDS constant table bound (6 * 8 * 8 floats)
constexpr int kWave = 64;
constexpr int kHistBins = 256;

struct __align__(32) Bead {
    double x, y, z;
    int type;
    int rep;      // replica index
};

// One replica system, device resident, so a captured step needs no host decision.
struct DevRep {
    double L[3];
    double cs[3];
    int nc[3];
    int first, count;          // bead range (molecule order == sorted order)
    int cfirst, ccount;        // chain range
    int sigma_fixed;           // 1: sigma does not follow temperature changes (gammaType)
    long long cellbase;        // first global cell key
    long long capacity;        // cells reserved for this replica
    unsigned long long seed;   // counter-RNG key
    double temperature, gamma, sigma, delta_lxy, tension;
    // MC trial state
    double scale[3];
    double newL[3];
    double fluct;
    double u_accept;
    double dU_pair, dU_chain;
    int accept;
    int pad1;
    long long trials, accepted;
    double hist_accept[kHistBins];
    double hist_reject[kHistBins];
};

// Generic BOND / BEND / BALL record as seen from one of its beads (global bead indices):
// kind 0 = bond (p1, p2), kind 1 = bend (p1, p2, p3), kind 2 = ball (centre p1, member p2).  Each bead lists the records it
// belongs to (CSR: gptr[i] .. gptr[i+1]); the record's two constants sit beside it.
struct __align__(16) GenEnt {
    int kind;
    int p1, p2, p3;
};

// Context-wide device state.
struct DevGlobal {
    int flags;
    int pad;
    long long step;            // step index of the step in progress (Philox counter)
};

enum : int { FLAG_BAD_POSITION = 1, FLAG_CELL_OVERFLOW = 2, FLAG_SORT_TIMEOUT = 4, FLAG_SORT_COUNT = 8 };

struct Params {
    int n;          // beads, all replicas
    int nrep;
    int nT;
    double cutoff;
    double dt;
    // the two-body tables' (a, b) and (b, a) rows differ somewhere: the pair kernels then
    // take each pair's row from the particle the reference's traversal makes first
    int asym;
};

// Counter-RNG stream tags (word 3 of the Philox counter).
enum : unsigned { TAG_LANGEVIN = 0, TAG_MC = 1, TAG_LANGEVIN_INIT = 2 };

// Deterministic segmented FP64 sums: a segment set is a list of ranges (one per replica)
// cut into tiles of kSegTile elements that start at the range start, so a replica's sum
// does not depend on the other replicas.
constexpr int kSegTile = 2048;
struct SegTiles {
    int2 *tile;        // (begin, end) per tile, device
    int *rep_first;    // first tile of each replica (R + 1 entries), device
    int ntiles;
    int nrep;
    double *tilesum;   // ntiles, device
};
// out[r * stride] = sum of x over replica r's range (add: out[r * stride] += that sum)
void launch_seg_sum(hipStream_t st, const SegTiles &S, const double *x, double *out, int stride, bool add = false);

// ---- launchers (sm_kernels.hip) ---------------------------------------------------------
// step_set >= 0: the device step counter is set to step_set; -1: it is advanced by one
// (the form a captured graph replays); <= -2: left as it is.  If ghist != NULL the digit
// histograms of the first `passes` 8-bit digits of the keys are added to it.
// The incremental re-sort's bookkeeping in the key kernels (cnt_b == NULL: none; the
// histograms then count every key): a bead moves when its key differs from kold (every
// bead when kold is NULL); the movers' keys go into the histograms and each block of 256
// beads writes its mover count to cnt_b[block].
struct MoveBook {
    const uint32_t *kold;
    uint32_t *cnt_b;
};
// The window sort (sort_mode 2: one replica, every key sorted, keys of 17 to 24 bits, at
// most kWinMaxTiles tiles).  Each sort has a window of cells: the previous sort's occupied
// cell range widened by one cell on each side, used when it holds at most 2^18 cells.  A
// key whose cell lies in it is stored as kWinTag | the cell's index in the window's own
// (z, y, x) order, which orders keys as the cell index does, and two 9-bit passes sort it.
// A key outside the window raises the sort's fallback flag; the same two launches then sort
// the cell indices in three 8-bit passes.  desc == NULL: no window sort.
constexpr uint32_t kWinTag = 1u << 31;
constexpr int kWinMaxTiles = 256;
struct WinBook {
    const uint32_t *desc;   // this sort's window {x0, y0, z0, wx, wy, wz}; wx == 0: none
    uint32_t *ghw;          // window digit histograms [16 copies][2 digits][512]
    uint32_t *bbox;         // [16 copies][8]: max of cx + 1, cy + 1, cz + 1, 2^20 - cx, 2^20 - cy, 2^20 - cz
    uint32_t *flag;         // nonzero: some key fell outside the window
};
// the window sort's scratch, after the sort_scratch_words(n, 4) words of the every-key sort:
// ghw, bbox, {flag, 3 tile counters, 2 barrier counters, pad}, then 3 passes x tiles x 512
// look-back words; the key kernel writes the first kWinKeyWords of it
constexpr size_t kWinKeyWords = 16 * 1024 + 16 * 8 + 8;
size_t sortw_words(int n);
uint32_t *win_scratch(uint32_t *sort_scratch, int n);   // where it starts in a sort's scratch
WinBook win_book(uint32_t *sort_scratch, int n, const uint32_t *desc);
bool sortw_fits(int n);   // the tiles of n keys are few enough for the window sort's launches
void launch_first(hipStream_t st, const Params &P, Bead *beads, float4 *vel, const float4 *acc, const DevRep *reps,
                  DevGlobal *glob, uint32_t *key, uint32_t *idx, int integrate, long long step_set, uint32_t *ghist,
                  int passes, const MoveBook &book, const WinBook &wb = WinBook{nullptr, nullptr, nullptr, nullptr});
// also zeroes clear[0 .. clear_words) (the next sort's scratch)
void launch_reorder(hipStream_t st, int n, const uint32_t *skey, const uint32_t *sidx, const Bead *beads,
                    Bead *sbeads, int *cstart, int *cend, uint32_t *clear, size_t clear_words);
// The window sort's reorder: skey holds the sorted keys in either form, skey_full receives
// the cell indices; it also derives the next sort's window (desc_next) from this sort's
// occupied range.  The cell grid is replica 0's, read on the device.
void launch_reorder_win(hipStream_t st, int n, const uint32_t *skey, const uint32_t *sidx, const Bead *beads,
                        Bead *sbeads, int *cstart, int *cend, uint32_t *clear, size_t clear_words,
                        const WinBook &wb, uint32_t *skey_full, uint32_t *desc_next, const DevRep *reps);
// variant 0: one phase (FP64 test, force law under the branch) over per-quarter-cell culled
// candidate lists; 1: FP32 screen into a per-lane hit list, then FP64 test + force law on
// the hits
void launch_pair_force(hipStream_t st, const Params &P, int chunk, int variant, const Bead *sbeads,
                       const uint32_t *skey, const uint32_t *sidx, const int *cstart, const int *cend,
                       const DevRep *reps, const float *tab, float4 *fpair);
// The one-phase pair force of a box-move step, fused with the FP64 dU gather: fpair as
// launch_pair_force writes it, and energy[] per sorted slot as GATHER_DU does, under each
// replica's trial scale (set by launch_mc_propose before).
void launch_pair_force_du(hipStream_t st, const Params &P, int chunk, const Bead *sbeads, const uint32_t *skey,
                          const uint32_t *sidx, const int *cstart, const int *cend, const DevRep *reps,
                          const float *fconst, const double *uconst, float4 *fpair, double *energy);
// gptr == NULL: CHAIN terms only; else also the generic records of each bead
struct GenTables {
    const int *gptr;
    const GenEnt *gent;
    const double2 *gconst;
};
void launch_bonded_kick(hipStream_t st, const Params &P, const Bead *beads, const int2 *chain,
                        const double4 *cconst, const GenTables &gen, const float4 *fpair, float4 *vel, float4 *acc,
                        const DevRep *reps, const DevGlobal *glob, int mode, float4 *fchain_out);
enum : int { KICK_STEP = 0, KICK_INIT = 1, KICK_NONE = 2 };
// bonded + Langevin + second half kick of `step`, then Verlet::first, keys and digit
// histograms of step + 1 (positions written to beads_out; the device step becomes step + 1)
void launch_kick_first(hipStream_t st, const Params &P, const Bead *beads, Bead *beads_out, const int2 *chain,
                       const double4 *cconst, const GenTables &gen, const float4 *fpair, float4 *vel, float4 *acc,
                       const DevRep *reps, DevGlobal *glob, long long step, uint32_t *key, uint32_t *idx,
                       uint32_t *ghist, int passes, const MoveBook &book,
                       const WinBook &wb = WinBook{nullptr, nullptr, nullptr, nullptr});
void launch_second(hipStream_t st, int n, const Bead *beads, const float4 *acc, float4 *vel, float dt);

// FP64 gathers over the 27-cell stencil.  GATHER_U / GATHER_DU write one FP64 value per
// sorted slot into energy[] (the slot's share of U, or of U(old) - U(scaled) with each
// replica's trial scale); GATHER_COUNT writes neighbour counts in molecule order;
// GATHER_LIST appends pairs; GATHER_FCOUNT runs the one-phase force kernel's traversal
// (its staging, culling and FP64 test) and writes, in molecule order, the count of its
// accepted pairs whose force term is not identically zero (uconst = the FP64 force rows).
enum : int { GATHER_U = 0, GATHER_DU = 1, GATHER_COUNT = 2, GATHER_LIST = 3, GATHER_FCOUNT = 4 };
int cells_blocks(int n, int chunk);       // workgroups of the cell traversals (sm_pair.hip)
void launch_pair_gather(hipStream_t st, const Params &P, int chunk, int mode, const Bead *sbeads,
                        const uint32_t *skey, const uint32_t *sidx, const int *cstart, const int *cend,
                        const DevRep *reps, const float *tab, const double *uconst, double *energy, int *counts,
                        int *pairs, unsigned long long *npairs, long long cap);
// One FP64 value per chain: its energy, or U(old) - U(scaled) when delta != 0.
void launch_chain_energy(hipStream_t st, int nchains, const int2 *chains, const double4 *cconst, const Bead *beads,
                         const DevRep *reps, int delta, double *e);
// Generic-record energies, one FP64 value per bead: the records whose first bead it is,
// U, or U(old) - U(scaled) when delta != 0.
void launch_gen_energy(hipStream_t st, int n, const GenTables &gen, const Bead *beads, const DevRep *reps,
                       int delta, double *e);
// Generic-record statistics per bead (what 0: bond length, 1: bend cos theta, 2 and 3: the
// bend's arm lengths), over the records whose first bead it is.
void launch_gen_stats(hipStream_t st, int n, const GenTables &gen, const Bead *beads, const DevRep *reps, int what,
                      double *e);
// One FP64 kinetic energy per bead.
void launch_kinetic(hipStream_t st, int n, const float4 *vel, double *e);
// Vesicle-radius terms per bead (what 0..2: imaged head coordinate, 3: head count, 4: head
// distance to com[3r..]); sums[4r..4r+3] -> com[3r..3r+2].
void launch_head_terms(hipStream_t st, int n, const Bead *beads, const DevRep *reps, int head, int what,
                       const double *com, double *e);
void launch_head_com(hipStream_t st, int nrep, const double *sums, double *com);
void launch_mc_propose(hipStream_t st, int nrep, DevRep *reps, const DevGlobal *glob);
// du[r*2 + 0] = dU_pair, du[r*2 + 1] = dU_chain of replica r
void launch_mc_decide(hipStream_t st, const Params &P, DevRep *reps, DevGlobal *glob, const double *du);
void launch_rescale(hipStream_t st, int n, Bead *beads, const DevRep *reps);
// fill kernel (graph-capturable replacement for hipMemsetAsync)
void launch_fill_u32(hipStream_t st, uint32_t *p, size_t count, uint32_t value);
void launch_set_step(hipStream_t st, DevGlobal *glob, long long step);

// ---- fixed scratch of the entry points that allocate nothing (sm_noalloc.hip) ----------
constexpr int kNoAllocMaxReplicas = 1024;
constexpr size_t kNoAllocScratchDoubles = 8 * (size_t)kNoAllocMaxReplicas;   // 64 KiB per GPU
// Looks up `device`'s copy of the scratch array once (that device must be current).
hipError_t noalloc_scratch_init(int device);
// The device's scratch address, or nullptr before noalloc_scratch_init(device).
double *noalloc_scratch(int device);
// radius[r] = dist[r] / sums[4r + 3], all device pointers
void launch_head_radius(hipStream_t st, int nrep, const double *sums, const double *dist, double *radius);

// ---- radix sort (sm_sort.hip) ---------------------------------------------------------
struct SortBuffers {
    uint32_t *k0, *v0, *k1, *v1;
    uint32_t *scratch[2];   // sort_scratch_words(n, 4) words each, double-buffered
    int *errflag;           // receives FLAG_SORT_TIMEOUT / FLAG_SORT_COUNT
    int n;
    int tiles;
};
size_t sort_scratch_words(int n, int max_passes);
size_t sort_hist_words();   // the digit histograms at the start of the scratch
// the incremental re-sort's tiles: kMoveTile beads (or previous slots) each; the key
// kernel's blocks of 256 beads write one mover count each (move_blocks of them)
constexpr int kMoveTile = 4096;
int move_tiles(int n);
int move_blocks(int n);
int sort_passes(int bits);
// Stable LSD sort of (k0,v0) by the low `bits` bits using scratch[parity], whose digit
// histograms k_first has filled and whose other words are zero; returns which buffer pair
// holds the result (0 -> k0/v0, 1 -> k1/v1).
int radix_sort(hipStream_t st, SortBuffers &b, int bits, int parity);
// Incremental re-sort: the new sorted order from the previous one and the beads whose key
// changed (whose digit histograms and per-tile counts the key kernel has filled into
// scratch[parity]); (k0,v0) and (k1,v1) hold the movers.  The result is the one radix_sort
// gives for the same keys.
struct MoveSort {
    const uint32_t *knew;                    // keys by bead index, this build
    const uint32_t *kold;                    // keys by bead index at the previous sort, or NULL: all move
    const uint32_t *skey_old, *sidx_old;     // the previous sorted order (unused when kold is NULL)
    const int *cstart, *cend;                // the previous dense cell table
    const uint32_t *cnt_b;                   // movers per key-kernel block (move_blocks)
    uint32_t *cnt_i, *cnt_s;                 // movers per index tile and per previous-slot tile (move_tiles)
    unsigned long long *mflags;              // per 64 previous slots: which beads moved
    uint32_t *mbefore;                       // per previous slot: movers before it, bit 31 = it moves
    int *mcount;                             // the mover count
    uint32_t *skey, *sidx;                   // the new sorted order
};
void sort_moved(hipStream_t st, SortBuffers &b, const MoveSort &m, int bits, int parity);
// The window sort of the keys k0 by bead index (from a key kernel given
// win_book(scratch[parity], n, desc); v0 is not read) into (k0,v0), through (k1,v1) and
// (k2,v2); the cell grid is replica 0's, read on the
// device.  Needs sortw_fits(n) and scratch of sort_scratch_words(n, 4) + sortw_words(n) words.
void sort_window(hipStream_t st, SortBuffers &b, uint32_t *k2, uint32_t *v2, int parity, const uint32_t *desc,
                 const DevRep *reps);

}  // namespace smdev
