// Shared device data layout and kernel launchers of the MI355X MD-step solver.
//
// HBM layout (one system, N beads):
//   beads   Bead[N]      FP64 position + type, original (molecule) order      32 B/bead
//   vel     float4[N]    FP32 velocity                                        16 B/bead
//   acc     float4[N]    FP32 acceleration a (= force, unit mass)             16 B/bead
//   fpair   float4[N]    FP32 pair force of the current step                   16 B/bead
//   key/idx uint32[N]x2  cell key and particle index, sorted by key (ping-pong)
//   sbeads  Bead[N]      positions in sorted (cell) order                      32 B/bead
//   cstart/cend int32[C] dense cell table over C = nx*ny*nz cells, never cleared:
//                        an entry is valid only when skey[cstart[k]] == k, so
//                        stale entries of emptied cells are rejected without a reset.
//   chain   int2[N]      per bead: first bead of its CHAIN (or -1), length | const id << 16
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace smdev {

constexpr int kMaxTypes = 8;           // LDS constant table bound (6 * 8 * 8 floats)
constexpr int kWave = 64;

struct __align__(32) Bead {
    double x, y, z;
    int type;
    int pad;
};

// Box and MC state, device resident so a captured step needs no host decision.
struct DevBox {
    double L[3];
    double cs[3];
    int nc[3];
    int flags;                 // bit0 non-finite/out-of-box position, bit1 cell table overflow
    long long capacity;        // cells the dense table holds
    // MC trial state
    double scale[3];
    double newL[3];
    double fluct;
    double u_accept;
    double dU_pair, dU_chain;
    int accept;
    int pad2;
    long long trials, accepted;
    double hist_accept[256];
    double hist_reject[256];
    // measurement results {U_pair, U_chain, K}
    double meas[3];
};

enum : int { FLAG_BAD_POSITION = 1, FLAG_CELL_OVERFLOW = 2, FLAG_SORT_TIMEOUT = 4 };

struct Params {
    int n;
    int nT;
    double cutoff;
    double dt;
    double gamma;
    double temperature;
    double delta_lxy;
    double tension;
    unsigned long long seed;
};

// Counter-RNG stream tags (word 3 of the Philox counter).
enum : unsigned { TAG_LANGEVIN = 0, TAG_MC = 1, TAG_LANGEVIN_INIT = 2 };

// ---- launchers (sm_kernels.hip) ---------------------------------------------------------
void launch_first(hipStream_t st, const Params &P, Bead *beads, float4 *vel, const float4 *acc,
                  const DevBox *box, uint32_t *key, uint32_t *idx, int integrate);
void launch_reorder(hipStream_t st, int n, const uint32_t *skey, const uint32_t *sidx, const Bead *beads,
                    Bead *sbeads, int *cstart, int *cend);
// variant 0: one phase (FP64 test, force law under the branch); 1: FP32 screen into a
// per-lane hit list, then FP64 test + force law on the hits
void launch_pair_force(hipStream_t st, const Params &P, int chunk, int variant, const Bead *sbeads,
                       const uint32_t *skey, const uint32_t *sidx, const int *cstart, const int *cend,
                       const DevBox *box, const float *fconst, float4 *fpair);
void launch_bonded_kick(hipStream_t st, const Params &P, const Bead *beads, const int2 *chain,
                        const double4 *cconst, const float4 *fpair, float4 *vel, float4 *acc, const DevBox *box,
                        long long step, int mode, float4 *fchain_out);
enum : int { KICK_STEP = 0, KICK_INIT = 1, KICK_NONE = 2 };
void launch_second(hipStream_t st, int n, const Bead *beads, const float4 *acc, float4 *vel, float dt);

// FP64 gathers over the 27-cell stencil; results per block in partial[], reduced by
// launch_reduce into out[0].
enum : int { GATHER_U = 0, GATHER_DU = 1, GATHER_COUNT = 2, GATHER_LIST = 3 };
int gather_blocks(int n);                 // 256-thread blocks over n items
int cells_blocks(int n, int chunk);       // workgroups of the cell traversals (sm_pair.hip)
void launch_pair_gather(hipStream_t st, const Params &P, int chunk, int mode, const Bead *sbeads,
                        const uint32_t *skey, const uint32_t *sidx, const int *cstart, const int *cend,
                        const DevBox *box, const double *uconst, double *partial, int *counts, int *pairs,
                        unsigned long long *npairs, long long cap);
int chain_blocks(int nchains);
void launch_chain_energy(hipStream_t st, int nchains, const int2 *chains, const double4 *cconst,
                         const Bead *beads, const DevBox *box, int delta, double *partial);
void launch_kinetic(hipStream_t st, int n, const float4 *vel, double *partial);
// out[k] = sum(partial[k*stride .. k*stride+count_k)) for k < nsum, deterministic order.
void launch_reduce(hipStream_t st, const double *partial, int stride, const int *counts_host_const,
                   int nsum, double *out);
void launch_mc_propose(hipStream_t st, const Params &P, DevBox *box, long long step);
void launch_mc_decide(hipStream_t st, const Params &P, DevBox *box, const double *partial_pair, int npair,
                      const double *partial_chain, int nchain);
void launch_rescale(hipStream_t st, int n, Bead *beads, const DevBox *box);

// ---- radix sort (sm_sort.hip) ---------------------------------------------------------
struct SortBuffers {
    uint32_t *k0, *v0, *k1, *v1;
    uint32_t *scratch;   // sort_scratch_words(n, 4) words
    int *errflag;        // receives FLAG_SORT_TIMEOUT
    int n;
    int tiles;
};
size_t sort_scratch_words(int n, int max_passes);
// Stable LSD sort of (k0,v0) by the low `bits` bits; returns which buffer pair holds the
// result (0 -> k0/v0, 1 -> k1/v1).
int radix_sort(hipStream_t st, SortBuffers &b, int bits);

}  // namespace smdev
