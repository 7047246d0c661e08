// Host side of the MI355X MD-step solver: the device context (one or many replica
// systems), step sequencing, hipGraph capture of the step, the launch-shape autotuner,
// and the context part of the C ABI declared in include/softmold_hip.h.
#include "../../include/softmold_hip.h"
#include "sm_adjoint.hpp"
#include "sm_errors.hpp"
#include "sm_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace smdev;
using namespace smapi;

namespace smapi {
thread_local std::string g_err;
}

namespace {

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ApiError(SM_ERR_HIP, std::string(hipGetErrorName(e_)) + " in " #expr " at " __FILE__ \
                                       ":" + std::to_string(__LINE__));                               \
    } while (0)

template <class T>
T *dalloc(size_t count)
{
    void *p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc(&p, count * sizeof(T)));
    return static_cast<T *>(p);
}

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

enum KClass { K_FIRST, K_SORT, K_REORDER, K_PAIR, K_BONDED, K_GATHER, K_CHAIN_E, K_KIN, K_MC, K_KICK_FIRST, K_NCLASS };
const char *kClassName[K_NCLASS] = {"integrate_first+key", "radix_sort", "reorder+cell_table", "pair_force",
                                    "bonded+langevin+kick", "pair_gather_fp64", "chain_energy_fp64",
                                    "kinetic_fp64", "mc_move", "bonded+kick+next_first+key"};

// The pair kernels' constant table: each (home type, candidate type) row of the FP32 force
// constants padded to 8 floats {r_m, C1, C2, r_c, C4, C5, 0, 0}, then the culling reach
// r^2 of each candidate type for the force (from those FP32 constants) and for the FP64
// energy gathers (from the FP64 potential constants), kMaxTypes floats each.  A candidate
// type's reach is the largest over the other types, and over both orders of the pair, of
// r_c when the row has an outer branch (C4 or C5 nonzero) and r_m otherwise, capped at the
// cutoff, plus 1e-4.
std::vector<float> pair_table(const std::vector<double> &hf, const std::vector<double> &hu, int nT, double cutoff)
{
    const size_t rows = (size_t)nT * nT;
    std::vector<float> tab(8 * rows + 2 * kMaxTypes, 0.f);
    for (size_t r = 0; r < rows; r++)
        for (int k = 0; k < 6; k++) tab[8 * r + k] = (float)hf[6 * r + k];
    const double rcap = std::sqrt(cutoff * cutoff);
    for (int t = 0; t < nT; t++) {
        double rF = 0.0, rE = 0.0;
        for (int h = 0; h < nT; h++)
            for (const size_t row : {(size_t)h * nT + t, (size_t)t * nT + h}) {
                const float *f = &tab[8 * row];
                const double fm = f[0], fc = f[3], f4 = f[4], f5 = f[5];
                rF = std::fmax(rF, (f4 != 0.0 || f5 != 0.0) ? fc : fm);
                const double *u = &hu[6 * row];
                rE = std::fmax(rE, (u[4] != 0.0 || u[5] != 0.0) ? u[3] : u[0]);
            }
        rF = std::fmin(rF, rcap) + 1e-4;
        rE = std::fmin(rE, rcap) + 1e-4;
        tab[8 * rows + t] = (float)(rF * rF);
        tab[8 * rows + kMaxTypes + t] = (float)(rE * rE);
    }
    return tab;
}

// Whether some (a, b) row of the force or potential table differs from its (b, a) row.  The
// reference evaluates each pair once, with the row of the particle its traversal makes
// first (MD.h:811-819, cellOpt.h:712-926); the pair kernels then pick that row per pair.
bool tables_asymmetric(const std::vector<double> &hf, const std::vector<double> &hu, int nT)
{
    for (int a = 0; a < nT; a++)
        for (int b = a + 1; b < nT; b++)
            for (int k = 0; k < 6; k++)
                if (hf[6 * (a * nT + b) + k] != hf[6 * (b * nT + a) + k] ||
                    hu[6 * (a * nT + b) + k] != hu[6 * (b * nT + a) + k])
                    return true;
    return false;
}

// k_first's step argument: set the device step counter, advance it by one, or leave it
constexpr long long kStepAdvance = -1, kStepKeep = -2;

// host mirror of a replica's layout
struct RepHost {
    int first = 0, count = 0, cfirst = 0, ccount = 0;
    int nbond = 0, nbend = 0;   // generic BOND / BEND records
    long long capacity = 0, cellbase = 0;
};

// Tiles of a segment set: ranges [begin_r, begin_r + len_r) cut into kSegTile pieces that
// start at each range's start.
struct SegHost {
    std::vector<int2> tile;
    std::vector<int> rep_first;
};

SegHost make_tiles(const std::vector<int> &begin, const std::vector<int> &len)
{
    SegHost h;
    for (size_t r = 0; r < begin.size(); r++) {
        h.rep_first.push_back((int)h.tile.size());
        for (int t = 0; t < len[r]; t += kSegTile)
            h.tile.push_back(make_int2(begin[r] + t, begin[r] + std::min(len[r], t + kSegTile)));
    }
    h.rep_first.push_back((int)h.tile.size());
    return h;
}

}  // namespace

struct sm_ctx {
    hipStream_t st = nullptr;
    int device = 0;   // the device current at creation, whose stream this is
    Params P{};
    int n = 0;
    int nrep = 0;
    int nchains = 0;
    bool any_mc = false;   // some replica has delta_lxy != 0
    std::vector<RepHost> rh;
    // device state
    Bead *beads = nullptr, *sbeads = nullptr;
    Bead *beads_alt = nullptr;   // the fused kick + next first half writes positions here, then swaps
    float4 *vel = nullptr, *acc = nullptr, *fpair = nullptr, *fscratch = nullptr;
    uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
    uint32_t *hist[2] = {nullptr, nullptr};   // sort scratch, double-buffered
    size_t hist_words = 0;
    int sort_parity = 0;                      // scratch the next sort uses
    bool hist_pending = false;                // its histograms hold keys not sorted yet
    // Incremental re-sort (sort_mode 1; 0: every build sorts all keys from index order): the
    // keys by index of this build (kidx[kcur]) and of the previous sort (kidx[1 - kcur]),
    // the sorted order of the previous sort in (sk, sv)[scur] (the new one goes into the
    // other pair), the mover counts and flags, the movers-before words and the mover count.
    // order_valid: those describe the last sort, which was incremental.  Measured on one
    // MI355X: 38.4 against 48.9 us per step at 240k beads (the movers' passes and the
    // merge are latency-bound launches), 126 against 115 us at 3M, 156 against 118 us for
    // the 64-replica sweep of 2.8M; contexts of a million beads or more start in mode 1,
    // single replicas that the window sort fits start in mode 2, and the autotuner tries
    // every mode the context allows.
    int sort_mode = 0;
    uint32_t *kidx[2] = {nullptr, nullptr};
    uint32_t *sk[2] = {nullptr, nullptr}, *sv[2] = {nullptr, nullptr};
    uint32_t *mbefore = nullptr, *cnt_b = nullptr, *cnt_i = nullptr, *cnt_s = nullptr;
    unsigned long long *mflags = nullptr;
    int *mcount = nullptr;
    int kcur = 0, scur = 0;
    bool order_valid = false;
    // Window sort (sort_mode 2, when win_ok: a single replica of at most 256 sort tiles; it
    // runs while the keys have 17 to 24 bits, the every-key sort otherwise): wdesc + 8p is
    // the window of the sort with scratch parity p, which the previous sort's reorder
    // derived on the device; (k2, v2) is the fallback's third buffer pair and then holds the
    // sorted cell indices; win_last: the last sort was a window sort.
    bool win_ok = false, win_last = false;
    uint32_t *k2 = nullptr, *v2 = nullptr, *wdesc = nullptr;
    bool win_now() const { return sort_mode == 2 && win_ok && sort_bits >= 17 && sort_bits <= 24; }
    int *cstart = nullptr, *cend = nullptr;
    long long capacity = 0;   // cells in the dense table, all replicas
    int sort_bits = 0;
    DevRep *reps = nullptr;
    DevGlobal *glob = nullptr;
    float *fconstF = nullptr;   // pair table: padded FP32 force rows, then the reach tables
    double *uconstD = nullptr;
    double *fconstD = nullptr;   // FP64 force rows (the force kernel's pair-count hook)
    int2 *chainPerBead = nullptr, *chainList = nullptr;
    double4 *cconstD = nullptr;
    // generic BOND / BEND records, per bead (CSR); gptr == nullptr when there are none
    int *gptr = nullptr;
    GenEnt *gent = nullptr;
    double2 *gconstD = nullptr;
    GenTables gen() const { return GenTables{gptr, gent, gconstD}; }
    // host copies of the constants (the adjoint changes them)
    std::vector<double> hfconst, huconst;
    std::vector<double4> hcconst;
    double *adjF = nullptr, *adjDF = nullptr, *adjDU = nullptr;   // FP64 F table, d/dtheta tables
    // sm_param_derivative: the U table's d/dUmin and d/dUmax forms of one type pair, their
    // culling tables, and the CHAIN constant sets with (k_bond, k_bend) = (1, 0) and (0, 1)
    double *pdU = nullptr;
    float *pdTab = nullptr;
    double4 *pdC = nullptr;
    int pd_a = -1, pd_b = -1;
    size_t pd_tab = 0;
    double *ebead = nullptr, *echain = nullptr, *dres = nullptr;   // dres: meas [R][3], then dU [R][2]
    SegTiles segB{}, segC{};
    int *dcounts = nullptr, *dpairs = nullptr;
    long long pairs_cap = 0;
    unsigned long long *dnp = nullptr;
    int chunk = 8;
    int pair_variant = 0;
    // box-move steps: the pair dU in the same traversal as the step's pair force (1) or in
    // a gather of its own after the kick (0, as the adjoint uses)
    bool mc_fused = true;
    int sorted_sel = 0;   // which sort buffer pair holds the sorted result
    // captured 8-step graph
    bool use_graph = false;
    // one graph per sort-scratch parity at the start of the block (8 sorts return to it)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // profiling
    bool prof = false;
    struct Rec { int cls; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double prof_ms[K_NCLASS] = {};
    long long prof_cnt[K_NCLASS] = {};

    const uint32_t *skey() const { return sort_mode == 1 ? sk[scur] : (win_last ? k2 : (sorted_sel ? k1 : k0)); }
    const uint32_t *sidx() const { return sort_mode == 1 ? sv[scur] : (win_last ? v0 : (sorted_sel ? v1 : v0)); }
    // the key kernel's outputs and bookkeeping for the next sort
    uint32_t *key_out() const { return sort_mode == 1 ? kidx[kcur] : k0; }
    // the window sort's first pass takes the bead indices from the keys' positions
    uint32_t *idx_out() const { return sort_mode == 1 || win_now() ? nullptr : v0; }
    WinBook win_keys() const
    {
        if (!win_now()) return WinBook{nullptr, nullptr, nullptr, nullptr};
        return win_book(hist[sort_parity], n, wdesc + 8 * sort_parity);
    }
    // a key kernel ran without its sort: zero what it added to the sort's scratch
    void restart_hist(uint32_t *gh)
    {
        launch_fill_u32(st, gh, sort_hist_words(), 0u);
        if (win_now()) launch_fill_u32(st, win_scratch(gh, n), kWinKeyWords, 0u);
    }
    MoveBook move_book() const
    {
        if (sort_mode != 1) return MoveBook{nullptr, nullptr};
        return MoveBook{order_valid ? kidx[1 - kcur] : nullptr, cnt_b};
    }
    double *meas_dev() const { return dres; }
    double *du_dev() const { return dres + 3 * (size_t)nrep; }

    ~sm_ctx()
    {
        if (st) (void)hipStreamSynchronize(st);
        drop_graph();
        for (auto &r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
        dfree(beads); dfree(sbeads); dfree(beads_alt); dfree(vel); dfree(acc); dfree(fpair); dfree(fscratch);
        dfree(k0); dfree(v0); dfree(k1); dfree(v1); dfree(hist[0]); dfree(hist[1]); dfree(cstart); dfree(cend);
        dfree(k2); dfree(v2); dfree(wdesc);
        for (int p = 0; p < 2; p++) { dfree(kidx[p]); dfree(sk[p]); dfree(sv[p]); }
        dfree(mbefore); dfree(cnt_b); dfree(cnt_i); dfree(cnt_s); dfree(mflags); dfree(mcount);
        dfree(reps); dfree(glob); dfree(fconstF); dfree(uconstD); dfree(fconstD); dfree(chainPerBead); dfree(chainList);
        dfree(gptr); dfree(gent); dfree(gconstD);
        dfree(cconstD); dfree(ebead); dfree(echain); dfree(dres); dfree(dcounts); dfree(dpairs); dfree(dnp);
        dfree(adjF); dfree(adjDF); dfree(adjDU); dfree(pdU); dfree(pdTab); dfree(pdC);
        dfree(segB.tile); dfree(segB.rep_first); dfree(segB.tilesum);
        dfree(segC.tile); dfree(segC.rep_first); dfree(segC.tilesum);
        if (st) (void)hipStreamDestroy(st);
    }

    void drop_graph()
    {
        for (int p = 0; p < 2; p++) {
            if (gexec[p]) (void)hipGraphExecDestroy(gexec[p]);
            if (graph[p]) (void)hipGraphDestroy(graph[p]);
            gexec[p] = nullptr;
            graph[p] = nullptr;
        }
    }

    // ---- profiling helpers ----
    hipEvent_t ev()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    template <class F>
    void timed(int cls, F &&f)
    {
        if (!prof) {
            f();
            return;
        }
        Rec r{cls, ev(), ev()};
        HIPCHK(hipEventRecord(r.a, st));
        f();
        HIPCHK(hipEventRecord(r.b, st));
        recs.push_back(r);
        if (recs.size() >= 8192) harvest();
    }
    void harvest()
    {
        HIPCHK(hipStreamSynchronize(st));
        for (auto &r : recs) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            prof_ms[r.cls] += ms;
            prof_cnt[r.cls]++;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }

    void check_launch() { HIPCHK(hipGetLastError()); }

    std::vector<DevRep> read_reps()
    {
        std::vector<DevRep> h(nrep);
        HIPCHK(hipMemcpyAsync(h.data(), reps, sizeof(DevRep) * nrep, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return h;
    }
    void write_reps(const std::vector<DevRep> &h)
    {
        HIPCHK(hipMemcpyAsync(reps, h.data(), sizeof(DevRep) * nrep, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    DevGlobal read_glob()
    {
        DevGlobal g;
        HIPCHK(hipMemcpyAsync(&g, glob, sizeof g, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return g;
    }

    void check_flags()
    {
        const DevGlobal g = read_glob();
        if (g.flags & FLAG_BAD_POSITION)
            throw ApiError(SM_ERR_STATE, "a particle position left the box or became non-finite");
        if (g.flags & FLAG_SORT_TIMEOUT)
            throw ApiError(SM_ERR_STATE, "cell-key sort look-back timed out");
        if (g.flags & FLAG_SORT_COUNT)
            throw ApiError(SM_ERR_STATE, "incremental cell sort: inconsistent mover count");
        if (g.flags & FLAG_CELL_OVERFLOW)
            throw ApiError(SM_ERR_STATE, "an accepted box move would exceed the cell table or make fewer than 3 cells per axis");
    }

    // Dense cell table sized for every box the next `moves` MC moves can reach, in every
    // replica; a replica that outgrows its reservation moves every cell base (keys are
    // recomputed each step, so only the table is reallocated).
    void ensure_capacity(long long moves)
    {
        std::vector<DevRep> hr = read_reps();
        bool relayout = capacity == 0;
        const double rc = P.cutoff;
        for (int r = 0; r < nrep; r++) {
            const double dL = hr[r].delta_lxy;
            const double grow = (double)moves * std::fabs(dL);
            double Lmax[3], Lmin[2];
            for (int d = 0; d < 2; d++) {
                Lmax[d] = hr[r].L[d] + grow;
                Lmin[d] = std::max(hr[r].L[d] - grow, 1e-300);
            }
            Lmax[2] = (moves > 0 && dL != 0) ? hr[r].L[2] * (hr[r].L[0] * hr[r].L[1]) / (Lmin[0] * Lmin[1]) : hr[r].L[2];
            long long need = 1;
            for (int d = 0; d < 3; d++) need *= (long long)(Lmax[d] / rc) + 1;
            if (need > rh[r].capacity) {
                rh[r].capacity = need + need / 8;
                relayout = true;
            }
        }
        if (!relayout) return;
        long long base = 0;
        for (int r = 0; r < nrep; r++) {
            rh[r].cellbase = base;
            base += rh[r].capacity;
            hr[r].cellbase = rh[r].cellbase;
            hr[r].capacity = rh[r].capacity;
        }
        if (base > 0xFFFFFFFFll) unsupported("cell tables of all replicas exceed 2^32 cells");
        drop_graph();
        dfree(cstart);
        dfree(cend);
        cstart = dalloc<int>((size_t)base);
        cend = dalloc<int>((size_t)base);
        HIPCHK(hipMemsetAsync(cstart, 0xFF, sizeof(int) * (size_t)base, st));
        HIPCHK(hipMemsetAsync(cend, 0xFF, sizeof(int) * (size_t)base, st));
        capacity = base;
        write_reps(hr);
        int bits = 1;
        while ((1ll << bits) < base) bits++;
        sort_bits = bits;
    }

    // keys (optionally after Verlet::first) with the next sort's digit histograms, sort,
    // reorder + dense cell table (clearing the other sort scratch)
    void first_and_keys(int integrate, long long step_set, bool for_sort = true)
    {
        uint32_t *gh = for_sort ? hist[sort_parity] : nullptr;
        // keys computed twice before a sort (sm_step_first called twice): start over
        if (for_sort && hist_pending) restart_hist(gh);
        hist_pending = hist_pending || for_sort;
        const int passes = sort_passes(sort_bits);
        timed(K_FIRST, [&] {
            if (for_sort)
                launch_first(st, P, beads, vel, acc, reps, glob, key_out(), idx_out(), integrate, step_set, gh, passes,
                             move_book(), win_keys());
            else   // keys only (no sort follows): into the mover buffer, which no sort state holds
                launch_first(st, P, beads, vel, acc, reps, glob, k0, nullptr, integrate, step_set, nullptr, passes,
                             MoveBook{nullptr, nullptr});
        });
    }
    void sort_and_reorder()
    {
        const int p = sort_parity;
        const bool win = win_now();
        if (sort_mode == 1) {
            timed(K_SORT, [&] {
                SortBuffers sb{k0, v0, k1, v1, {hist[0], hist[1]}, &glob->flags, n, 0};
                const MoveSort m{kidx[kcur], order_valid ? kidx[1 - kcur] : nullptr, sk[scur], sv[scur], cstart, cend,
                                 cnt_b, cnt_i, cnt_s, mflags, mbefore, mcount, sk[1 - scur], sv[1 - scur]};
                sort_moved(st, sb, m, sort_bits, p);
            });
            scur = 1 - scur;
            kcur = 1 - kcur;
            order_valid = true;
        } else if (win) {
            timed(K_SORT, [&] {
                SortBuffers sb{k0, v0, k1, v1, {hist[0], hist[1]}, &glob->flags, n, 0};
                sort_window(st, sb, k2, v2, p, wdesc + 8 * p, reps);
            });
            order_valid = false;
        } else {
            timed(K_SORT, [&] {
                SortBuffers sb{k0, v0, k1, v1, {hist[0], hist[1]}, &glob->flags, n, 0};
                sorted_sel = radix_sort(st, sb, sort_bits, p);
            });
            order_valid = false;
        }
        win_last = win;
        timed(K_REORDER, [&] {
            if (win)
                launch_reorder_win(st, n, k0, v0, beads, sbeads, cstart, cend, hist[1 - p], hist_words,
                                   win_book(hist[p], n, wdesc + 8 * p), k2, wdesc + 8 * (1 - p), reps);
            else
                launch_reorder(st, n, skey(), sidx(), beads, sbeads, cstart, cend, hist[1 - p], hist_words);
        });
        sort_parity = 1 - p;
        hist_pending = false;
    }
    void build(int integrate, long long step_set)
    {
        first_and_keys(integrate, step_set);
        sort_and_reorder();
    }
    void pair_force()
    {
        timed(K_PAIR, [&] {
            launch_pair_force(st, P, chunk, pair_variant, sbeads, skey(), sidx(), cstart, cend, reps, fconstF, fpair);
        });
    }
    void bonded(int mode, float4 *fchain_out)
    {
        timed(K_BONDED, [&] {
            launch_bonded_kick(st, P, beads, chainPerBead, cconstD, gen(), fpair, vel, acc, reps, glob, mode,
                               fchain_out);
        });
    }
    // {U_pair, U_chain, K} of every replica into dres[3r + 0..2] (device)
    void measure_device()
    {
        timed(K_GATHER, [&] {
            launch_pair_gather(st, P, chunk, GATHER_U, sbeads, skey(), sidx(), cstart, cend, reps, fconstF, uconstD, ebead,
                               nullptr, nullptr, nullptr, 0);
            launch_seg_sum(st, segB, ebead, meas_dev() + 0, 3);
        });
        timed(K_CHAIN_E, [&] {
            launch_chain_energy(st, nchains, chainList, cconstD, beads, reps, 0, echain);
            launch_seg_sum(st, segC, echain, meas_dev() + 1, 3);
            if (gptr) {
                launch_gen_energy(st, n, gen(), beads, reps, 0, ebead);
                launch_seg_sum(st, segB, ebead, meas_dev() + 1, 3, true);
            }
        });
        timed(K_KIN, [&] {
            launch_kinetic(st, n, vel, ebead);
            launch_seg_sum(st, segB, ebead, meas_dev() + 2, 3);
        });
    }
    // U(old) - U(scaled) of every replica with its trial scale into dres[3R + 2r + 0..1]
    void delta_device()
    {
        timed(K_GATHER, [&] {
            launch_pair_gather(st, P, chunk, GATHER_DU, sbeads, skey(), sidx(), cstart, cend, reps, fconstF, uconstD, ebead,
                               nullptr, nullptr, nullptr, 0);
            launch_seg_sum(st, segB, ebead, du_dev() + 0, 2);
        });
        delta_bonded();
    }
    // the CHAIN and generic-record part of the dU
    void delta_bonded()
    {
        timed(K_CHAIN_E, [&] {
            launch_chain_energy(st, nchains, chainList, cconstD, beads, reps, 1, echain);
            launch_seg_sum(st, segC, echain, du_dev() + 1, 2);
            if (gptr) {
                launch_gen_energy(st, n, gen(), beads, reps, 1, ebead);
                launch_seg_sum(st, segB, ebead, du_dev() + 1, 2, true);
            }
        });
    }
    void mc_move()
    {
        timed(K_MC, [&] { launch_mc_propose(st, nrep, reps, glob); });
        delta_device();
        mc_decide();
    }
    void mc_decide()
    {
        timed(K_MC, [&] {
            launch_mc_decide(st, P, reps, glob, du_dev());
            launch_rescale(st, n, beads, reps);
        });
    }
    // A box-move step's pair force with its pair dU: the trial scale is drawn first (it
    // depends on the step's key only), then one traversal gives the forces and each owned
    // pair's U(old) - U(scaled) at the positions the force sees, which the kick leaves as
    // they are.  The bonded dU and the decision follow the kick (mc_finish).
    void pair_force_mc()
    {
        timed(K_MC, [&] { launch_mc_propose(st, nrep, reps, glob); });
        pair_force_du();
    }
    // the pair force with the pair dU under each replica's current trial scale
    void pair_force_du()
    {
        timed(K_PAIR, [&] {
            launch_pair_force_du(st, P, chunk, sbeads, skey(), sidx(), cstart, cend, reps, fconstF, uconstD, fpair,
                                 ebead);
            launch_seg_sum(st, segB, ebead, du_dev() + 0, 2);
        });
    }
    void pair_step(bool mc)
    {
        if (mc && mc_fused)
            pair_force_mc();
        else
            pair_force();
    }
    void mc_finish()
    {
        if (mc_fused) {
            delta_bonded();
            mc_decide();
        } else {
            mc_move();
        }
    }
    bool mc_step(long long step) const { return step % 8 == 0 && step != 0 && any_mc; }

    // Step s's bonded + Langevin forces and second half kick, with step s + 1's first half,
    // keys and digit histograms, as one kernel (same arithmetic as bonded() then
    // step_first(s + 1)); the new positions land in the other bead buffer.
    void kick_first(long long s)
    {
        if (gexec[0] || gexec[1]) drop_graph();   // captured graphs hold the bead buffer
        uint32_t *gh = hist[sort_parity];
        if (hist_pending) restart_hist(gh);
        hist_pending = true;
        const int passes = sort_passes(sort_bits);
        timed(K_KICK_FIRST, [&] {
            launch_kick_first(st, P, beads, beads_alt, chainPerBead, cconstD, gen(), fpair, vel, acc, reps, glob, s,
                              key_out(), idx_out(), gh, passes, move_book(), win_keys());
        });
        std::swap(beads, beads_alt);
    }

    void step_first(long long step_set) { first_and_keys(1, step_set); }
    void step_finish(bool measure, bool mc)
    {
        sort_and_reorder();
        pair_step(mc);
        bonded(KICK_STEP, nullptr);
        if (measure) measure_device();
        if (mc) mc_finish();
    }
    void eager_step(long long step)
    {
        step_first(step);
        step_finish(false, mc_step(step));
    }

    // Steps s .. s+7 with s % 8 == 1 as one graph: the step counter advances on the
    // device, so the same graph serves every such block; the MC move (if any replica
    // has one) falls on the block's last step.
    void graph_block(long long s)
    {
        if (sort_mode == 1 && !order_valid) {
            // the graph would replay a first sort without a previous order: step this block
            // eagerly, after which the order is there to capture against
            for (int k = 0; k < 8; k++) eager_step(s + k);
            return;
        }
        const int p = sort_parity;   // 8 sorts later the parity is p again
        if (!gexec[p]) {
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < 8; k++) {
                step_first(kStepAdvance);
                step_finish(false, k == 7 && any_mc);
            }
            hipError_t e = hipStreamEndCapture(st, &graph[p]);
            if (e != hipSuccess) {
                graph[p] = nullptr;
                HIPCHK(e);
            }
            HIPCHK(hipGraphInstantiate(&gexec[p], graph[p], nullptr, nullptr, 0));
        }
        sort_parity = p;   // capture only recorded the toggles
        // the block advances from step s - 1
        launch_set_step(st, glob, s - 1);
        HIPCHK(hipGraphLaunch(gexec[p], st));
    }

    void run(long long first, long long nsteps)
    {
        long long moves = 0;
        for (long long s = first; s < first + nsteps; s++)
            if (mc_step(s)) moves++;
        if (moves) ensure_capacity(moves);
        long long s = first;
        const long long end = first + nsteps;
        if (use_graph && !prof) {
            while (s < end && s % 8 != 1) eager_step(s++);
            while (end - s >= 8) {
                graph_block(s);
                s += 8;
            }
            while (s < end) eager_step(s++);
            return;
        }
        // eager: a step without a box move ends in the fused kernel that also does the next
        // step's first half; a box-move step and the run's last step end unfused
        bool first_done = false;
        for (; s < end; s++) {
            if (!first_done) step_first(s);
            const bool mc = mc_step(s);
            sort_and_reorder();
            pair_step(mc);
            if (!mc && s + 1 < end) {
                kick_first(s);
                first_done = true;
            } else {
                bonded(KICK_STEP, nullptr);
                if (mc) mc_finish();
                first_done = false;
            }
        }
    }
};

// ---------------------------------------------------------------------------------------
// validation and upload
// ---------------------------------------------------------------------------------------
namespace {

void validate(const sm_system *s)
{
    if (!s) invalid("null system");
    const int n = s->n_particles;
    if (n <= 0) invalid("n_particles must be positive");
    if (!s->pos || !s->type || !s->vel) invalid("pos, type and vel are required");
    if (s->n_types < 1 || s->n_types > kMaxTypes)
        unsupported("n_types must be in [1, " + std::to_string(kMaxTypes) + "]");
    if (!s->fconst || !s->uconst) invalid("fconst and uconst are required");
    if (!(s->cutoff > 0)) invalid("cutoff must be positive");
    if (!(s->dt > 0)) invalid("dt must be positive");
    if (!(s->gamma > 0)) invalid("gamma must be > 0");
    if (!(s->sigma_temperature >= 0)) invalid("sigma_temperature must be >= 0");
    for (int d = 0; d < 3; d++) {
        if (!(s->box[d] > 0)) invalid("box lengths must be positive");
        if ((int)(s->box[d] / s->cutoff) < 3) unsupported("fewer than 3 cells along an axis");
    }
    const int nT = s->n_types;
    for (int i = 0; i < n; i++) {
        if (s->type[i] < 0 || s->type[i] >= nT) invalid("particle " + std::to_string(i) + " has a type outside [0, nTypes)");
        for (int d = 0; d < 3; d++) {
            double p = s->pos[3 * (size_t)i + d];
            if (!(p >= 0 && p <= s->box[d]))
                invalid("particle " + std::to_string(i) + " position is out of bounds");
        }
    }
    // the reference's force takes the branch at row offset int(r / r_m) * 3 (MD.h:819),
    // which stays inside the row for every r below the cutoff only while 2 r_m >= cutoff
    for (int r = 0; r < nT * nT; r++)
        if (!(2.0 * s->fconst[6 * r] >= s->cutoff))
            unsupported("two-body row " + std::to_string(r / nT) + "," + std::to_string(r % nT) +
                        ": r_m below half the cutoff (the reference's branch offset int(r/r_m)*3 leaves the row)");
    if (s->n_chain_blocks < 0) invalid("negative chain block count");
    if (s->n_chain_blocks > 0 && (!s->chain || !s->chain_const)) invalid("chain blocks without data");
    // CHAIN blocks (start, nChains, length): inside the particles, each bead in one chain
    // at most (the reference checks its chains' bonds against nParticles, system.h:515-535)
    std::vector<char> chained(s->n_chain_blocks > 0 ? n : 0, 0);
    for (int b = 0; b < s->n_chain_blocks; b++) {
        const int start = s->chain[3 * b], cnt = s->chain[3 * b + 1], len = s->chain[3 * b + 2];
        if (len < 3 || len > 0xFFFF) unsupported("CHAIN length must be in [3, 65535]");
        if (cnt < 0 || start < 0 || (long long)start + (long long)cnt * len > n) invalid("CHAIN block out of range");
        for (int i = start; i < start + cnt * len; i++) {
            if (chained[i]) unsupported("a bead belongs to two CHAIN records");
            chained[i] = 1;
        }
    }
    if (s->n_bonds < 0 || s->n_bends < 0 || s->n_balls < 0) invalid("negative BOND / BEND / BALL record count");
    if ((s->n_bonds > 0 && (!s->bond || !s->bond_const)) || (s->n_bends > 0 && (!s->bend || !s->bend_const)) ||
        (s->n_balls > 0 && (!s->ball || !s->ball_const)))
        invalid("BOND / BEND / BALL records without data");
    auto in_range = [&](const int32_t *t, int m) {
        for (int k = 0; k < m; k++)
            if (t[k] < 0 || t[k] >= n) invalid("a BOND / BEND / BALL record names a particle out of range");
        for (int k = 0; k < m; k++)
            for (int l = k + 1; l < m; l++)
                if (t[k] == t[l]) invalid("a BOND / BEND / BALL record repeats a particle");
    };
    for (int j = 0; j < s->n_bonds; j++) in_range(s->bond + 2 * j, 2);
    for (int j = 0; j < s->n_bends; j++) in_range(s->bend + 3 * j, 3);
    for (int j = 0; j < s->n_balls; j++) in_range(s->ball + 2 * j, 2);
}

// replicas of one context share the interaction constants, cutoff and time step
void validate_batch(const sm_system *s, int nrep)
{
    long long total = 0, blocks = 0;
    for (int r = 0; r < nrep; r++) {
        validate(&s[r]);
        total += s[r].n_particles;
        blocks += s[r].n_chain_blocks;
        if (r == 0) continue;
        const sm_system &a = s[0], &b = s[r];
        if (b.n_types != a.n_types || b.cutoff != a.cutoff || b.dt != a.dt)
            unsupported("replicas must share n_types, cutoff and dt");
        const size_t nt = 6 * (size_t)a.n_types * a.n_types;
        if (std::memcmp(a.fconst, b.fconst, nt * sizeof(double)) != 0 ||
            std::memcmp(a.uconst, b.uconst, nt * sizeof(double)) != 0)
            unsupported("replicas must share the two-body constant tables");
    }
    if (total > 0x7FFFFFFF) unsupported("more than 2^31-1 beads in one context");
    if (blocks > 0x8000) unsupported("too many CHAIN constant sets");
}

sm_ctx *create(const sm_system *s, int nrep)
{
    validate_batch(s, nrep);
    sm_ctx *c = new sm_ctx();
    try {
        c->nrep = nrep;
        c->rh.resize(nrep);
        int n = 0, nch = 0;
        for (int r = 0; r < nrep; r++) {
            c->rh[r].first = n;
            c->rh[r].count = s[r].n_particles;
            n += s[r].n_particles;
        }
        c->n = n;
        c->P.n = n;
        c->P.nrep = nrep;
        c->P.nT = s[0].n_types;
        c->P.cutoff = s[0].cutoff;
        c->P.dt = s[0].dt;
        HIPCHK(hipStreamCreate(&c->st));
        // the device's fixed scratch for the entry points that allocate nothing, looked up
        // once per device while it is current
        HIPCHK(hipGetDevice(&c->device));
        HIPCHK(noalloc_scratch_init(c->device));

        // chains: per-bead role table and per-chain list, replica after replica
        std::vector<int2> perBead(n, make_int2(-1, 0));
        std::vector<int2> list;
        std::vector<double4> ccD;
        for (int r = 0; r < nrep; r++) {
            const int off = c->rh[r].first;
            c->rh[r].cfirst = (int)list.size();
            for (int b = 0; b < s[r].n_chain_blocks; b++) {
                // ranges, lengths, overlaps and the constant-set count checked by validate_batch()
                const int start = s[r].chain[3 * b], cnt = s[r].chain[3 * b + 1], len = s[r].chain[3 * b + 2];
                const double *cc = s[r].chain_const + 4 * b;
                const int cid = (int)ccD.size();
                ccD.push_back(make_double4(cc[0], cc[1], cc[2], cc[3]));
                for (int k = 0; k < cnt; k++) {
                    const int st0 = off + start + k * len;
                    list.push_back(make_int2(st0, len | (cid << 16)));
                    for (int j = 0; j < len; j++) perBead[st0 + j] = make_int2(st0, len | (cid << 16));
                }
            }
            c->rh[r].ccount = (int)list.size() - c->rh[r].cfirst;
        }
        nch = (int)list.size();
        c->nchains = nch;

        // generic BOND / BEND records: each record listed under each of its beads
        std::vector<int> gcount(n + 1, 0);
        std::vector<GenEnt> grec;
        std::vector<double2> gcst;
        for (int r = 0; r < nrep; r++) {
            // indices checked by validate()
            const int off = c->rh[r].first;
            auto bead = [&](int32_t b) { return off + b; };
            c->rh[r].nbond = s[r].n_bonds;
            c->rh[r].nbend = s[r].n_bends;
            for (int j = 0; j < s[r].n_bonds; j++) {
                grec.push_back(GenEnt{0, bead(s[r].bond[2 * j]), bead(s[r].bond[2 * j + 1]), -1});
                gcst.push_back(make_double2(s[r].bond_const[2 * j], s[r].bond_const[2 * j + 1]));
            }
            for (int j = 0; j < s[r].n_balls; j++) {
                grec.push_back(GenEnt{2, bead(s[r].ball[2 * j]), bead(s[r].ball[2 * j + 1]), -1});
                gcst.push_back(make_double2(s[r].ball_const[2 * j], s[r].ball_const[2 * j + 1]));
            }
            for (int j = 0; j < s[r].n_bends; j++) {
                grec.push_back(GenEnt{1, bead(s[r].bend[3 * j]), bead(s[r].bend[3 * j + 1]), bead(s[r].bend[3 * j + 2])});
                gcst.push_back(make_double2(s[r].bend_const[2 * j], s[r].bend_const[2 * j + 1]));
            }
        }
        std::vector<int> gptrH;
        std::vector<GenEnt> gentH;
        std::vector<double2> gcstH;
        if (!grec.empty()) {
            auto members = [](const GenEnt &g, int out[3]) {
                out[0] = g.p1;
                out[1] = g.p2;
                out[2] = g.p3;
                return g.kind == 1 ? 3 : 2;
            };
            for (const GenEnt &g : grec) {
                int mb[3];
                const int m = members(g, mb);
                for (int k = 0; k < m; k++) gcount[mb[k] + 1]++;
            }
            gptrH.assign(n + 1, 0);
            for (int i = 0; i < n; i++) gptrH[i + 1] = gptrH[i] + gcount[i + 1];
            gentH.resize(gptrH[n]);
            gcstH.resize(gptrH[n]);
            std::vector<int> fill(gptrH.begin(), gptrH.end() - 1);
            for (size_t e = 0; e < grec.size(); e++) {
                int mb[3];
                const int m = members(grec[e], mb);
                for (int k = 0; k < m; k++) {
                    gentH[fill[mb[k]]] = grec[e];
                    gcstH[fill[mb[k]]++] = gcst[e];
                }
            }
        }

        // host staging
        std::vector<Bead> hb(n);
        std::vector<float4> hv(n), ha(n);
        std::vector<DevRep> hr(nrep);
        for (int r = 0; r < nrep; r++) {
            const sm_system &q = s[r];
            const int off = c->rh[r].first;
            for (int i = 0; i < q.n_particles; i++) {
                const size_t k = 3 * (size_t)i;
                Bead &b = hb[off + i];
                b.x = q.pos[k];
                b.y = q.pos[k + 1];
                b.z = q.pos[k + 2];
                b.type = q.type[i];
                b.rep = r;
                hv[off + i] = make_float4((float)q.vel[k], (float)q.vel[k + 1], (float)q.vel[k + 2], 0.f);
                ha[off + i] = q.acc ? make_float4((float)q.acc[k], (float)q.acc[k + 1], (float)q.acc[k + 2], 0.f)
                                    : make_float4(0.f, 0.f, 0.f, 0.f);
            }
            DevRep &R = hr[r];
            std::memset(&R, 0, sizeof R);
            for (int d = 0; d < 3; d++) {
                R.L[d] = q.box[d];
                R.nc[d] = (int)(q.box[d] / q.cutoff);
                R.cs[d] = q.box[d] / R.nc[d];
                R.scale[d] = 1.0;
            }
            R.first = c->rh[r].first;
            R.count = c->rh[r].count;
            R.cfirst = c->rh[r].cfirst;
            R.ccount = c->rh[r].ccount;
            R.seed = q.seed;
            R.temperature = q.temperature;
            R.gamma = q.gamma;
            const double Ts = q.sigma_temperature > 0 ? q.sigma_temperature : q.temperature;
            R.sigma = std::sqrt((6.0 * Ts * q.gamma) / q.dt);
            R.sigma_fixed = q.sigma_temperature > 0 ? 1 : 0;
            R.delta_lxy = q.delta_lxy;
            R.tension = q.tension;
            if (q.delta_lxy != 0) c->any_mc = true;
        }
        const int nT = s[0].n_types;
        const size_t nconst = 6 * (size_t)nT * nT;
        c->hfconst.assign(s[0].fconst, s[0].fconst + nconst);
        c->huconst.assign(s[0].uconst, s[0].uconst + nconst);
        c->P.asym = tables_asymmetric(c->hfconst, c->huconst, nT) ? 1 : 0;
        const std::vector<float> fF = pair_table(c->hfconst, c->huconst, nT, s[0].cutoff);
        c->hcconst = ccD;

        // segment tiles: beads (molecule order == sorted order per replica) and chains
        std::vector<int> bb(nrep), bl(nrep), cb(nrep), cl(nrep);
        for (int r = 0; r < nrep; r++) {
            bb[r] = c->rh[r].first; bl[r] = c->rh[r].count;
            cb[r] = c->rh[r].cfirst; cl[r] = c->rh[r].ccount;
        }
        const SegHost tb = make_tiles(bb, bl), tc = make_tiles(cb, cl);

        c->beads = dalloc<Bead>(n);
        c->sbeads = dalloc<Bead>(n);
        c->beads_alt = dalloc<Bead>(n);
        c->vel = dalloc<float4>(n);
        c->acc = dalloc<float4>(n);
        c->fpair = dalloc<float4>(n);
        c->fscratch = dalloc<float4>(n);
        c->k0 = dalloc<uint32_t>(n);
        c->v0 = dalloc<uint32_t>(n);
        c->k1 = dalloc<uint32_t>(n);
        c->v1 = dalloc<uint32_t>(n);
        for (int p = 0; p < 2; p++) {
            c->kidx[p] = dalloc<uint32_t>(n);
            c->sk[p] = dalloc<uint32_t>(n);
            c->sv[p] = dalloc<uint32_t>(n);
        }
        c->win_ok = nrep == 1 && sortw_fits(n);
        c->sort_mode = n >= 1000000 ? 1 : (c->win_ok ? 2 : 0);
        if (c->win_ok) {
            c->k2 = dalloc<uint32_t>(n);
            c->v2 = dalloc<uint32_t>(n);
            c->wdesc = dalloc<uint32_t>(16);
            HIPCHK(hipMemsetAsync(c->wdesc, 0, sizeof(uint32_t) * 16, c->st));
        }
        c->mbefore = dalloc<uint32_t>(n);
        c->cnt_b = dalloc<uint32_t>(move_blocks(n));
        c->cnt_i = dalloc<uint32_t>(move_tiles(n));
        c->cnt_s = dalloc<uint32_t>(move_tiles(n));
        c->mflags = dalloc<unsigned long long>(((size_t)n + 63) / 64);
        c->mcount = dalloc<int>(1);
        c->hist_words = sort_scratch_words(n, 4) + (c->win_ok ? sortw_words(n) : 0);
        for (int p = 0; p < 2; p++) {
            c->hist[p] = dalloc<uint32_t>(c->hist_words);
            HIPCHK(hipMemsetAsync(c->hist[p], 0, sizeof(uint32_t) * c->hist_words, c->st));
        }
        c->reps = dalloc<DevRep>(nrep);
        c->glob = dalloc<DevGlobal>(1);
        c->fconstF = dalloc<float>(fF.size());
        c->uconstD = dalloc<double>(nconst);
        c->fconstD = dalloc<double>(nconst);
        c->chainPerBead = dalloc<int2>(n);
        c->chainList = dalloc<int2>(list.size());
        c->cconstD = dalloc<double4>(ccD.size());
        c->ebead = dalloc<double>(n);
        c->echain = dalloc<double>(nch);
        c->dres = dalloc<double>(5 * (size_t)nrep);
        c->dcounts = dalloc<int>(n);
        c->dnp = dalloc<unsigned long long>(1);
        auto upseg = [&](SegTiles &S, const SegHost &h) {
            S.ntiles = (int)h.tile.size();
            S.nrep = nrep;
            S.tile = dalloc<int2>(h.tile.size());
            S.rep_first = dalloc<int>(h.rep_first.size());
            S.tilesum = dalloc<double>(h.tile.size());
            if (!h.tile.empty())
                HIPCHK(hipMemcpyAsync(S.tile, h.tile.data(), sizeof(int2) * h.tile.size(), hipMemcpyHostToDevice, c->st));
            HIPCHK(hipMemcpyAsync(S.rep_first, h.rep_first.data(), sizeof(int) * h.rep_first.size(),
                                  hipMemcpyHostToDevice, c->st));
        };
        upseg(c->segB, tb);
        upseg(c->segC, tc);

        DevGlobal g{};
        hipStream_t st = c->st;
        HIPCHK(hipMemcpyAsync(c->beads, hb.data(), sizeof(Bead) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->vel, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->acc, ha.data(), sizeof(float4) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(c->fpair, 0, sizeof(float4) * n, st));
        HIPCHK(hipMemcpyAsync(c->reps, hr.data(), sizeof(DevRep) * nrep, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->glob, &g, sizeof g, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->fconstF, fF.data(), sizeof(float) * fF.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->uconstD, s[0].uconst, sizeof(double) * nconst, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->fconstD, s[0].fconst, sizeof(double) * nconst, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->chainPerBead, perBead.data(), sizeof(int2) * n, hipMemcpyHostToDevice, st));
        if (!gentH.empty()) {
            c->gptr = dalloc<int>(gptrH.size());
            c->gent = dalloc<GenEnt>(gentH.size());
            c->gconstD = dalloc<double2>(gcstH.size());
            HIPCHK(hipMemcpyAsync(c->gptr, gptrH.data(), sizeof(int) * gptrH.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(c->gent, gentH.data(), sizeof(GenEnt) * gentH.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(c->gconstD, gcstH.data(), sizeof(double2) * gcstH.size(), hipMemcpyHostToDevice,
                                  st));
        }
        if (!list.empty()) {
            HIPCHK(hipMemcpyAsync(c->chainList, list.data(), sizeof(int2) * list.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(c->cconstD, ccD.data(), sizeof(double4) * ccD.size(), hipMemcpyHostToDevice, st));
        }
        // sorted slots per wave in the pair kernel: short chunks keep small systems filling
        // the chip, longer ones cut the per-cell overhead of large ones (measured on the
        // 15k, 240k and 3M-bead systems; every chunk gives the same forces)
        c->chunk = n >= 50000 ? 8 : 2;
        c->ensure_capacity(0);
        HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
        delete c;
        throw;
    }
    return c;
}

template <class T>
void download_f4(sm_ctx *c, const float4 *src, T *dst)
{
    std::vector<float4> h(c->n);
    HIPCHK(hipMemcpyAsync(h.data(), src, sizeof(float4) * c->n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (int i = 0; i < c->n; i++) {
        dst[3 * (size_t)i] = h[i].x;
        dst[3 * (size_t)i + 1] = h[i].y;
        dst[3 * (size_t)i + 2] = h[i].z;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
extern "C" {

const char *sm_last_error(void) { return g_err.c_str(); }
const char *sm_version(void) { return "softmold-mi355x 0.2 (gfx950)"; }

int sm_create(const sm_system *s, sm_ctx **out)
{
    if (!out) {
        g_err = "null output pointer";
        return SM_ERR_INVALID;
    }
    *out = nullptr;
    return guarded([&] { *out = create(s, 1); });
}

int sm_create_batch(const sm_system *systems, int32_t n_replicas, sm_ctx **out)
{
    if (!out) {
        g_err = "null output pointer";
        return SM_ERR_INVALID;
    }
    *out = nullptr;
    return guarded([&] {
        if (!systems || n_replicas < 1) invalid("need at least one replica");
        *out = create(systems, n_replicas);
    });
}

int sm_destroy(sm_ctx *c)
{
    return guarded([&] { delete c; });
}

int sm_replicas(sm_ctx *c, int32_t *n_replicas, int32_t *first, int32_t *count)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (n_replicas) *n_replicas = c->nrep;
        for (int r = 0; r < c->nrep; r++) {
            if (first) first[r] = c->rh[r].first;
            if (count) count[r] = c->rh[r].count;
        }
    });
}

int sm_sync(sm_ctx *c)
{
    return guarded([&] {
        if (!c) invalid("null context");
        HIPCHK(hipStreamSynchronize(c->st));
        c->check_flags();
    });
}

int sm_download_state(sm_ctx *c, double *pos, double *vel, double *acc, double *box)
{
    return guarded([&] {
        if (!c) invalid("null context");
        const int n = c->n;
        if (pos) {
            std::vector<Bead> hb(n);
            HIPCHK(hipMemcpyAsync(hb.data(), c->beads, sizeof(Bead) * n, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            for (int i = 0; i < n; i++) {
                pos[3 * (size_t)i] = hb[i].x;
                pos[3 * (size_t)i + 1] = hb[i].y;
                pos[3 * (size_t)i + 2] = hb[i].z;
            }
        }
        if (vel) download_f4(c, c->vel, vel);
        if (acc) download_f4(c, c->acc, acc);
        std::vector<DevRep> hr = c->read_reps();
        if (box)
            for (int r = 0; r < c->nrep; r++)
                for (int d = 0; d < 3; d++) box[3 * r + d] = hr[r].L[d];
        c->check_flags();
    });
}

int sm_upload_state(sm_ctx *c, const double *pos, const double *vel, const double *acc, const double *box)
{
    return guarded([&] {
        if (!c) invalid("null context");
        const int n = c->n;
        std::vector<DevRep> hr = c->read_reps();
        if (box) {
            for (int r = 0; r < c->nrep; r++)
                for (int d = 0; d < 3; d++) {
                    const double L = box[3 * r + d];
                    if (!(L > 0)) invalid("box lengths must be positive");
                    const int nc = (int)(L / c->P.cutoff);
                    if (nc < 3) unsupported("fewer than 3 cells along an axis");
                    hr[r].L[d] = L;
                    hr[r].nc[d] = nc;
                    hr[r].cs[d] = L / nc;
                }
        }
        if (pos) {
            std::vector<Bead> hb(n);
            HIPCHK(hipMemcpyAsync(hb.data(), c->beads, sizeof(Bead) * n, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            for (int i = 0; i < n; i++) {
                const DevRep &R = hr[hb[i].rep];
                for (int d = 0; d < 3; d++) {
                    const double p = pos[3 * (size_t)i + d];
                    if (!(p >= 0 && p <= R.L[d])) invalid("particle " + std::to_string(i) + " position is out of bounds");
                }
                hb[i].x = pos[3 * (size_t)i];
                hb[i].y = pos[3 * (size_t)i + 1];
                hb[i].z = pos[3 * (size_t)i + 2];
            }
            HIPCHK(hipMemcpyAsync(c->beads, hb.data(), sizeof(Bead) * n, hipMemcpyHostToDevice, c->st));
        }
        auto up3 = [&](float4 *dst, const double *src) {
            std::vector<float4> h(n);
            for (int i = 0; i < n; i++)
                h[i] = make_float4((float)src[3 * (size_t)i], (float)src[3 * (size_t)i + 1], (float)src[3 * (size_t)i + 2], 0.f);
            HIPCHK(hipMemcpyAsync(dst, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        };
        if (vel) up3(c->vel, vel);
        if (acc) up3(c->acc, acc);
        c->write_reps(hr);
        c->ensure_capacity(0);
    });
}

int sm_set_temperature(sm_ctx *c, double T)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (!(T >= 0)) invalid("temperature must be >= 0");
        std::vector<DevRep> hr = c->read_reps();
        for (auto &R : hr) {
            R.temperature = T;
            if (!R.sigma_fixed) R.sigma = std::sqrt((6.0 * T * R.gamma) / c->P.dt);
        }
        c->write_reps(hr);
    });
}

int sm_init_forces(sm_ctx *c)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->build(0, 0);
        c->pair_force();
        c->bonded(KICK_INIT, nullptr);
        c->check_launch();
    });
}

int sm_second_half(sm_ctx *c)
{
    return guarded([&] {
        if (!c) invalid("null context");
        launch_second(c->st, c->n, c->beads, c->acc, c->vel, (float)(c->P.dt * 0.5));
        c->check_launch();
    });
}

int sm_step_first(sm_ctx *c, int64_t step)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (step < 0) invalid("negative step index");
        // the cell table must hold this step's box move before the keys are computed
        if (c->mc_step(step)) c->ensure_capacity(1);
        c->step_first(step);
        c->check_launch();
    });
}

int sm_step_finish(sm_ctx *c, int64_t step, double *measure)
{
    return guarded([&] {
        if (!c) invalid("null context");
        const bool mc = c->mc_step(step);
        c->step_finish(measure != nullptr, mc);
        c->check_launch();
        if (measure) {
            HIPCHK(hipMemcpyAsync(measure, c->meas_dev(), 3 * sizeof(double) * c->nrep, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            c->check_flags();
        }
    });
}

int sm_run(sm_ctx *c, int64_t first, int32_t nsteps)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (nsteps < 0 || first < 0) invalid("negative step index or count");
        c->run(first, nsteps);
        c->check_launch();
    });
}

int sm_mc_stats(sm_ctx *c, int64_t *trials, int64_t *accepted)
{
    return guarded([&] {
        if (!c) invalid("null context");
        std::vector<DevRep> hr = c->read_reps();
        for (int r = 0; r < c->nrep; r++) {
            if (trials) trials[r] = hr[r].trials;
            if (accepted) accepted[r] = hr[r].accepted;
        }
    });
}

int sm_mc_hist(sm_ctx *c, int32_t replica, int32_t cap, double *accepted, double *rejected, int32_t *n_out)
{
    return guarded([&] {
        if (!c || !n_out || cap < 0) invalid("bad argument");
        if (replica < 0 || replica >= c->nrep) invalid("replica index out of range");
        std::vector<DevRep> hr = c->read_reps();
        const DevRep &R = hr[replica];
        const double dL = R.delta_lxy;
        int nb = 0;
        if (dL != 0) nb = (int)(2.0 * dL / (dL / 101.0)) + 1;
        if (nb > kHistBins) nb = kHistBins;
        *n_out = nb;
        const int m = nb < cap ? nb : cap;
        for (int i = 0; i < m; i++) {
            if (accepted) accepted[i] = R.hist_accept[i];
            if (rejected) rejected[i] = R.hist_reject[i];
        }
    });
}

int sm_measure(sm_ctx *c, double *out)
{
    return guarded([&] {
        if (!c || !out) invalid("null argument");
        c->build(0, kStepKeep);
        c->measure_device();
        c->check_launch();
        HIPCHK(hipMemcpyAsync(out, c->meas_dev(), 3 * sizeof(double) * c->nrep, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->check_flags();
    });
}

int sm_bond_stats(sm_ctx *c, double *out)
{
    return guarded([&] {
        if (!c || !out) invalid("null argument");
        const int R = c->nrep;
        std::vector<double> h((size_t)4 * R, 0.0);
        if (c->gptr) {
            double *w = dalloc<double>((size_t)4 * R);
            try {
                for (int what = 0; what < 4; what++) {
                    launch_gen_stats(c->st, c->n, c->gen(), c->beads, c->reps, what, c->ebead);
                    launch_seg_sum(c->st, c->segB, c->ebead, w + what, 4);
                }
                c->check_launch();
                HIPCHK(hipMemcpyAsync(h.data(), w, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
                HIPCHK(hipStreamSynchronize(c->st));
            } catch (...) {
                dfree(w);
                throw;
            }
            dfree(w);
        }
        for (int r = 0; r < R; r++) {
            // the divisions on the host, as the reference writes them: lBond / nBond even
            // when nBond is 0; the bend sums divided only when nBend > 1
            const int nb = c->rh[r].nbond, ne = c->rh[r].nbend;
            double *o = out + 4 * (size_t)r;
            const double zero = 0.0;
            o[0] = h[4 * r] / (nb ? (double)nb : zero);
            o[1] = h[4 * r + 1];
            o[2] = h[4 * r + 2];
            o[3] = h[4 * r + 3];
            if (ne > 1)
                for (int k = 1; k < 4; k++) o[k] /= (double)ne;
        }
    });
}

int sm_vesicle_radius(sm_ctx *c, int32_t head_type, double *radius)
{
    return guarded([&] {
        if (!c || !radius) invalid("null argument");
        const int R = c->nrep;
        // [R][4] head sums, [R][3] centres, [R] distance sums
        double *w = dalloc<double>((size_t)8 * R);
        try {
            double *sums = w, *com = w + 4 * R, *dist = w + 7 * R;
            for (int what = 0; what < 4; what++) {
                launch_head_terms(c->st, c->n, c->beads, c->reps, head_type, what, nullptr, c->ebead);
                launch_seg_sum(c->st, c->segB, c->ebead, sums + what, 4);
            }
            launch_head_com(c->st, R, sums, com);
            launch_head_terms(c->st, c->n, c->beads, c->reps, head_type, 4, com, c->ebead);
            launch_seg_sum(c->st, c->segB, c->ebead, dist, 1);
            c->check_launch();
            std::vector<double> h((size_t)8 * R);
            HIPCHK(hipMemcpyAsync(h.data(), w, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            for (int r = 0; r < R; r++) radius[r] = h[7 * R + r] / h[4 * r + 3];
        } catch (...) {
            dfree(w);
            throw;
        }
        dfree(w);
    });
}

int sm_vesicle_radius_noalloc(sm_ctx *c, int32_t head_type, double *radius)
{
    return guarded([&] {
        if (!c || !radius) invalid("null argument");
        const int R = c->nrep;
        if (R > kNoAllocMaxReplicas)
            unsupported("sm_vesicle_radius_noalloc: " + std::to_string(R) + " replicas exceed its fixed scratch of " +
                        std::to_string(kNoAllocMaxReplicas) + "; use sm_vesicle_radius");
        double *w = noalloc_scratch(c->device);
        if (!w) throw ApiError(SM_ERR_STATE, "fixed scratch of this device was not set up");
        // same kernels and layout as sm_vesicle_radius: [R][4] head sums, [R][3] centres,
        // [R] distance sums; the radii then replace the centres.  Every word read is
        // written first by this call, so nothing is cleared.
        double *sums = w, *com = w + 4 * (size_t)R, *dist = w + 7 * (size_t)R;
        for (int what = 0; what < 4; what++) {
            launch_head_terms(c->st, c->n, c->beads, c->reps, head_type, what, nullptr, c->ebead);
            launch_seg_sum(c->st, c->segB, c->ebead, sums + what, 4);
        }
        launch_head_com(c->st, R, sums, com);
        launch_head_terms(c->st, c->n, c->beads, c->reps, head_type, 4, com, c->ebead);
        launch_seg_sum(c->st, c->segB, c->ebead, dist, 1);
        launch_head_radius(c->st, R, sums, dist, com);
        c->check_launch();
        HIPCHK(hipMemcpyAsync(radius, com, sizeof(double) * (size_t)R, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    });
}

int sm_compute_forces(sm_ctx *c, double *pair_out, double *chain_out)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->build(0, kStepKeep);
        c->pair_force();
        c->bonded(KICK_NONE, c->fscratch);
        c->check_launch();
        if (pair_out) download_f4(c, c->fpair, pair_out);
        if (chain_out) download_f4(c, c->fscratch, chain_out);
        c->check_flags();
    });
}

int sm_cell_keys(sm_ctx *c, int32_t *keys)
{
    return guarded([&] {
        if (!c || !keys) invalid("null argument");
        c->first_and_keys(0, kStepKeep, false);   // no sort follows: leave its histograms alone
        c->check_launch();
        std::vector<uint32_t> h(c->n);
        HIPCHK(hipMemcpyAsync(h.data(), c->k0, sizeof(uint32_t) * c->n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int r = 0; r < c->nrep; r++)
            for (int i = c->rh[r].first; i < c->rh[r].first + c->rh[r].count; i++)
                keys[i] = (int32_t)((long long)h[i] - c->rh[r].cellbase);
        c->check_flags();
    });
}

int sm_pair_count(sm_ctx *c, int32_t *counts, int64_t *total)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->build(0, kStepKeep);
        launch_pair_gather(c->st, c->P, c->chunk, GATHER_COUNT, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend,
                           c->reps, c->fconstF, c->uconstD, nullptr, c->dcounts, nullptr, nullptr, 0);
        c->check_launch();
        std::vector<int32_t> h(c->n);
        HIPCHK(hipMemcpyAsync(h.data(), c->dcounts, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        long long t = 0;
        for (int i = 0; i < c->n; i++) t += h[i];
        if (counts) std::memcpy(counts, h.data(), sizeof(int32_t) * c->n);
        if (total) *total = t / 2;
        c->check_flags();
    });
}

int sm_force_pair_count(sm_ctx *c, int32_t *counts, int64_t *total)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (c->pair_variant != 0) unsupported("the force pair count follows the one-phase force kernel (pair_variant 0)");
        c->build(0, kStepKeep);
        launch_pair_gather(c->st, c->P, c->chunk, GATHER_FCOUNT, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend,
                           c->reps, c->fconstF, c->fconstD, nullptr, c->dcounts, nullptr, nullptr, 0);
        c->check_launch();
        std::vector<int32_t> h(c->n);
        HIPCHK(hipMemcpyAsync(h.data(), c->dcounts, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        long long t = 0;
        for (int i = 0; i < c->n; i++) t += h[i];
        if (counts) std::memcpy(counts, h.data(), sizeof(int32_t) * c->n);
        if (total) *total = t / 2;
        c->check_flags();
    });
}

int sm_pair_list(sm_ctx *c, int32_t *pairs, int64_t cap, int64_t *n_out)
{
    return guarded([&] {
        if (!c || cap < 0) invalid("bad argument");
        if (cap > c->pairs_cap) {
            dfree(c->dpairs);
            c->dpairs = dalloc<int>(2 * (size_t)cap);
            c->pairs_cap = cap;
        }
        c->build(0, kStepKeep);
        HIPCHK(hipMemsetAsync(c->dnp, 0, sizeof(unsigned long long), c->st));
        launch_pair_gather(c->st, c->P, c->chunk, GATHER_LIST, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend,
                           c->reps, c->fconstF, c->uconstD, nullptr, c->dcounts, c->dpairs, c->dnp, cap);
        c->check_launch();
        unsigned long long np = 0;
        HIPCHK(hipMemcpyAsync(&np, c->dnp, sizeof np, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        const long long keep = std::min<long long>((long long)np, cap);
        if (pairs && keep > 0) {
            HIPCHK(hipMemcpyAsync(pairs, c->dpairs, sizeof(int32_t) * 2 * (size_t)keep, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        }
        if (n_out) *n_out = (int64_t)np;
        c->check_flags();
    });
}

int sm_potential_delta(sm_ctx *c, const double *scale, double *du_pair, double *du_chain)
{
    return guarded([&] {
        if (!c || !scale) invalid("null argument");
        std::vector<DevRep> hr = c->read_reps();
        for (int r = 0; r < c->nrep; r++)
            for (int d = 0; d < 3; d++) hr[r].scale[d] = scale[3 * r + d];
        c->write_reps(hr);
        c->build(0, kStepKeep);
        if (c->mc_fused) {   // the traversal a box-move step runs (it also leaves the pair forces)
            c->pair_force_du();
            c->delta_bonded();
        } else {
            c->delta_device();
        }
        c->check_launch();
        std::vector<double> h(2 * (size_t)c->nrep);
        HIPCHK(hipMemcpyAsync(h.data(), c->du_dev(), sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int r = 0; r < c->nrep; r++) {
            if (du_pair) du_pair[r] = h[2 * r];
            if (du_chain) du_chain[r] = h[2 * r + 1];
        }
        c->check_flags();
    });
}

int sm_tune(sm_ctx *c, const char *knob, int32_t value)
{
    return guarded([&] {
        if (!c || !knob) invalid("null argument");
        if (std::strcmp(knob, "pair_chunk") == 0) {
            if (value < 1 || value > 4096) invalid("pair_chunk must be in [1, 4096]");
            c->chunk = value;
        } else if (std::strcmp(knob, "pair_variant") == 0) {
            if (value < 0 || value > 1) invalid("pair_variant must be 0 or 1");
            c->pair_variant = value;
        } else if (std::strcmp(knob, "graph") == 0) {
            if (value < 0 || value > 1) invalid("graph must be 0 or 1");
            c->use_graph = value != 0;
        } else if (std::strcmp(knob, "sort_mode") == 0) {
            if (value < 0 || value > 2) invalid("sort_mode must be 0, 1 or 2");
            if (value == 2 && !c->win_ok) invalid("sort_mode 2 needs a single replica of at most 524288 beads");
            c->sort_mode = value;
            c->order_valid = false;
        } else if (std::strcmp(knob, "mc_fused") == 0) {
            if (value < 0 || value > 1) invalid("mc_fused must be 0 or 1");
            c->mc_fused = value != 0;
        } else {
            invalid(std::string("unknown knob ") + knob);
        }
        c->drop_graph();
    });
}

int sm_tune_get(sm_ctx *c, const char *knob, int32_t *value)
{
    return guarded([&] {
        if (!c || !knob || !value) invalid("null argument");
        if (std::strcmp(knob, "pair_chunk") == 0) *value = c->chunk;
        else if (std::strcmp(knob, "pair_variant") == 0) *value = c->pair_variant;
        else if (std::strcmp(knob, "graph") == 0) *value = c->use_graph ? 1 : 0;
        else if (std::strcmp(knob, "sort_mode") == 0) *value = c->sort_mode;
        else if (std::strcmp(knob, "mc_fused") == 0) *value = c->mc_fused ? 1 : 0;
        else if (std::strcmp(knob, "sort_window_path") == 0) {
            *value = 0;
            if (c->win_last) {
                // the last sort's window and fallback flag: its scratch is cleared only by the
                // next sort's reorder
                const int lp = 1 - c->sort_parity;
                uint32_t wx = 0, flag = 0;
                HIPCHK(hipMemcpyAsync(&wx, c->wdesc + 8 * lp + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, c->st));
                HIPCHK(hipMemcpyAsync(&flag, win_book(c->hist[lp], c->n, nullptr).flag, sizeof(uint32_t),
                                      hipMemcpyDeviceToHost, c->st));
                HIPCHK(hipStreamSynchronize(c->st));
                *value = (wx != 0 && flag == 0) ? 1 : 0;
            }
        } else invalid(std::string("unknown knob ") + knob);
    });
}

int sm_autotune(sm_ctx *c, int32_t steps, double *best_ms_per_step)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (steps < 8 || steps > 100000) invalid("steps per trial must be in [8, 100000]");
        if (c->prof) invalid("profiling is on");
        const long long s0 = c->read_glob().step + 1;
        // every trial starts from the same state; reserve cells for the trials' moves
        // first so that no trial moves the cell bases
        long long moves = 0;
        for (long long s = s0; s < s0 + steps; s++)
            if (c->mc_step(s)) moves++;
        c->ensure_capacity(moves);
        const int n = c->n;
        Bead *b0 = dalloc<Bead>(n);
        float4 *v0 = dalloc<float4>(n), *a0 = dalloc<float4>(n);
        DevRep *r0 = dalloc<DevRep>(c->nrep);
        DevGlobal *g0 = dalloc<DevGlobal>(1);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto cleanup = [&] {
            dfree(b0); dfree(v0); dfree(a0); dfree(r0); dfree(g0);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        };
        try {
            hipStream_t st = c->st;
            auto copy = [&](void *d, const void *s, size_t bytes) {
                HIPCHK(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, st));
            };
            auto save = [&] {
                copy(b0, c->beads, sizeof(Bead) * n); copy(v0, c->vel, sizeof(float4) * n);
                copy(a0, c->acc, sizeof(float4) * n); copy(r0, c->reps, sizeof(DevRep) * c->nrep);
                copy(g0, c->glob, sizeof(DevGlobal));
            };
            auto restore = [&] {
                copy(c->beads, b0, sizeof(Bead) * n); copy(c->vel, v0, sizeof(float4) * n);
                copy(c->acc, a0, sizeof(float4) * n); copy(c->reps, r0, sizeof(DevRep) * c->nrep);
                copy(c->glob, g0, sizeof(DevGlobal));
            };
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            save();
            // launch shapes that leave every result bit-identical: the chunk of the one-phase
            // pair kernel (a cell is processed the same way whatever chunk holds it), the
            // graph replay (the same kernels on the same stream) and the sort mode (both
            // give the same order)
            struct Cand { int chunk, variant, graph, sort; };
            std::vector<Cand> cands;
            for (int ch : {2, 4, 8, 16, 32, 64}) cands.push_back({ch, 0, 1, c->sort_mode});
            double best = 1e300;
            Cand bestc{c->chunk, c->pair_variant, c->use_graph ? 1 : 0, c->sort_mode};
            auto trial = [&](const Cand &k) {
                c->chunk = k.chunk;
                c->pair_variant = k.variant;
                c->use_graph = k.graph != 0;
                if (c->sort_mode != k.sort) {
                    c->sort_mode = k.sort;
                    c->order_valid = false;
                }
                c->drop_graph();
                restore();
                c->run(s0, 8);          // warm: first launches, graph capture
                restore();
                HIPCHK(hipEventRecord(e0, st));
                c->run(s0, steps);
                HIPCHK(hipEventRecord(e1, st));
                HIPCHK(hipEventSynchronize(e1));
                c->check_launch();
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                const double per = ms / steps;
                if (per < best) {
                    best = per;
                    bestc = k;
                }
            };
            for (const auto &k : cands) trial(k);
            const Cand b0 = bestc;
            for (int m = 0; m <= 2; m++)   // the other sort modes
                if (m != b0.sort && (m != 2 || c->win_ok)) trial({b0.chunk, b0.variant, 1, m});
            trial({bestc.chunk, bestc.variant, 0, bestc.sort});       // is the graph worth it here
            c->chunk = bestc.chunk;
            c->pair_variant = bestc.variant;
            c->use_graph = bestc.graph != 0;
            if (c->sort_mode != bestc.sort) {
                c->sort_mode = bestc.sort;
                c->order_valid = false;
            }
            c->drop_graph();
            restore();
            HIPCHK(hipStreamSynchronize(st));
            c->check_flags();
            if (best_ms_per_step) *best_ms_per_step = best;
        } catch (...) {
            (void)hipStreamSynchronize(c->st);
            cleanup();
            throw;
        }
        cleanup();
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// interaction constants as adjoint parameters
// ---------------------------------------------------------------------------------------
namespace {

void check_pair(const sm_ctx *c, int a, int b)
{
    if (a < 0 || b < 0 || a >= c->P.nT || b >= c->P.nT) invalid("type index out of range");
}

// laradjiRevaleeFC / laradjiRevaleePC rows (potentials/laradjiRevalee.h:84-152) of the
// type pair (a, b) and (b, a) from Umin, Umax, keeping each row's r_m and r_c
void set_pair_rows(sm_ctx *c, int a, int b, double umin, double umax)
{
    const int nT = c->P.nT;
    for (int pass = 0; pass < 2; pass++) {
        const size_t row = 6 * (size_t)(pass ? b * nT + a : a * nT + b);
        double *F = &c->hfconst[row], *U = &c->huconst[row];
        const double rm = F[0], rc = F[3];
        F[1] = (2.0 * (umax - umin)) / (rm * rm);
        F[2] = 0;
        F[4] = (6.0 * umin) / ((rc - rm) * (rc - rm));
        F[5] = (6.0 * umin) / ((rc - rm) * (rc - rm) * (rc - rm));
        U[0] = rm;
        U[1] = (umax - umin) / (rm * rm);
        U[2] = umin;
        U[3] = rc;
        U[4] = (3.0 * umin) / ((rc - rm) * (rc - rm));
        U[5] = (2.0 * umin) / ((rc - rm) * (rc - rm) * (rc - rm));
    }
}

void upload_constants(sm_ctx *c)
{
    const size_t nt = c->hfconst.size();
    // a captured graph holds the pair kernels of the tables' symmetry
    const int asym = tables_asymmetric(c->hfconst, c->huconst, c->P.nT) ? 1 : 0;
    if (asym != c->P.asym) {
        c->P.asym = asym;
        c->drop_graph();
    }
    const std::vector<float> fF = pair_table(c->hfconst, c->huconst, c->P.nT, c->P.cutoff);
    HIPCHK(hipMemcpyAsync(c->fconstF, fF.data(), sizeof(float) * fF.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->uconstD, c->huconst.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->fconstD, c->hfconst.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->st));
    if (!c->hcconst.empty())
        HIPCHK(hipMemcpyAsync(c->cconstD, c->hcconst.data(), sizeof(double4) * c->hcconst.size(),
                              hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
}

// FP64 F table and the d/d(Umin), d/d(Umax) tables of the (a, b) rows
void upload_adjoint_tables(sm_ctx *c, int a, int b)
{
    const int nT = c->P.nT;
    const size_t nt = 6 * (size_t)nT * nT;
    std::vector<double> dF(2 * nt, 0.0), dU(2 * nt, 0.0);
    for (int pass = 0; pass < 2; pass++) {
        const size_t row = 6 * (size_t)(pass ? b * nT + a : a * nT + b);
        const double rm = c->hfconst[row], rc = c->hfconst[row + 3];
        const double w = rc - rm;
        double *Fmin = &dF[row], *Fmax = &dF[nt + row], *Umin = &dU[row], *Umax = &dU[nt + row];
        Fmin[1] = -2.0 / (rm * rm);
        Fmin[4] = 6.0 / (w * w);
        Fmin[5] = 6.0 / (w * w * w);
        Fmax[1] = 2.0 / (rm * rm);
        Umin[1] = -1.0 / (rm * rm);
        Umin[2] = 1.0;
        Umin[4] = 3.0 / (w * w);
        Umin[5] = 2.0 / (w * w * w);
        Umax[1] = 1.0 / (rm * rm);
    }
    if (!c->adjF) {
        c->adjF = dalloc<double>(nt);
        c->adjDF = dalloc<double>(2 * nt);
        c->adjDU = dalloc<double>(2 * nt);
    }
    HIPCHK(hipMemcpyAsync(c->adjF, c->hfconst.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->adjDF, dF.data(), sizeof(double) * 2 * nt, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->adjDU, dU.data(), sizeof(double) * 2 * nt, hipMemcpyHostToDevice, c->st));
}

// The tables sm_param_derivative runs the energy kernels with.  U is linear in each
// laradjiRevaleePC row's entries {r_m, C1, C2, r_c, C4, C5} with r_m and r_c fixed
// (laradjiRevalee.h:124-152), and those entries are linear in Umin and Umax, so the FP64
// potential gather run on a table whose (a, b) and (b, a) rows hold the entries'
// derivatives (every row keeping its r_m and r_c, every other coefficient 0) sums
// dU_pair/dUmin or dU_pair/dUmax.  harmonicP and bendP are linear in k_bond and k_bend
// (MD.h:422-431,769-789), so the CHAIN energy with (k_bond, k_bend) = (1, 0) or (0, 1) is
// dU_chain/dk_bond or dU_chain/dk_bend.  Depends only on r_m, r_c, l0 and the cos-theta
// term, which the fitting parameters leave alone; rebuilt when the pair changes.
void upload_param_derivative_tables(sm_ctx *c, int a, int b)
{
    if (c->pdU && c->pd_a == a && c->pd_b == b) return;
    const int nT = c->P.nT;
    const size_t nt = 6 * (size_t)nT * nT;
    std::vector<double> dU(2 * nt, 0.0);
    for (size_t row = 0; row < nt; row += 6)
        for (int p = 0; p < 2; p++) {
            dU[p * nt + row] = c->huconst[row];
            dU[p * nt + row + 3] = c->huconst[row + 3];
        }
    for (int pass = 0; pass < 2; pass++) {
        const size_t row = 6 * (size_t)(pass ? b * nT + a : a * nT + b);
        const double rm = c->huconst[row], rc = c->huconst[row + 3];
        const double w = rc - rm;
        double *Umin = &dU[row], *Umax = &dU[nt + row];
        Umin[1] = -1.0 / (rm * rm);
        Umin[2] = 1.0;
        Umin[4] = 3.0 / (w * w);
        Umin[5] = 2.0 / (w * w * w);
        Umax[1] = 1.0 / (rm * rm);
    }
    std::vector<double> dUmin(dU.begin(), dU.begin() + nt), dUmax(dU.begin() + nt, dU.end());
    const std::vector<float> t0 = pair_table(c->hfconst, dUmin, nT, c->P.cutoff);
    const std::vector<float> t1 = pair_table(c->hfconst, dUmax, nT, c->P.cutoff);
    std::vector<double4> cc;
    for (const double4 &k : c->hcconst) cc.push_back(make_double4(k.x, 1.0, k.z, 0.0));
    for (const double4 &k : c->hcconst) cc.push_back(make_double4(k.x, 0.0, k.z, 1.0));
    if (!c->pdU) {
        c->pdU = dalloc<double>(2 * nt);
        c->pd_tab = t0.size();
        c->pdTab = dalloc<float>(2 * c->pd_tab);
        c->pdC = dalloc<double4>(cc.size());
    }
    HIPCHK(hipMemcpyAsync(c->pdU, dU.data(), sizeof(double) * 2 * nt, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pdTab, t0.data(), sizeof(float) * c->pd_tab, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->pdTab + c->pd_tab, t1.data(), sizeof(float) * c->pd_tab, hipMemcpyHostToDevice, c->st));
    if (!cc.empty())
        HIPCHK(hipMemcpyAsync(c->pdC, cc.data(), sizeof(double4) * cc.size(), hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->pd_a = a;
    c->pd_b = b;
}

// device copy of the evolving state
struct StateCopy {
    Bead *beads = nullptr;
    float4 *vel = nullptr, *acc = nullptr;
    DevRep *reps = nullptr;
    DevGlobal *glob = nullptr;
    void alloc(const sm_ctx *c)
    {
        beads = dalloc<Bead>(c->n);
        vel = dalloc<float4>(c->n);
        acc = dalloc<float4>(c->n);
        reps = dalloc<DevRep>(c->nrep);
        glob = dalloc<DevGlobal>(1);
    }
    void release() { dfree(beads); dfree(vel); dfree(acc); dfree(reps); dfree(glob); }
    void save(const sm_ctx *c)
    {
        HIPCHK(hipMemcpyAsync(beads, c->beads, sizeof(Bead) * c->n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(vel, c->vel, sizeof(float4) * c->n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(acc, c->acc, sizeof(float4) * c->n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(reps, c->reps, sizeof(DevRep) * c->nrep, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(glob, c->glob, sizeof(DevGlobal), hipMemcpyDeviceToDevice, c->st));
    }
    void load(sm_ctx *c) const
    {
        HIPCHK(hipMemcpyAsync(c->beads, beads, sizeof(Bead) * c->n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->vel, vel, sizeof(float4) * c->n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->acc, acc, sizeof(float4) * c->n, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->reps, reps, sizeof(DevRep) * c->nrep, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->glob, glob, sizeof(DevGlobal), hipMemcpyDeviceToDevice, c->st));
    }
};

}  // namespace

extern "C" {

int sm_set_params(sm_ctx *c, int32_t type_a, int32_t type_b, const double *theta)
{
    return guarded([&] {
        if (!c || !theta) invalid("null argument");
        check_pair(c, type_a, type_b);
        set_pair_rows(c, type_a, type_b, theta[0], theta[1]);
        for (auto &k : c->hcconst) {
            k.y = theta[2];
            k.w = theta[3];
        }
        upload_constants(c);
    });
}

int sm_get_params(sm_ctx *c, int32_t type_a, int32_t type_b, double *theta)
{
    return guarded([&] {
        if (!c || !theta) invalid("null argument");
        check_pair(c, type_a, type_b);
        const double *F = &c->hfconst[6 * (size_t)(type_a * c->P.nT + type_b)];
        const double rm = F[0], rc = F[3];
        const double umin = F[4] * (rc - rm) * (rc - rm) / 6.0;
        theta[0] = umin;
        theta[1] = F[1] * rm * rm / 2.0 + umin;
        theta[2] = c->hcconst.empty() ? 0.0 : c->hcconst[0].y;
        theta[3] = c->hcconst.empty() ? 0.0 : c->hcconst[0].w;
    });
}

int sm_param_derivative(sm_ctx *c, int32_t type_a, int32_t type_b, double *out)
{
    return guarded([&] {
        if (!c || !out) invalid("null argument");
        check_pair(c, type_a, type_b);
        upload_param_derivative_tables(c, type_a, type_b);
        c->build(0, kStepKeep);
        const int R = c->nrep;
        const size_t nt = 6 * (size_t)c->P.nT * c->P.nT;
        double *w = c->dres;   // 5 words per replica: the four derivatives, then free
        for (int p = 0; p < 2; p++) {
            launch_pair_gather(c->st, c->P, c->chunk, GATHER_U, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend,
                               c->reps, c->pdTab + p * c->pd_tab, c->pdU + p * nt, c->ebead, nullptr, nullptr, nullptr,
                               0);
            launch_seg_sum(c->st, c->segB, c->ebead, w + p, 5);
        }
        const size_t ncs = c->hcconst.size();
        // the measurement words are free between calls (each measurement copies its own out)
        if (c->nchains > 0)
            for (int p = 0; p < 2; p++) {
                launch_chain_energy(c->st, c->nchains, c->chainList, c->pdC + p * ncs, c->beads, c->reps, 0, c->echain);
                launch_seg_sum(c->st, c->segC, c->echain, w + 2 + p, 5);
            }
        c->check_launch();
        std::vector<double> h((size_t)5 * R);
        HIPCHK(hipMemcpyAsync(h.data(), w, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int r = 0; r < R; r++)
            for (int p = 0; p < 4; p++) out[4 * (size_t)r + p] = (p >= 2 && c->nchains == 0) ? 0.0 : h[5 * r + p];
        c->check_flags();
    });
}

int sm_adjoint(sm_ctx *c, int32_t type_a, int32_t type_b, int32_t nsteps, int32_t checkpoint, double *J,
               double *grad)
{
    return guarded([&] {
        if (!c || !J) invalid("null argument");
        if (nsteps < 1 || checkpoint < 1 || checkpoint > nsteps) invalid("need 1 <= checkpoint <= nsteps");
        const int R = c->nrep;
        check_pair(c, type_a, type_b);
        const int n = c->n, S = nsteps, k = checkpoint;
        const double obj = 1.0 / S;
        long long moves = 0;
        for (long long s = 1; s <= S; s++)
            if (c->mc_step(s)) moves++;
        c->ensure_capacity(moves);
        upload_adjoint_tables(c, type_a, type_b);

        const int ncp = (S + k - 1) / k;   // checkpoints at steps 0, k, 2k, ...
        // the forward run and the recomputed segments take their box-move dU the same way
        // (the gather after the kick), so they make the same decisions
        struct KeepFused {
            sm_ctx *c;
            bool v;
            ~KeepFused() { c->mc_fused = v; }
        } keep{c, c->mc_fused};
        c->mc_fused = false;
        StateCopy entry;
        std::vector<StateCopy> cps(ncp);
        Bead *segBeads = nullptr;
        DevRep *segPre = nullptr, *segPost = nullptr;
        double4 *X = nullptr, *V = nullptr, *A = nullptr, *F = nullptr, *HV = nullptr;
        double *g = nullptr, *gs = nullptr;
        auto cleanup = [&] {
            entry.release();
            for (auto &cp : cps) cp.release();
            dfree(segBeads); dfree(segPre); dfree(segPost);
            dfree(X); dfree(V); dfree(A); dfree(F); dfree(HV); dfree(g); dfree(gs);
        };
        try {
            hipStream_t st = c->st;
            entry.alloc(c);
            entry.save(c);
            for (auto &cp : cps) cp.alloc(c);

            // forward: init forces, then steps 1..S, J = mean of U after each kick
            c->build(0, 0);
            c->pair_force();
            c->bonded(KICK_INIT, nullptr);
            cps[0].save(c);
            std::vector<double> Jsum(R, 0.0), m(3 * (size_t)R);
            for (int s = 1; s <= S; s++) {
                c->step_first(s);
                c->step_finish(true, c->mc_step(s));
                HIPCHK(hipMemcpyAsync(m.data(), c->meas_dev(), sizeof(double) * m.size(), hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                for (int r = 0; r < R; r++) Jsum[r] += m[3 * r] + m[3 * r + 1];
                if (s % k == 0 && s / k < ncp) cps[s / k].save(c);
            }
            c->check_launch();
            c->check_flags();
            for (int r = 0; r < R; r++) J[r] = Jsum[r] * obj;

            if (grad) {
                segBeads = dalloc<Bead>((size_t)k * n);
                segPre = dalloc<DevRep>((size_t)k * c->nrep);
                segPost = dalloc<DevRep>((size_t)k * c->nrep);
                X = dalloc<double4>(n); V = dalloc<double4>(n); A = dalloc<double4>(n);
                F = dalloc<double4>(n); HV = dalloc<double4>(n);
                g = dalloc<double>(kNTheta * (size_t)n);
                gs = dalloc<double>(kNTheta * (size_t)R);
                for (double4 *p : {X, V, A, F, HV}) HIPCHK(hipMemsetAsync(p, 0, sizeof(double4) * n, st));
                HIPCHK(hipMemsetAsync(g, 0, sizeof(double) * kNTheta * n, st));
                const AdjTables T{c->adjF, c->adjDF, c->adjDU};
                auto adjoint_forces = [&](double o, int hv) {
                    launch_adj_pair(st, c->P, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend, c->reps, T, A, F,
                                    HV, g, o, hv);
                    launch_adj_chain(st, c->P, c->beads, c->chainPerBead, c->cconstD, c->reps, A, F, HV, g, o, hv);
                    if (c->gptr) launch_adj_gen(st, c->P, c->gen(), c->beads, c->reps, A, F, HV, hv);
                };
                for (int seg = ncp - 1; seg >= 0; seg--) {
                    const int s0 = seg * k, s1 = std::min(S, s0 + k);
                    // recompute the segment, keeping each step's positions before its MC move
                    cps[seg].load(c);
                    for (int s = s0 + 1; s <= s1; s++) {
                        const int t = s - s0 - 1;
                        c->step_first(s);
                        c->step_finish(false, false);
                        HIPCHK(hipMemcpyAsync(segBeads + (size_t)t * n, c->beads, sizeof(Bead) * n,
                                              hipMemcpyDeviceToDevice, st));
                        HIPCHK(hipMemcpyAsync(segPre + (size_t)t * c->nrep, c->reps, sizeof(DevRep) * c->nrep,
                                              hipMemcpyDeviceToDevice, st));
                        if (c->mc_step(s)) c->mc_move();
                        HIPCHK(hipMemcpyAsync(segPost + (size_t)t * c->nrep, c->reps, sizeof(DevRep) * c->nrep,
                                              hipMemcpyDeviceToDevice, st));
                    }
                    for (int s = s1; s > s0; s--) {
                        const int t = s - s0 - 1;
                        HIPCHK(hipMemcpyAsync(c->beads, segBeads + (size_t)t * n, sizeof(Bead) * n,
                                              hipMemcpyDeviceToDevice, st));
                        HIPCHK(hipMemcpyAsync(c->reps, segPre + (size_t)t * c->nrep, sizeof(DevRep) * c->nrep,
                                              hipMemcpyDeviceToDevice, st));
                        c->build(0, kStepKeep);
                        adjoint_forces(obj, 1);
                        launch_adj_step(st, c->P, c->beads, segPost + (size_t)t * c->nrep, c->mc_step(s) ? 1 : 0,
                                        X, V, A, F, HV, obj, s - 1 >= 1 ? 1 : 0);
                    }
                }
                // a_0 = F(x_0) + ...: its adjoint meets dF/dtheta at the initial positions
                cps[0].load(c);
                c->build(0, kStepKeep);
                adjoint_forces(0.0, 0);
                // per replica: grad[r * kNTheta + p]
                for (int p = 0; p < kNTheta; p++) launch_seg_sum(st, c->segB, g + (size_t)p * n, gs + p, kNTheta);
                c->check_launch();
                HIPCHK(hipMemcpyAsync(grad, gs, sizeof(double) * kNTheta * R, hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
            }
            entry.load(c);
            HIPCHK(hipStreamSynchronize(st));
            c->drop_graph();
        } catch (...) {
            (void)hipStreamSynchronize(c->st);
            cleanup();
            throw;
        }
        cleanup();
    });
}

int sm_profile_enable(sm_ctx *c, int32_t on)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->harvest();
        c->prof = on != 0;
        for (int k = 0; k < K_NCLASS; k++) { c->prof_ms[k] = 0; c->prof_cnt[k] = 0; }
    });
}

int sm_profile_read(sm_ctx *c, int32_t cap, const char **names, double *avg_ms, int64_t *count, int32_t *n_out)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->harvest();
        int m = 0;
        for (int k = 0; k < K_NCLASS && m < cap; k++, m++) {
            if (names) names[m] = kClassName[k];
            if (avg_ms) avg_ms[m] = c->prof_cnt[k] ? c->prof_ms[k] / (double)c->prof_cnt[k] : 0.0;
            if (count) count[m] = c->prof_cnt[k];
        }
        if (n_out) *n_out = m;
    });
}

}  // extern "C"
