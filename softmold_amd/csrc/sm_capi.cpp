// Host side of the MI355X MD-step solver: device context, step sequencing and the C ABI
// declared in include/softmold_hip.h.
#include "../../include/softmold_hip.h"
#include "mpd_io.hpp"
#include "sm_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

using namespace smdev;

namespace smdev {
void launch_reduce_sum(hipStream_t st, const double *partial, int count, double *out);
}

namespace {

thread_local std::string g_err;

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ApiError(SM_ERR_HIP, std::string(hipGetErrorName(e_)) + " in " #expr " at " __FILE__ \
                                       ":" + std::to_string(__LINE__));                               \
    } while (0)

template <class F>
int guarded(F &&f)
{
    try {
        f();
        return SM_OK;
    } catch (const ApiError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return SM_ERR_INVALID;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SM_ERR_INVALID;
    }
}

[[noreturn]] void invalid(const std::string &m) { throw ApiError(SM_ERR_INVALID, m); }
[[noreturn]] void unsupported(const std::string &m) { throw ApiError(SM_ERR_UNSUPPORTED, m); }

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

enum KClass { K_FIRST, K_SORT, K_REORDER, K_PAIR, K_BONDED, K_GATHER, K_CHAIN_E, K_KIN, K_MC, K_NCLASS };
const char *kClassName[K_NCLASS] = {"integrate_first+key", "radix_sort", "reorder+cell_table", "pair_force",
                                    "bonded+langevin+kick", "pair_gather_fp64", "chain_energy_fp64",
                                    "kinetic_fp64", "mc_move"};

}  // namespace

struct sm_ctx {
    hipStream_t st = nullptr;
    Params P{};
    int n = 0;
    int nchains = 0;
    double box0[3] = {0, 0, 0};
    // device state
    Bead *beads = nullptr, *sbeads = nullptr;
    float4 *vel = nullptr, *acc = nullptr, *fpair = nullptr, *fscratch = nullptr;
    uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr, *hist = nullptr;
    int *cstart = nullptr, *cend = nullptr;
    long long capacity = 0;
    int sort_bits = 0;
    DevBox *box = nullptr;
    float *fconstF = nullptr;
    double *uconstD = nullptr;
    int2 *chainPerBead = nullptr, *chainList = nullptr;
    float4 *cconstF = nullptr;
    double4 *cconstD = nullptr;
    double *partA = nullptr, *partB = nullptr, *partC = nullptr, *dres = nullptr;
    int *dcounts = nullptr, *dpairs = nullptr;
    long long pairs_cap = 0;
    unsigned long long *dnp = nullptr;
    int chunk = 64;
    int pair_variant = 0;
    int sorted_sel = 0;  // which sort buffer pair holds the sorted result
    // profiling
    bool prof = false;
    struct Rec { int cls; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double prof_ms[K_NCLASS] = {};
    long long prof_cnt[K_NCLASS] = {};

    const uint32_t *skey() const { return sorted_sel ? k1 : k0; }
    const uint32_t *sidx() const { return sorted_sel ? v1 : v0; }

    ~sm_ctx()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (auto &r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
        dfree(beads); dfree(sbeads); dfree(vel); dfree(acc); dfree(fpair); dfree(fscratch);
        dfree(k0); dfree(v0); dfree(k1); dfree(v1); dfree(hist); dfree(cstart); dfree(cend); dfree(box);
        dfree(fconstF); dfree(uconstD); dfree(chainPerBead); dfree(chainList); dfree(cconstF); dfree(cconstD);
        dfree(partA); dfree(partB); dfree(partC); dfree(dres); dfree(dcounts); dfree(dpairs); dfree(dnp);
        if (st) (void)hipStreamDestroy(st);
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

    void check_launch()
    {
        HIPCHK(hipGetLastError());
    }

    DevBox read_box()
    {
        DevBox hb;
        HIPCHK(hipMemcpyAsync(&hb, box, sizeof(DevBox), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return hb;
    }

    void check_flags()
    {
        DevBox hb = read_box();
        if (hb.flags & FLAG_BAD_POSITION)
            throw ApiError(SM_ERR_STATE, "a particle position left the box or became non-finite");
        if (hb.flags & FLAG_SORT_TIMEOUT)
            throw ApiError(SM_ERR_STATE, "cell-key sort look-back timed out");
        if (hb.flags & FLAG_CELL_OVERFLOW)
            throw ApiError(SM_ERR_STATE, "an accepted box move would exceed the cell table or make fewer than 3 cells per axis");
    }

    // Dense cell table sized for every box the next `moves` MC moves can reach.
    void ensure_capacity(long long moves)
    {
        DevBox hb = read_box();
        const double rc = P.cutoff, dL = P.delta_lxy;
        const double grow = (double)moves * std::fabs(dL);
        double Lmax[3], Lmin[2];
        for (int d = 0; d < 2; d++) {
            Lmax[d] = hb.L[d] + grow;
            Lmin[d] = std::max(hb.L[d] - grow, 1e-300);
        }
        Lmax[2] = (moves > 0 && dL != 0) ? hb.L[2] * (hb.L[0] * hb.L[1]) / (Lmin[0] * Lmin[1]) : hb.L[2];
        long long need = 1;
        for (int d = 0; d < 3; d++) need *= (long long)(Lmax[d] / rc) + 1;
        if (need <= capacity) return;
        long long cap = need + need / 8;
        dfree(cstart);
        dfree(cend);
        cstart = dalloc<int>((size_t)cap);
        cend = dalloc<int>((size_t)cap);
        HIPCHK(hipMemsetAsync(cstart, 0xFF, sizeof(int) * (size_t)cap, st));
        HIPCHK(hipMemsetAsync(cend, 0xFF, sizeof(int) * (size_t)cap, st));
        capacity = cap;
        HIPCHK(hipMemcpyAsync(&box->capacity, &capacity, sizeof(long long), hipMemcpyHostToDevice, st));
        int bits = 1;
        while ((1ll << bits) < cap) bits++;
        sort_bits = bits;
    }

    // keys (optionally after Verlet::first), sort, reorder + dense cell table
    void first_and_keys(int integrate)
    {
        timed(K_FIRST, [&] { launch_first(st, P, beads, vel, acc, box, k0, v0, integrate); });
    }
    void sort_and_reorder()
    {
        timed(K_SORT, [&] {
            SortBuffers sb{k0, v0, k1, v1, hist, &box->flags, n, 0};
            sorted_sel = radix_sort(st, sb, sort_bits);
        });
        timed(K_REORDER, [&] { launch_reorder(st, n, skey(), sidx(), beads, sbeads, cstart, cend); });
    }
    void build(int integrate)
    {
        first_and_keys(integrate);
        sort_and_reorder();
    }
    void pair_force()
    {
        timed(K_PAIR, [&] {
            launch_pair_force(st, P, chunk, pair_variant, sbeads, skey(), sidx(), cstart, cend, box, fconstF, fpair);
        });
    }
    void bonded(long long step, int mode, float4 *fchain_out)
    {
        timed(K_BONDED, [&] {
            launch_bonded_kick(st, P, beads, chainPerBead, cconstD, fpair, vel, acc, box, step, mode, fchain_out);
        });
    }
    // {U_pair, U_chain, K} into dres[0..2] (device)
    void measure_device()
    {
        timed(K_GATHER, [&] {
            launch_pair_gather(st, P, chunk, GATHER_U, sbeads, skey(), sidx(), cstart, cend, box, uconstD, partA, nullptr,
                               nullptr, nullptr, 0);
        });
        timed(K_CHAIN_E, [&] { launch_chain_energy(st, nchains, chainList, cconstD, beads, box, 0, partB); });
        timed(K_KIN, [&] { launch_kinetic(st, n, vel, partC); });
        launch_reduce_sum(st, partA, cells_blocks(n, chunk), dres + 0);
        launch_reduce_sum(st, partB, chain_blocks(nchains), dres + 1);
        launch_reduce_sum(st, partC, gather_blocks(n), dres + 2);
    }
    void delta_device()
    {
        timed(K_GATHER, [&] {
            launch_pair_gather(st, P, chunk, GATHER_DU, sbeads, skey(), sidx(), cstart, cend, box, uconstD, partA, nullptr,
                               nullptr, nullptr, 0);
        });
        timed(K_CHAIN_E, [&] { launch_chain_energy(st, nchains, chainList, cconstD, beads, box, 1, partB); });
    }
    void mc_move(long long step)
    {
        timed(K_MC, [&] { launch_mc_propose(st, P, box, step); });
        delta_device();
        timed(K_MC, [&] {
            launch_mc_decide(st, P, box, partA, cells_blocks(n, chunk), partB, chain_blocks(nchains));
            launch_rescale(st, n, beads, box);
        });
    }
    bool mc_step(long long step) const { return step % 8 == 0 && step != 0 && P.delta_lxy != 0; }

    void step_first(long long step)
    {
        (void)step;
        first_and_keys(1);
    }
    void step_finish(long long step, bool measure)
    {
        sort_and_reorder();
        pair_force();
        bonded(step, KICK_STEP, nullptr);
        if (measure) measure_device();
        if (mc_step(step)) mc_move(step);
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
    if (!(s->gamma > 0)) unsupported("only a scalar gamma > 0 is implemented (gammaType is not)");
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
    // The full-shell gather evaluates (t1,t2) from one side and (t2,t1) from the other;
    // the reference's half shell evaluates one order only, so asymmetric tables differ.
    for (int a = 0; a < nT; a++)
        for (int b = 0; b < nT; b++)
            for (int k = 0; k < 6; k++)
                if (s->fconst[6 * (a * nT + b) + k] != s->fconst[6 * (b * nT + a) + k] ||
                    s->uconst[6 * (a * nT + b) + k] != s->uconst[6 * (b * nT + a) + k])
                    unsupported("non-symmetric two-body constant tables are not supported");
    if (s->n_chain_blocks < 0) invalid("negative chain block count");
    if (s->n_chain_blocks > 0 && (!s->chain || !s->chain_const)) invalid("chain blocks without data");
}

}  // namespace

extern "C" {

const char *sm_last_error(void) { return g_err.c_str(); }
const char *sm_version(void) { return "softmold-mi355x 0.1 (gfx950)"; }

int sm_create(const sm_system *s, sm_ctx **out)
{
    if (!out) {
        g_err = "null output pointer";
        return SM_ERR_INVALID;
    }
    *out = nullptr;
    sm_ctx *c = nullptr;
    int rc = guarded([&] {
        validate(s);
        c = new sm_ctx();
        const int n = s->n_particles;
        c->n = n;
        c->P.n = n;
        c->P.nT = s->n_types;
        c->P.cutoff = s->cutoff;
        c->P.dt = s->dt;
        c->P.gamma = s->gamma;
        c->P.temperature = s->temperature;
        c->P.delta_lxy = s->delta_lxy;
        c->P.tension = s->tension;
        c->P.seed = s->seed;
        for (int d = 0; d < 3; d++) c->box0[d] = s->box[d];
        HIPCHK(hipStreamCreate(&c->st));

        // chains: per-bead role table and per-chain list
        std::vector<int2> perBead(n, make_int2(-1, 0));
        std::vector<int2> list;
        std::vector<float4> ccF;
        std::vector<double4> ccD;
        for (int b = 0; b < s->n_chain_blocks; b++) {
            const int start = s->chain[3 * b], nch = s->chain[3 * b + 1], len = s->chain[3 * b + 2];
            if (len < 3 || len > 0xFFFF) unsupported("CHAIN length must be in [3, 65535]");
            if (nch < 0 || start < 0 || (long long)start + (long long)nch * len > n) invalid("CHAIN block out of range");
            const double *cc = s->chain_const + 4 * b;
            const int cid = (int)ccF.size();
            if (cid > 0x7FFF) unsupported("too many CHAIN constant sets");
            ccF.push_back(make_float4((float)cc[0], (float)cc[1], (float)cc[2], (float)cc[3]));
            ccD.push_back(make_double4(cc[0], cc[1], cc[2], cc[3]));
            for (int k = 0; k < nch; k++) {
                const int st0 = start + k * len;
                list.push_back(make_int2(st0, len | (cid << 16)));
                for (int j = 0; j < len; j++) {
                    if (perBead[st0 + j].x != -1) unsupported("a bead belongs to two CHAIN records");
                    perBead[st0 + j] = make_int2(st0, len | (cid << 16));
                }
            }
        }
        c->nchains = (int)list.size();

        // host staging
        std::vector<Bead> hb(n);
        std::vector<float4> hv(n), ha(n);
        for (int i = 0; i < n; i++) {
            hb[i].x = s->pos[3 * (size_t)i];
            hb[i].y = s->pos[3 * (size_t)i + 1];
            hb[i].z = s->pos[3 * (size_t)i + 2];
            hb[i].type = s->type[i];
            hb[i].pad = 0;
            hv[i] = make_float4((float)s->vel[3 * (size_t)i], (float)s->vel[3 * (size_t)i + 1],
                                (float)s->vel[3 * (size_t)i + 2], 0.f);
            if (s->acc)
                ha[i] = make_float4((float)s->acc[3 * (size_t)i], (float)s->acc[3 * (size_t)i + 1],
                                    (float)s->acc[3 * (size_t)i + 2], 0.f);
            else
                ha[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const int nT = s->n_types;
        std::vector<float> fF(6 * nT * nT);
        for (size_t k = 0; k < fF.size(); k++) fF[k] = (float)s->fconst[k];

        c->beads = dalloc<Bead>(n);
        c->sbeads = dalloc<Bead>(n);
        c->vel = dalloc<float4>(n);
        c->acc = dalloc<float4>(n);
        c->fpair = dalloc<float4>(n);
        c->fscratch = dalloc<float4>(n);
        c->k0 = dalloc<uint32_t>(n);
        c->v0 = dalloc<uint32_t>(n);
        c->k1 = dalloc<uint32_t>(n);
        c->v1 = dalloc<uint32_t>(n);
        c->hist = dalloc<uint32_t>(sort_scratch_words(n, 4));
        c->box = dalloc<DevBox>(1);
        c->fconstF = dalloc<float>(fF.size());
        c->uconstD = dalloc<double>(fF.size());
        c->chainPerBead = dalloc<int2>(n);
        c->chainList = dalloc<int2>(list.size());
        c->cconstF = dalloc<float4>(ccF.size());
        c->cconstD = dalloc<double4>(ccD.size());
        const int gb = gather_blocks(n);
        c->partA = dalloc<double>(cells_blocks(n, 1));
        c->partB = dalloc<double>(chain_blocks(c->nchains));
        c->partC = dalloc<double>(gb);
        c->dres = dalloc<double>(4);
        c->dcounts = dalloc<int>(n);
        c->dnp = dalloc<unsigned long long>(1);

        DevBox db;
        std::memset(&db, 0, sizeof db);
        for (int d = 0; d < 3; d++) {
            db.L[d] = s->box[d];
            db.nc[d] = (int)(s->box[d] / s->cutoff);
            db.cs[d] = s->box[d] / db.nc[d];
            db.scale[d] = 1.0;
        }
        hipStream_t st = c->st;
        HIPCHK(hipMemcpyAsync(c->beads, hb.data(), sizeof(Bead) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->vel, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->acc, ha.data(), sizeof(float4) * n, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(c->fpair, 0, sizeof(float4) * n, st));
        HIPCHK(hipMemcpyAsync(c->box, &db, sizeof(DevBox), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->fconstF, fF.data(), sizeof(float) * fF.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->uconstD, s->uconst, sizeof(double) * fF.size(), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(c->chainPerBead, perBead.data(), sizeof(int2) * n, hipMemcpyHostToDevice, st));
        if (!list.empty()) {
            HIPCHK(hipMemcpyAsync(c->chainList, list.data(), sizeof(int2) * list.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(c->cconstF, ccF.data(), sizeof(float4) * ccF.size(), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(c->cconstD, ccD.data(), sizeof(double4) * ccD.size(), hipMemcpyHostToDevice, st));
        }
        // small systems: shorter chunks so the pair kernel still fills the chip
        c->chunk = (n >= 100000) ? 16 : 8;
        c->ensure_capacity(0);
        HIPCHK(hipStreamSynchronize(st));
    });
    if (rc != SM_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return SM_OK;
}

int sm_destroy(sm_ctx *c)
{
    return guarded([&] { delete c; });
}

int sm_sync(sm_ctx *c)
{
    return guarded([&] {
        if (!c) invalid("null context");
        HIPCHK(hipStreamSynchronize(c->st));
        c->check_flags();
    });
}

int sm_download_state(sm_ctx *c, double *pos, double *vel, double *acc, double box[3])
{
    return guarded([&] {
        if (!c) invalid("null context");
        const int n = c->n;
        std::vector<Bead> hb(pos ? n : 0);
        std::vector<float4> hv(vel ? n : 0), ha(acc ? n : 0);
        if (pos) HIPCHK(hipMemcpyAsync(hb.data(), c->beads, sizeof(Bead) * n, hipMemcpyDeviceToHost, c->st));
        if (vel) HIPCHK(hipMemcpyAsync(hv.data(), c->vel, sizeof(float4) * n, hipMemcpyDeviceToHost, c->st));
        if (acc) HIPCHK(hipMemcpyAsync(ha.data(), c->acc, sizeof(float4) * n, hipMemcpyDeviceToHost, c->st));
        DevBox b = c->read_box();
        for (int i = 0; i < n; i++) {
            if (pos) { pos[3 * (size_t)i] = hb[i].x; pos[3 * (size_t)i + 1] = hb[i].y; pos[3 * (size_t)i + 2] = hb[i].z; }
            if (vel) { vel[3 * (size_t)i] = hv[i].x; vel[3 * (size_t)i + 1] = hv[i].y; vel[3 * (size_t)i + 2] = hv[i].z; }
            if (acc) { acc[3 * (size_t)i] = ha[i].x; acc[3 * (size_t)i + 1] = ha[i].y; acc[3 * (size_t)i + 2] = ha[i].z; }
        }
        if (box) for (int d = 0; d < 3; d++) box[d] = b.L[d];
        if (b.flags) c->check_flags();
    });
}

int sm_upload_state(sm_ctx *c, const double *pos, const double *vel, const double *acc, const double box[3])
{
    return guarded([&] {
        if (!c) invalid("null context");
        const int n = c->n;
        DevBox b = c->read_box();
        if (box) {
            for (int d = 0; d < 3; d++) {
                if (!(box[d] > 0)) invalid("box lengths must be positive");
                b.L[d] = box[d];
                b.nc[d] = (int)(box[d] / c->P.cutoff);
                if (b.nc[d] < 3) unsupported("fewer than 3 cells along an axis");
                b.cs[d] = box[d] / b.nc[d];
            }
        }
        if (pos) {
            std::vector<Bead> hb(n);
            HIPCHK(hipMemcpyAsync(hb.data(), c->beads, sizeof(Bead) * n, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            for (int i = 0; i < n; i++) {
                for (int d = 0; d < 3; d++) {
                    double p = pos[3 * (size_t)i + d];
                    if (!(p >= 0 && p <= b.L[d])) invalid("particle " + std::to_string(i) + " position is out of bounds");
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
        HIPCHK(hipMemcpyAsync(c->box, &b, sizeof(DevBox), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->ensure_capacity(0);
    });
}

int sm_set_temperature(sm_ctx *c, double T)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->P.temperature = T;
    });
}

int sm_init_forces(sm_ctx *c)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->build(0);
        c->pair_force();
        c->bonded(0, KICK_INIT, nullptr);
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
        c->step_first(step);
        c->check_launch();
    });
}

int sm_step_finish(sm_ctx *c, int64_t step, double *measure)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (c->mc_step(step)) c->ensure_capacity(1);
        c->step_finish(step, measure != nullptr);
        c->check_launch();
        if (measure) {
            HIPCHK(hipMemcpyAsync(measure, c->dres, 3 * sizeof(double), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            c->check_flags();
        }
    });
}

int sm_run(sm_ctx *c, int64_t first, int32_t nsteps)
{
    return guarded([&] {
        if (!c) invalid("null context");
        if (nsteps < 0) invalid("negative step count");
        long long moves = 0;
        for (long long s = first; s < first + nsteps; s++)
            if (c->mc_step(s)) moves++;
        if (moves) c->ensure_capacity(moves);
        for (long long s = first; s < first + nsteps; s++) {
            c->step_first(s);
            c->step_finish(s, false);
        }
        c->check_launch();
    });
}

int sm_mc_stats(sm_ctx *c, int64_t *trials, int64_t *accepted)
{
    return guarded([&] {
        if (!c) invalid("null context");
        DevBox b = c->read_box();
        if (trials) *trials = b.trials;
        if (accepted) *accepted = b.accepted;
    });
}

int sm_mc_hist(sm_ctx *c, int32_t cap, double *accepted, double *rejected, int32_t *n_out)
{
    return guarded([&] {
        if (!c || !n_out || cap < 0) invalid("bad argument");
        const double dL = c->P.delta_lxy;
        int nb = 0;
        if (dL != 0) nb = (int)(2.0 * dL / (dL / 101.0)) + 1;
        if (nb > 256) nb = 256;
        *n_out = nb;
        if (nb == 0) return;
        DevBox b = c->read_box();
        const int m = nb < cap ? nb : cap;
        for (int i = 0; i < m; i++) {
            if (accepted) accepted[i] = b.hist_accept[i];
            if (rejected) rejected[i] = b.hist_reject[i];
        }
    });
}

int sm_measure(sm_ctx *c, double out[3])
{
    return guarded([&] {
        if (!c || !out) invalid("null argument");
        c->build(0);
        c->measure_device();
        c->check_launch();
        HIPCHK(hipMemcpyAsync(out, c->dres, 3 * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->check_flags();
    });
}

int sm_compute_forces(sm_ctx *c, double *pair_out, double *chain_out)
{
    return guarded([&] {
        if (!c) invalid("null context");
        const int n = c->n;
        c->build(0);
        c->pair_force();
        c->bonded(0, KICK_NONE, c->fscratch);
        c->check_launch();
        std::vector<float4> h(n);
        auto out3 = [&](const float4 *src, double *dst) {
            HIPCHK(hipMemcpyAsync(h.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            for (int i = 0; i < n; i++) {
                dst[3 * (size_t)i] = h[i].x;
                dst[3 * (size_t)i + 1] = h[i].y;
                dst[3 * (size_t)i + 2] = h[i].z;
            }
        };
        if (pair_out) out3(c->fpair, pair_out);
        if (chain_out) out3(c->fscratch, chain_out);
        c->check_flags();
    });
}

int sm_cell_keys(sm_ctx *c, int32_t *keys)
{
    return guarded([&] {
        if (!c || !keys) invalid("null argument");
        c->first_and_keys(0);
        c->check_launch();
        HIPCHK(hipMemcpyAsync(keys, c->k0, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->check_flags();
    });
}

int sm_pair_count(sm_ctx *c, int32_t *counts, int64_t *total)
{
    return guarded([&] {
        if (!c) invalid("null context");
        c->build(0);
        launch_pair_gather(c->st, c->P, c->chunk, GATHER_COUNT, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend, c->box,
                           c->uconstD, c->partA, c->dcounts, nullptr, nullptr, 0);
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
        c->build(0);
        HIPCHK(hipMemsetAsync(c->dnp, 0, sizeof(unsigned long long), c->st));
        launch_pair_gather(c->st, c->P, c->chunk, GATHER_LIST, c->sbeads, c->skey(), c->sidx(), c->cstart, c->cend, c->box,
                           c->uconstD, c->partA, c->dcounts, c->dpairs, c->dnp, cap);
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

int sm_potential_delta(sm_ctx *c, const double scale[3], double *du_pair, double *du_chain)
{
    return guarded([&] {
        if (!c || !scale) invalid("null argument");
        HIPCHK(hipMemcpyAsync(&c->box->scale, scale, 3 * sizeof(double), hipMemcpyHostToDevice, c->st));
        c->build(0);
        c->delta_device();
        launch_reduce_sum(c->st, c->partA, cells_blocks(c->n, c->chunk), c->dres + 0);
        launch_reduce_sum(c->st, c->partB, chain_blocks(c->nchains), c->dres + 1);
        c->check_launch();
        double r[2];
        HIPCHK(hipMemcpyAsync(r, c->dres, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (du_pair) *du_pair = r[0];
        if (du_chain) *du_chain = r[1];
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
        } else {
            invalid(std::string("unknown knob ") + knob);
        }
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

// ---------------------------------------------------------------------------------------
// .mpd and generators
// ---------------------------------------------------------------------------------------
}  // extern "C"

struct sm_mpd {
    smio::System sys;
    std::vector<int32_t> chain;
    std::vector<double> chain_const;
};

namespace {

const struct { const char *name; smio::Cmd cmd; } kScalars[] = {
    {"gamma", smio::GAMMA}, {"initialTemp", smio::INITIALTEMP}, {"finalTemp", smio::FINALTEMP},
    {"seed", smio::SEED}, {"cutoff", smio::CUTOFF}, {"initialTime", smio::INITIALTIME},
    {"finalTime", smio::FINALTIME}, {"deltaT", smio::DELTAT}, {"storeInterval", smio::STOREINTERVAL},
    {"measureInterval", smio::MEASUREINTERVAL}, {"deltaLXY", smio::DELTALXY},
    {"removeSolvent", smio::REMOVESOLVENT}, {"tempStepInterval", smio::TEMPSTEPINTERVAL},
    {"tension", smio::TENSION}, {"nTypes", smio::NTYPES}, {"nParticles", smio::NPARTICLES},
    {"nMolecules", smio::NMOLECULES}};

double *scalar_ref(smio::System &s, smio::Cmd c)
{
    switch (c) {
    case smio::GAMMA: return &s.gamma;
    case smio::INITIALTEMP: return &s.initialTemp;
    case smio::FINALTEMP: return &s.finalTemp;
    case smio::CUTOFF: return &s.cutoff;
    case smio::INITIALTIME: return &s.initialTime;
    case smio::FINALTIME: return &s.finalTime;
    case smio::DELTAT: return &s.deltaT;
    case smio::STOREINTERVAL: return &s.storeInterval;
    case smio::MEASUREINTERVAL: return &s.measureInterval;
    case smio::DELTALXY: return &s.deltaLXY;
    case smio::REMOVESOLVENT: return &s.removeSolvent;
    case smio::TEMPSTEPINTERVAL: return &s.tempStepInterval;
    case smio::TENSION: return &s.tension;
    default: return nullptr;
    }
}

int io_guard(const std::function<void()> &f)
{
    try {
        f();
        return SM_OK;
    } catch (const ApiError &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return SM_ERR_IO;
    }
}

}  // namespace

extern "C" {

int sm_mpd_read(const char *name, sm_mpd **out)
{
    if (!name || !out) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    *out = nullptr;
    return io_guard([&] {
        auto *m = new sm_mpd();
        try {
            m->sys = smio::read_file(std::string(name) + ".mpd");
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int sm_mpd_write(const sm_mpd *m, const char *name)
{
    if (!m || !name) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    return io_guard([&] { smio::write_file(m->sys, std::string(name) + ".mpd"); });
}

int sm_mpd_free(sm_mpd *m)
{
    delete m;
    return SM_OK;
}

int sm_mpd_system(const sm_mpd *mc, sm_system *out)
{
    if (!mc || !out) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    sm_mpd *m = const_cast<sm_mpd *>(mc);
    return guarded([&] {
        const smio::System &s = m->sys;
        m->chain.clear();
        m->chain_const.clear();
        for (const auto &mo : s.mol) {
            if (mo.type != smio::CHAIN)
                unsupported("molecule type " + std::to_string(mo.type) + " is not implemented (CHAIN only)");
            for (const auto &b : mo.bonds) {
                m->chain.insert(m->chain.end(), {b[0], b[1], b[2]});
                m->chain_const.insert(m->chain_const.end(), mo.constants.begin(), mo.constants.end());
            }
        }
        if (s.gamma <= 0) unsupported("gamma must be > 0 (per-type gammaType is not implemented)");
        std::memset(out, 0, sizeof *out);
        out->n_particles = s.nParticles;
        out->n_types = s.nTypes;
        out->pos = s.pos.data();
        out->type = s.type.data();
        out->vel = s.vel.data();
        out->acc = nullptr;
        out->n_chain_blocks = (int32_t)(m->chain.size() / 3);
        out->chain = m->chain.data();
        out->chain_const = m->chain_const.data();
        out->fconst = s.fconst.data();
        out->uconst = s.uconst.data();
        for (int d = 0; d < 3; d++) out->box[d] = s.size[d];
        out->cutoff = s.cutoff;
        out->dt = s.deltaT;
        out->gamma = s.gamma;
        out->temperature = s.initialTemp;
        out->delta_lxy = s.deltaLXY;
        out->tension = s.tension;
        out->seed = (uint64_t)(uint32_t)s.seed;
    });
}

int sm_mpd_get(const sm_mpd *m, const char *key, double *value)
{
    if (!m || !key || !value) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    for (const auto &k : kScalars) {
        if (std::strcmp(k.name, key) != 0) continue;
        smio::System &s = const_cast<sm_mpd *>(m)->sys;
        switch (k.cmd) {
        case smio::SEED: *value = s.seed; return SM_OK;
        case smio::NTYPES: *value = s.nTypes; return SM_OK;
        case smio::NPARTICLES: *value = s.nParticles; return SM_OK;
        case smio::NMOLECULES: *value = s.nMolecules; return SM_OK;
        default: *value = *scalar_ref(s, k.cmd); return SM_OK;
        }
    }
    g_err = std::string("unknown scalar key ") + key;
    return SM_ERR_INVALID;
}

int sm_mpd_set(sm_mpd *m, const char *key, double value)
{
    if (!m || !key) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    for (const auto &k : kScalars) {
        if (std::strcmp(k.name, key) != 0) continue;
        double *p = scalar_ref(m->sys, k.cmd);
        if (k.cmd == smio::SEED) m->sys.seed = (int)value;
        else if (!p) {
            g_err = std::string(key) + " is not settable";
            return SM_ERR_INVALID;
        } else *p = value;
        m->sys.present[k.cmd] = true;
        return SM_OK;
    }
    g_err = std::string("unknown scalar key ") + key;
    return SM_ERR_INVALID;
}

int sm_mpd_set_state(sm_mpd *m, const double *pos, const double *vel, const double box[3])
{
    if (!m) {
        g_err = "null argument";
        return SM_ERR_INVALID;
    }
    const size_t n3 = 3 * (size_t)m->sys.nParticles;
    if (pos) std::memcpy(m->sys.pos.data(), pos, n3 * sizeof(double));
    if (vel) std::memcpy(m->sys.vel.data(), vel, n3 * sizeof(double));
    if (box) for (int d = 0; d < 3; d++) m->sys.size[d] = box[d];
    return SM_OK;
}

int sm_gen_liposome(int32_t seed, int32_t n_lipids, double rho, sm_mpd **out)
{
    if (!out || n_lipids <= 0 || !(rho > 0)) {
        g_err = "bad argument";
        return SM_ERR_INVALID;
    }
    return io_guard([&] {
        auto *m = new sm_mpd();
        m->sys = smio::gen_liposome(seed, n_lipids, rho);
        *out = m;
    });
}

int sm_gen_bilayer(int32_t seed, int32_t n_lipids, double rho, sm_mpd **out)
{
    if (!out || n_lipids <= 0 || !(rho > 0)) {
        g_err = "bad argument";
        return SM_ERR_INVALID;
    }
    return io_guard([&] {
        auto *m = new sm_mpd();
        m->sys = smio::gen_bilayer(seed, n_lipids, rho);
        *out = m;
    });
}

}  // extern "C"
