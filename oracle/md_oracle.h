/* Test infrastructure: FP64 CPU restatement of the SoftMold MD step (the parity oracle).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this library,
 * and only as the checker; the product never links it.  Every function cites the
 * reference file:line it restates (paths relative to the reference repository root).
 *
 * Conventions: positions/velocities/accelerations are N x 3 row-major doubles; types are
 * int32; CHAIN blocks are (start, nChains, length) int32 triples each with 4 constants
 * {bond length, k_bond, cos-theta term, k_bend} (models/lipidModels.h:1332-1350).
 */
#ifndef SOFTMOLD_MD_ORACLE_H
#define SOFTMOLD_MD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t n;              /* particles */
    int32_t ntypes;
    double *pos;            /* n*3, in/out */
    int32_t *type;          /* n */
    double *vel;            /* n*3, in/out */
    double *acc;            /* n*3, in/out */
    int32_t nchain;         /* CHAIN blocks */
    const int32_t *chain;   /* nchain*3: start, nChains, length */
    const double *chain_const; /* nchain*4 */
    const double *fconst;   /* 6*ntypes*ntypes */
    const double *uconst;   /* 6*ntypes*ntypes */
    double box[3];          /* in/out (MC move) */
    double cutoff, dt, gamma, temp, delta_lxy, tension;
    uint32_t seed;
    int32_t nbond;              /* generic BOND records: pairs ... */
    const int32_t *bond;        /* nbond*2 */
    const double *bond_const;   /* nbond*2: bond length, k_bond (harmonicF, MD.h:384-403) */
    int32_t nbend;              /* generic BEND records: triples ... */
    const int32_t *bend;        /* nbend*3 */
    const double *bend_const;   /* nbend*2: cos-theta term, k_bend (bendF, MD.h:683-723) */
    int32_t nball;              /* BALL records: (centre, member) ... */
    const int32_t *ball;        /* nball*2 */
    const double *ball_const;   /* nball*2: r0, k (harmonicHalfF, MD.h:1164-1187) */
} orc_sys;

/* OpenMP thread count for the oracle's parallel loops (1 gives the reference's
 * single-thread arithmetic order). */
void orc_set_threads(int32_t n);

/* MT19937 (Matsumoto & Nishimura 1998), as MTRand(seed) / rand53 use it
 * (include/algorithms/MersenneTwister.h:165-205,310-341). */
typedef struct { uint32_t s[624]; int left; int next; } orc_mt;
void orc_mt_seed(orc_mt *m, uint32_t seed);
uint32_t orc_mt_u32(orc_mt *m);
double orc_mt_rand53(orc_mt *m);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the counter-RNG mapping shared
 * bit-for-bit with the GPU path. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_langevin(uint64_t seed, uint64_t step, uint32_t particle, double psi[3]);
void orc_philox_mc(uint64_t seed, uint64_t step, double *psi_fluct, double *u_accept);

/* Cell grid (include/algorithms/cellOpt.h:194-209, resize :1526-1570). */
void orc_cell_grid(const double box[3], double cutoff, int32_t nc[3], double cs[3]);
/* Cell key per particle (cellOpt.h:530-556). */
void orc_cell_keys(const orc_sys *s, int32_t *keys);

/* Non-bonded pair terms over the half-shell cell traversal (cellOpt.h:712-1180 with
 * Force/Potential from include/MD.h:795-930).  Single-thread order equals the
 * reference's single-thread order. */
void orc_pair_force(const orc_sys *s, double *acc);          /* acc += F_pair */
void orc_pair_count(const orc_sys *s, int32_t *counts);      /* FORCE_COUNT analogue */
/* pairs within the cutoff whose force term is not identically zero: r^2 < r_m^2, or a row
 * with a nonzero outer coefficient (F row entries 4 or 5) */
void orc_pair_count_active(const orc_sys *s, int32_t *counts);
int64_t orc_pair_list(const orc_sys *s, int32_t *pairs, int64_t cap); /* (i,j), i<j */
double orc_pair_potential(const orc_sys *s);
double orc_pair_dpotential(const orc_sys *s, const double scale[3]);

/* CHAIN bonded terms (include/system.h:1782-1866, 2487-2568, 3280-3384). */
void orc_chain_force(const orc_sys *s, double *acc);
double orc_chain_potential(const orc_sys *s);
double orc_chain_dpotential(const orc_sys *s, const double scale[3]);

/* Langevin (include/algorithms/langevin.h:233-331) and velocity Verlet
 * (include/algorithms/verlet.h:289-356, 464-477). */
void orc_langevin_mt(orc_sys *s, orc_mt *streams, int32_t nstreams, double T);
void orc_langevin_philox(orc_sys *s, uint64_t step, double T);
void orc_verlet_first(orc_sys *s);
void orc_verlet_second(orc_sys *s);
double orc_kinetic(const orc_sys *s);

/* The MD.cpp driver loop without file output (MD.cpp:186-262, 335-721).
 * rng_mode 0: MT19937 streams exactly as the reference seeds them (parity with the
 * compiled reference needs nthreads == 1); rng_mode 1: the counter RNG of the GPU path. */
typedef struct orc_md orc_md;
orc_md *orc_md_create(orc_sys *s, int32_t rng_mode, int32_t mt_streams);
void orc_md_destroy(orc_md *m);
void orc_md_init_forces(orc_md *m);            /* MD.cpp:186-262 */
int32_t orc_md_run(orc_md *m, int32_t start, int32_t end); /* MD.cpp:335-729; returns #accepted MC moves */
int32_t orc_md_trials(const orc_md *m);
/* accept (1) / reject (0) bits of the MC trials so far, the latest in bit 0 */
uint64_t orc_md_decisions(const orc_md *m);
double orc_md_objective(orc_md *m, int32_t nsteps);
/* analysis/vesicleRadius.cpp:71-175, one frame */
double orc_vesicle_radius(const double *pos, const int32_t *type, int32_t n, const double box[3], int32_t head);         /* mean U after each kick, steps 1..nsteps */
void orc_md_measure(orc_md *m, double *U, double *K);      /* dataExtraction.h:827-1050 */

#ifdef __cplusplus
}
#endif
#endif
