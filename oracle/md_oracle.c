/* Test infrastructure: FP64 CPU restatement of the SoftMold MD step (the parity oracle).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this file's
 * library, and only as the checker or the timed CPU baseline; the product never links
 * it.  Each function names the reference file:line it restates (paths relative to the
 * reference repository root).  Pinning (what each function is checked against) lives in
 * tests/test_oracle.py: golden vectors written by the compiled reference, closed forms,
 * finite differences, invariants and brute force.
 */
#include "md_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int32_t n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------------------
 * MT19937 -- published algorithm (Matsumoto & Nishimura, ACM TOMACS 8(1), 1998); the
 * reference wraps it as MTRand (include/algorithms/MersenneTwister.h:165-205 seeding and
 * reload, :275-288 tempering, :337-341 rand53 = genrand_res53).
 * ---------------------------------------------------------------------------------- */
void orc_mt_seed(orc_mt *m, uint32_t seed)
{
    m->s[0] = seed;
    for (int i = 1; i < 624; i++)
        m->s[i] = 1812433253u * (m->s[i - 1] ^ (m->s[i - 1] >> 30)) + (uint32_t)i;
    m->left = 0;
    m->next = 624;
}

static void mt_reload(orc_mt *m)
{
    for (int i = 0; i < 624; i++) {
        uint32_t y = (m->s[i] & 0x80000000u) | (m->s[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = m->s[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        m->s[i] = v;
    }
    m->left = 624;
    m->next = 0;
}

uint32_t orc_mt_u32(orc_mt *m)
{
    if (m->left == 0) mt_reload(m);
    m->left--;
    uint32_t y = m->s[m->next++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
}

double orc_mt_rand53(orc_mt *m)
{
    uint32_t a = orc_mt_u32(m) >> 5, b = orc_mt_u32(m) >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2,
 * 3", SC'11): multipliers 0xD2511F53/0xCD9E8D57, Weyl key increments 0x9E3779B9/
 * 0xBB67AE85, ten rounds.  The reference has no counter RNG (it uses per-thread MT19937,
 * include/algorithms/langevin.h:98-107); this is the GPU path's RNG, restated here so
 * the GPU trajectory can be replayed on the CPU bit-for-bit in its noise.
 * ---------------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream tags in counter word 3. */
enum { TAG_LANGEVIN = 0, TAG_MC = 1, TAG_LANGEVIN_INIT = 2 };

/* psi in (-1,1): ((x>>8) * 2^-23 - 1) + 2^-24.  Every intermediate is exact in binary32,
 * so the FP32 GPU kernel and this FP64 code produce the identical value.  It plays the
 * role of 2u-1 in langevin.h:259 and MD.cpp:595. */
static double psi_from_u32(uint32_t x)
{
    return ((double)(x >> 8) * (1.0 / 8388608.0) - 1.0) + (1.0 / 16777216.0);
}

static void philox_tagged(uint64_t seed, uint64_t step, uint32_t particle, uint32_t tag,
                          uint32_t out[4])
{
    uint32_t ctr[4] = { particle, (uint32_t)step, (uint32_t)(step >> 32), tag };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    orc_philox4x32_10(ctr, key, out);
}

void orc_philox_langevin(uint64_t seed, uint64_t step, uint32_t particle, double psi[3])
{
    uint32_t o[4];
    philox_tagged(seed, step, particle, TAG_LANGEVIN, o);
    psi[0] = psi_from_u32(o[0]);
    psi[1] = psi_from_u32(o[1]);
    psi[2] = psi_from_u32(o[2]);
}

static void philox_langevin_init(uint64_t seed, uint32_t particle, double psi[3])
{
    uint32_t o[4];
    philox_tagged(seed, 0, particle, TAG_LANGEVIN_INIT, o);
    psi[0] = psi_from_u32(o[0]);
    psi[1] = psi_from_u32(o[1]);
    psi[2] = psi_from_u32(o[2]);
}

void orc_philox_mc(uint64_t seed, uint64_t step, double *psi_fluct, double *u_accept)
{
    uint32_t o[4];
    philox_tagged(seed, step, 0, TAG_MC, o);
    *psi_fluct = psi_from_u32(o[0]);
    /* 53-bit uniform in [0,1), the genrand_res53 construction (MersenneTwister.h:337-341) */
    *u_accept = ((o[1] >> 5) * 67108864.0 + (o[2] >> 6)) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------------------
 * Cell grid and keys.
 * ---------------------------------------------------------------------------------- */
void orc_cell_grid(const double box[3], double cutoff, int32_t nc[3], double cs[3])
{
    /* cellOpt.h:194-196 (nCells = int(size/cutoff)), :207-209 (cellSize = size/nCells) */
    for (int d = 0; d < 3; d++) {
        nc[d] = (int32_t)(box[d] / cutoff);
        cs[d] = box[d] / nc[d];
    }
}

static inline int32_t cell_of(double p, double cs, int32_t n)
{
    /* cellOpt.h:532-539: int() truncation, then the one-cell correction at the top edge */
    int32_t c = (int32_t)(p / cs);
    return (c >= n) ? c - 1 : c;
}

void orc_cell_keys(const orc_sys *s, int32_t *keys)
{
    int32_t nc[3]; double cs[3];
    orc_cell_grid(s->box, s->cutoff, nc, cs);
    for (int i = 0; i < s->n; i++) {
        const double *p = s->pos + 3 * (size_t)i;
        int32_t x = cell_of(p[0], cs[0], nc[0]);
        int32_t y = cell_of(p[1], cs[1], nc[1]);
        int32_t z = cell_of(p[2], cs[2], nc[2]);
        keys[i] = x + y * nc[0] + z * nc[0] * nc[1];  /* cellOpt.h:555 */
    }
}

/* ------------------------------------------------------------------------------------
 * Cell lists.  The reference pushes particles 0..N-1 onto per-cell linked lists, so a
 * cell's list runs in descending particle index (cellOpt.h:572-585), and visits occupied
 * cells in ascending key (LOW_DENSITY full-cell list, cellOpt.h:657-709).  The same
 * visiting order is reproduced without a dense grid: particles sorted by (key asc,
 * index desc), occupied cells located by binary search.
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int32_t nc[3];
    double cs[3];
    int32_t n;
    int32_t *order;    /* particles in traversal order */
    int32_t ncell;     /* occupied cells */
    int32_t *ckey;     /* occupied cell keys, ascending */
    int32_t *cbeg;     /* run start in order[] */
} cells_t;

typedef struct { int32_t key, idx; } kv_t;

static int kv_cmp(const void *a, const void *b)
{
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->key != y->key) return (x->key < y->key) ? -1 : 1;
    return (x->idx > y->idx) ? -1 : (x->idx < y->idx);  /* index descending */
}

static void cells_build(const orc_sys *s, cells_t *c)
{
    orc_cell_grid(s->box, s->cutoff, c->nc, c->cs);
    c->n = s->n;
    kv_t *kv = (kv_t *)malloc(sizeof(kv_t) * (size_t)(s->n > 0 ? s->n : 1));
    int32_t *keys = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n > 0 ? s->n : 1));
    orc_cell_keys(s, keys);
    for (int i = 0; i < s->n; i++) { kv[i].key = keys[i]; kv[i].idx = i; }
    qsort(kv, (size_t)s->n, sizeof(kv_t), kv_cmp);
    c->order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n + 1));
    c->ckey = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n + 1));
    c->cbeg = (int32_t *)malloc(sizeof(int32_t) * (size_t)(s->n + 2));
    c->ncell = 0;
    for (int i = 0; i < s->n; i++) {
        c->order[i] = kv[i].idx;
        if (i == 0 || kv[i].key != kv[i - 1].key) {
            c->ckey[c->ncell] = kv[i].key;
            c->cbeg[c->ncell] = i;
            c->ncell++;
        }
    }
    c->cbeg[c->ncell] = s->n;
    free(kv);
    free(keys);
}

static void cells_free(cells_t *c)
{
    free(c->order); free(c->ckey); free(c->cbeg);
}

static int32_t cells_find(const cells_t *c, int32_t key)
{
    int32_t lo = 0, hi = c->ncell - 1;
    while (lo <= hi) {
        int32_t mid = (lo + hi) >> 1;
        if (c->ckey[mid] == key) return mid;
        if (c->ckey[mid] < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* The 13 forward neighbours of the half shell, in the reference's order (cellOpt.h:715). */
static const int NEXT_BOX[13][3] = {
    {1,0,0},{0,1,0},{1,1,0},{0,0,1},{1,0,1},{0,1,1},{1,1,1},
    {-1,0,1},{0,-1,1},{-1,1,1},{1,-1,1},{-1,-1,1},{1,-1,0}};

/* Neighbour cell (ordinal, or -1 if empty) and its periodic image shift (cellOpt.h:806-823). */
static int32_t neighbour(const cells_t *c, const double box[3], int32_t key, int j, double img[3])
{
    int32_t nx = c->nc[0], ny = c->nc[1], nz = c->nc[2];
    int32_t x = key % nx, y = (key / nx) % ny, z = key / (nx * ny);
    int32_t X = x + NEXT_BOX[j][0], Y = y + NEXT_BOX[j][1], Z = z + NEXT_BOX[j][2];
    img[0] = img[1] = img[2] = 0.0;
    if (X < 0) img[0] -= box[0];
    if (X >= nx) img[0] += box[0];
    if (Y < 0) img[1] -= box[1];
    if (Y >= ny) img[1] += box[1];
    if (Z < 0) img[2] -= box[2];
    if (Z >= nz) img[2] += box[2];
    int32_t k = ((X + nx) % nx) + ((Y + ny) % ny) * nx + ((Z + nz) % nz) * nx * ny;
    return cells_find(c, k);
}

/* Force<double> (include/MD.h:795-848); count 1 gives the FORCE_COUNT variant
 * (MD.h:838-845): every pair within the cutoff.  count 2 counts only the pairs whose force
 * term is not identically zero: r^2 < r_m^2 (the inner branch), or an outer branch with a
 * nonzero coefficient (row entries 4, 5; Umin != 0 in laradjiRevalee.h:97-110). */
static inline void pair_force(double cut2, int nT, const double *p1, int t1, const double *p2,
                              int t2, double *a1, double *a2, const double *fC, int count)
{
    double dx = p1[0] - p2[0];
    double dy = p1[1] - p2[1];
    double dz = p1[2] - p2[2];
    double dr = dx * dx + dy * dy + dz * dz;
    if (dr < cut2) {
        if (count == 2) {
            const int ci = 6 * (t1 * nT + t2);
            if (!(fC[ci + 4] != 0.0 || fC[ci + 5] != 0.0 || dr < fC[ci] * fC[ci])) return;
        }
        if (count) {
            a1[0] += 1; a1[1] += 1; a1[2] += 1;
            a2[0] += 1; a2[1] += 1; a2[2] += 1;
            return;
        }
        dr = sqrt(dr);
        int ci = 6 * (t1 * nT + t2);
        ci += ((int)(dr / fC[ci])) * 3;
        double mag = fC[ci] - dr;
        mag = ((fC[ci + 1] - fC[ci + 2] * mag) * mag) / dr;
        dx *= mag; dy *= mag; dz *= mag;
        a1[0] += dx; a1[1] += dy; a1[2] += dz;
        a2[0] -= dx; a2[1] -= dy; a2[2] -= dz;
    }
}

/* Potential<double> (include/MD.h:895-930). */
static inline double pair_potential(double cut2, int nT, const double *p1, int t1,
                                    const double *p2, int t2, const double *uC)
{
    double dx = p1[0] - p2[0];
    double dy = p1[1] - p2[1];
    double dz = p1[2] - p2[2];
    double dr = dx * dx + dy * dy + dz * dz;
    double pot = 0;
    if (dr < cut2) {
        dr = sqrt(dr);
        int ci = 6 * (t1 * nT + t2);
        if (dr <= uC[ci]) {
            pot = uC[ci] - dr;
            pot = uC[ci + 1] * pot * pot + uC[ci + 2];
        } else {
            pot = uC[ci + 3] - dr;
            pot = pot * pot * (uC[ci + 4] - pot * uC[ci + 5]);
        }
    }
    return pot;
}

/* Shared traversal of CellOpt::computeForce/computePotential/computeDPotential
 * (cellOpt.h:712-926, 928-1041, 1043-1180): per occupied cell, the intra-cell triangle
 * with local copies, then the 13 forward neighbours with the periodic image added to the
 * neighbour's position.  mode 0: forces, 1: counts, 2: potential, 3: dPotential,
 * 4: pair list, 5: counts of the pairs with a nonzero force term.  Forces and counts run
 * cells in parallel with per-cell locks like the
 * reference; with one thread the arithmetic order is the reference's single-thread order. */
typedef struct {
    int mode;
    double *acc;
    const double *scale;
    int32_t *pairs;
    int64_t cap, npairs;
} trav_t;

static double traverse(const orc_sys *s, trav_t *t)
{
    cells_t c;
    cells_build(s, &c);
    const double cut2 = s->cutoff * s->cutoff;
    const int nT = s->ntypes;
    const int fcount = t->mode == 1 ? 1 : (t->mode == 5 ? 2 : 0);
    double total = 0.0;
    int maxn = 1;
    for (int i = 0; i < c.ncell; i++) {
        int n = c.cbeg[i + 1] - c.cbeg[i];
        if (n > maxn) maxn = n;
    }
#ifdef _OPENMP
    omp_lock_t *locks = NULL;
    int use_locks = (t->mode <= 1 || t->mode == 5);
    if (use_locks) {
        locks = (omp_lock_t *)malloc(sizeof(omp_lock_t) * (size_t)(c.ncell > 0 ? c.ncell : 1));
        for (int i = 0; i < c.ncell; i++) omp_init_lock(&locks[i]);
    }
    int parallel = (t->mode <= 3 || t->mode == 5);
#pragma omp parallel if (parallel) reduction(+ : total)
#endif
    {
        double *tp = (double *)malloc(sizeof(double) * 3 * (size_t)maxn);
        int *tt = (int *)malloc(sizeof(int) * (size_t)maxn);
        int *ti = (int *)malloc(sizeof(int) * (size_t)maxn);
        double *ta = (double *)malloc(sizeof(double) * 3 * (size_t)maxn);
        const double *S = t->scale;
#ifdef _OPENMP
#pragma omp for schedule(dynamic)
#endif
        for (int ci = 0; ci < c.ncell; ci++) {
            int ni = 0;
            for (int q = c.cbeg[ci]; q < c.cbeg[ci + 1]; q++, ni++) {
                int m = c.order[q];
                memcpy(tp + 3 * ni, s->pos + 3 * (size_t)m, 3 * sizeof(double));
                tt[ni] = s->type[m];
                ti[ni] = m;
                ta[3 * ni] = ta[3 * ni + 1] = ta[3 * ni + 2] = 0.0;
                for (int j = 0; j < ni; j++) {
                    if (t->mode <= 1 || t->mode == 5) {
                        pair_force(cut2, nT, tp + 3 * ni, tt[ni], tp + 3 * j, tt[j], ta + 3 * ni,
                                   ta + 3 * j, s->fconst, fcount);
                    } else if (t->mode == 2) {
                        total += pair_potential(cut2, nT, tp + 3 * ni, tt[ni], tp + 3 * j, tt[j], s->uconst);
                    } else if (t->mode == 3) {
                        double pA[3], pB[3];
                        double oldP = pair_potential(cut2, nT, tp + 3 * ni, tt[ni], tp + 3 * j, tt[j], s->uconst);
                        for (int d = 0; d < 3; d++) { pA[d] = tp[3 * ni + d] * S[d]; pB[d] = tp[3 * j + d] * S[d]; }
                        double newP = pair_potential(cut2, nT, pA, tt[ni], pB, tt[j], s->uconst);
                        total += (oldP - newP);
                    } else {
                        const double *a = tp + 3 * ni, *b = tp + 3 * j;
                        double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
                        if (dx * dx + dy * dy + dz * dz < cut2) {
                            if (t->npairs < t->cap) {
                                int32_t u = ti[ni], v = ti[j];
                                t->pairs[2 * t->npairs] = u < v ? u : v;
                                t->pairs[2 * t->npairs + 1] = u < v ? v : u;
                            }
                            t->npairs++;
                        }
                    }
                }
            }
            int32_t key = c.ckey[ci];
            for (int j = 0; j < 13; j++) {
                double img[3];
                int32_t nb = neighbour(&c, s->box, key, j, img);
                if (nb < 0) continue;
#ifdef _OPENMP
                if (use_locks) omp_set_lock(&locks[nb]);
#endif
                for (int q = c.cbeg[nb]; q < c.cbeg[nb + 1]; q++) {
                    int n = c.order[q];
                    double cp[3];
                    cp[0] = s->pos[3 * (size_t)n] + img[0];
                    cp[1] = s->pos[3 * (size_t)n + 1] + img[1];
                    cp[2] = s->pos[3 * (size_t)n + 2] + img[2];
                    int tn = s->type[n];
                    if (t->mode <= 1 || t->mode == 5) {
                        double ca[3] = {0.0, 0.0, 0.0};
                        for (int m = 0; m < ni; m++)
                            pair_force(cut2, nT, tp + 3 * m, tt[m], cp, tn, ta + 3 * m, ca, s->fconst, fcount);
                        t->acc[3 * (size_t)n] += ca[0];
                        t->acc[3 * (size_t)n + 1] += ca[1];
                        t->acc[3 * (size_t)n + 2] += ca[2];
                    } else if (t->mode == 2) {
                        for (int m = 0; m < ni; m++)
                            total += pair_potential(cut2, nT, tp + 3 * m, tt[m], cp, tn, s->uconst);
                    } else if (t->mode == 3) {
                        for (int m = 0; m < ni; m++) {
                            double pA[3], pB[3];
                            double oldP = pair_potential(cut2, nT, tp + 3 * m, tt[m], cp, tn, s->uconst);
                            for (int d = 0; d < 3; d++) { pA[d] = tp[3 * m + d] * S[d]; pB[d] = cp[d] * S[d]; }
                            double newP = pair_potential(cut2, nT, pA, tt[m], pB, tn, s->uconst);
                            total += (oldP - newP);
                        }
                    } else {
                        for (int m = 0; m < ni; m++) {
                            const double *a = tp + 3 * m;
                            double dx = a[0] - cp[0], dy = a[1] - cp[1], dz = a[2] - cp[2];
                            if (dx * dx + dy * dy + dz * dz < cut2) {
                                if (t->npairs < t->cap) {
                                    int32_t u = ti[m], v = n;
                                    t->pairs[2 * t->npairs] = u < v ? u : v;
                                    t->pairs[2 * t->npairs + 1] = u < v ? v : u;
                                }
                                t->npairs++;
                            }
                        }
                    }
                }
#ifdef _OPENMP
                if (use_locks) omp_unset_lock(&locks[nb]);
#endif
            }
            if (t->mode <= 1 || t->mode == 5) {
#ifdef _OPENMP
                if (use_locks) omp_set_lock(&locks[ci]);
#endif
                for (int m = 0; m < ni; m++) {
                    t->acc[3 * (size_t)ti[m]] += ta[3 * m];
                    t->acc[3 * (size_t)ti[m] + 1] += ta[3 * m + 1];
                    t->acc[3 * (size_t)ti[m] + 2] += ta[3 * m + 2];
                }
#ifdef _OPENMP
                if (use_locks) omp_unset_lock(&locks[ci]);
#endif
            }
        }
        free(tp); free(tt); free(ti); free(ta);
    }
#ifdef _OPENMP
    if (locks) {
        for (int i = 0; i < c.ncell; i++) omp_destroy_lock(&locks[i]);
        free(locks);
    }
#endif
    cells_free(&c);
    return total;
}

void orc_pair_force(const orc_sys *s, double *acc)
{
    trav_t t = {0, acc, NULL, NULL, 0, 0};
    traverse(s, &t);
}

void orc_pair_count(const orc_sys *s, int32_t *counts)
{
    double *acc = (double *)calloc(3 * (size_t)s->n + 1, sizeof(double));
    trav_t t = {1, acc, NULL, NULL, 0, 0};
    traverse(s, &t);
    for (int i = 0; i < s->n; i++) counts[i] = (int32_t)acc[3 * (size_t)i];
    free(acc);
}

void orc_pair_count_active(const orc_sys *s, int32_t *counts)
{
    double *acc = (double *)calloc(3 * (size_t)s->n + 1, sizeof(double));
    trav_t t = {5, acc, NULL, NULL, 0, 0};
    traverse(s, &t);
    for (int i = 0; i < s->n; i++) counts[i] = (int32_t)acc[3 * (size_t)i];
    free(acc);
}

int64_t orc_pair_list(const orc_sys *s, int32_t *pairs, int64_t cap)
{
    trav_t t = {4, NULL, NULL, pairs, cap, 0};
    traverse(s, &t);
    return t.npairs;
}

double orc_pair_potential(const orc_sys *s)
{
    trav_t t = {2, NULL, NULL, NULL, 0, 0};
    return traverse(s, &t);
}

double orc_pair_dpotential(const orc_sys *s, const double scale[3])
{
    trav_t t = {3, NULL, scale, NULL, 0, 0};
    return traverse(s, &t);
}

/* ------------------------------------------------------------------------------------
 * CHAIN bonded terms.
 * ---------------------------------------------------------------------------------- */
/* harmonicF (include/MD.h:384-403) */
static inline void harmonic_f(double d[3], double *a1, double *a2, const double *c)
{
    double dr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double mag = dr - c[0];
    mag = -mag * c[1] / dr;
    double x = d[0] * mag, y = d[1] * mag, z = d[2] * mag;
    a1[0] += x; a1[1] += y; a1[2] += z;
    a2[0] -= x; a2[1] -= y; a2[2] -= z;
}

/* harmonicP (include/MD.h:422-431) */
static inline double harmonic_p(const double d[3], const double *c)
{
    double dr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double pot = dr - c[0];
    return 0.5 * c[1] * pot * pot;
}

/* bendF (include/MD.h:683-723) */
static inline void bend_f(const double da_in[3], const double db_in[3], double *a1, double *a2,
                          double *a3, const double *c)
{
    double da[3] = {da_in[0], da_in[1], da_in[2]}, db[3] = {db_in[0], db_in[1], db_in[2]};
    double dra = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
    double drb = sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]);
    for (int d = 0; d < 3; d++) { da[d] /= dra; db[d] /= drb; }
    double cost = (da[0] * db[0]) + (da[1] * db[1]) + (da[2] * db[2]);
    double mag = c[0] - cost;
    mag *= c[1];
    for (int d = 0; d < 3; d++) {
        double bufa = mag * (db[d] - (da[d] * cost)) / dra;
        double bufb = mag * (da[d] - (db[d] * cost)) / drb;
        a1[d] += bufa;
        a2[d] += (bufb - bufa);
        a3[d] -= bufb;
    }
}

/* bendP (include/MD.h:769-789) */
static inline double bend_p(const double da_in[3], const double db_in[3], const double *c)
{
    double da[3] = {da_in[0], da_in[1], da_in[2]}, db[3] = {db_in[0], db_in[1], db_in[2]};
    double dra = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
    double drb = sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]);
    for (int d = 0; d < 3; d++) { da[d] /= dra; db[d] /= drb; }
    double cost = (da[0] * db[0]) + (da[1] * db[1]) + (da[2] * db[2]);
    double pot = c[0] - cost;
    return c[1] * pot * pot * 0.5;
}

/* Minimum-image difference as doChainForce writes it (system.h:1809-1827): strict '>'
 * against half the box on each side. */
static inline void min_image_diff(const double *p1, const double *p2, const double box[3], double d[3])
{
    for (int k = 0; k < 3; k++) {
        d[k] = p1[k] - p2[k];
        if (d[k] > box[k] / 2.0) d[k] -= box[k];
        if (d[k] < -box[k] / 2.0) d[k] += box[k];
    }
}

/* mode 0: forces into acc; 1: potential; 2: dPotential under scale. */
static double chain_terms(const orc_sys *s, int mode, double *acc, const double *S)
{
    double total = 0.0;
    for (int b = 0; b < s->nchain; b++) {
        const int32_t start = s->chain[3 * b], nch = s->chain[3 * b + 1], len = s->chain[3 * b + 2];
        const double *cb = s->chain_const + 4 * b;      /* CHAINBOND = 0 (MD.h:65) */
        const double *ce = cb + 2;                      /* CHAINBEND = 2 (MD.h:66) */
        for (int k = start; k < start + len * nch; k += len) {
            double da[3], db[3];
            /* interior triples (system.h:1803-1832) */
            for (int l = k; l < k + len - 3; l++) {
                const double *p1 = s->pos + 3 * (size_t)l, *p2 = p1 + 3, *p3 = p1 + 6;
                min_image_diff(p1, p2, s->box, da);
                min_image_diff(p2, p3, s->box, db);
                if (mode == 0) {
                    double d2[3] = {da[0], da[1], da[2]};
                    harmonic_f(d2, acc + 3 * (size_t)l, acc + 3 * (size_t)(l + 1), cb);
                    bend_f(da, db, acc + 3 * (size_t)l, acc + 3 * (size_t)(l + 1), acc + 3 * (size_t)(l + 2), ce);
                } else if (mode == 1) {
                    total += harmonic_p(da, cb);
                    total += bend_p(da, db, ce);
                } else {
                    double oldP = harmonic_p(da, cb);
                    oldP += bend_p(da, db, ce);
                    for (int d = 0; d < 3; d++) { da[d] *= S[d]; db[d] *= S[d]; }
                    double newP = harmonic_p(da, cb);
                    newP += bend_p(da, db, ce);
                    total += (oldP - newP);
                }
            }
            /* last triple carries both bonds (system.h:1833-1862) */
            int f = k + len - 3, g = k + len - 2, h = k + len - 1;
            const double *p1 = s->pos + 3 * (size_t)f, *p2 = s->pos + 3 * (size_t)g, *p3 = s->pos + 3 * (size_t)h;
            min_image_diff(p1, p2, s->box, da);
            min_image_diff(p2, p3, s->box, db);
            if (mode == 0) {
                double d2[3] = {da[0], da[1], da[2]}, d3[3] = {db[0], db[1], db[2]};
                harmonic_f(d2, acc + 3 * (size_t)f, acc + 3 * (size_t)g, cb);
                harmonic_f(d3, acc + 3 * (size_t)g, acc + 3 * (size_t)h, cb);
                bend_f(da, db, acc + 3 * (size_t)f, acc + 3 * (size_t)g, acc + 3 * (size_t)h, ce);
            } else if (mode == 1) {
                total += harmonic_p(da, cb);
                total += harmonic_p(db, cb);
                total += bend_p(da, db, ce);
            } else {
                double oldP = harmonic_p(da, cb);
                oldP += harmonic_p(db, cb);
                oldP += bend_p(da, db, ce);
                for (int d = 0; d < 3; d++) { da[d] *= S[d]; db[d] *= S[d]; }
                double newP = harmonic_p(da, cb);
                newP += harmonic_p(db, cb);
                newP += bend_p(da, db, ce);
                total += (oldP - newP);
            }
        }
    }
    return total;
}

/* Generic BOND and BEND records, each with its own two constants: doBondForce and
 * doBendForce (system.h:1881-2035), doBondPotential / doBendPotential (:2718-2780,
 * :2782-2830), doBondDPotential / doBendDPotential (:3399-3432, :3477-3530; the imaged
 * differences are scaled, not the positions).  The reference visits molecules in file
 * order; here CHAIN blocks come first, then BOND records, then BEND records, which only
 * changes the FP64 summation order of a bead's bonded force. */
static double generic_terms(const orc_sys *s, int mode, double *acc, const double *S)
{
    double total = 0.0;
    for (int j = 0; j < s->nbond; j++) {
        const int32_t f = s->bond[2 * j], g = s->bond[2 * j + 1];
        const double *c = s->bond_const + 2 * j;
        double d[3];
        min_image_diff(s->pos + 3 * (size_t)f, s->pos + 3 * (size_t)g, s->box, d);
        if (mode == 0) {
            harmonic_f(d, acc + 3 * (size_t)f, acc + 3 * (size_t)g, c);
        } else if (mode == 1) {
            total += harmonic_p(d, c);
        } else {
            const double oldP = harmonic_p(d, c);
            for (int k = 0; k < 3; k++) d[k] *= S[k];
            const double newP = harmonic_p(d, c);
            total += (oldP - newP);
        }
    }
    /* BALL (doBallForce system.h:1937-1971, doBallPotential :2749-2780): harmonicHalfF /
     * harmonicHalfP (MD.h:1164-1222), active only for r < r0; the box move adds nothing
     * for them (MD.cpp:664 drops doBallDPotential's value) */
    for (int j = 0; j < s->nball; j++) {
        const int32_t f = s->ball[2 * j], g = s->ball[2 * j + 1];
        const double *c = s->ball_const + 2 * j;
        double d[3];
        min_image_diff(s->pos + 3 * (size_t)f, s->pos + 3 * (size_t)g, s->box, d);
        const double dr2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (dr2 >= c[0] * c[0] || mode == 2) continue;
        if (mode == 0) harmonic_f(d, acc + 3 * (size_t)f, acc + 3 * (size_t)g, c);
        else total += harmonic_p(d, c);
    }
    for (int j = 0; j < s->nbend; j++) {
        const int32_t f = s->bend[3 * j], g = s->bend[3 * j + 1], h = s->bend[3 * j + 2];
        const double *c = s->bend_const + 2 * j;
        double da[3], db[3];
        min_image_diff(s->pos + 3 * (size_t)f, s->pos + 3 * (size_t)g, s->box, da);
        min_image_diff(s->pos + 3 * (size_t)g, s->pos + 3 * (size_t)h, s->box, db);
        if (mode == 0) {
            bend_f(da, db, acc + 3 * (size_t)f, acc + 3 * (size_t)g, acc + 3 * (size_t)h, c);
        } else if (mode == 1) {
            total += bend_p(da, db, c);
        } else {
            const double oldP = bend_p(da, db, c);
            for (int k = 0; k < 3; k++) { da[k] *= S[k]; db[k] *= S[k]; }
            const double newP = bend_p(da, db, c);
            total += (oldP - newP);
        }
    }
    return total;
}

/* all bonded terms: CHAIN blocks, then the generic records */
void orc_chain_force(const orc_sys *s, double *acc)
{
    chain_terms(s, 0, acc, NULL);
    generic_terms(s, 0, acc, NULL);
}
double orc_chain_potential(const orc_sys *s)
{
    double u = chain_terms(s, 1, NULL, NULL);
    return u + generic_terms(s, 1, NULL, NULL);
}
double orc_chain_dpotential(const orc_sys *s, const double scale[3])
{
    double u = chain_terms(s, 2, NULL, scale);
    return u + generic_terms(s, 2, NULL, scale);
}

/* ------------------------------------------------------------------------------------
 * Thermostat and integrator.
 * ---------------------------------------------------------------------------------- */
/* Langevin::compute with scalar gamma (langevin.h:233-331): a += -g v + sigma (2u-1),
 * sigma = sqrt(6 T g / dt).  Particles are split into nstreams contiguous chunks the way
 * an OpenMP static schedule splits them; chunk t draws from streams[t]. */
void orc_langevin_mt(orc_sys *s, orc_mt *streams, int32_t nstreams, double T)
{
    const double g = s->gamma;
    const double sigma = sqrt((6.0 * T * g) / s->dt);
    const int n = s->n;
    for (int t = 0; t < nstreams; t++) {
        int chunk = n / nstreams, rem = n % nstreams;
        int lo = t * chunk + (t < rem ? t : rem);
        int hi = lo + chunk + (t < rem ? 1 : 0);
        for (int i = lo; i < hi; i++) {
            double *a = s->acc + 3 * (size_t)i;
            const double *v = s->vel + 3 * (size_t)i;
            for (int d = 0; d < 3; d++) {
                double psi = 2.0 * orc_mt_rand53(&streams[t]) - 1.0;
                a[d] += (-g * v[d] + sigma * psi);
            }
        }
    }
}

static void langevin_psi(orc_sys *s, double T, int init, uint64_t step)
{
    const double g = s->gamma;
    const double sigma = sqrt((6.0 * T * g) / s->dt);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < s->n; i++) {
        double psi[3];
        if (init) philox_langevin_init(s->seed, (uint32_t)i, psi);
        else orc_philox_langevin(s->seed, step, (uint32_t)i, psi);
        double *a = s->acc + 3 * (size_t)i;
        const double *v = s->vel + 3 * (size_t)i;
        for (int d = 0; d < 3; d++) a[d] += (-g * v[d] + sigma * psi[d]);
    }
}

void orc_langevin_philox(orc_sys *s, uint64_t step, double T) { langevin_psi(s, T, 0, step); }

/* Verlet::first, no half carry (verlet.h:289-356): only type != 0 moves; wrap p > L. */
void orc_verlet_first(orc_sys *s)
{
    const double halfDt = 0.5 * s->dt, dt = s->dt;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < s->n; i++) {
        if (s->type[i] == 0) continue;
        double *p = s->pos + 3 * (size_t)i, *v = s->vel + 3 * (size_t)i;
        const double *a = s->acc + 3 * (size_t)i;
        for (int d = 0; d < 3; d++) v[d] += (a[d] * halfDt);
        for (int d = 0; d < 3; d++) p[d] += v[d] * dt;
        for (int d = 0; d < 3; d++) {
            if (p[d] > s->box[d]) p[d] -= s->box[d];
            if (p[d] < 0) p[d] += s->box[d];
        }
    }
}

/* Verlet::second (verlet.h:464-477) */
void orc_verlet_second(orc_sys *s)
{
    const double halfDt = s->dt * 0.5;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < s->n; i++) {
        if (s->type[i] == 0) continue;
        double *v = s->vel + 3 * (size_t)i;
        const double *a = s->acc + 3 * (size_t)i;
        for (int d = 0; d < 3; d++) v[d] += (a[d] * halfDt);
    }
}

/* Kinetic::compute (include/algorithms/dataCollection.h:667-673) */
double orc_kinetic(const orc_sys *s)
{
    double e = 0.0;
    for (int i = 0; i < s->n; i++) {
        const double *v = s->vel + 3 * (size_t)i;
        e += ((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / 2.0);
    }
    return e;
}

/* ------------------------------------------------------------------------------------
 * The MD.cpp loop.
 * ---------------------------------------------------------------------------------- */
struct orc_md {
    orc_sys *s;
    int rng_mode;
    int nstreams;
    orc_mt *lstreams;   /* Langevin: MTRand(seed + t) per thread (langevin.h:98-107) */
    orc_mt mc;          /* MC box move: MTRand(seed) (MD.cpp:149) */
    int32_t trials, accepted;
    uint64_t decisions;   /* accept bits of the trials so far, latest in bit 0 */
};

orc_md *orc_md_create(orc_sys *s, int32_t rng_mode, int32_t mt_streams)
{
    orc_md *m = (orc_md *)calloc(1, sizeof(orc_md));
    m->s = s;
    m->rng_mode = rng_mode;
    m->nstreams = mt_streams > 0 ? mt_streams : 1;
    m->lstreams = (orc_mt *)malloc(sizeof(orc_mt) * (size_t)m->nstreams);
    for (int t = 0; t < m->nstreams; t++) orc_mt_seed(&m->lstreams[t], s->seed + (uint32_t)t);
    orc_mt_seed(&m->mc, s->seed);
    return m;
}

void orc_md_destroy(orc_md *m)
{
    if (!m) return;
    free(m->lstreams);
    free(m);
}

static void zero_acc(orc_sys *s) { memset(s->acc, 0, sizeof(double) * 3 * (size_t)s->n); }

static void thermostat(orc_md *m, int init, uint64_t step)
{
    if (m->rng_mode == 0) orc_langevin_mt(m->s, m->lstreams, m->nstreams, m->s->temp);
    else langevin_psi(m->s, m->s->temp, init, step);
}

/* MD.cpp:186-262: zero a, pair forces, thermostat, molecule forces. */
void orc_md_init_forces(orc_md *m)
{
    zero_acc(m->s);
    orc_pair_force(m->s, m->s->acc);
    thermostat(m, 1, 0);
    orc_chain_force(m->s, m->s->acc);
}

/* MD.cpp:589-721: Metropolis constant-tension box move. */
static void mc_move(orc_md *m, int step)
{
    orc_sys *s = m->s;
    double size[3] = {s->box[0], s->box[1], s->box[2]};
    double old[3] = {s->box[0], s->box[1], s->box[2]};
    double psi, u;
    if (m->rng_mode == 0) {
        psi = 2.0 * orc_mt_rand53(&m->mc) - 1.0;
        u = NAN;
    } else {
        orc_philox_mc(s->seed, (uint64_t)step, &psi, &u);
    }
    double fx = s->delta_lxy * psi;
    double fy = fx;
    double fz = (size[0] * size[1]) / ((size[0] + fx) * (size[1] + fy));
    size[0] += fx;
    size[1] += fy;
    size[2] *= fz;
    double aS[3] = {size[0] / old[0], size[1] / old[1], size[2] / old[2]};
    double dP = orc_pair_dpotential(s, aS);
    dP += orc_chain_dpotential(s, aS);
    if (s->tension != 0) dP += s->tension * ((size[0] * size[1]) - (old[0] * old[1]));
    double D = dP / s->temp;
    D = exp(D);
    if (m->rng_mode == 0) u = orc_mt_rand53(&m->mc);
    const int acc = (D >= u || -dP <= 0);
    if (acc) {
        for (int i = 0; i < s->n; i++) {
            s->pos[3 * (size_t)i] *= aS[0];
            s->pos[3 * (size_t)i + 1] *= aS[1];
            s->pos[3 * (size_t)i + 2] *= aS[2];
        }
        s->box[0] = size[0]; s->box[1] = size[1]; s->box[2] = size[2];
        m->accepted++;
    }
    m->decisions = (m->decisions << 1) | (uint64_t)acc;
    m->trials++;
}

int32_t orc_md_run(orc_md *m, int32_t start, int32_t end)
{
    orc_sys *s = m->s;
    int32_t acc0 = m->accepted;
    for (int i = start; i <= end; i++) {
        orc_verlet_first(s);
        zero_acc(s);
        thermostat(m, 0, (uint64_t)i);
        orc_pair_force(s, s->acc);
        orc_chain_force(s, s->acc);
        orc_verlet_second(s);
        if (i % 8 == 0 && i != 0 && s->delta_lxy != 0) mc_move(m, i);
    }
    return m->accepted - acc0;
}

int32_t orc_md_trials(const orc_md *m) { return m->trials; }
uint64_t orc_md_decisions(const orc_md *m) { return m->decisions; }

/* The objective of the constant-fitting adjoint: init forces, steps 1..nsteps of the loop
 * above, J = mean over the steps of U_pair + U_chain taken after the kick and before the
 * box move (where MD.cpp:513-548 measures). */
double orc_md_objective(orc_md *m, int32_t nsteps)
{
    orc_sys *s = m->s;
    orc_md_init_forces(m);
    double J = 0.0;
    for (int i = 1; i <= nsteps; i++) {
        orc_verlet_first(s);
        zero_acc(s);
        thermostat(m, 0, (uint64_t)i);
        orc_pair_force(s, s->acc);
        orc_chain_force(s, s->acc);
        orc_verlet_second(s);
        J += orc_pair_potential(s) + orc_chain_potential(s);
        if (i % 8 == 0 && s->delta_lxy != 0) mc_move(m, i);
    }
    return J / nsteps;
}

/* dataExtraction::compute potential and Kinetic (dataExtraction.h:840-842,942-948;
 * dataCollection.h:667-673). */
void orc_md_measure(orc_md *m, double *U, double *K)
{
    double u = orc_pair_potential(m->s);
    u += orc_chain_potential(m->s);
    *U = u;
    *K = orc_kinetic(m->s);
}

/* ------------------------------------------------------------------------------------
 * Analysis.
 * ---------------------------------------------------------------------------------- */
static double vr_image(double d, double L)
{
    double m = 0.0;
    if (d > L / 2.0) m = L;
    if (d < -L / 2.0) m = -L;
    return m;
}

/* One frame of analysis/vesicleRadius.cpp:71-165: the centre of mass of the beads of type
 * `head`, each moved by minImg towards particle 0 (subtracted, :88-110), then the mean
 * distance from it with the image term ADDED to the difference (:137-151). */
double orc_vesicle_radius(const double *pos, const int32_t *type, int32_t n, const double box[3], int32_t head)
{
    double com[3] = {0.0, 0.0, 0.0}, nh = 0.0;
    for (int i = 0; i < n; i++) {
        if (type[i] != head) continue;
        for (int k = 0; k < 3; k++) {
            const double d = pos[3 * (size_t)i + k] - pos[k];
            com[k] += pos[3 * (size_t)i + k] - vr_image(d, box[k]);
        }
        nh++;
    }
    for (int k = 0; k < 3; k++) com[k] /= nh;
    double radius = 0.0;
    for (int i = 0; i < n; i++) {
        if (type[i] != head) continue;
        double d[3];
        for (int k = 0; k < 3; k++) {
            d[k] = pos[3 * (size_t)i + k] - com[k];
            d[k] += vr_image(d[k], box[k]);
        }
        radius += sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    }
    return radius / nh;
}
