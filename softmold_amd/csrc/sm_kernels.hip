// gfx950 kernels of the MD step apart from the cell traversals (sm_pair.hip) and the sort
// (sm_sort.hip): integrate + cell key, reorder + dense cell table, CHAIN bonds/bends fused
// with the Langevin force and the second half-kick, FP64 chain energy and kinetic terms,
// deterministic segmented FP64 sums, and the device-side Monte-Carlo box move.  Every
// kernel serves all replicas of a context in one launch; a bead finds its replica's box,
// thermostat and MC state through Bead::rep.
#include "sm_sort_dev.hpp"

namespace smdev {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------------------
// counter RNG: Philox4x32-10 (Salmon et al., SC'11); counter = (particle, step lo, step hi,
// tag), key = (seed lo, seed hi).  The particle word is the bead's index inside its
// replica, so a replica draws the same numbers alone or batched.
// ---------------------------------------------------------------------------------------
__device__ inline uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    return make_uint4(c0, c1, c2, c3);
}

// psi in (-1,1) = ((x>>8) 2^-23 - 1) + 2^-24; every step is exact in binary32.
__device__ inline float psi32(uint32_t x)
{
    return ((float)(x >> 8) * (1.0f / 8388608.0f) - 1.0f) + (1.0f / 16777216.0f);
}

__device__ inline double psi64(uint32_t x)
{
    return ((double)(x >> 8) * (1.0 / 8388608.0) - 1.0) + (1.0 / 16777216.0);
}

// Cell index along one axis: int() truncation of the FP64 quotient, then clamped into
// [0, n-1] (the reference's one-cell correction at the top edge, cellOpt.h:532-539).
__device__ inline int cell_axis(double p, double cs, int n)
{
    int c = (int)(p / cs);
    c = c >= n ? n - 1 : c;
    return c < 0 ? 0 : c;
}

__device__ inline bool in_box(double p, double L) { return p >= 0.0 && p <= L; }

// ---------------------------------------------------------------------------------------
// Verlet::first (verlet.h:289-356) + cell key (cellOpt.h:530-556)
// ---------------------------------------------------------------------------------------
// Half kick v += a dt/2 (Verlet::first's and Verlet::second's velocity update); every
// kernel that kicks uses this, so a kick rounds the same wherever it runs.
__device__ inline void half_kick(float4 &v, const float4 &a, float halfdt)
{
    v.x = fmaf(a.x, halfdt, v.x);
    v.y = fmaf(a.y, halfdt, v.y);
    v.z = fmaf(a.z, halfdt, v.z);
}

// Verlet::first's drift x += v dt and the wrap into [0, L] (verlet.h:289-356)
__device__ inline void drift_wrap(Bead &b, const float4 &v, double dt, double L0, double L1, double L2)
{
    b.x = fma((double)v.x, dt, b.x);
    b.y = fma((double)v.y, dt, b.y);
    b.z = fma((double)v.z, dt, b.z);
    if (b.x > L0) b.x -= L0;
    if (b.x < 0) b.x += L0;
    if (b.y > L1) b.y -= L1;
    if (b.y < 0) b.y += L1;
    if (b.z > L2) b.z -= L2;
    if (b.z < 0) b.z += L2;
}

// cell key of a bead (cellOpt.h:530-556), flagging a position outside the box
__device__ inline uint32_t bead_key(const Bead &b, const DevRep &R, DevGlobal *__restrict__ glob, int &cx, int &cy,
                                    int &cz)
{
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    if (!(in_box(b.x, R.L[0]) && in_box(b.y, R.L[1]) && in_box(b.z, R.L[2]))) atomicOr(&glob->flags, FLAG_BAD_POSITION);
    cx = cell_axis(b.x, R.cs[0], nx);
    cy = cell_axis(b.y, R.cs[1], ny);
    cz = cell_axis(b.z, R.cs[2], nz);
    return (uint32_t)(R.cellbase + (long long)(cx + cy * nx + cz * nx * ny));
}
__device__ inline uint32_t bead_key(const Bead &b, const DevRep &R, DevGlobal *__restrict__ glob)
{
    int cx, cy, cz;
    return bead_key(b, R, glob, cx, cy, cz);
}

// Verlet::first of bead i and its cell key
__device__ inline uint32_t first_one(int i, Bead *__restrict__ beads, float4 *__restrict__ vel,
                                     const float4 *__restrict__ acc, const DevRep *__restrict__ reps,
                                     DevGlobal *__restrict__ glob, float halfdt, double dt, int integrate, int &cx,
                                     int &cy, int &cz)
{
    Bead b = beads[i];
    const DevRep &R = reps[b.rep];
    if (integrate && b.type != 0) {
        float4 v = vel[i];
        half_kick(v, acc[i], halfdt);
        drift_wrap(b, v, dt, R.L[0], R.L[1], R.L[2]);
        vel[i] = v;
        beads[i] = b;
    }
    return bead_key(b, R, glob, cx, cy, cz);
}

// The incremental re-sort's mover test for bead i with key k (see MoveBook); returns
// whether the key goes into the histograms.
__device__ inline bool note_move(const MoveBook &book, int i, uint32_t k, bool valid)
{
    if (!valid || !book.cnt_b) return valid;
    return book.kold == nullptr || book.kold[i] != k;
}

// the block's mover count (every block writes one)
__device__ inline void flush_moves(const MoveBook &book, int nmoved)
{
    if (book.cnt_b && threadIdx.x == 0) book.cnt_b[blockIdx.x] = (uint32_t)nmoved;
}

__global__ __launch_bounds__(kBlock) void k_first(int n, Bead *__restrict__ beads, float4 *__restrict__ vel,
                                                  const float4 *__restrict__ acc, const DevRep *__restrict__ reps,
                                                  DevGlobal *__restrict__ glob, uint32_t *__restrict__ key,
                                                  uint32_t *__restrict__ idx, float halfdt, double dt, int integrate,
                                                  long long step_set, uint32_t *__restrict__ ghist, int passes,
                                                  MoveBook book, WinBook wb)
{
    __shared__ uint32_t hs[1024];
    uint32_t (*h)[256] = reinterpret_cast<uint32_t (*)[256]>(hs);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool win = wb.desc != nullptr;
    WinDesc w{};
    if (win) w = load_win(wb.desc);
    if (i == 0 && step_set >= -1) glob->step = step_set >= 0 ? step_set : glob->step + 1;   // no other reader here
    if (ghist) {
        if (win)
            for (int j = 0; j < 4; j++) hs[threadIdx.x + 256 * j] = 0u;
        else
            for (int p = 0; p < passes; p++) h[p][threadIdx.x] = 0u;
        __syncthreads();
    }
    uint32_t k = 0u;
    int cx = 0, cy = 0, cz = 0;
    if (i < n) k = first_one(i, beads, vel, acc, reps, glob, halfdt, dt, integrate, cx, cy, cz);
    if (ghist && win) {
        k = win_key_end(wb, w, hs, k, cx, cy, cz, i < n);
    } else if (ghist) {
        // the sort's digit histograms while the key is in a register (all lanes: ballots)
        const bool counted = note_move(book, i, k, i < n);
        hist_add(h, k, counted, passes);
        const int nmoved = __syncthreads_count(counted);
        hist_flush(h, ghist, passes);
        flush_moves(book, nmoved);
    }
    if (i < n) {
        key[i] = k;
        if (idx) idx[i] = (uint32_t)i;
    }
}

// ---------------------------------------------------------------------------------------
// reorder into cell order; cell runs written into the dense (never cleared) table; the
// next sort's scratch zeroed on the side
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_reorder(int n, const uint32_t *__restrict__ skey,
                                                    const uint32_t *__restrict__ sidx, const Bead *__restrict__ beads,
                                                    Bead *__restrict__ sbeads, int *__restrict__ cstart,
                                                    int *__restrict__ cend, uint32_t *__restrict__ clear,
                                                    size_t clear_words)
{
    const int s = blockIdx.x * kBlock + threadIdx.x;
    for (size_t w = (size_t)s; w < clear_words; w += (size_t)gridDim.x * kBlock) clear[w] = 0u;
    if (s >= n) return;
    // every load issued at once (the neighbours' indices clamped in bounds)
    const uint32_t k = skey[s], kprev = skey[s > 0 ? s - 1 : 0], knext = skey[s < n - 1 ? s + 1 : s];
    sbeads[s] = beads[sidx[s]];
    if (s == 0 || kprev != k) cstart[k] = s;
    if (s == n - 1 || knext != k) cend[k] = s + 1;
}

// k_reorder after a window sort: the sorted keys are window keys (the window path) or cell
// indices (the fallback); either way equal keys are equal cells.  The cell indices go to
// skf for the pair kernels.  Block 0 also sets the next sort's window: this sort's after
// the window path; after the fallback, the occupied cell range the fallback measured,
// widened by one cell on each side and clipped to the grid, or none (wx = 0) when that
// holds more than 2^18 cells.
__global__ __launch_bounds__(kBlock) void k_reorder_win(int n, const uint32_t *__restrict__ skey,
                                                        const uint32_t *__restrict__ sidx,
                                                        const Bead *__restrict__ beads, Bead *__restrict__ sbeads,
                                                        int *__restrict__ cstart, int *__restrict__ cend,
                                                        uint32_t *__restrict__ clear, size_t clear_words, WinBook wb,
                                                        uint32_t *__restrict__ skf, uint32_t *__restrict__ desc_next,
                                                        const DevRep *__restrict__ reps)
{
    const int nx = reps[0].nc[0], ny = reps[0].nc[1], nz = reps[0].nc[2];
    const int s = blockIdx.x * kBlock + threadIdx.x;
    for (size_t w = (size_t)s; w < clear_words; w += (size_t)gridDim.x * kBlock) clear[w] = 0u;
    const WinDesc w = load_win(wb.desc);
    const bool fallback = w.wx == 0u || *wb.flag != 0u;
    if (blockIdx.x == 0 && threadIdx.x < kWave) {
        const int t = threadIdx.x;
        uint32_t v = 0u;
        if (fallback && t < 6)
            for (int c = 0; c < kHistCopies; c++) v = max(v, wb.bbox[c * 8 + t]);
        uint32_t f[6];
#pragma unroll
        for (int j = 0; j < 6; j++) f[j] = (uint32_t)__shfl((int)v, j);
        if (t == 0) {
            uint32_t d[6] = {w.x0, w.y0, w.z0, w.wx, w.wy, w.wz};
            if (fallback) {
                for (int j = 0; j < 6; j++) d[j] = 0u;
                if (f[0] != 0u) {
                    const int ncell[3] = {nx, ny, nz};
                    unsigned long long cells = 1ull;
                    for (int a = 0; a < 3; a++) {
                        const int lo = max((int)(kBoxEnc - f[3 + a]) - 1, 0);
                        const int hi = min((int)f[a], ncell[a] - 1);   // f = highest cell + 1
                        d[a] = (uint32_t)lo;
                        d[3 + a] = (uint32_t)(hi - lo + 1);
                        cells *= (unsigned long long)(hi - lo + 1);
                    }
                    if (cells > (1ull << 18))
                        for (int j = 0; j < 6; j++) d[j] = 0u;
                }
            }
            for (int j = 0; j < 6; j++) desc_next[j] = d[j];
        }
    }
    if (s >= n) return;
    const uint32_t k = skey[s], kprev = skey[s > 0 ? s - 1 : 0], knext = skey[s < n - 1 ? s + 1 : s];
    sbeads[s] = beads[sidx[s]];
    uint32_t full = k;
    if (!fallback) {
        const uint32_t wk = k & ~kWinTag;
        const uint32_t q = wk / w.wx, x = wk - q * w.wx, z = q / w.wy, y = q - z * w.wy;
        full = ((z + w.z0) * (uint32_t)ny + (y + w.y0)) * (uint32_t)nx + (x + w.x0);
    }
    skf[s] = full;
    if (s == 0 || kprev != k) cstart[full] = s;
    if (s == n - 1 || knext != k) cend[full] = s + 1;
}

// ---------------------------------------------------------------------------------------
// CHAIN bonded forces by bead role (bond: harmonicF MD.h:384-403; bend: bendF
// MD.h:683-723; chain walk system.h:1782-1866), fused with the Langevin force
// (langevin.h:233-331) and Verlet::second (verlet.h:464-477).
// ---------------------------------------------------------------------------------------
// Bonded terms are evaluated in FP64: a bead touches at most 2 bonds and 3 bends, FP64
// add/mul/fma issue at the FP32 rate on gfx950, and the kernel is bandwidth bound.
__device__ inline double3 min_image(const Bead &p, const Bead &q, double L0, double L1, double L2)
{
    double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    if (dx > L0 / 2.0) dx -= L0;
    if (dx < -L0 / 2.0) dx += L0;
    if (dy > L1 / 2.0) dy -= L1;
    if (dy < -L1 / 2.0) dy += L1;
    if (dz > L2 / 2.0) dz -= L2;
    if (dz < -L2 / 2.0) dz += L2;
    return make_double3(dx, dy, dz);
}

__device__ inline double len3(const double3 &d) { return sqrt(d.x * d.x + d.y * d.y + d.z * d.z); }

// harmonic force on the first particle of the bond (a1 += d * mag), with |d| given
__device__ inline double3 bond_fl(double3 d, double dr, double l0, double kb)
{
    const double mag = -(dr - l0) * kb / dr;
    return make_double3(d.x * mag, d.y * mag, d.z * mag);
}

__device__ inline double3 bond_f(double3 d, double l0, double kb) { return bond_fl(d, len3(d), l0, kb); }

// bend: the force on the particle with role 0, 1 or 2 of (p1, p2, p3), given the two
// differences and their inverse lengths
__device__ inline double3 bend_fl(double3 da, double ia, double3 db, double ib, double c0, double kt, int role)
{
    da.x *= ia; da.y *= ia; da.z *= ia;
    db.x *= ib; db.y *= ib; db.z *= ib;
    const double cost = da.x * db.x + da.y * db.y + da.z * db.z;
    const double mag = (c0 - cost) * kt;
    const double ta = mag * ia, tb = mag * ib;
    const double3 A = make_double3(ta * (db.x - da.x * cost), ta * (db.y - da.y * cost), ta * (db.z - da.z * cost));
    const double3 B = make_double3(tb * (da.x - db.x * cost), tb * (da.y - db.y * cost), tb * (da.z - db.z * cost));
    if (role == 0) return A;
    if (role == 1) return make_double3(B.x - A.x, B.y - A.y, B.z - A.z);
    return make_double3(-B.x, -B.y, -B.z);
}

__device__ inline double3 bend_f(double3 da, double3 db, double c0, double kt, int role)
{
    return bend_fl(da, 1.0 / len3(da), db, 1.0 / len3(db), c0, kt, role);
}

// CHAIN and generic bonded force on bead i (FP64)
__device__ inline double3 bonded_force(int i, int n, const Bead &p0, const Bead *__restrict__ beads,
                                       const int2 *__restrict__ chain, const double4 *__restrict__ cconst,
                                       const GenTables &gen, double L0, double L1, double L2)
{
    double3 fc = make_double3(0.0, 0.0, 0.0);
    // the four beads around i, loaded with the chain record rather than after it (indices
    // clamped in bounds; a bead outside this one's chain is loaded and not used)
    const Bead pm2 = beads[max(i - 2, 0)], pm1 = beads[max(i - 1, 0)];
    const Bead pp1 = beads[min(i + 1, n - 1)], pp2 = beads[min(i + 2, n - 1)];
    const int2 ci = chain[i];
    // evaluated for every bead and kept only for a chain bead (no branch for the compiler to
    // sink the neighbour loads into); a bead outside any chain reads constant set 0
    {
        const bool inchain = ci.x >= 0;
        const int start = inchain ? ci.x : i, len = inchain ? (ci.y & 0xFFFF) : 1, cid = inchain ? (ci.y >> 16) : 0;
        const double4 K = cconst[cid];
        const int j = i - start;
        // the chain differences around this bead, each with its length and inverse length
        // computed once (the bonds and bends below share them)
        const double3 z3 = make_double3(0.0, 0.0, 0.0);
        const double3 dm2 = j >= 2 ? min_image(pm2, pm1, L0, L1, L2) : z3;        // p(i-2) - p(i-1)
        const double3 dm1 = j >= 1 ? min_image(pm1, p0, L0, L1, L2) : z3;         // p(i-1) - p(i)
        const double3 d0 = j <= len - 2 ? min_image(p0, pp1, L0, L1, L2) : z3;    // p(i) - p(i+1)
        const double3 d1 = j <= len - 3 ? min_image(pp1, pp2, L0, L1, L2) : z3;   // p(i+1) - p(i+2)
        const double rm2 = j >= 2 ? len3(dm2) : 1.0, rm1 = j >= 1 ? len3(dm1) : 1.0;
        const double r0 = j <= len - 2 ? len3(d0) : 1.0, r1 = j <= len - 3 ? len3(d1) : 1.0;
        if (j >= 1) {  // bond (i-1, i): this bead is a2
            const double3 f = bond_fl(dm1, rm1, K.x, K.y);
            fc.x -= f.x; fc.y -= f.y; fc.z -= f.z;
        }
        if (j <= len - 2) {  // bond (i, i+1): this bead is a1
            const double3 f = bond_fl(d0, r0, K.x, K.y);
            fc.x += f.x; fc.y += f.y; fc.z += f.z;
        }
        const double im2 = 1.0 / rm2, im1 = 1.0 / rm1, i0 = 1.0 / r0, i1 = 1.0 / r1;
        if (j <= len - 3) {  // bend (i, i+1, i+2), role 0
            const double3 f = bend_fl(d0, i0, d1, i1, K.z, K.w, 0);
            fc.x += f.x; fc.y += f.y; fc.z += f.z;
        }
        if (j >= 1 && j <= len - 2) {  // bend (i-1, i, i+1), role 1
            const double3 f = bend_fl(dm1, im1, d0, i0, K.z, K.w, 1);
            fc.x += f.x; fc.y += f.y; fc.z += f.z;
        }
        if (j >= 2) {  // bend (i-2, i-1, i), role 2
            const double3 f = bend_fl(dm2, im2, dm1, im1, K.z, K.w, 2);
            fc.x += f.x; fc.y += f.y; fc.z += f.z;
        }
    }
    if (gen.gptr) {
        // generic BOND / BEND records (doBondForce, doBendForce, system.h:1881-2035):
        // the same per-record force, taken from this bead's side
        const int e1 = gen.gptr[i + 1];
        for (int e = gen.gptr[i]; e < e1; e++) {
            const GenEnt g = gen.gent[e];
            const double2 K = gen.gconst[e];
            double3 f;
            if (g.kind == 0) {
                f = bond_f(min_image(beads[g.p1], beads[g.p2], L0, L1, L2), K.x, K.y);
                if (i != g.p1) { f.x = -f.x; f.y = -f.y; f.z = -f.z; }
            } else if (g.kind == 2) {
                // harmonicHalfF (MD.h:1164-1187): only inside r0
                const double3 d = min_image(beads[g.p1], beads[g.p2], L0, L1, L2);
                f = (d.x * d.x + d.y * d.y + d.z * d.z < K.x * K.x) ? bond_f(d, K.x, K.y) : make_double3(0.0, 0.0, 0.0);
                if (i != g.p1) { f.x = -f.x; f.y = -f.y; f.z = -f.z; }
            } else {
                const Bead q1 = beads[g.p1], q2 = beads[g.p2], q3 = beads[g.p3];
                f = bend_f(min_image(q1, q2, L0, L1, L2), min_image(q2, q3, L0, L1, L2), K.x, K.y,
                           i == g.p1 ? 0 : (i == g.p2 ? 1 : 2));
            }
            fc.x += f.x; fc.y += f.y; fc.z += f.z;
        }
    }
    return fc;
}

// total acceleration of bead i at `step`: pair + bonded + Langevin (-gamma v + sigma xi)
__device__ inline float4 total_accel(int i, const DevRep &R, const double3 &fc, const float4 &fp, const float4 &v,
                                     long long step, uint32_t tag)
{
    const uint4 o = philox((uint32_t)(i - R.first), (uint32_t)step, (uint32_t)((unsigned long long)step >> 32), tag,
                           (uint32_t)R.seed, (uint32_t)(R.seed >> 32));
    const float gamma = (float)R.gamma, sigma = (float)R.sigma;
    float4 a;
    a.x = (-gamma * v.x + sigma * psi32(o.x));
    a.y = (-gamma * v.y + sigma * psi32(o.y));
    a.z = (-gamma * v.z + sigma * psi32(o.z));
    a.x += fp.x; a.y += fp.y; a.z += fp.z;
    a.x += (float)fc.x; a.y += (float)fc.y; a.z += (float)fc.z;
    a.w = 0.f;
    return a;
}

__global__ __launch_bounds__(kBlock) void k_bonded_kick(int n, const Bead *__restrict__ beads,
                                                        const int2 *__restrict__ chain,
                                                        const double4 *__restrict__ cconst, GenTables gen,
                                                        const float4 *__restrict__ fpair, float4 *__restrict__ vel,
                                                        float4 *__restrict__ acc, const DevRep *__restrict__ reps,
                                                        const DevGlobal *__restrict__ glob, float halfdt, int mode,
                                                        float4 *__restrict__ fchain_out)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Bead p0 = beads[i];
    const DevRep &R = reps[p0.rep];
    const double3 fc = bonded_force(i, n, p0, beads, chain, cconst, gen, R.L[0], R.L[1], R.L[2]);
    if (mode == KICK_NONE) {
        fchain_out[i] = make_float4((float)fc.x, (float)fc.y, (float)fc.z, 0.f);
        return;
    }
    float4 v = vel[i];
    const float4 a = total_accel(i, R, fc, fpair[i], v, glob->step, mode == KICK_INIT ? TAG_LANGEVIN_INIT : TAG_LANGEVIN);
    acc[i] = a;
    if (mode == KICK_STEP && p0.type != 0) {
        half_kick(v, a, halfdt);
        vel[i] = v;
    }
}

// Step `step`'s bonded forces, Langevin force and second half kick, then Verlet::first and
// the cell keys of step + 1, in one pass over the beads: the same arithmetic as
// k_bonded_kick followed by k_first.  The drifted positions go to beads_out, since other
// beads' bonded forces still read this step's positions; the step counter is taken as an
// argument (no thread reads the device counter, which thread 0 advances to step + 1).
// ONE: a single-replica context, whose replica constants are read from reps[0] without
// waiting for the bead's replica index
// WIN: keys for a window sort (wb)
template <bool ONE, bool WIN>
__global__ __launch_bounds__(kBlock) void k_kick_first(int n, const Bead *__restrict__ beads,
                                                       Bead *__restrict__ beads_out,
                                                       const int2 *__restrict__ chain,
                                                       const double4 *__restrict__ cconst, GenTables gen,
                                                       const float4 *__restrict__ fpair, float4 *__restrict__ vel,
                                                       float4 *__restrict__ acc, const DevRep *__restrict__ reps,
                                                       DevGlobal *__restrict__ glob, float halfdt, double dt,
                                                       long long step, uint32_t *__restrict__ key,
                                                       uint32_t *__restrict__ idx, uint32_t *__restrict__ ghist,
                                                       int passes, MoveBook book, WinBook wb)
{
    __shared__ uint32_t hs[1024];
    uint32_t (*h)[256] = reinterpret_cast<uint32_t (*)[256]>(hs);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    WinDesc w{};
    if (WIN) w = load_win(wb.desc);
    if (i == 0) glob->step = step + 1;
    if (WIN)
        for (int j = 0; j < 4; j++) hs[threadIdx.x + 256 * j] = 0u;
    else
        for (int p = 0; p < passes; p++) h[p][threadIdx.x] = 0u;
    __syncthreads();
    uint32_t k = 0u;
    int cx = 0, cy = 0, cz = 0;
    if (i < n) {
        Bead b = beads[i];
        const DevRep &R = reps[ONE ? 0 : b.rep];
        const double3 fc = bonded_force(i, n, b, beads, chain, cconst, gen, R.L[0], R.L[1], R.L[2]);
        float4 v = vel[i];
        const float4 a = total_accel(i, R, fc, fpair[i], v, step, TAG_LANGEVIN);
        acc[i] = a;
        if (b.type != 0) {
            half_kick(v, a, halfdt);   // Verlet::second of this step
            half_kick(v, a, halfdt);   // Verlet::first of the next
            drift_wrap(b, v, dt, R.L[0], R.L[1], R.L[2]);
            vel[i] = v;
        }
        beads_out[i] = b;
        k = bead_key(b, R, glob, cx, cy, cz);
    }
    if (WIN) {
        k = win_key_end(wb, w, hs, k, cx, cy, cz, i < n);
    } else {
        const bool counted = note_move(book, i, k, i < n);
        hist_add(h, k, counted, passes);
        const int nmoved = __syncthreads_count(counted);
        hist_flush(h, ghist, passes);
        flush_moves(book, nmoved);
    }
    if (i < n) {
        key[i] = k;
        if (idx) idx[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_second(int n, const Bead *__restrict__ beads,
                                                   const float4 *__restrict__ acc, float4 *__restrict__ vel,
                                                   float halfdt)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (beads[i].type == 0) return;
    float4 v = vel[i];
    half_kick(v, acc[i], halfdt);
    vel[i] = v;
}

// ---------------------------------------------------------------------------------------
// Segmented FP64 sums with fixed-shape trees: the result is independent of scheduling
// and of the other replicas.
// ---------------------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *lds)
{
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) lds[threadIdx.x] += lds[threadIdx.x + off];
        __syncthreads();
    }
    const double r = lds[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void k_seg_tiles(const double *__restrict__ x, const int2 *__restrict__ tile,
                                                      double *__restrict__ tilesum)
{
    __shared__ double lds[kBlock];
    const int2 t = tile[blockIdx.x];
    double s = 0.0;
    for (int i = t.x + (int)threadIdx.x; i < t.y; i += kBlock) s += x[i];
    const double b = block_sum(s, lds);
    if (threadIdx.x == 0) tilesum[blockIdx.x] = b;
}

__global__ __launch_bounds__(kBlock) void k_seg_final(const double *__restrict__ tilesum,
                                                      const int *__restrict__ rep_first, double *__restrict__ out,
                                                      int stride, int add)
{
    __shared__ double lds[kBlock];
    const int r = blockIdx.x;
    double s = 0.0;
    for (int t = rep_first[r] + (int)threadIdx.x; t < rep_first[r + 1]; t += kBlock) s += tilesum[t];
    const double b = block_sum(s, lds);
    if (threadIdx.x == 0) out[(size_t)r * stride] = add ? out[(size_t)r * stride] + b : b;
}

// CHAIN energies in FP64 (harmonicP MD.h:422-431, bendP MD.h:769-789, walk
// system.h:2487-2568 / dPotential system.h:3280-3384), one thread per chain.
__device__ inline void mi3(const Bead &p, const Bead &q, const double *L, double d[3])
{
    d[0] = p.x - q.x; d[1] = p.y - q.y; d[2] = p.z - q.z;
    for (int k = 0; k < 3; k++) {
        if (d[k] > L[k] / 2.0) d[k] -= L[k];
        if (d[k] < -L[k] / 2.0) d[k] += L[k];
    }
}

__device__ inline double harm_p(const double d[3], double l0, double kb)
{
    const double dr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double p = dr - l0;
    return 0.5 * kb * p * p;
}

__device__ inline double bend_p(const double a[3], const double b[3], double c0, double kt)
{
    const double ra = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double rb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    const double cost = (a[0] / ra) * (b[0] / rb) + (a[1] / ra) * (b[1] / rb) + (a[2] / ra) * (b[2] / rb);
    const double p = c0 - cost;
    return kt * p * p * 0.5;
}

__global__ __launch_bounds__(kBlock) void k_chain_energy(int nch, const int2 *__restrict__ chains,
                                                         const double4 *__restrict__ cconst,
                                                         const Bead *__restrict__ beads,
                                                         const DevRep *__restrict__ reps, int delta,
                                                         double *__restrict__ e)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= nch) return;
    const int2 ch = chains[c];
    const int start = ch.x, len = ch.y & 0xFFFF;
    const DevRep &R = reps[beads[start].rep];
    const double L[3] = {R.L[0], R.L[1], R.L[2]};
    const double S[3] = {R.scale[0], R.scale[1], R.scale[2]};
    const double4 K = cconst[ch.y >> 16];
    double accu = 0.0;
    for (int l = start; l <= start + len - 3; l++) {
        const bool last = (l == start + len - 3);
        double da[3], db[3];
        mi3(beads[l], beads[l + 1], L, da);
        mi3(beads[l + 1], beads[l + 2], L, db);
        double u = harm_p(da, K.x, K.y);
        if (last) u += harm_p(db, K.x, K.y);
        u += bend_p(da, db, K.z, K.w);
        if (delta) {
            for (int d = 0; d < 3; d++) { da[d] *= S[d]; db[d] *= S[d]; }
            double un = harm_p(da, K.x, K.y);
            if (last) un += harm_p(db, K.x, K.y);
            un += bend_p(da, db, K.z, K.w);
            accu += (u - un);
        } else {
            accu += u;
        }
    }
    e[c] = accu;
}

// Generic BOND / BEND energies (doBondPotential / doBendPotential, system.h:2718-2830;
// dPotential with the imaged differences scaled, system.h:3399-3530), one term per bead:
// the records whose first bead it is.
__global__ __launch_bounds__(kBlock) void k_gen_energy(int n, GenTables gen, const Bead *__restrict__ beads,
                                                       const DevRep *__restrict__ reps, int delta,
                                                       double *__restrict__ e)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const DevRep &R = reps[beads[i].rep];
    const double L[3] = {R.L[0], R.L[1], R.L[2]};
    const double S[3] = {R.scale[0], R.scale[1], R.scale[2]};
    double accu = 0.0;
    const int e1 = gen.gptr[i + 1];
    for (int k = gen.gptr[i]; k < e1; k++) {
        const GenEnt g = gen.gent[k];
        if (g.p1 != i) continue;
        const double2 K = gen.gconst[k];
        double da[3], db[3], u, un = 0.0;
        mi3(beads[g.p1], beads[g.p2], L, da);
        if (g.kind == 0) {
            u = harm_p(da, K.x, K.y);
            if (delta) {
                for (int d = 0; d < 3; d++) da[d] *= S[d];
                un = harm_p(da, K.x, K.y);
            }
        } else if (g.kind == 2) {
            // harmonicHalfP (MD.h:1210-1222) in U; nothing in dU (MD.cpp:664)
            u = (!delta && da[0] * da[0] + da[1] * da[1] + da[2] * da[2] < K.x * K.x) ? harm_p(da, K.x, K.y) : 0.0;
        } else {
            mi3(beads[g.p2], beads[g.p3], L, db);
            u = bend_p(da, db, K.x, K.y);
            if (delta) {
                for (int d = 0; d < 3; d++) { da[d] *= S[d]; db[d] *= S[d]; }
                un = bend_p(da, db, K.x, K.y);
            }
        }
        accu += delta ? (u - un) : u;
    }
    e[i] = accu;
}

// Bond and bend statistics of the generic records (dataExtraction.h:861-930), one term per
// bead over the records whose first bead it is: what 0: bond length; 1: cos theta of a
// bend; 2, 3: its two arm lengths.  These image with >= / <= against half the box, unlike
// the forces.
__device__ inline void mi3_closed(const Bead &p, const Bead &q, const double *L, double d[3])
{
    d[0] = p.x - q.x; d[1] = p.y - q.y; d[2] = p.z - q.z;
    for (int k = 0; k < 3; k++) {
        if (d[k] >= L[k] / 2.0) d[k] -= L[k];
        if (d[k] <= -L[k] / 2.0) d[k] += L[k];
    }
}

__global__ __launch_bounds__(kBlock) void k_gen_stats(int n, GenTables gen, const Bead *__restrict__ beads,
                                                      const DevRep *__restrict__ reps, int what,
                                                      double *__restrict__ e)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const DevRep &R = reps[beads[i].rep];
    const double L[3] = {R.L[0], R.L[1], R.L[2]};
    double accu = 0.0;
    const int e1 = gen.gptr[i + 1];
    for (int k = gen.gptr[i]; k < e1; k++) {
        const GenEnt g = gen.gent[k];
        if (g.p1 != i || g.kind == 2 || (g.kind == 0) != (what == 0)) continue;
        double da[3], db[3];
        mi3_closed(beads[g.p1], beads[g.p2], L, da);
        const double ra = sqrt(da[0] * da[0] + da[1] * da[1] + da[2] * da[2]);
        if (what == 0 || what == 2) {
            accu += ra;
            continue;
        }
        mi3_closed(beads[g.p2], beads[g.p3], L, db);
        const double rb = sqrt(db[0] * db[0] + db[1] * db[1] + db[2] * db[2]);
        accu += what == 3 ? rb : (da[0] * db[0] + da[1] * db[1] + da[2] * db[2]) / (ra * rb);
    }
    e[i] = accu;
}

// Kinetic::compute (dataCollection.h:667-673), one term per bead
__global__ __launch_bounds__(kBlock) void k_kinetic(int n, const float4 *__restrict__ vel, double *__restrict__ e)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 v = vel[i];
    const double x = v.x, y = v.y, z = v.z;
    e[i] = (x * x + y * y + z * z) / 2.0;
}

// ---------------------------------------------------------------------------------------
// Vesicle radius of each replica from its head beads (analysis/vesicleRadius.cpp:71-175),
// one FP64 term per bead for the segmented sums:
//   what 0..2: coordinate `what` of a head, moved by one box length towards the replica's
//              first bead when farther than half a box from it (the centre-of-mass loop);
//   what 3:    1 per head (the head count);
//   what 4:    the head's distance to the centre com[3r..3r+2], with the second loop's
//              image term, which the reference ADDS to the difference (d += minImg, :147-151):
//              the two agree with a true minimum image whenever no head lies more than half
//              a box from the centre, and differ, as the reference does, otherwise.
// Non-head beads give 0.
// ---------------------------------------------------------------------------------------
__device__ inline double image_shift(double d, double L)
{
    double m = 0.0;
    if (d > L / 2.0) m = L;
    if (d < -L / 2.0) m = -L;
    return m;
}

__global__ __launch_bounds__(kBlock) void k_head_terms(int n, const Bead *__restrict__ beads,
                                                       const DevRep *__restrict__ reps, int head, int what,
                                                       const double *__restrict__ com, double *__restrict__ e)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Bead b = beads[i];
    double v = 0.0;
    if (b.type == head) {
        const DevRep &R = reps[b.rep];
        if (what < 3) {
            const Bead p0 = beads[R.first];
            const double x = what == 0 ? b.x : (what == 1 ? b.y : b.z);
            const double x0 = what == 0 ? p0.x : (what == 1 ? p0.y : p0.z);
            v = x - image_shift(x - x0, R.L[what]);
        } else if (what == 3) {
            v = 1.0;
        } else {
            const double *c = com + 3 * b.rep;
            double dx = b.x - c[0], dy = b.y - c[1], dz = b.z - c[2];
            dx += image_shift(dx, R.L[0]);
            dy += image_shift(dy, R.L[1]);
            dz += image_shift(dz, R.L[2]);
            v = sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    e[i] = v;
}

// sums[4r + 0..3] = (sum x, sum y, sum z, heads) -> com[3r + 0..2]
__global__ void k_head_com(int nrep, const double *__restrict__ sums, double *__restrict__ com)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    for (int k = 0; k < 3; k++) com[3 * r + k] = sums[4 * r + k] / sums[4 * r + 3];
}

// ---------------------------------------------------------------------------------------
// Monte-Carlo constant-tension move (MD.cpp:589-721), decided on the device, one thread
// per replica.  A replica with delta_lxy == 0 makes no trial.
// ---------------------------------------------------------------------------------------
__global__ void k_mc_propose(int nrep, DevRep *reps, const DevGlobal *__restrict__ glob)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    DevRep &R = reps[r];
    R.accept = 0;
    if (R.delta_lxy == 0) {
        R.scale[0] = R.scale[1] = R.scale[2] = 1.0;
        return;
    }
    const long long step = glob->step;
    const uint4 o = philox(0u, (uint32_t)step, (uint32_t)((unsigned long long)step >> 32), TAG_MC,
                           (uint32_t)R.seed, (uint32_t)(R.seed >> 32));
    const double psi = psi64(o.x);
    R.u_accept = ((o.y >> 5) * 67108864.0 + (o.z >> 6)) * (1.0 / 9007199254740992.0);
    const double Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
    const double fx = R.delta_lxy * psi, fy = fx;
    const double fz = (Lx * Ly) / ((Lx + fx) * (Ly + fy));
    R.fluct = fx;
    R.newL[0] = Lx + fx;
    R.newL[1] = Ly + fy;
    R.newL[2] = Lz * fz;
    R.scale[0] = R.newL[0] / Lx;
    R.scale[1] = R.newL[1] / Ly;
    R.scale[2] = R.newL[2] / Lz;
}

__global__ void k_mc_decide(int nrep, DevRep *reps, DevGlobal *glob, const double *__restrict__ du, double cutoff)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrep) return;
    DevRep &R = reps[r];
    if (R.delta_lxy == 0) return;
    const double dup = du[2 * r], duc = du[2 * r + 1];
    const double dL = R.delta_lxy;
    const double Lx = R.L[0], Ly = R.L[1];
    double dP = dup;
    dP += duc;
    if (R.tension != 0) dP += R.tension * ((R.newL[0] * R.newL[1]) - (Lx * Ly));
    double D = dP / R.temperature;
    D = exp(D);
    bool acc = (D >= R.u_accept) || (-dP <= 0);
    int nc3[3];
    long long cells = 1;
    for (int d = 0; d < 3; d++) {
        nc3[d] = (int)(R.newL[d] / cutoff);
        cells *= nc3[d];
    }
    if (acc && (cells > R.capacity || nc3[0] < 3 || nc3[1] < 3 || nc3[2] < 3)) {
        atomicOr(&glob->flags, FLAG_CELL_OVERFLOW);
        acc = false;
    }
    R.dU_pair = dup;
    R.dU_chain = duc;
    const double interval = dL / 101.0;
    const long long bin = (long long)((R.fluct + dL) / interval);
    if (acc) {
        for (int d = 0; d < 3; d++) {
            R.L[d] = R.newL[d];
            R.nc[d] = nc3[d];
            R.cs[d] = R.newL[d] / nc3[d];
        }
        R.accepted++;
        if (bin >= 0 && bin < kHistBins) R.hist_accept[bin] += 1.0;
    } else if (bin >= 0 && bin < kHistBins) {
        R.hist_reject[bin] += 1.0;
    }
    R.accept = acc ? 1 : 0;
    R.trials++;
}

__global__ __launch_bounds__(kBlock) void k_rescale(int n, Bead *__restrict__ beads, const DevRep *__restrict__ reps)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Bead b = beads[i];
    const DevRep &R = reps[b.rep];
    if (!R.accept) return;
    b.x *= R.scale[0];
    b.y *= R.scale[1];
    b.z *= R.scale[2];
    beads[i] = b;
}

__global__ void k_set_step(DevGlobal *glob, long long step) { glob->step = step; }

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, size_t count, uint32_t value)
{
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < count) p[i] = value;
}

inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
void launch_first(hipStream_t st, const Params &P, Bead *beads, float4 *vel, const float4 *acc, const DevRep *reps,
                  DevGlobal *glob, uint32_t *key, uint32_t *idx, int integrate, long long step_set, uint32_t *ghist,
                  int passes, const MoveBook &book, const WinBook &wb)
{
    hipLaunchKernelGGL(k_first, dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads, vel, acc, reps, glob, key, idx,
                       (float)(0.5 * P.dt), P.dt, integrate, step_set, ghist, passes, book, wb);
}

void launch_reorder(hipStream_t st, int n, const uint32_t *skey, const uint32_t *sidx, const Bead *beads,
                    Bead *sbeads, int *cstart, int *cend, uint32_t *clear, size_t clear_words)
{
    hipLaunchKernelGGL(k_reorder, dim3(blocks(n)), dim3(kBlock), 0, st, n, skey, sidx, beads, sbeads, cstart, cend,
                       clear, clear_words);
}

void launch_reorder_win(hipStream_t st, int n, const uint32_t *skey, const uint32_t *sidx, const Bead *beads,
                        Bead *sbeads, int *cstart, int *cend, uint32_t *clear, size_t clear_words,
                        const WinBook &wb, uint32_t *skey_full, uint32_t *desc_next, const DevRep *reps)
{
    hipLaunchKernelGGL(k_reorder_win, dim3(blocks(n)), dim3(kBlock), 0, st, n, skey, sidx, beads, sbeads, cstart,
                       cend, clear, clear_words, wb, skey_full, desc_next, reps);
}

void launch_bonded_kick(hipStream_t st, const Params &P, const Bead *beads, const int2 *chain,
                        const double4 *cconst, const GenTables &gen, const float4 *fpair, float4 *vel, float4 *acc,
                        const DevRep *reps, const DevGlobal *glob, int mode, float4 *fchain_out)
{
    hipLaunchKernelGGL(k_bonded_kick, dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads, chain, cconst, gen, fpair,
                       vel, acc, reps, glob, (float)(P.dt * 0.5), mode, fchain_out);
}

void launch_kick_first(hipStream_t st, const Params &P, const Bead *beads, Bead *beads_out, const int2 *chain,
                       const double4 *cconst, const GenTables &gen, const float4 *fpair, float4 *vel, float4 *acc,
                       const DevRep *reps, DevGlobal *glob, long long step, uint32_t *key, uint32_t *idx,
                       uint32_t *ghist, int passes, const MoveBook &book, const WinBook &wb)
{
    const float halfdt = (float)(P.dt * 0.5);
    if (P.nrep == 1 && wb.desc)
        hipLaunchKernelGGL((k_kick_first<true, true>), dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads, beads_out,
                           chain, cconst, gen, fpair, vel, acc, reps, glob, halfdt, P.dt, step, key, idx, ghist,
                           passes, book, wb);
    else if (P.nrep == 1)
        hipLaunchKernelGGL((k_kick_first<true, false>), dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads,
                           beads_out, chain, cconst, gen, fpair, vel, acc, reps, glob, halfdt, P.dt, step, key, idx,
                           ghist, passes, book, wb);
    else
        hipLaunchKernelGGL((k_kick_first<false, false>), dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads,
                           beads_out, chain, cconst, gen, fpair, vel, acc, reps, glob, halfdt, P.dt, step, key, idx,
                           ghist, passes, book, wb);
}

void launch_second(hipStream_t st, int n, const Bead *beads, const float4 *acc, float4 *vel, float halfdt)
{
    hipLaunchKernelGGL(k_second, dim3(blocks(n)), dim3(kBlock), 0, st, n, beads, acc, vel, halfdt);
}

void launch_seg_sum(hipStream_t st, const SegTiles &S, const double *x, double *out, int stride, bool add)
{
    if (S.ntiles > 0)
        hipLaunchKernelGGL(k_seg_tiles, dim3(S.ntiles), dim3(kBlock), 0, st, x, S.tile, S.tilesum);
    hipLaunchKernelGGL(k_seg_final, dim3(S.nrep), dim3(kBlock), 0, st, S.tilesum, S.rep_first, out, stride,
                       add ? 1 : 0);
}

void launch_gen_stats(hipStream_t st, int n, const GenTables &gen, const Bead *beads, const DevRep *reps, int what,
                      double *e)
{
    hipLaunchKernelGGL(k_gen_stats, dim3(blocks(n)), dim3(kBlock), 0, st, n, gen, beads, reps, what, e);
}

void launch_gen_energy(hipStream_t st, int n, const GenTables &gen, const Bead *beads, const DevRep *reps,
                       int delta, double *e)
{
    hipLaunchKernelGGL(k_gen_energy, dim3(blocks(n)), dim3(kBlock), 0, st, n, gen, beads, reps, delta, e);
}

void launch_chain_energy(hipStream_t st, int nchains, const int2 *chains, const double4 *cconst, const Bead *beads,
                         const DevRep *reps, int delta, double *e)
{
    if (nchains <= 0) return;
    hipLaunchKernelGGL(k_chain_energy, dim3(blocks(nchains)), dim3(kBlock), 0, st, nchains, chains, cconst, beads,
                       reps, delta, e);
}

void launch_kinetic(hipStream_t st, int n, const float4 *vel, double *e)
{
    hipLaunchKernelGGL(k_kinetic, dim3(blocks(n)), dim3(kBlock), 0, st, n, vel, e);
}

void launch_head_terms(hipStream_t st, int n, const Bead *beads, const DevRep *reps, int head, int what,
                       const double *com, double *e)
{
    hipLaunchKernelGGL(k_head_terms, dim3(blocks(n)), dim3(kBlock), 0, st, n, beads, reps, head, what, com, e);
}

void launch_head_com(hipStream_t st, int nrep, const double *sums, double *com)
{
    hipLaunchKernelGGL(k_head_com, dim3(blocks(nrep)), dim3(kBlock), 0, st, nrep, sums, com);
}

void launch_mc_propose(hipStream_t st, int nrep, DevRep *reps, const DevGlobal *glob)
{
    hipLaunchKernelGGL(k_mc_propose, dim3(blocks(nrep)), dim3(kBlock), 0, st, nrep, reps, glob);
}

void launch_mc_decide(hipStream_t st, const Params &P, DevRep *reps, DevGlobal *glob, const double *du)
{
    hipLaunchKernelGGL(k_mc_decide, dim3(blocks(P.nrep)), dim3(kBlock), 0, st, P.nrep, reps, glob, du, P.cutoff);
}

void launch_rescale(hipStream_t st, int n, Bead *beads, const DevRep *reps)
{
    hipLaunchKernelGGL(k_rescale, dim3(blocks(n)), dim3(kBlock), 0, st, n, beads, reps);
}

void launch_set_step(hipStream_t st, DevGlobal *glob, long long step)
{
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, st, glob, step);
}

void launch_fill_u32(hipStream_t st, uint32_t *p, size_t count, uint32_t value)
{
    if (count == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((count + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, p, count,
                       value);
}

}  // namespace smdev
