// Cell-stencil traversals: the FP32 Laradji-Revalee pair force (the hot kernel) and the
// FP64 gathers (potential, dU of a box trial, neighbour counts, pair list).
//
// Reference: CellOpt::computeForce / computePotential / computeDPotential
// (include/algorithms/cellOpt.h:712-1180) with Force/Potential (include/MD.h:795-930);
// GPU twins collideForce/Potential/DPotential_kernel (include/cuda/cell.h:564-805).
//
// Work decomposition: the particles are sorted by cell key; one wave64 takes a chunk of
// `chunk` sorted slots and owns every cell whose run starts inside it.  Per cell:
//   1. lanes 0..26 look up the 27 stencil cells in the dense cell table (validated
//      against the sorted keys, so the table is never cleared);
//   2. a wave scan of the 27 counts lays the candidates out contiguously and every lane
//      stages its share into LDS (the stencil cell of an item from a per-pass LDS map of
//      the cells' first slots and a DPP max-scan; its image shift and sorted slot from
//      per-cell LDS tables), so all global loads of a pass are in flight together; a
//      pass holds kMaxC candidates and runs only the 64-slot rounds it fills;
//   3. tpp = 64/H lanes serve each of the H home particles, each lane walking every
//      tpp-th candidate; a home's partial results are combined by a fixed-shape tree of
//      shuffles (at step s, lane sub adds lane sub + s when sub is a multiple of 2s).
// Force, one phase (M_FORCE1, the default): candidates are staged as FP64 imaged
// positions, and each staged candidate is entered into the list of each quarter of the home
// cell (x below / above the cell centre times y below / above) whose box it comes within
// its type's reach of (rc, or r_m for a type whose every pair has zero force and potential
// beyond r_m); a home particle walks the list of its quarter only, and no pair with a
// nonzero term is skipped.  A lane takes the FP64 difference, applies the exact FP64 test
// r^2 < rc^2 (no contraction, as MD.h:798-810) and under that branch evaluates the force
// law in FP32 with a Newton-corrected r.  Force, two phase (M_FORCE, kept as a tuning variant):
// candidates are staged a second time as FP32 positions relative to the home cell
// centre; a lane first screens kList candidates with the FP32 copy against a cutoff
// widened by more than the FP32 error, keeping the hits in a per-lane LDS list, then runs
// the exact FP64 test and the force law on the hits only.  Gathers: FP64 throughout, each pair's
// difference taken in the reference's direction -- the particle of the cell that owns
// the pair in the 13-cell half shell (cellOpt.h:715) minus the imaged other.  The energy
// gathers (U, dU) stage only the home cell and its forward half shell and use the same
// quarter-cell culling, with the reach radius widened by the trial scale for dU.
#include "sm_internal.hpp"

namespace smdev {

namespace {

constexpr int kWaves = 1;   // waves per workgroup: one, so a finished wave frees its slot at once
constexpr int kList = 32;      // candidates a lane screens before evaluating its hits

enum : int { M_FORCE = 0, M_U = 1, M_DU = 2, M_COUNT = 3, M_LIST = 4, M_FORCE1 = 5, M_FCOUNT = 6, M_FDU = 7 };

struct __align__(16) Cand {
    double x, y, z;
    int t, c;
};

__device__ inline int wave_incl_scan(int v, int lane)
{
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int t = __shfl_up(v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// Inclusive prefix maximum over the wave of non-negative values, in DPP steps: shifts by
// 1, 2, 4, 8 within each 16-lane row, then row 0's and row 2's last lanes into rows 1 and
// 3, then lane 31 into rows 2 and 3.  Lanes a step does not reach keep their value.
__device__ inline int wave_max_scan(int v)
{
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true));   // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true));   // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true));   // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true));   // row_shr:8
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}

// Image shift of one stencil cell: the box length added to a candidate's coordinates
// where the stencil crosses the periodic boundary.
struct __align__(16) Shift {
    double x, y, z, pad;
};

__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dispatched round-robin over the 8 XCDs.  G == 0: each XCD gets one
// contiguous range of the sorted particles (shared stencil cells stay in one L2); G > 0:
// the XCDs take turns at groups of G consecutive chunks (the last partial round of groups
// in order), which balances systems whose work is uneven along the sort order -- the
// half-shell energy gathers over a flat bilayer ran 19 % faster, while the force over a
// vesicle ran 5 % slower.
template <int G>
__device__ inline int xcd_block(int b, int nblocks)
{
    if (G == 0) {
        const int x = b & 7, q = nblocks >> 3, r = nblocks & 7;
        return x * q + min(x, r) + (b >> 3);
    }
    const int full = (nblocks / (8 * G)) * (8 * G);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / (G > 0 ? G : 1)) * 8 + x) * G + (k % (G > 0 ? G : 1));
}

// r^2 rounded as the reference computes it, (x*x + y*y) + z*z with every product and sum
// rounded separately: HIP contracts a*b + c into one FMA by default, which would decide
// pairs at the cutoff differently from the reference (MD.h:798-810 on x86-64 has no FMA).
__device__ inline double sq3(double x, double y, double z)
{
#pragma clang fp contract(off)
    return (x * x + y * y) + z * z;
}



// sqrt of a positive, normal double: v_rsq_f64 and the two Goldschmidt corrections the
// device library uses, without its range scaling and special-value fix-ups (r^2 here lies
// in (0, 4 rc^2]); 0 maps to 0.
__device__ inline double sqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double r = x * y, h = 0.5 * y;
    const double e = fma(-h, r, 0.5);
    r = fma(r, e, r);
    h = fma(h, e, h);
    double d = fma(-r, r, x);
    r = fma(d, h, r);
    d = fma(-r, r, x);
    r = fma(d, h, r);
    return x > 0.0 ? r : 0.0;
}

// sqrt_pos's sequence returning r and 1/r (twice the refined half-reciprocal)
__device__ inline void sqrt_inv_pos(double x, double &r, double &ri)
{
    const double y = __builtin_amdgcn_rsq(x);
    double rr = x * y, h = 0.5 * y;
    const double e = fma(-h, rr, 0.5);
    rr = fma(rr, e, rr);
    h = fma(h, e, h);
    double d = fma(-rr, rr, x);
    rr = fma(d, h, rr);
    d = fma(-rr, rr, x);
    rr = fma(d, h, rr);
    r = rr;
    ri = 2.0 * h;
}

// Potential<double> (MD.h:895-930) at a distance dr already taken, both branches evaluated
// and one selected
__device__ inline double pair_u_r(double dr, const double *C)
{
#pragma clang fp contract(off)
    const double a = C[0] - dr;
    const double inner = C[1] * a * a + C[2];
    const double b = C[3] - dr;
    const double outer = b * b * (C[4] - b * C[5]);
    return dr <= C[0] ? inner : outer;
}

// Potential<double> (MD.h:895-930) with both branches evaluated and one selected, so a
// wave does not run the two paths one after the other; each branch keeps the reference's
// operation order.
__device__ inline double pair_u_any(double r2, const double *C)
{
#pragma clang fp contract(off)
    const double dr = sqrt_pos(r2);
    const double a = C[0] - dr;
    const double inner = C[1] * a * a + C[2];
    const double b = C[3] - dr;
    const double outer = b * b * (C[4] - b * C[5]);
    return dr <= C[0] ? inner : outer;
}

// Forward half shell of the reference (cellOpt.h:715) over stencil index
// (dx+1) + 3(dy+1) + 9(dz+1).
__device__ constexpr uint32_t forward_mask()
{
    const int nb[13][3] = {{1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1},
                           {-1, 0, 1}, {0, -1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}, {1, -1, 0}};
    uint32_t m = 0;
    for (int j = 0; j < 13; j++) m |= 1u << ((nb[j][0] + 1) + 3 * (nb[j][1] + 1) + 9 * (nb[j][2] + 1));
    return m;
}

// A staged candidate read as two 16-byte LDS loads, so its type arrives with its position.
__device__ inline Cand load_cand(const Cand *p)
{
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const uint4 b = reinterpret_cast<const uint4 *>(p)[1];
    Cand c;
    c.x = a.x;
    c.y = a.y;
    c.z = __hiloint2double((int)b.y, (int)b.x);
    c.t = (int)b.z;
    c.c = (int)b.w;
    return c;
}

// One owned pair's U(old) - U(scaled) from its imaged FP64 difference (dx, dy, dz) and its
// potential row C.  Under the trial scale every coordinate difference scales by s_k, so the
// scaled r^2 is r^2 + sum (s_k^2 - 1) d_k^2 from the differences at hand (s2 = s_k^2 - 1);
// it differs from differencing the scaled positions (cellOpt.h:1098-1105) by a rounding,
// where the potential is continuous.  dseries: every |s_k^2 - 1| <= 1e-2, so r' - r comes
// from the binomial series of sqrt(1 + delta).
__device__ inline double du_pair(double dx, double dy, double dz, double r2, const double *C, double cut2,
                                 double s2x, double s2y, double s2z, bool dseries)
{
    const double dr2 = fma(s2x, dx * dx, fma(s2y, dy * dy, s2z * (dz * dz)));
    const double r2n = r2 + dr2;
    const bool in0 = r2 < cut2, in1 = r2n < cut2;
    if (!(in0 || in1)) return 0.0;
    if (!dseries) return (in0 ? pair_u_any(r2, C) : 0.0) - (in1 ? pair_u_any(r2n, C) : 0.0);
    // r' - r = r (sqrt(1 + delta) - 1) by the binomial series through delta^5: with
    // |delta| <= 1e-2 the rest is below 5e-12 of dr
    double r, ri;
    sqrt_inv_pos(r2, r, ri);
    const double del = dr2 * (ri * ri);
    const double dr = (r * del) * fma(del, fma(del, fma(del, fma(del, 0.02734375, -0.0390625), 0.0625), -0.125), 0.5);
    const double rn = r + dr;
    const bool innerO = r <= C[0], innerN = rn <= C[0];
    if (in0 && in1 && innerO == innerN) {
        // both distances on one polynomial branch: U(r) - U(r + dr) in closed form, with no
        // cancellation between the two potentials
        const double a = C[0] - r;
        const double uin = C[1] * dr * fma(2.0, a, -dr);
        const double b = C[3] - r;
        const double bn = b - dr;
        const double uout = dr * (C[4] * fma(2.0, b, -dr) - C[5] * fma(3.0 * b, bn, dr * dr));
        return innerO ? uin : uout;
    }
    // a branch or the cutoff lies between the two distances
    return (in0 ? pair_u_r(r, C) : 0.0) - (in1 ? pair_u_r(rn, C) : 0.0);
}

// One candidate of the force kernels: the exact FP64 test r^2 < rc^2 (MD.h:798-810), then
// the force law (MD.h:811-848) in FP32, its magnitude times the FP64 difference added to
// FP64 sums.  With r0 = r2f * rsq(r2f) (exact as an FP64
// product), r = r0 + corr is one Newton step against the FP64 residual r^2 - r0^2, and
// m = C - r is formed as fma(-r2f, rsq, C) - corr, so it carries no FP32 rounding of r^2
// or r0 (near r_m the force is C1 m with C1 ~ 400).  1/r is rsq(r2f) refined by one Newton
// step against r.  row is the pair's row of the constant table padded to 8 floats,
// {r_m, C1, C2, r_c, C4, C5, 0, 0}, read as two float4 in one LDS trip; the branch
// r >= r_m selects (r_c, C4, C5) over (r_m, C1, C2), which is the reference's row offset
// int(r/r_m)*3 clamped (MD.h:819).
__device__ inline void force_law(double dxd, double dyd, double dzd, double r2, const float4 *row, double &fx,
                                 double &fy, double &fz)
{
    const float4 A = row[0], B = row[1];
    const float r2f = (float)r2;
    const float ri0 = __builtin_amdgcn_rsqf(r2f);
    const double r0d = (double)r2f * (double)ri0;
    const float corr = (float)fma(-r0d, r0d, r2) * (0.5f * ri0);
    const float r = (float)r0d + corr;
    const float ri = ri0 * (2.0f - r * ri0);
    const bool outer = r >= A.x;
    const float c0 = outer ? A.w : A.x, c1 = outer ? B.x : A.y, c2 = outer ? B.y : A.z;
    const float m = fmaf(-r2f, ri0, c0) - corr;
    const float mag = ((c1 - c2 * m) * m) * ri;
    // the pair's force, mag * (FP64 difference), accumulated in FP64: a bead sums ~60
    // terms of up to a few hundred, so FP32 partial sums would round at 2^-24 of that
    // each time, more than the terms' own FP32 error
    const double magd = (double)mag;
    fx = fma(dxd, magd, fx);
    fy = fma(dyd, magd, fy);
    fz = fma(dzd, magd, fz);
}

// force_term over a two-body table whose (a, b) and (b, a) rows agree: C8 holds the home
// type's rows, one per candidate type.
__device__ inline void force_term(const Bead &hb, const Cand &q, bool other, double cut2, const float4 *C8,
                                  double &fx, double &fy, double &fz)
{
    const double dxd = hb.x - q.x;
    const double dyd = hb.y - q.y;
    const double dzd = hb.z - q.z;
    const double r2 = sq3(dxd, dyd, dzd);
    // the type bound keeps the type's LDS read ahead of the branch, with the position
    if (r2 < cut2 && other && (unsigned)q.t < (unsigned)kMaxTypes)
        force_law(dxd, dyd, dzd, r2, C8 + 2 * q.t, fx, fy, fz);
}

// Which side of a pair the reference's half-shell traversal makes its first particle p1,
// whose type is the row of the two-body table (MD.h:811-819): across cells the particle
// of the cell whose forward half shell holds the other's cell (cellOpt.h:715,800-854);
// within a cell the lower particle index, the cell's list running in descending index
// (cellOpt.h:572-585,780-795).  c is the candidate's stencil cell seen from the home
// particle, me and other the two original indices.
__device__ inline bool home_first(int c, uint32_t me, uint32_t other)
{
    return c == 13 ? me < other : ((forward_mask() >> c) & 1u) != 0;
}

// force_term for two-body tables whose (a, b) and (b, a) rows differ: the candidate's
// type carries its stencil cell above bit 8, its slot gives its original index through
// sidx, and the row is the first particle's.  The same law on either side of a pair takes
// the same row, so the two sides' forces are equal and opposite as in the reference.
__device__ inline void force_term_asym(const Bead &hb, const Cand &q, bool other, double cut2, const float4 *C8all,
                                       int nT, uint32_t me, const uint32_t *__restrict__ sidx, double &fx,
                                       double &fy, double &fz)
{
    const double dxd = hb.x - q.x;
    const double dyd = hb.y - q.y;
    const double dzd = hb.z - q.z;
    const double r2 = sq3(dxd, dyd, dzd);
    if (r2 < cut2 && other) {
        const int tq = q.t & 0xff, c = q.t >> 8;
        const int row = home_first(c, me, c == 13 ? sidx[q.c] : 0u) ? hb.type * nT + tq : tq * nT + hb.type;
        force_law(dxd, dyd, dzd, r2, C8all + 2 * row, fx, fy, fz);
    }
}

// The force kernel's candidate as a count (M_FCOUNT): the same FP64 test as force_term,
// then the pair counted when its force term is not identically zero -- r^2 < r_m^2, or a row
// with a nonzero outer coefficient (Umin != 0, laradjiRevalee.h:97-110).  Cd holds the home
// type's FP64 force rows {r_m, C1, C2, r_c, C4, C5}.
template <bool ASYM>
__device__ inline void count_term(const Bead &hb, const Cand &q, bool other, double cut2, const double *Cd,
                                  const double *Call, int nT, uint32_t me, const uint32_t *__restrict__ sidx, int &cnt)
{
    const double r2 = sq3(hb.x - q.x, hb.y - q.y, hb.z - q.z);
    if (r2 < cut2 && other && (ASYM || (unsigned)q.t < (unsigned)kMaxTypes)) {
        const double *C = Cd + 6 * q.t;
        if (ASYM) {
            const int tq = q.t & 0xff, c = q.t >> 8;
            C = Call + 6 * (home_first(c, me, c == 13 ? sidx[q.c] : 0u) ? hb.type * nT + tq : tq * nT + hb.type);
        }
        if (C[4] != 0.0 || C[5] != 0.0 || r2 < C[0] * C[0]) cnt++;
    }
}

struct Out {
    float4 *force;            // M_FORCE, original order
    double *energy;           // M_U, M_DU: one FP64 value per sorted slot
    int *counts;              // M_COUNT, original order
    int *pairs;               // M_LIST
    unsigned long long *npairs;
    long long cap;
};

// ASYM: the two-body table's (a, b) and (b, a) rows may differ, and each pair takes the row
// of the particle the reference's traversal makes first (home_first).  ONE: a
// single-replica context, whose box and grid are read from reps[0] once rather than per
// cell after the home bead's replica index.
template <int MODE, bool ASYM, bool ONE = false>
__global__ __launch_bounds__(kWaves * kWave) __attribute__((amdgpu_waves_per_eu(4))) void k_cells(
    int n, int nT, int chunk, int nblocks, double cut2, float cut2pre, const Bead *__restrict__ sb,
    const uint32_t *__restrict__ skey, const uint32_t *__restrict__ sidx, const int *__restrict__ cstart,
    const int *__restrict__ cend, const DevRep *__restrict__ reps, const float *__restrict__ tab,
    const double *__restrict__ uconst, Out out)
{
    // the one-phase force of a box-move step with the FP64 U(old) - U(scaled) of each pair
    // the home particle owns, in the same traversal (the box move's dU at the positions the
    // force sees)
    constexpr bool kFDU = MODE == M_FDU;
    constexpr bool kF = MODE == M_FORCE || MODE == M_FORCE1 || kFDU;   // force (two-phase / one-phase)
    constexpr bool k2P = MODE == M_FORCE;
    constexpr bool kF1 = MODE == M_FORCE1 || kFDU;
    constexpr bool kE = MODE == M_U || MODE == M_DU;   // FP64 energy gathers
    // the one-phase force traversal counting its accepted pairs with a nonzero term
    // instead of evaluating them: same staging, culling, lists and FP64 test
    constexpr bool kFC = MODE == M_FCOUNT;
    constexpr bool kCull = kF1 || kE || kFC;           // candidates culled per quarter of the home cell
    // staged candidates per pass: the force kernel measured fastest at 192; the two-phase
    // force and the FP64 energy gathers at 128 (their larger constant tables and registers
    // leave less LDS per resident wave)
    constexpr int kMaxC = (k2P || kE || kFDU) ? 128 : 192;
    constexpr int kStageIters = kMaxC / kWave;
    static_assert(kMaxC % kWave == 0, "staging rounds cover whole waves");
    __shared__ float4 sC8[kF ? 2 * kMaxTypes * kMaxTypes : 1];   // rows padded to 8 floats
    __shared__ double sCd[(kF && !kFDU) ? 1 : 6 * kMaxTypes * kMaxTypes];
    __shared__ Cand sCand[kWaves][kMaxC];   // staged candidates of the current pass
    __shared__ float4 sRel[k2P ? kWaves : 1][k2P ? kMaxC : 1];
    __shared__ uint8_t sList[k2P ? kWaves : 1][k2P ? kList : 1][kWave];   // per-lane hit lists, lane-minor
    // per stencil cell: image shift, and slot in the sorted order minus slot in the pass
    // layout
    __shared__ Shift sShift[kWaves][27];
    __shared__ int sDelta[kWaves][27];
    // per pass: the stencil cell at each cell's first staged slot (0 elsewhere), then, in
    // the same memory once the slots are staged, the candidate lists of the four quarters
    // of the home cell packed one after another (list q at q * kMaxC) and padded so the
    // force loop's reads one trip ahead (at most 3 x 64 entries past the end of a list)
    // stay inside the array
    constexpr int kListBytes = kCull ? 4 * kMaxC + 3 * kWave : kMaxC;
    static_assert(kListBytes % 4 == 0, "the map is cleared in 32-bit words");
    __shared__ uint32_t sLists[kWaves][kListBytes / 4];

    // constant rows and culling reach from the pair table that the host builds: the FP32
    // force rows padded to 8 floats, then the reach r^2 per candidate type for the force
    // and for the FP64 energy gathers.  The reach of a candidate type is r_c where some
    // pair with it has an outer branch (Umin != 0 in laradjiRevalee.h:97-110), else r_m:
    // beyond r_m such pairs have zero force and potential, so the candidates culled by it
    // contribute exactly zero.
    __shared__ float sReach2[kMaxTypes];
    const int lane = threadIdx.x & (kWave - 1);
    // wave index, wave-uniform (the compiler then keeps the chunk's cell setup scalar)
    const int wave = kWaves == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
    const long long a0 = ((long long)xcd_block<kE ? 64 : 0>(blockIdx.x, nblocks) * kWaves + wave) * chunk;
    const int a = (int)min(a0, (long long)n);
    const int aend = min(a + chunk, n);
    // the first window's keys (this slot's and the one before), loaded before the constant
    // tables are staged so that the two memory trips overlap
    uint32_t kc = 0u, kp = 0u;
    if (a + lane < aend) {
        kc = skey[a + lane];
        if (a + lane > 0) kp = skey[a + lane - 1];
    }
    // the constant tables: every load first (indices clamped in bounds, so no branch holds
    // one back), then the LDS stores, so the staging costs one memory trip
    constexpr int kThreads = kWaves * kWave;
    constexpr int kC8Per = (2 * kMaxTypes * kMaxTypes + kThreads - 1) / kThreads;
    constexpr int kCdPer = (6 * kMaxTypes * kMaxTypes + kThreads - 1) / kThreads;
    float4 c8[kF ? kC8Per : 1];
    double cd[(!kF || kFDU) ? kCdPer : 1];
    float rf = 0.f, re = 0.f;
    if (kF) {
        const float4 *t8 = reinterpret_cast<const float4 *>(tab);
#pragma unroll
        for (int k = 0; k < kC8Per; k++) {
            const int t = threadIdx.x + k * kThreads;
            c8[k] = t8[min(t, 2 * nT * nT - 1)];
        }
    }
    if (!kF || kFDU) {
#pragma unroll
        for (int k = 0; k < kCdPer; k++) {
            const int t = threadIdx.x + k * kThreads;
            cd[k] = uconst[min(t, 6 * nT * nT - 1)];
        }
    }
    const int treach = min((int)threadIdx.x, nT - 1);   // threads past nT repeat the last type
    if (kCull) {
        rf = tab[8 * nT * nT + treach];
        re = tab[8 * nT * nT + kMaxTypes + treach];
    }
    // stored whole (entries past the table hold copies of its last one and are never read),
    // so that no branch sinks a load down to its store
    static_assert(kC8Per * kThreads <= 2 * kMaxTypes * kMaxTypes && kCdPer * kThreads <= 6 * kMaxTypes * kMaxTypes,
                  "the tables are stored whole");
    if (kF) {
#pragma unroll
        for (int k = 0; k < kC8Per; k++) sC8[threadIdx.x + k * kThreads] = c8[k];
    }
    if (!kF || kFDU) {
#pragma unroll
        for (int k = 0; k < kCdPer; k++) sCd[threadIdx.x + k * kThreads] = cd[k];
    }
    // every thread stores (those past nT the last type's value again, the same bits)
    if (kCull) sReach2[treach] = kFDU ? fmaxf(rf, re) : (kE ? re : rf);
    __syncthreads();

    const unsigned long long lanes_lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    constexpr uint32_t fwd = forward_mask();
    Cand *Cw = sCand[wave];
    float4 *Rw = sRel[k2P ? wave : 0];

    for (int base = a; base < aend; base += kWave) {
        const int s = base + lane;
        if (base != a && s < aend) {
            kc = skey[s];
            kp = skey[s - 1];
        }
        const bool st = s < aend && (s == 0 || kp != kc);
        unsigned long long starts = __ballot(st);
        while (starts) {
            const int s0 = base + __ffsll((long long)starts) - 1;
            starts &= starts - 1;
            // the home cell's key from the lane that loaded it
            const uint32_t kg = (uint32_t)__builtin_amdgcn_readlane((int)kc, s0 - base);
            // the home cell's replica: box, grid and key offset (wave-uniform)
            const DevRep &R = reps[ONE ? 0 : sb[s0].rep];
            const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
            const double Lx = R.L[0], Ly = R.L[1], Lz = R.L[2];
            const double csx = R.cs[0], csy = R.cs[1], csz = R.cs[2];
            const float hcx = (float)(0.5 * csx), hcy = (float)(0.5 * csy), hcz = (float)(0.5 * csz);
            const double ax = R.scale[0], ay = R.scale[1], az = R.scale[2];
            // dU: s^2 - 1 per axis as (s - 1)(s + 1), exact in its leading bits since s - 1 is
            // exact for a trial scale within a factor 2 of 1
            const double s2x = (ax - 1.0) * (ax + 1.0), s2y = (ay - 1.0) * (ay + 1.0), s2z = (az - 1.0) * (az + 1.0);
            // the scaled distance by a series in delta = (r'^2 - r^2) / r^2, a weighted mean of
            // the s_k^2 - 1 (wave-uniform; larger trial scales take two square roots)
            const bool dseries = fmax(fabs(s2x), fmax(fabs(s2y), fabs(s2z))) <= 1e-2;
            // energy gathers: a candidate farther than (rc + margin) / min(1, scale) from the
            // home quarter-cell is beyond rc both before and after the trial scaling
            const double smin = fmin(1.0, fmin(ax, fmin(ay, az)));
            const float reachS = (MODE == M_DU || kFDU) ? (float)(1.0 / (smin * smin)) : 1.f;   // dU: widened by the trial scale
            const uint32_t kbase = (uint32_t)R.cellbase;
            const uint32_t k = kg - kbase;
            const int kx = (int)(k % (uint32_t)nx);
            const int ky = (int)((k / (uint32_t)nx) % (uint32_t)ny);
            const int kz = (int)(k / (uint32_t)(nx * ny));
            // home cell centre: FP32 positions relative to it stay within 1.5 cells
            const double ox0 = (kx + 0.5) * csx, oy0 = (ky + 0.5) * csy, oz0 = (kz + 0.5) * csz;

            // 1. stencil lookups, one cell per lane; the energy gathers take each pair from
            // the side that owns it, so they stage only the home cell and its forward half
            // shell
            constexpr bool kOwnedOnly = MODE == M_U || MODE == M_DU;
            int nst = 0, ncnt = 0;
            Shift shl;
            shl.x = shl.y = shl.z = shl.pad = 0.0;
            if (lane < 27) {
                const int X0 = kx + lane % 3 - 1, Y0 = ky + (lane / 3) % 3 - 1, Z0 = kz + lane / 9 - 1;
                shl.x = X0 < 0 ? -Lx : (X0 >= nx ? Lx : 0.0);
                shl.y = Y0 < 0 ? -Ly : (Y0 >= ny ? Ly : 0.0);
                shl.z = Z0 < 0 ? -Lz : (Z0 >= nz ? Lz : 0.0);
            }
            if (lane < 27 && (!kOwnedOnly || lane == 13 || ((fwd >> lane) & 1u))) {
                int X0 = kx + lane % 3 - 1, Y0 = ky + (lane / 3) % 3 - 1, Z0 = kz + lane / 9 - 1;
                X0 = X0 < 0 ? X0 + nx : (X0 >= nx ? X0 - nx : X0);
                Y0 = Y0 < 0 ? Y0 + ny : (Y0 >= ny ? Y0 - ny : Y0);
                Z0 = Z0 < 0 ? Z0 + nz : (Z0 >= nz ? Z0 - nz : Z0);
                const uint32_t kk = kbase + (uint32_t)(X0 + Y0 * nx + Z0 * nx * ny);
                // both ends in one trip, then the run's key (index clamped in bounds): no
                // branch for the compiler to sink a load into
                const int c0 = cstart[kk];
                const int c1 = cend[kk];
                const bool ok = (unsigned)c0 < (unsigned)n && skey[min(max(c0, 0), n - 1)] == kk;
                nst = ok ? c0 : 0;
                ncnt = ok ? c1 - c0 : 0;
            }
            // 2. layout of the candidates
            const int incl = wave_incl_scan(ncnt, lane);
            const int excl = incl - ncnt;
            const int total = __builtin_amdgcn_readlane(incl, 26);
            const int selfOff = __builtin_amdgcn_readlane(excl, 13);
            // home particles: the home cell's own lookup (stencil cell 13 holds it, and the
            // cell table entry of the cell being sorted into is current)
            const int nh = __builtin_amdgcn_readlane(ncnt, 13);
            Shift *ShW = sShift[wave];
            if (lane < 27) {
                ShW[lane] = shl;
                sDelta[wave][lane] = nst - excl;
            }
            uint8_t *mapW = reinterpret_cast<uint8_t *>(sLists[wave]);   // map, then lists

            for (int h0 = 0; h0 < nh; h0 += kWave) {
                const int H = min(kWave, nh - h0);
                const int tpp = kWave / H;
                const int hi = lane / tpp;
                const int sub = lane - hi * tpp;
                const bool act = hi < H;
                // the home particle and its index: loaded for every lane (an idle lane takes the
                // cell's first) and first used after the pass below has issued its loads, so
                // the trips overlap; every use is under `act`
                const int self = selfOff + h0 + hi;
                const int hslot = s0 + h0 + hi;
                const Bead hb0 = sb[act ? hslot : s0];
                const uint32_t me = sidx[act ? hslot : s0];
                double fx = 0.0, fy = 0.0, fz = 0.0;
                double accu = 0.0;
                int cnt = 0;
                for (int cb = 0; cb < total; cb += kMaxC) {
                    const int ce = min(total, cb + kMaxC);
                    // staging rounds this pass needs (wave-uniform): the unrolled rounds
                    // past it are skipped, so a short last pass costs what it holds
                    const int iters = (ce - cb + kWave - 1) / kWave;
                    // stage [cb, ce): addresses first, then all loads, then the LDS stores
                    // map of the pass: every stencil cell with slots in [cb, ce) marks its
                    // first one with its index (the cell holding slot cb marks position 0),
                    // and a max-scan over the positions gives each slot its cell
                    if (lane < kMaxC / 4) sLists[wave][lane] = 0u;
                    if (lane < 27 && ncnt > 0 && incl > cb && excl < ce) mapW[max(excl, cb) - cb] = (uint8_t)lane;
                    wave_lds_sync();
                    int src[kStageIters], cel[kStageIters];
                    int carry = 0;
#pragma unroll
                    for (int it = 0; it < kStageIters; it++) {
                        src[it] = -1;
                        cel[it] = 0;
                        if (it >= iters) continue;
                        const int t = cb + lane + it * kWave;
                        const int c = max(wave_max_scan(t < ce ? (int)mapW[t - cb] : 0), carry);
                        carry = __builtin_amdgcn_readlane(c, kWave - 1);
                        cel[it] = c;
                        src[it] = (t < ce) ? t + sDelta[wave][c] : -1;
                    }
                    Bead qb[kStageIters];
#pragma unroll
                    for (int it = 0; it < kStageIters; it++)
                        if (it < iters && src[it] >= 0) qb[it] = sb[src[it]];
                    int reach[kStageIters];
#pragma unroll
                    for (int it = 0; it < kStageIters; it++) {
                        reach[it] = 0;
                        if (it >= iters || src[it] < 0) continue;
                        const int c = cel[it];
                        Cand cd;
                        const Shift sh = ShW[c];
                        const double ix = qb[it].x + sh.x;
                        const double iy = qb[it].y + sh.y;
                        const double iz = qb[it].z + sh.z;
                        if (k2P)
                            Rw[lane + it * kWave] = make_float4((float)(ix - ox0), (float)(iy - oy0), (float)(iz - oz0), 0.f);
                        if (kCull) {
                            // which quarters of the home cell (x below / above the centre,
                            // times y below / above) the candidate can reach: box distance
                            // below the reach radius; bit qx + 2 qy
                            const float rx = (float)(ix - ox0), ry = (float)(iy - oy0), rz = (float)(iz - oz0);
                            const float gz = fmaxf(fabsf(rz) - hcz, 0.f);
                            const float gx0 = fmaxf(fmaxf(rx, -hcx - rx), 0.f), gx1 = fmaxf(fmaxf(-rx, rx - hcx), 0.f);
                            const float gy0 = fmaxf(fmaxf(ry, -hcy - ry), 0.f), gy1 = fmaxf(fmaxf(-ry, ry - hcy), 0.f);
                            const float z2 = gz * gz;
                            const float a0 = fmaf(gx0, gx0, z2), a1 = fmaf(gx1, gx1, z2);
                            const float b0 = gy0 * gy0, b1 = gy1 * gy1;
                            const float r2max = sReach2[qb[it].type] * reachS;
                            reach[it] = (a0 + b0 < r2max ? 1 : 0) | (a1 + b0 < r2max ? 2 : 0) |
                                        (a0 + b1 < r2max ? 4 : 0) | (a1 + b1 < r2max ? 8 : 0);
                        }
                        // forces and energy gathers stage imaged positions (q + shift, the
                        // one rounding the reference's minImg makes of the same operands);
                        // the count and list gathers the stored ones, since they take each
                        // pair's difference from either side
                        cd.x = (kF || kE || kFC) ? ix : qb[it].x;
                        cd.y = (kF || kE || kFC) ? iy : qb[it].y;
                        cd.z = (kF || kE || kFC) ? iz : qb[it].z;
                        // force: the candidate's sorted slot (the self test reads it with the
                        // type, in the same 16-byte load); energy gathers: its stencil cell;
                        // count and list gathers, and the force over an asymmetric table: the
                        // slot, with the stencil cell above the type's bits (types are below
                        // kMaxTypes <= 256)
                        constexpr bool kSlotCell = MODE == M_COUNT || MODE == M_LIST || kFDU || (ASYM && (kF || kFC));
                        cd.t = kSlotCell ? (qb[it].type | (c << 8)) : qb[it].type;
                        cd.c = (kF || kFC || kSlotCell) ? src[it] : c;
                        Cw[lane + it * kWave] = cd;
                    }
                    // per quarter of the home cell, the staged candidates it can reach
                    int nq[4] = {0, 0, 0, 0};
                    if (kCull) {
#pragma unroll
                        for (int it = 0; it < kStageIters; it++) {
                            if (it >= iters) break;
#pragma unroll
                            for (int q = 0; q < 4; q++) {
                                const bool r = (reach[it] >> q) & 1;
                                const unsigned long long b = __ballot(r);
                                if (r) mapW[q * kMaxC + nq[q] + __popcll(b & lanes_lt)] = (uint8_t)(lane + it * kWave);
                                nq[q] += __popcll(b);
                            }
                        }
                    }
                    wave_lds_sync();
                    // the home particle for this pass: the empty asm keeps every use of it
                    // (its quarter, its table rows) here, after the pass's loads, where the
                    // compiler would otherwise hoist them ahead of the pass loop and wait there
                    Bead hb = hb0;
                    asm volatile("" : "+v"(hb.x), "+v"(hb.y), "+v"(hb.z), "+v"(hb.type));
                    // home position relative to the cell centre (its quarter; the two-phase screen)
                    const float hfx = (float)(hb.x - ox0), hfy = (float)(hb.y - oy0), hfz = (float)(hb.z - oz0);
                    if (kFDU && act) {
                        // the force of every candidate and, for each pair this side owns
                        // (forward half shell; the higher slot within the home cell), the
                        // FP64 dU term with the row of the particle the reference makes first
                        const int quad = (hfx >= 0.f ? 1 : 0) + (hfy >= 0.f ? 2 : 0);
                        const uint8_t *lst = mapW + quad * kMaxC;
                        const int nl = quad == 0 ? nq[0] : (quad == 1 ? nq[1] : (quad == 2 ? nq[2] : nq[3]));
                        // the next entry is read one iteration ahead (entries past nl fall in
                        // the next list or the padding and are never used)
                        int inext = lst[sub];
                        for (int jj = sub; jj < nl; jj += tpp) {
                            const int i0 = inext;
                            inext = lst[jj + tpp];
                            const Cand q = load_cand(Cw + i0);
                            const int tq = q.t & 0xff, c = q.t >> 8;
                            const double dxd = hb.x - q.x, dyd = hb.y - q.y, dzd = hb.z - q.z;
                            const double r2 = sq3(dxd, dyd, dzd);
                            // the row of the pair's first particle (home_first; symmetric
                            // tables take the home's row either way)
                            bool hf = true;
                            if (ASYM) hf = c == 13 ? me < sidx[q.c] : ((fwd >> c) & 1u) != 0;
                            const int row = hf ? hb.type * nT + tq : tq * nT + hb.type;
                            if (r2 < cut2 && q.c != hslot) force_law(dxd, dyd, dzd, r2, sC8 + 2 * row, fx, fy, fz);
                            // owned: a forward half-shell cell, or the higher slot in the home cell
                            const bool owned = c == 13 ? q.c > hslot : ((fwd >> c) & 1u) != 0;
                            if (owned) accu += du_pair(dxd, dyd, dzd, r2, sCd + 6 * row, cut2, s2x, s2y, s2z, dseries);
                        }
                    } else if ((MODE == M_FORCE1 || kFC) && act) {
                        // one phase over the home's quarter list: FP64 test, force law under
                        // the branch; two candidates per trip so both LDS reads are in
                        // flight before either is used
                        const float4 *C8 = sC8 + 2 * hb.type * nT;
                        const double *Cd = sCd + 6 * hb.type * nT;
                        const int quad = (hfx >= 0.f ? 1 : 0) + (hfy >= 0.f ? 2 : 0);
                        const uint8_t *lst = mapW + quad * kMaxC;
                        const int nl = quad == 0 ? nq[0] : (quad == 1 ? nq[1] : (quad == 2 ? nq[2] : nq[3]));
                        // the next trip's list entries are read one trip ahead (entries past
                        // nl fall in the next list or the padding and are never used)
                        int j = sub;
                        const uint8_t *lp = lst + sub;
                        int i0 = lp[0], i1 = lp[tpp];
                        for (; j + tpp < nl; j += 2 * tpp) {
                            lp += 2 * tpp;
                            const int n0 = lp[0], n1 = lp[tpp];
                            const Cand q0 = load_cand(Cw + i0);
                            const Cand q1 = load_cand(Cw + i1);
                            if (kFC) {
                                count_term<ASYM>(hb, q0, q0.c != hslot, cut2, Cd, sCd, nT, me, sidx, cnt);
                                count_term<ASYM>(hb, q1, q1.c != hslot, cut2, Cd, sCd, nT, me, sidx, cnt);
                            } else if (ASYM) {
                                force_term_asym(hb, q0, q0.c != hslot, cut2, sC8, nT, me, sidx, fx, fy, fz);
                                force_term_asym(hb, q1, q1.c != hslot, cut2, sC8, nT, me, sidx, fx, fy, fz);
                            } else {
                                force_term(hb, q0, q0.c != hslot, cut2, C8, fx, fy, fz);
                                force_term(hb, q1, q1.c != hslot, cut2, C8, fx, fy, fz);
                            }
                            i0 = n0;
                            i1 = n1;
                        }
                        if (j < nl) {
                            const Cand q0 = load_cand(Cw + i0);
                            if (kFC)
                                count_term<ASYM>(hb, q0, q0.c != hslot, cut2, Cd, sCd, nT, me, sidx, cnt);
                            else if (ASYM)
                                force_term_asym(hb, q0, q0.c != hslot, cut2, sC8, nT, me, sidx, fx, fy, fz);
                            else
                                force_term(hb, q0, q0.c != hslot, cut2, C8, fx, fy, fz);
                        }
                    } else if (k2P) {
                        const float4 *C8 = sC8 + 2 * hb.type * nT;
                        uint8_t *lst = &sList[k2P ? wave : 0][0][lane];
                        for (int jb = cb + sub; __ballot(act && jb < ce); jb += tpp * kList) {
                            // phase 1: FP32 screen, hits kept in the lane's list
                            int nhit = 0;
                            if (act) {
#pragma unroll 2
                                for (int kk = 0; kk < kList; kk++) {
                                    const int j = jb + kk * tpp;
                                    if (j < ce) {
                                        const float4 q = Rw[j - cb];
                                        const float dx = hfx - q.x, dy = hfy - q.y, dz = hfz - q.z;
                                        const float r2 = dx * dx + dy * dy + dz * dz;
                                        if (r2 < cut2pre && j != self) lst[kWave * nhit++] = (uint8_t)(j - cb);
                                    }
                                }
                            }
                            int most = nhit;
#pragma unroll
                            for (int o = 1; o < kWave; o <<= 1) most = max(most, __shfl_xor(most, o));
                            // phase 2: exact FP64 test and the force law on the hits
                            for (int kk = 0; kk < most; kk++)
                                if (kk < nhit) {
                                    const Cand q = load_cand(Cw + lst[kWave * kk]);
                                    if (ASYM)
                                        force_term_asym(hb, q, true, cut2, sC8, nT, me, sidx, fx, fy, fz);
                                    else
                                        force_term(hb, q, true, cut2, C8, fx, fy, fz);
                                }
                        }
                    } else if (kE && act) {
                        // each pair once, from the side that owns it: only the home cell and
                        // its forward half shell are staged, and within the home cell the
                        // lower slot owns the pair
                        const int quad = (hfx >= 0.f ? 1 : 0) + (hfy >= 0.f ? 2 : 0);
                        const uint8_t *lst = mapW + quad * kMaxC;
                        const int nl = quad == 0 ? nq[0] : (quad == 1 ? nq[1] : (quad == 2 ? nq[2] : nq[3]));
                        const double *Ch = sCd + 6 * hb.type * nT;
                        // one owned pair's share: U, or U(old) - U(scaled) (du_pair)
                        auto term = [&](const Cand &q, int j) -> double {
                            if (j == self || (q.c == 13 && j < self)) return 0.0;
                            // the staged forward half shell's pairs have the home particle
                            // first; a home-cell pair the lower original index
                            const double *C = Ch + 6 * q.t;
                            if constexpr (ASYM) {
                                if (q.c == 13 && sidx[s0 + (j - selfOff)] < me) C = sCd + 6 * (q.t * nT + hb.type);
                            }
                            const double dx = hb.x - q.x, dy = hb.y - q.y, dz = hb.z - q.z;   // staged imaged
                            const double r2 = sq3(dx, dy, dz);
                            if (MODE == M_U) return r2 < cut2 ? pair_u_any(r2, C) : 0.0;
                            return du_pair(dx, dy, dz, r2, C, cut2, s2x, s2y, s2z, dseries);
                        };
                        int inext = lst[sub];   // read one iteration ahead, as above
                        for (int jj = sub; jj < nl; jj += tpp) {
                            const int i0 = inext;
                            inext = lst[jj + tpp];
                            accu += term(Cw[i0], cb + i0);
                        }
                    } else if (!kF && act) {
                        // neighbour counts and pair lists see both sides of every pair
                        for (int j = cb + sub; j < ce; j += tpp) {
                            if (j == self) continue;
                            const Cand q = Cw[j - cb];
                            const int c = q.t >> 8;
                            const Shift sh = ShW[c];
                            const double sx = sh.x, sy = sh.y, sz = sh.z;
                            const bool own = (c == 13) || ((fwd >> c) & 1u);
                            double ox, oy, oz, px, py, pz;
                            if (own) {
                                ox = hb.x; oy = hb.y; oz = hb.z;
                                px = q.x + sx; py = q.y + sy; pz = q.z + sz;
                            } else {
                                ox = q.x; oy = q.y; oz = q.z;
                                px = hb.x - sx; py = hb.y - sy; pz = hb.z - sz;
                            }
                            const double r2 = sq3(ox - px, oy - py, oz - pz);
                            if (r2 < cut2) {
                                cnt++;
                                if (MODE == M_LIST) {
                                    const uint32_t other = sidx[q.c];
                                    if (me < other) {
                                        const unsigned long long slot = atomicAdd(out.npairs, 1ull);
                                        if ((long long)slot < out.cap) {
                                            out.pairs[2 * slot] = (int)me;
                                            out.pairs[2 * slot + 1] = (int)other;
                                        }
                                    }
                                }
                            }
                        }
                    }
                    wave_lds_sync();
                }
                if (kF) {
                    // a home's tpp lanes combined by a fixed-shape tree: at step s, lane sub
                    // adds lane sub + s when sub is a multiple of 2s
                    for (int st = 1; st < tpp; st <<= 1) {
                        const double ox = __shfl(fx, lane + st), oy = __shfl(fy, lane + st), oz = __shfl(fz, lane + st);
                        if ((sub & (2 * st - 1)) == 0 && sub + st < tpp) {
                            fx += ox;
                            fy += oy;
                            fz += oz;
                        }
                    }
                    if (act && sub == 0) out.force[me] = make_float4((float)fx, (float)fy, (float)fz, 0.f);
                    if (kFDU) {
                        for (int st = 1; st < tpp; st <<= 1) {
                            const double o = __shfl(accu, lane + st);
                            if ((sub & (2 * st - 1)) == 0 && sub + st < tpp) accu += o;
                        }
                        if (act && sub == 0) out.energy[s0 + h0 + hi] = accu;
                    }
                } else if (MODE == M_COUNT || kFC) {
                    for (int st = 1; st < tpp; st <<= 1) {
                        const int o = __shfl(cnt, lane + st);
                        if ((sub & (2 * st - 1)) == 0 && sub + st < tpp) cnt += o;
                    }
                    if (act && sub == 0) out.counts[me] = cnt;
                } else if (MODE == M_U || MODE == M_DU) {
                    // the home particle's share, its lanes combined by the same fixed tree
                    for (int st = 1; st < tpp; st <<= 1) {
                        const double o = __shfl(accu, lane + st);
                        if ((sub & (2 * st - 1)) == 0 && sub + st < tpp) accu += o;
                    }
                    if (act && sub == 0) out.energy[s0 + h0 + hi] = accu;
                }
            }
        }
    }
}

template <int MODE, bool ASYM = false, bool ONE = false>
void launch_cells(hipStream_t st, const Params &P, int chunk, const Bead *sbeads, const uint32_t *skey,
                  const uint32_t *sidx, const int *cstart, const int *cend, const DevRep *reps,
                  const float *fconst, const double *uconst, const Out &out)
{
    const int nb = cells_blocks(P.n, chunk);
    const double cut2 = P.cutoff * P.cutoff;
    // FP32 screen: positions relative to the home cell centre are within 1.5 cells, so
    // each coordinate carries at most 2^-23 * 1.5 * csmax rounding; widen r^2 by far more.
    // With at least 3 cells per axis, a cell is shorter than 4/3 of the cutoff.
    const double csmax = P.cutoff * 4.0 / 3.0;
    const double slack = 64.0 * P.cutoff * 1.5 * csmax * (1.0 / 8388608.0) + 1e-6 * cut2;
    // the two-phase kernel's FP32 screening radius (the culled modes take theirs per
    // candidate type from the constant tables, in the kernel)
    const float cut2pre = (float)(cut2 + slack);
    hipLaunchKernelGGL((k_cells<MODE, ASYM, ONE>), dim3(nb), dim3(kWaves * kWave), 0, st, P.n, P.nT, chunk, nb, cut2, cut2pre,
                       sbeads, skey, sidx, cstart, cend, reps, fconst, uconst, out);
}

}  // namespace

int cells_blocks(int n, int chunk)
{
    const long long waves = ((long long)(n > 0 ? n : 1) + chunk - 1) / chunk;
    return (int)((waves + kWaves - 1) / kWaves);
}

void launch_pair_force(hipStream_t st, const Params &P, int chunk, int variant, const Bead *sbeads,
                       const uint32_t *skey, const uint32_t *sidx, const int *cstart, const int *cend,
                       const DevRep *reps, const float *fconst, float4 *fpair)
{
    Out o{fpair, nullptr, nullptr, nullptr, nullptr, 0};
    if (variant == 1 && P.asym)
        launch_cells<M_FORCE, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, nullptr, o);
    else if (variant == 1)
        launch_cells<M_FORCE>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, nullptr, o);
    else if (P.asym && P.nrep == 1)
        launch_cells<M_FORCE1, true, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, nullptr, o);
    else if (P.asym)
        launch_cells<M_FORCE1, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, nullptr, o);
    else if (P.nrep == 1)
        launch_cells<M_FORCE1, false, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, nullptr, o);
    else
        launch_cells<M_FORCE1>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, nullptr, o);
}

void launch_pair_force_du(hipStream_t st, const Params &P, int chunk, const Bead *sbeads, const uint32_t *skey,
                          const uint32_t *sidx, const int *cstart, const int *cend, const DevRep *reps,
                          const float *fconst, const double *uconst, float4 *fpair, double *energy)
{
    Out o{fpair, energy, nullptr, nullptr, nullptr, 0};
    if (P.asym && P.nrep == 1)
        launch_cells<M_FDU, true, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, uconst, o);
    else if (P.asym)
        launch_cells<M_FDU, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, uconst, o);
    else if (P.nrep == 1)
        launch_cells<M_FDU, false, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, uconst, o);
    else
        launch_cells<M_FDU>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, fconst, uconst, o);
}

void launch_pair_gather(hipStream_t st, const Params &P, int chunk, int mode, const Bead *sbeads,
                        const uint32_t *skey, const uint32_t *sidx, const int *cstart, const int *cend,
                        const DevRep *reps, const float *tab, const double *uconst, double *energy, int *counts,
                        int *pairs, unsigned long long *npairs, long long cap)
{
    Out o{nullptr, energy, counts, pairs, npairs, cap};
    switch (mode) {
    case GATHER_U:
        if (P.asym)
            launch_cells<M_U, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        else
            launch_cells<M_U>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        break;
    case GATHER_DU:
        if (P.asym)
            launch_cells<M_DU, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        else
            launch_cells<M_DU>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        break;
    case GATHER_COUNT:
        launch_cells<M_COUNT>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        break;
    case GATHER_FCOUNT:
        if (P.asym)
            launch_cells<M_FCOUNT, true>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        else
            launch_cells<M_FCOUNT>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o);
        break;
    default: launch_cells<M_LIST>(st, P, chunk, sbeads, skey, sidx, cstart, cend, reps, tab, uconst, o); break;
    }
}

}  // namespace smdev
