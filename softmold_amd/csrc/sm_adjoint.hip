// Discrete adjoint of the MD step with respect to the interaction constants
// theta = (Umin, Umax of one type pair; k_bond, k_bend of the CHAIN sets), FP64 throughout.
//
// The forward step (sm_kernels.hip) is, per bead of type != 0 (c = dt/2):
//   v_h = v + c a;  x' = x + dt v_h (wrapped);  a' = F(x'; theta) - gamma v_h + sigma xi;
//   v' = v_h + c a'
// (type 0 beads keep x and get no kicks, verlet.h:296).  With the counter RNG the noise xi
// is a fixed function of (seed, step, particle), so the trajectory is a differentiable
// function of theta away from the Monte-Carlo accept/reject decisions, which the adjoint
// holds fixed (an accepted move is the linear map x -> S x).  The reverse step needs
// (dF/dx) w, the Hessian-vector product of the conservative forces: for pairs it is
// written out from f(r) and f'(r); for CHAIN bonds and bends it is the exact directional
// derivative of the force code evaluated on dual numbers.
//
// These kernels serve short fitting trajectories of small systems: stencil cells are read
// straight from global memory, no LDS staging.  The pair term gives each bead 32 lanes,
// one per stencil cell, and combines them by a fixed shuffle tree; with one thread per
// bead the 12k-bead bilayer filled 47 workgroups and the pair term took 465 us a call.
//
// Reference: force law Force<double> (MD.h:795-848), potential (MD.h:895-930), constant
// tables laradjiRevaleeFC/PC (potentials/laradjiRevalee.h:84-152), harmonicF/bendF
// (MD.h:384-403,683-723), harmonicP/bendP (MD.h:422-431,769-789), the CHAIN walk
// (system.h:1782-1866).
#include "sm_adjoint.hpp"

namespace smdev {

namespace {

constexpr int kBlock = 256;

inline int blocks(long long n) { return (int)((n + kBlock - 1) / kBlock); }

// ---- dual numbers: value and derivative along one direction ----------------------------
struct Dual {
    double v, d;
};
__device__ inline Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator-(Dual a) { return Dual{-a.v, -a.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return Dual{q, (a.d - q * b.d) / b.v};
}
__device__ inline Dual sqrt_t(Dual a)
{
    const double r = sqrt(a.v);
    return Dual{r, a.d / (2.0 * r)};
}
__device__ inline double sqrt_t(double a) { return sqrt(a); }
__device__ inline double val(double x) { return x; }
__device__ inline double val(Dual x) { return x.v; }
__device__ inline Dual cst(double x, Dual) { return Dual{x, 0.0}; }
__device__ inline double cst(double x, double) { return x; }

template <class T>
struct V3 {
    T x, y, z;
};

template <class T>
__device__ inline V3<T> sub_mi(const V3<T> &p, const V3<T> &q, const double *L)
{
    V3<T> d{p.x - q.x, p.y - q.y, p.z - q.z};
    const T z0 = cst(0.0, d.x);
    // the image shift depends on the value only, as in the force code
    if (val(d.x) > L[0] / 2.0) d.x = d.x - cst(L[0], z0);
    if (val(d.x) < -L[0] / 2.0) d.x = d.x + cst(L[0], z0);
    if (val(d.y) > L[1] / 2.0) d.y = d.y - cst(L[1], z0);
    if (val(d.y) < -L[1] / 2.0) d.y = d.y + cst(L[1], z0);
    if (val(d.z) > L[2] / 2.0) d.z = d.z - cst(L[2], z0);
    if (val(d.z) < -L[2] / 2.0) d.z = d.z + cst(L[2], z0);
    return d;
}

// harmonic bond force on the first particle (a1 += d * mag)
template <class T>
__device__ inline V3<T> bond_f(const V3<T> &d, double l0, double kb)
{
    const T dr = sqrt_t(d.x * d.x + d.y * d.y + d.z * d.z);
    const T mag = (cst(0.0, dr) - (dr - cst(l0, dr)) * cst(kb, dr)) / dr;
    return V3<T>{d.x * mag, d.y * mag, d.z * mag};
}

// force on the particle with role 0, 1 or 2 of the bend (p1, p2, p3); da = p1 - p2,
// db = p2 - p3
template <class T>
__device__ inline V3<T> bend_f(V3<T> da, V3<T> db, double c0, double kt, int role)
{
    const T one = cst(1.0, da.x);
    const T dra = sqrt_t(da.x * da.x + da.y * da.y + da.z * da.z);
    const T drb = sqrt_t(db.x * db.x + db.y * db.y + db.z * db.z);
    const T ia = one / dra, ib = one / drb;
    da = V3<T>{da.x * ia, da.y * ia, da.z * ia};
    db = V3<T>{db.x * ib, db.y * ib, db.z * ib};
    const T cost = da.x * db.x + da.y * db.y + da.z * db.z;
    const T mag = (cst(c0, one) - cost) * cst(kt, one);
    const T ta = mag * ia, tb = mag * ib;
    const V3<T> A{ta * (db.x - da.x * cost), ta * (db.y - da.y * cost), ta * (db.z - da.z * cost)};
    const V3<T> B{tb * (da.x - db.x * cost), tb * (da.y - db.y * cost), tb * (da.z - db.z * cost)};
    if (role == 0) return A;
    if (role == 1) return V3<T>{B.x - A.x, B.y - A.y, B.z - A.z};
    return V3<T>{-B.x, -B.y, -B.z};
}

// the CHAIN force on bead i from its bonds and bends (k_bonded_kick's walk)
template <class T>
__device__ inline V3<T> chain_force(const V3<T> *p, int j, int len, const double *L, double l0, double kb,
                                    double c0, double kt)
{
    // p[0..4] = beads i-2 .. i+2 (only the ones inside the chain are read)
    const T z = cst(0.0, p[2].x);
    V3<T> f{z, z, z};
    auto add = [&](const V3<T> &g, double s) {
        const T ss = cst(s, z);
        f.x = f.x + g.x * ss;
        f.y = f.y + g.y * ss;
        f.z = f.z + g.z * ss;
    };
    if (j >= 1) add(bond_f(sub_mi(p[1], p[2], L), l0, kb), -1.0);
    if (j <= len - 2) add(bond_f(sub_mi(p[2], p[3], L), l0, kb), 1.0);
    if (j <= len - 3) add(bend_f(sub_mi(p[2], p[3], L), sub_mi(p[3], p[4], L), c0, kt, 0), 1.0);
    if (j >= 1 && j <= len - 2) add(bend_f(sub_mi(p[1], p[2], L), sub_mi(p[2], p[3], L), c0, kt, 1), 1.0);
    if (j >= 2) add(bend_f(sub_mi(p[0], p[1], L), sub_mi(p[1], p[2], L), c0, kt, 2), 1.0);
    return f;
}

__device__ inline double sq3(double x, double y, double z)
{
#pragma clang fp contract(off)
    return (x * x + y * y) + z * z;
}

// The reference's forward half shell (cellOpt.h:715) as a mask over the stencil index
// (dx+1) + 3(dy+1) + 9(dz+1).
__host__ __device__ constexpr uint32_t forward_cells()
{
    const int nb[13][3] = {{1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1},
                           {-1, 0, 1}, {0, -1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}, {1, -1, 0}};
    uint32_t m = 0;
    for (int j = 0; j < 13; j++) m |= 1u << ((nb[j][0] + 1) + 3 * (nb[j][1] + 1) + 9 * (nb[j][2] + 1));
    return m;
}

// lanes of a bead's stencil in k_adj_pair: one per stencil cell (27 used)
constexpr int kAdjLanes = 32;

__device__ inline double lane_sum(double x)
{
#pragma unroll
    for (int o = kAdjLanes / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kAdjLanes);
    return x;
}

__global__ __launch_bounds__(kBlock) void k_adj_pair(int n, int nT, double cut2, const Bead *__restrict__ sb,
                                                     const uint32_t *__restrict__ skey,
                                                     const uint32_t *__restrict__ sidx,
                                                     const int *__restrict__ cstart, const int *__restrict__ cend,
                                                     const DevRep *__restrict__ reps, const double *__restrict__ fc,
                                                     const double *__restrict__ dfc, const double *__restrict__ duc,
                                                     const double4 *__restrict__ w, double4 *__restrict__ fout,
                                                     double4 *__restrict__ hvout, double *__restrict__ g, double obj,
                                                     int want_hv)
{
    const long long gt = (long long)blockIdx.x * kBlock + threadIdx.x;
    const int s = (int)(gt / kAdjLanes);   // sorted slot of the bead (uniform over its lanes)
    const int c = (int)(gt % kAdjLanes);   // its stencil cell
    const bool live = s < n;
    double Fx = 0, Fy = 0, Fz = 0, Hx = 0, Hy = 0, Hz = 0;
    double gw[2] = {0.0, 0.0}, du[2] = {0.0, 0.0};
    uint32_t me = 0;
    if (live) {
        const Bead hb = sb[s];
        me = sidx[s];
        const DevRep &R = reps[hb.rep];
        const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
        const uint32_t kbase = (uint32_t)R.cellbase;
        const uint32_t k = skey[s] - kbase;
        const int kx = (int)(k % (uint32_t)nx), ky = (int)((k / (uint32_t)nx) % (uint32_t)ny),
                  kz = (int)(k / (uint32_t)(nx * ny));
        const double4 wi = w[me];
        const size_t tsz = 6 * (size_t)nT * nT;
        if (c < 27) {
            int X0 = kx + c % 3 - 1, Y0 = ky + (c / 3) % 3 - 1, Z0 = kz + c / 9 - 1;
            const double sx = X0 < 0 ? -R.L[0] : (X0 >= nx ? R.L[0] : 0.0);
            const double sy = Y0 < 0 ? -R.L[1] : (Y0 >= ny ? R.L[1] : 0.0);
            const double sz = Z0 < 0 ? -R.L[2] : (Z0 >= nz ? R.L[2] : 0.0);
            X0 = X0 < 0 ? X0 + nx : (X0 >= nx ? X0 - nx : X0);
            Y0 = Y0 < 0 ? Y0 + ny : (Y0 >= ny ? Y0 - ny : Y0);
            Z0 = Z0 < 0 ? Z0 + nz : (Z0 >= nz ? Z0 - nz : Z0);
            const uint32_t kk = kbase + (uint32_t)(X0 + Y0 * nx + Z0 * nx * ny);
            const int c0 = cstart[kk];
            const int c1 = ((unsigned)c0 < (unsigned)n && skey[c0] == kk) ? cend[kk] : c0;
            for (int j = c0; j < c1; j++) {
                if (j == s) continue;
                const Bead q = sb[j];
                const double dx = hb.x - (q.x + sx), dy = hb.y - (q.y + sy), dz = hb.z - (q.z + sz);
                const double r2 = sq3(dx, dy, dz);
                if (!(r2 < cut2)) continue;
                const double r = sqrt(r2);
                // the row of the particle the reference's traversal makes first: across
                // cells the one whose forward half shell (cellOpt.h:715) holds the other's
                // cell, within a cell the lower index (cellOpt.h:572-585,780-795)
                constexpr uint32_t fwd = forward_cells();
                const bool first = c == 13 ? me < sidx[j] : ((fwd >> c) & 1u) != 0;
                const size_t row = 6 * (first ? (size_t)hb.type * nT + q.type : (size_t)q.type * nT + hb.type);
                const double *C = fc + row;
                const int o = (r >= C[0]) ? 3 : 0;
                const double m = C[o] - r;
                const double f = (C[o + 1] - C[o + 2] * m) * m;        // force along d/r
                const double fr = f / r;
                Fx += dx * fr;
                Fy += dy * fr;
                Fz += dz * fr;
                if (want_hv) {
                    const double4 wj = w[sidx[j]];
                    const double ux = wi.x - wj.x, uy = wi.y - wj.y, uz = wi.z - wj.z;
                    const double fp = -(C[o + 1] - 2.0 * C[o + 2] * m);  // df/dr
                    const double e = (dx * ux + dy * uy + dz * uz) / r;  // dr along the direction
                    const double t = (fp / r - f / r2) * e;
                    Hx += ux * fr + dx * t;
                    Hy += uy * fr + dy * t;
                    Hz += uz * fr + dz * t;
                }
                const double wd = (wi.x * dx + wi.y * dy + wi.z * dz) / r;
                const double *U = duc + row;   // only C1, C2, C4, C5 of the U row depend on theta
                for (int p = 0; p < 2; p++) {
                    const double *D = dfc + p * tsz + row;
                    gw[p] += wd * ((D[o + 1] - D[o + 2] * m) * m);
                    if (obj != 0.0) {
                        const double *E = U + p * tsz;
                        double d;
                        if (r <= C[0]) {
                            const double mu = C[0] - r;
                            d = E[1] * mu * mu + E[2];
                        } else {
                            const double mu = C[3] - r;
                            d = mu * mu * (E[4] - mu * E[5]);
                        }
                        du[p] += 0.5 * d;
                    }
                }
            }
        }
    }
    // every lane of the group takes part in the tree, live or not
    Fx = lane_sum(Fx);
    Fy = lane_sum(Fy);
    Fz = lane_sum(Fz);
    Hx = lane_sum(Hx);
    Hy = lane_sum(Hy);
    Hz = lane_sum(Hz);
    const double g0 = lane_sum(gw[0] + obj * du[0]);
    const double g1 = lane_sum(gw[1] + obj * du[1]);
    if (live && c == 0) {
        fout[me] = make_double4(Fx, Fy, Fz, 0.0);
        if (want_hv) hvout[me] = make_double4(Hx, Hy, Hz, 0.0);
        g[me] += g0;
        g[(size_t)n + me] += g1;
    }
}

__global__ __launch_bounds__(kBlock) void k_adj_chain(int n, const Bead *__restrict__ beads,
                                                      const int2 *__restrict__ chain,
                                                      const double4 *__restrict__ cconst,
                                                      const DevRep *__restrict__ reps,
                                                      const double4 *__restrict__ w, double4 *__restrict__ fout,
                                                      double4 *__restrict__ hvout, double *__restrict__ g, double obj,
                                                      int want_hv)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int2 ci = chain[i];
    if (ci.x < 0) return;
    const Bead b0 = beads[i];
    const DevRep &R = reps[b0.rep];
    const double L[3] = {R.L[0], R.L[1], R.L[2]};
    const int start = ci.x, len = ci.y & 0xFFFF, j = i - start;
    const double4 K = cconst[ci.y >> 16];
    V3<Dual> pd[5];
    V3<double> pv[5];
    for (int t = 0; t < 5; t++) {
        const int jj = j + t - 2;
        if (jj < 0 || jj >= len) {
            pd[t] = V3<Dual>{Dual{0, 0}, Dual{0, 0}, Dual{0, 0}};
            pv[t] = V3<double>{0, 0, 0};
            continue;
        }
        const Bead q = beads[i + t - 2];
        const double4 wq = want_hv ? w[i + t - 2] : make_double4(0, 0, 0, 0);
        pd[t] = V3<Dual>{Dual{q.x, wq.x}, Dual{q.y, wq.y}, Dual{q.z, wq.z}};
        pv[t] = V3<double>{q.x, q.y, q.z};
    }
    const V3<Dual> f = chain_force(pd, j, len, L, K.x, K.y, K.z, K.w);
    double4 fo = fout[i];
    fo.x += f.x.v; fo.y += f.y.v; fo.z += f.z.v;
    fout[i] = fo;
    if (want_hv) {
        double4 h = hvout[i];
        h.x += f.x.d; h.y += f.y.d; h.z += f.z.d;
        hvout[i] = h;
    }
    // theta derivatives: the bond part of the force is linear in k_bond, the bend part in
    // k_bend, so dF/dk is the force at k = 1 with the other stiffness off
    const V3<double> fb = chain_force(pv, j, len, L, K.x, 1.0, K.z, 0.0);
    const V3<double> ft = chain_force(pv, j, len, L, K.x, 0.0, K.z, 1.0);
    const double4 wi = w[i];
    double gb = wi.x * fb.x + wi.y * fb.y + wi.z * fb.z;
    double gt = wi.x * ft.x + wi.y * ft.y + wi.z * ft.z;
    if (obj != 0.0) {
        // dU/dk over the bond (i, i+1) and the bend (i, i+1, i+2) this bead owns
        if (j <= len - 2) {
            const V3<double> d = sub_mi(pv[2], pv[3], L);
            const double dr = sqrt(d.x * d.x + d.y * d.y + d.z * d.z) - K.x;
            gb += obj * 0.5 * dr * dr;
        }
        if (j <= len - 3) {
            const V3<double> a = sub_mi(pv[2], pv[3], L), c = sub_mi(pv[3], pv[4], L);
            const double ra = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
            const double rb = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
            const double cost = (a.x / ra) * (c.x / rb) + (a.y / ra) * (c.y / rb) + (a.z / ra) * (c.z / rb);
            const double p = K.z - cost;
            gt += obj * 0.5 * p * p;
        }
    }
    g[2 * (size_t)n + i] += gb;
    g[3 * (size_t)n + i] += gt;
}

// Generic BOND / BEND / BALL records (doBondForce, doBendForce, doBallForce,
// system.h:1881-2035): per bead, the force of each record it belongs to, from its side, as
// k_bonded_kick takes it, evaluated on dual numbers along w for (dF/dx) w.  Their constants
// are not among theta, so they add nothing to g.
__global__ __launch_bounds__(kBlock) void k_adj_gen(int n, GenTables gen, const Bead *__restrict__ beads,
                                                    const DevRep *__restrict__ reps, const double4 *__restrict__ w,
                                                    double4 *__restrict__ fout, double4 *__restrict__ hvout,
                                                    int want_hv)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int e0 = gen.gptr[i], e1 = gen.gptr[i + 1];
    if (e0 == e1) return;
    const DevRep &R = reps[beads[i].rep];
    const double L[3] = {R.L[0], R.L[1], R.L[2]};
    auto at = [&](int q) {
        const Bead b = beads[q];
        const double4 wq = want_hv ? w[q] : make_double4(0, 0, 0, 0);
        return V3<Dual>{Dual{b.x, wq.x}, Dual{b.y, wq.y}, Dual{b.z, wq.z}};
    };
    const Dual z{0.0, 0.0};
    V3<Dual> f{z, z, z};
    for (int e = e0; e < e1; e++) {
        const GenEnt g = gen.gent[e];
        const double2 K = gen.gconst[e];
        V3<Dual> fr;
        if (g.kind == 0 || g.kind == 2) {
            const V3<Dual> d = sub_mi(at(g.p1), at(g.p2), L);
            // BALL (kind 2, harmonicHalfF, MD.h:1164-1187): only inside r0
            const bool on = g.kind == 0 || d.x.v * d.x.v + d.y.v * d.y.v + d.z.v * d.z.v < K.x * K.x;
            fr = on ? bond_f(d, K.x, K.y) : V3<Dual>{z, z, z};
            if (i != g.p1) fr = V3<Dual>{-fr.x, -fr.y, -fr.z};
        } else {
            const V3<Dual> q1 = at(g.p1), q2 = at(g.p2), q3 = at(g.p3);
            fr = bend_f(sub_mi(q1, q2, L), sub_mi(q2, q3, L), K.x, K.y, i == g.p1 ? 0 : (i == g.p2 ? 1 : 2));
        }
        f = V3<Dual>{f.x + fr.x, f.y + fr.y, f.z + fr.z};
    }
    double4 fo = fout[i];
    fo.x += f.x.v; fo.y += f.y.v; fo.z += f.z.v;
    fout[i] = fo;
    if (want_hv) {
        double4 h = hvout[i];
        h.x += f.x.d; h.y += f.y.d; h.z += f.z.d;
        hvout[i] = h;
    }
}

// Reverse of one step s, given the adjoints (X, V, A) of (x_s, v_s, a_s), where A already
// holds c V_s for type != 0 (the v_s = v_h + c a_s term):
//   X~ = S X (accepted MC move) - obj F(x~_s) + (dF/dx) A     (objective obj U(x~_s))
//   V_h = V - gamma A (+ dt X~ for type != 0)
//   X_{s-1} = X~,  V_{s-1} = V_h,  A_{s-1} = c V_h (type != 0), plus c V_{s-1} when the
//   previous step is a full step (prep_next) rather than the initial force evaluation.
__global__ __launch_bounds__(kBlock) void k_adj_step(int n, const Bead *__restrict__ beads,
                                                     const DevRep *__restrict__ post, int mc, double4 *__restrict__ X,
                                                     double4 *__restrict__ V, double4 *__restrict__ A,
                                                     const double4 *__restrict__ F, const double4 *__restrict__ HV,
                                                     double obj, double c, double dt, int prep_next)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const Bead b = beads[i];
    const DevRep &R = post[b.rep];
    double4 x = X[i];
    if (mc && R.accept) {
        x.x *= R.scale[0];
        x.y *= R.scale[1];
        x.z *= R.scale[2];
    }
    const double4 f = F[i], h = HV[i], a = A[i], v = V[i];
    x.x += h.x - obj * f.x;
    x.y += h.y - obj * f.y;
    x.z += h.z - obj * f.z;
    const double gamma = R.gamma;
    double4 vh = make_double4(v.x - gamma * a.x, v.y - gamma * a.y, v.z - gamma * a.z, 0.0);
    double4 an = make_double4(0, 0, 0, 0);
    if (b.type != 0) {
        vh.x += dt * x.x;
        vh.y += dt * x.y;
        vh.z += dt * x.z;
        const double k = prep_next ? 2.0 * c : c;
        an = make_double4(k * vh.x, k * vh.y, k * vh.z, 0.0);
    }
    X[i] = x;
    V[i] = vh;
    A[i] = an;
}

}  // namespace

void launch_adj_pair(hipStream_t st, const Params &P, const Bead *sbeads, const uint32_t *skey, const uint32_t *sidx,
                     const int *cstart, const int *cend, const DevRep *reps, const AdjTables &T, const double4 *w,
                     double4 *fout, double4 *hvout, double *g, double obj, int want_hv)
{
    hipLaunchKernelGGL(k_adj_pair, dim3(blocks((long long)P.n * kAdjLanes)), dim3(kBlock), 0, st, P.n, P.nT,
                       P.cutoff * P.cutoff, sbeads,
                       skey, sidx, cstart, cend, reps, T.fc, T.dfc, T.duc, w, fout, hvout, g, obj, want_hv);
}

void launch_adj_gen(hipStream_t st, const Params &P, const GenTables &gen, const Bead *beads, const DevRep *reps,
                    const double4 *w, double4 *fout, double4 *hvout, int want_hv)
{
    hipLaunchKernelGGL(k_adj_gen, dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, gen, beads, reps, w, fout, hvout,
                       want_hv);
}

void launch_adj_chain(hipStream_t st, const Params &P, const Bead *beads, const int2 *chain, const double4 *cconst,
                      const DevRep *reps, const double4 *w, double4 *fout, double4 *hvout, double *g, double obj,
                      int want_hv)
{
    hipLaunchKernelGGL(k_adj_chain, dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads, chain, cconst, reps, w, fout,
                       hvout, g, obj, want_hv);
}

void launch_adj_step(hipStream_t st, const Params &P, const Bead *beads, const DevRep *post, int mc, double4 *X,
                     double4 *V, double4 *A, const double4 *F, const double4 *HV, double obj, int prep_next)
{
    hipLaunchKernelGGL(k_adj_step, dim3(blocks(P.n)), dim3(kBlock), 0, st, P.n, beads, post, mc, X, V, A, F, HV, obj,
                       0.5 * P.dt, P.dt, prep_next);
}

}  // namespace smdev
