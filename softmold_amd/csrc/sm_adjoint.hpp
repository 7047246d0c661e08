// Launchers of the adjoint kernels (sm_adjoint.hip).
#pragma once
#include "sm_internal.hpp"

namespace smdev {

// Adjoint parameters: theta = (Umin, Umax) of one type pair, then k_bond, k_bend.
constexpr int kNTheta = 4;

struct AdjTables {
    const double *fc;    // F constants, FP64, 6 nT^2
    const double *dfc;   // dF-constants / d(Umin), d(Umax): 2 x 6 nT^2
    const double *duc;   // dU-constants / d(Umin), d(Umax): 2 x 6 nT^2
};

// Per sorted slot: FP64 pair force F (=), (dF/dx) w (=, if want_hv), and into the per-bead
// accumulators g[p*n + i] (p = 0, 1): w_i . dF_i/dtheta_p + obj * dU_i/dtheta_p, where
// dU_i is the bead's half share of its pairs.
void launch_adj_pair(hipStream_t st, const Params &P, const Bead *sbeads, const uint32_t *skey, const uint32_t *sidx,
                     const int *cstart, const int *cend, const DevRep *reps, const AdjTables &T, const double4 *w,
                     double4 *fout, double4 *hvout, double *g, double obj, int want_hv);
// Per bead: CHAIN force (+=), its (dF/dx) w (+=), and into g[2n + i], g[3n + i] the k_bond and
// k_bend terms (w . dF/dk + obj * dU/dk over the bonds and bends the bead owns).
void launch_adj_chain(hipStream_t st, const Params &P, const Bead *beads, const int2 *chain, const double4 *cconst,
                      const DevRep *reps, const double4 *w, double4 *fout, double4 *hvout, double *g, double obj,
                      int want_hv);
// Per bead: the generic BOND / BEND / BALL records' force (+=) and its (dF/dx) w (+=).
void launch_adj_gen(hipStream_t st, const Params &P, const GenTables &gen, const Bead *beads, const DevRep *reps,
                    const double4 *w, double4 *fout, double4 *hvout, int want_hv);
// One reverse step (see sm_adjoint.hip); post = the replicas after this step's MC move.
void launch_adj_step(hipStream_t st, const Params &P, const Bead *beads, const DevRep *post, int mc, double4 *X,
                     double4 *V, double4 *A, const double4 *F, const double4 *HV, double obj, int prep_next);

}  // namespace smdev
