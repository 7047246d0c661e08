/* softmold_hip.h -- C ABI of the MI355X (gfx950) SoftMold MD-step solver.
 *
 * The library replaces the compute path of the reference `MD` executable (MD.cpp:72-754)
 * and its GPU twin (MDcuda.cu:77-495): cell-list build with a spatial sort, the
 * Laradji-Revalee pair force, CHAIN bond and bend forces, the Langevin velocity-Verlet
 * update and the Monte-Carlo constant-tension box move.  The reference exposes that path
 * as header-only C++ templates (include/cuda/mpdCuda.h:1-19, `mpd::*_device(...)`) driven
 * by MDcuda.cu; this header is the plain-pointer replacement a foreign host binds to.
 *
 * Conventions
 *  - Every function returns SM_OK (0) or an SM_ERR_* code; sm_last_error() gives the
 *    thread-local message of the last failure.  No C++ exception crosses this ABI.
 *  - Host buffers are caller-owned and copied; the library owns device memory.
 *  - Arrays are row-major: positions, velocities, accelerations are n x 3 doubles.
 *  - A context is single-threaded and stream-ordered; calls that return host data
 *    synchronise its stream.
 */
#ifndef SOFTMOLD_HIP_H
#define SOFTMOLD_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SM_OK 0
#define SM_ERR_INVALID 1     /* bad argument or input system (reference: Blob::errorChecking throws, system.h:438-580) */
#define SM_ERR_HIP 2         /* HIP runtime failure (reference: errors.h:7-30 runtime_error) */
#define SM_ERR_UNSUPPORTED 3 /* input outside the implemented path (e.g. a molecule type other than CHAIN) */
#define SM_ERR_STATE 4       /* device-side check failed (non-finite position, cell table overflow) */
#define SM_ERR_IO 5          /* file could not be read/written or parsed (reference: Script throws 0, scriptFormat.h:48-52) */

typedef struct sm_ctx sm_ctx;
typedef struct sm_mpd sm_mpd;

/* One MD system, the fields of a `.mpd` the MD step reads (include/system.h:362-393). */
typedef struct sm_system {
    int32_t n_particles;
    int32_t n_types;
    const double *pos;          /* n*3 */
    const int32_t *type;        /* n, 0 <= type < n_types; type 0 does not move (verlet.h:296) */
    const double *vel;          /* n*3 */
    const double *acc;          /* n*3 or NULL */
    int32_t n_chain_blocks;     /* CHAIN records: (start, nChains, length) (molecules.h:38-40) */
    const int32_t *chain;       /* 3 per block */
    const double *chain_const;  /* 4 per block: bond length, k_bond, cos-theta term, k_bend */
    /* Two-body tables, row t1*n_types + t2.  A pair takes the row of the particle the
     * reference's half-shell traversal makes first (MD.h:811-819, cellOpt.h:712-926),
     * which only matters where the (a,b) and (b,a) rows differ.  Rows with r_m below
     * cutoff/2 are refused (SM_ERR_UNSUPPORTED): the reference's branch offset
     * int(r/r_m)*3 (MD.h:819) would leave the row. */
    const double *fconst;       /* twoBodyFconst, 6*n_types^2 (potentials/laradjiRevalee.h:84-112) */
    const double *uconst;       /* twoBodyUconst, 6*n_types^2 (potentials/laradjiRevalee.h:124-152) */
    double box[3];
    double cutoff, dt, gamma, temperature;
    double delta_lxy, tension;  /* MC box move every 8 steps when delta_lxy != 0 (MD.cpp:589) */
    uint64_t seed;              /* counter-RNG key: noise = Philox4x32-10(seed, step, particle) */
    /* Generic BOND and BEND records (MDcuda.cu:162-184; forces system.h:1881-2035), each
     * with its own two constants; n_bonds / n_bends may be 0 with NULL arrays. */
    int32_t n_bonds;
    const int32_t *bond;        /* 2 per bond: particle indices */
    const double *bond_const;   /* 2 per bond: bond length, k_bond (harmonicF, MD.h:384-403) */
    int32_t n_bends;
    const int32_t *bend;        /* 3 per bend */
    const double *bend_const;   /* 2 per bend: cos-theta term, k_bend (bendF, MD.h:683-723) */
    /* BALL records (doBallForce, system.h:1937-1971): a half-harmonic repulsion between the
     * ball's centre particle and each member, active only inside r0 (harmonicHalfF,
     * MD.h:1164-1187); U adds harmonicHalfP (MD.h:1210-1222), the box move's dU does not
     * (MD.cpp:664 drops doBallDPotential's value). */
    int32_t n_balls;
    const int32_t *ball;        /* 2 per record: centre, member */
    const double *ball_const;   /* 2 per record: r0, k */
    /* > 0: the Langevin sigma = sqrt(6 T gamma / dt) is taken once at this temperature and
     * kept when the temperature changes, as the gammaType thermostat of MD.cpp:135-138 does
     * (langevin.h:235-241 computes its per-type sigma on the first call only); 0: sigma
     * follows the temperature (scalar gamma, MD.cpp:131-134). */
    double sigma_temperature;
} sm_system;

const char *sm_last_error(void);
const char *sm_version(void);

/* ---- context ---------------------------------------------------------------------- */
/* A context holds R >= 1 replica systems stepped together, every kernel serving all of
 * them in one launch.  Per-replica arrays below hold R entries (box: 3 per replica);
 * bead arrays hold all beads, replica after replica (sm_replicas gives the ranges).
 * A replica's results do not depend on the other replicas: its forces, energies and
 * trajectory equal those of the same system in a context of its own. */
/* Replaces state construction + toDevice (MDcuda.cu:218-278, include/cuda/state.h:11-97). */
int sm_create(const sm_system *sys, sm_ctx **out);
/* R replicas (the batched sweep; no reference counterpart -- the reference runs one MD
 * process per system).  Replicas must share n_types, cutoff, dt and the constant tables;
 * box, temperature, gamma, delta_lxy, tension and seed are per replica. */
int sm_create_batch(const sm_system *systems, int32_t n_replicas, sm_ctx **out);
int sm_destroy(sm_ctx *ctx);
/* Replica count and each replica's bead range; any pointer may be NULL. */
int sm_replicas(sm_ctx *ctx, int32_t *n_replicas, int32_t *first, int32_t *count);
int sm_sync(sm_ctx *ctx);
int sm_download_state(sm_ctx *ctx, double *pos, double *vel, double *acc, double *box);   /* state.toHost, state.h:77-97 */
int sm_upload_state(sm_ctx *ctx, const double *pos, const double *vel, const double *acc, const double *box);
int sm_set_temperature(sm_ctx *ctx, double T);   /* every replica; System.setInitialTemp (MD.cpp:369-371) */

/* ---- the step ----------------------------------------------------------------------- */
/* Step-0 forces (MD.cpp:186-262): a = F_pair + Langevin(T) + F_chain, no integration. */
int sm_init_forces(sm_ctx *ctx);
/* Verlet::second only (MD.cpp:307, restart from a mid-step store). */
int sm_second_half(sm_ctx *ctx);
/* First half of step `step`: Verlet::first + wrap (MD.cpp:356). The `.mpd` store of the
 * reference happens right after this (MD.cpp:373-381). */
int sm_step_first(sm_ctx *ctx, int64_t step);
/* Rest of step `step` (MD.cpp:410-511, 589-721): Langevin, cell build, pair and CHAIN forces,
 * Verlet::second; if measure != NULL it receives {U_pair, U_chain, K} of each replica after
 * the kick (dataExtraction.h:827-1050); then the MC box move when step%8==0, step!=0, in
 * every replica with delta_lxy != 0. With the "mc_fused" knob on (the default) a box-move
 * step draws its trial scale before the pair force and takes the pair dU in the same
 * traversal. Both paths sum the same FP64 pair terms, in different orders; the knob set to
 * 0 gives the separate dU gather. */
int sm_step_finish(sm_ctx *ctx, int64_t step, double *measure);
/* Whole steps first_step .. first_step+nsteps-1 with no measurement, all on the device.
 * With the "graph" knob on, aligned blocks of 8 steps replay one captured hipGraph. */
int sm_run(sm_ctx *ctx, int64_t first_step, int32_t nsteps);
/* MC statistics so far, per replica (MD.cpp:713-720,748). */
int sm_mc_stats(sm_ctx *ctx, int64_t *trials, int64_t *accepted);
/* Accepted/rejected trial histograms of one replica over the fluctuation
 * (MD.cpp:106-116,708-718): bin width delta_lxy/101, bin b covers b*width - delta_lxy;
 * n_out = int(2*101)+1 bins (0 when delta_lxy == 0); at most `cap` bins are copied. */
int sm_mc_hist(sm_ctx *ctx, int32_t replica, int32_t cap, double *accepted, double *rejected, int32_t *n_out);

/* ---- measurement and parity hooks --------------------------------------------------- */
/* {U_pair, U_chain, K} of each replica in FP64 at the current state
 * (dataExtraction.h:840-842,942-948); out holds 3 per replica. */
int sm_measure(sm_ctx *ctx, double *out);
/* Mean distance of each replica's beads of type head_type (2 = HEAD, MD.h:21) from their
 * centre of mass, computed on the device in FP64 from the resident positions: the
 * per-frame computation of analysis/vesicleRadius.cpp:71-175, with its image rules (the
 * centre-of-mass loop images heads towards the replica's first bead; the radius loop adds
 * its image term).  radius holds one value per replica (NaN for a replica without heads). */
int sm_vesicle_radius(sm_ctx *ctx, int32_t head_type, double *radius);
/* sm_vesicle_radius without any device allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under stream capture, tight measuring
 * loops of a sweep driver, harnesses that forbid hipMalloc).  Its partial sums live in a
 * fixed per-GPU device array of 8 doubles per replica, set up at context creation, so it
 * accepts contexts of at most 1024 replicas (SM_ERR_UNSUPPORTED beyond).  Results equal
 * sm_vesicle_radius's bit for bit.  All calls on one GPU share that array: calls that
 * overlap on one GPU (other host threads, other contexts' streams) race. */
int sm_vesicle_radius_noalloc(sm_ctx *ctx, int32_t head_type, double *radius);
/* Statistics of the generic records at the current state, 4 per replica, as MD writes them
 * to lBond_name.dat and bend_name.dat (dataExtraction.h:861-930,1009-1027): mean bond
 * length (0/0 = NaN without BOND records), then the bends' mean cos theta and mean arm
 * lengths (sums, not means, when there is at most one BEND record). */
int sm_bond_stats(sm_ctx *ctx, double *out);
/* Conservative step-0 forces with FP32 pair arithmetic: pair part (CellOpt::computeForce,
 * cellOpt.h:712-926) and CHAIN part (Blob::doChainForce, system.h:1782-1866), n*3 each;
 * either pointer may be NULL. */
int sm_compute_forces(sm_ctx *ctx, double *pair_out, double *chain_out);
/* Cell key per particle within its replica's grid (CellOpt::build, cellOpt.h:530-556). */
int sm_cell_keys(sm_ctx *ctx, int32_t *keys);
/* Neighbour count per particle with the r^2 < rc^2 test in FP64 (the FORCE_COUNT build
 * of Force, MD.h:838-845); total = number of pairs. */
int sm_pair_count(sm_ctx *ctx, int32_t *counts, int64_t *total);
/* The pair force kernel's own accepted set, counted per particle: its traversal, staging,
 * per-quarter-cell culling and FP64 test r^2 < rc^2, with a pair counted when its force
 * term is not identically zero -- r^2 < r_m^2, or a row with a nonzero outer coefficient
 * (Force, MD.h:811-848; Umin != 0 in laradjiRevalee.h:97-110).  A parity hook for the
 * culling: the same definition over all pairs of the FORCE_COUNT traversal (MD.h:838-845)
 * must give the same counts.  Needs the one-phase force kernel (pair_variant 0). */
int sm_force_pair_count(sm_ctx *ctx, int32_t *counts, int64_t *total);
/* Pairs (i<j) within the cutoff, FP64 test, in no particular order. */
int sm_pair_list(sm_ctx *ctx, int32_t *pairs, int64_t cap, int64_t *n_out);
/* dU = U(old) - U(scaled) over the current pair set, FP64 (CellOpt::computeDPotential
 * cellOpt.h:1043-1180 and Blob::doChainDPotential system.h:3280-3384); scale holds 3
 * per replica, du_pair and du_chain one per replica. */
int sm_potential_delta(sm_ctx *ctx, const double *scale, double *du_pair, double *du_chain);

/* ---- launch shapes ------------------------------------------------------------------ */
/* Set a launch-shape knob of this build; no counterpart in the reference (its launches are
 * fixed, include/cuda/other.h:6-11).  Knobs: "pair_chunk" = sorted slots per wave64 in
 * the pair-force kernel (1..4096); "pair_variant" = 0 one-phase force kernel with the
 * home cell's candidates culled per quarter cell (default), 1 two-phase (FP32 screen into
 * per-lane hit lists, then the FP64 test on the hits);
 * "graph" = 1 to replay sm_run's aligned 8-step blocks as one captured hipGraph;
 * "sort_mode" = 1 to re-sort the cell keys incrementally -- the beads whose cell changed
 * are sorted and merged into the previous order (the default from a million beads up) --
 * 0 to sort every key each build, or 2 (single-replica contexts of at most 524288 beads,
 * their default) to sort every key within a window of cells around the previous build's
 * occupied cells: two 9-bit passes when every key lies in it, else the cell indices' 8-bit
 * passes, decided on the device.  All give the same order, so none changes a result.
 * "mc_fused" = 1 (default) to take a box-move step's pair dU in the traversal of
 * that step's pair force, 0 in a gather of its own after the kick (sm_potential_delta
 * follows the same knob); the two sum the same FP64 terms in different orders.
 * sm_tune_get also reads "sort_window_path": 1 if the last build's sort took the window
 * path (it synchronises the context's stream). */
int sm_tune(sm_ctx *ctx, const char *knob, int32_t value);
int sm_tune_get(sm_ctx *ctx, const char *knob, int32_t *value);
/* Autotuner (no reference counterpart): times `steps` steps of this context's own work
 * for each pair_chunk in {2,4,8,16,32,64} with the one-phase kernel and the graph on, then
 * the best with each other sort_mode, then the best without the graph, from one saved
 * state; keeps the fastest setting, restores
 * the state, and returns the best ms per step.  These settings leave every result
 * bit-identical, so tuning never changes a trajectory. */
int sm_autotune(sm_ctx *ctx, int32_t steps, double *best_ms_per_step);

/* ---- adjoint with respect to the interaction constants -------------------------------- */
/* theta = {Umin, Umax} of the type pair (a, b) -- the laradjiRevaleeFC/PC rows (a,b) and
 * (b,a), each keeping its r_m and r_c (potentials/laradjiRevalee.h:84-152) -- then k_bond
 * and k_bend of every CHAIN constant set (molecule constants 1 and 3, system.h:1129-1158).
 * No counterpart in the reference, whose constants are fixed per run; the callers are
 * constant-fitting loops. */
int sm_set_params(sm_ctx *ctx, int32_t type_a, int32_t type_b, const double *theta);
/* dU/dtheta at the current state, 4 per replica, FP64: d(U_pair + U_chain)/dUmin and /dUmax
 * of the type pair (a, b), then /dk_bond and /dk_bend of the CHAIN sets.  U is linear in the
 * laradjiRevaleePC row entries with r_m and r_c fixed (potentials/laradjiRevalee.h:124-152,
 * Potential MD.h:895-930) and in the bond and bend stiffnesses (harmonicP MD.h:422-431,
 * bendP MD.h:769-789), so these are the energy sums over derivative tables.  Generic
 * BOND / BEND / BALL records carry their own constants and contribute nothing.  No
 * counterpart in the reference; the callers are equilibrium-response estimators (the
 * derivative of an ensemble average by its covariance with dU/dtheta). */
int sm_param_derivative(sm_ctx *ctx, int32_t type_a, int32_t type_b, double *out);
int sm_get_params(sm_ctx *ctx, int32_t type_a, int32_t type_b, double *theta);
/* From the context's current positions and velocities: initial forces, then steps
 * 1..nsteps with the counter RNG (box moves as configured), J = mean over the steps of
 * U_pair + U_chain (with the generic records' energy) after each step's kick, one value
 * per replica.  If grad != NULL it receives dJ/dtheta (4 values per replica, replica
 * after replica) from the discrete adjoint of the step in FP64, recomputing the trajectory
 * from checkpoints every `checkpoint` steps; MC accept/reject decisions are held fixed
 * (the pathwise derivative).  The context state is left as it was. */
int sm_adjoint(sm_ctx *ctx, int32_t type_a, int32_t type_b, int32_t nsteps, int32_t checkpoint, double *J,
               double *grad);

/* ---- profiling support -------------------------------------------------------------- */
/* Average device time (ms) of each kernel class over the launches since the last reset,
 * measured with HIP events on the context's stream.  names[] receives static strings. */
int sm_profile_enable(sm_ctx *ctx, int32_t on);
int sm_profile_read(sm_ctx *ctx, int32_t cap, const char **names, double *avg_ms, int64_t *count, int32_t *n_out);

/* ---- `.mpd` files and the reference generators (host only) -------------------------- */
/* Script<...>::read / Blob::input (scriptFormat.h:79-94, system.h:589-1313); `path` is
 * the file name without ".mpd", as MD.cpp's argument. */
int sm_mpd_read(const char *name, sm_mpd **out);
int sm_mpd_write(const sm_mpd *m, const char *name);   /* Script::write / Blob::output, system.h:1315-1749 */
int sm_mpd_free(sm_mpd *m);
/* Borrowed view of the system (pointers valid until the handle changes or is freed). */
int sm_mpd_system(const sm_mpd *m, sm_system *out);
int sm_mpd_get(const sm_mpd *m, const char *key, double *value);   /* scalar keys */
/* Scalar keys; marks the key present, except tension and deltaLXY at 0 and deltaT <= 0,
 * which leave the file as the Blob setters do (system.h:4461-4546). */
int sm_mpd_set(sm_mpd *m, const char *key, double value);
int sm_mpd_set_state(sm_mpd *m, const double *pos, const double *vel, const double box[3]);
/* generate/liposome.cpp:1-97 with models/lipidModels.h:1313-1469. */
int sm_gen_liposome(int32_t seed, int32_t n_lipids, double areal_density, sm_mpd **out);
/* generate/bilayer.cpp:1-125 (no anchors) with models/lipidModels.h:1740-1924. */
int sm_gen_bilayer(int32_t seed, int32_t n_lipids, double areal_density, sm_mpd **out);

#ifdef __cplusplus
}
#endif
#endif
