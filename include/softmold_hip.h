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
    const double *fconst;       /* twoBodyFconst, 6*n_types^2 (potentials/laradjiRevalee.h:84-112) */
    const double *uconst;       /* twoBodyUconst, 6*n_types^2 (potentials/laradjiRevalee.h:124-152) */
    double box[3];
    double cutoff, dt, gamma, temperature;
    double delta_lxy, tension;  /* MC box move every 8 steps when delta_lxy != 0 (MD.cpp:589) */
    uint64_t seed;              /* counter-RNG key: noise = Philox4x32-10(seed, step, particle) */
} sm_system;

const char *sm_last_error(void);
const char *sm_version(void);

/* ---- context ---------------------------------------------------------------------- */
/* Replaces state construction + toDevice (MDcuda.cu:218-278, include/cuda/state.h:11-97). */
int sm_create(const sm_system *sys, sm_ctx **out);
int sm_destroy(sm_ctx *ctx);
int sm_sync(sm_ctx *ctx);
int sm_download_state(sm_ctx *ctx, double *pos, double *vel, double *acc, double box[3]);   /* state.toHost, state.h:77-97 */
int sm_upload_state(sm_ctx *ctx, const double *pos, const double *vel, const double *acc, const double box[3]);
int sm_set_temperature(sm_ctx *ctx, double T);   /* System.setInitialTemp (MD.cpp:369-371) */

/* ---- the step ----------------------------------------------------------------------- */
/* Step-0 forces (MD.cpp:186-262): a = F_pair + Langevin(T) + F_chain, no integration. */
int sm_init_forces(sm_ctx *ctx);
/* Verlet::second only (MD.cpp:307, restart from a mid-step store). */
int sm_second_half(sm_ctx *ctx);
/* First half of step `step`: Verlet::first + wrap (MD.cpp:356). The `.mpd` store of the
 * reference happens right after this (MD.cpp:373-381). */
int sm_step_first(sm_ctx *ctx, int64_t step);
/* Rest of step `step` (MD.cpp:410-511, 589-721): Langevin, cell build, pair and CHAIN forces,
 * Verlet::second; if measure != NULL it receives {U_pair, U_chain, K} after the kick
 * (dataExtraction.h:827-1050); then the MC box move when step%8==0, step!=0, delta_lxy!=0. */
int sm_step_finish(sm_ctx *ctx, int64_t step, double *measure);
/* Whole steps first_step .. first_step+nsteps-1 with no measurement, all on the device. */
int sm_run(sm_ctx *ctx, int64_t first_step, int32_t nsteps);
/* MC statistics so far (MD.cpp:713-720,748). */
int sm_mc_stats(sm_ctx *ctx, int64_t *trials, int64_t *accepted);
/* Accepted/rejected trial histograms over the fluctuation (MD.cpp:106-116,708-718): bin
 * width delta_lxy/101, bin b covers b*width - delta_lxy; n_out = int(2*101)+1 bins
 * (0 when delta_lxy == 0); at most `cap` bins are copied. */
int sm_mc_hist(sm_ctx *ctx, int32_t cap, double *accepted, double *rejected, int32_t *n_out);

/* ---- measurement and parity hooks --------------------------------------------------- */
/* {U_pair, U_chain, K} in FP64 at the current state (dataExtraction.h:840-842,942-948). */
int sm_measure(sm_ctx *ctx, double out[3]);
/* Conservative step-0 forces with FP32 pair arithmetic: pair part (CellOpt::computeForce,
 * cellOpt.h:712-926) and CHAIN part (Blob::doChainForce, system.h:1782-1866), n*3 each;
 * either pointer may be NULL. */
int sm_compute_forces(sm_ctx *ctx, double *pair_out, double *chain_out);
/* Cell key per particle (CellOpt::build, cellOpt.h:530-556). */
int sm_cell_keys(sm_ctx *ctx, int32_t *keys);
/* Neighbour count per particle with the r^2 < rc^2 test in FP64 (the FORCE_COUNT build
 * of Force, MD.h:838-845); total = number of pairs. */
int sm_pair_count(sm_ctx *ctx, int32_t *counts, int64_t *total);
/* Pairs (i<j) within the cutoff, FP64 test, in no particular order. */
int sm_pair_list(sm_ctx *ctx, int32_t *pairs, int64_t cap, int64_t *n_out);
/* dU = U(old) - U(scaled) over the current pair set, FP64 (CellOpt::computeDPotential
 * cellOpt.h:1043-1180 and Blob::doChainDPotential system.h:3280-3384). */
int sm_potential_delta(sm_ctx *ctx, const double scale[3], double *du_pair, double *du_chain);

/* ---- launch shapes ------------------------------------------------------------------ */
/* Set a launch-shape knob of this build; no counterpart in the reference (its launches are
 * fixed, include/cuda/other.h:6-11).  Knobs: "pair_chunk" = sorted slots per wave64 in
 * the pair-force kernel (1..4096). */
int sm_tune(sm_ctx *ctx, const char *knob, int32_t value);

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
int sm_mpd_set(sm_mpd *m, const char *key, double value);          /* scalar keys; marks present */
int sm_mpd_set_state(sm_mpd *m, const double *pos, const double *vel, const double box[3]);
/* generate/liposome.cpp:1-97 with models/lipidModels.h:1313-1469. */
int sm_gen_liposome(int32_t seed, int32_t n_lipids, double areal_density, sm_mpd **out);
/* generate/bilayer.cpp:1-125 (no anchors) with models/lipidModels.h:1740-1924. */
int sm_gen_bilayer(int32_t seed, int32_t n_lipids, double areal_density, sm_mpd **out);

#ifdef __cplusplus
}
#endif
#endif
