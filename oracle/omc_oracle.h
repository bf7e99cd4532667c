/*
 * CPU oracle for the ompMC shower() path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference transport (src/ompmc.c:37-211,
 * 1102-2121, 3097-3199, 3606-5447, src/omc_random.c:58-187 and the
 * omc_dosxyz / omc_matrad user code it calls back into).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path (lib/libomc_hip.so) never does.
 *
 * Random-number modes:
 *   ORA_RNG_RANMAR  one sequential RANMAR stream, as the reference; with the
 *                   reference's serial loop this reproduces its output
 *                   bit for bit.
 *   ORA_RNG_PHILOX  one Philox4x32-10 stream per particle, the scheme the
 *                   device engine uses; results do not depend on the order
 *                   particles are processed in, so a device run and this
 *                   run follow the same histories.
 *   ORA_RNG_TAPE    24-bit integers from a caller-provided tape (per-function
 *                   parity vectors).
 */
#ifndef OMC_ORACLE_H
#define OMC_ORACLE_H

#include "omc_host.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_RNG_RANMAR = 0, ORA_RNG_PHILOX = 1, ORA_RNG_TAPE = 2 };

typedef struct {
    uint32_t k0, k1;    /* Philox key */
    uint32_t ctr;       /* draws consumed */
} ora_stream;

typedef struct {
    int    iq, ir;
    double e, x, y, z, u, v, w, dnear, wt;
    ora_stream rs;
} ora_particle;

typedef struct {
    int64_t n_hist;
    int64_t n_dep;            /* ausgab calls */
    int64_t n_cross_photon;   /* photon region changes */
    int64_t n_cross_charged;  /* charged-particle region changes */
    int64_t n_draws;          /* random numbers consumed */
    /* physics events, counted where the reference calls each routine (the
     * device engine counts the same events in the same places) */
    int64_t n_compton, n_photo, n_pair, n_rayleigh;
    int64_t n_brems, n_moller, n_bhabha, n_annih, n_rannih;
    int64_t n_fict;           /* discrete events rejected by sigf/sig0 */
    int64_t n_msdist, n_sscat;
    int64_t n_howfar_charged; /* howfar calls in electron() */
    int64_t n_photon_steps;   /* iterations of photon()'s transport loop */
    int64_t n_electron_steps; /* iterations of electron()'s ustep loop */
} ora_counters;

typedef struct ora_ctx ora_ctx;

/* Geometry + physics context.  `regions` and `phantom` describe the egsphant
 * voxel geometry; `nsplit` is the photon splitting number (1 = off). */
ora_ctx *ora_create(const omc_tables *t, const omc_regions *regions,
                    const omc_phantom *phantom, int nsplit, int rng_mode);
void     ora_destroy(ora_ctx *c);
/* RANMAR seeds as in the .inp "rng seeds" key (src/omc_random.c:58-119). */
void     ora_seed_ranmar(ora_ctx *c, int ixx, int jxx);
/* Philox seed shared by all histories of a run. */
void     ora_seed_philox(ora_ctx *c, uint32_t s0, uint32_t s1);
void     ora_set_tape(ora_ctx *c, const int32_t *tape, int64_t n);
int64_t  ora_tape_pos(const ora_ctx *c);
/* Next uniform in [0,1) of the active mode (exposed for known-answer tests). */
double   ora_random(ora_ctx *c);
void     ora_get_counters(const ora_ctx *c, ora_counters *out);
/* Optional per-history ledger, 4 entries per history id: energy deposited
 * in the phantom, deposited outside (region 0), carried out by photons, and
 * the path checksum sum(energy * region) over deposits in the phantom. */
void     ora_set_ledger(ora_ctx *c, double *ledger3, int64_t n);

/* omc_dosxyz batch loop (omc_dosxyz.c:1206-1263): nbatch batches of
 * nperbatch histories; after each batch the voxel tally is folded into
 * sum / sum2 ([nreg] each, region 0 included). */
int ora_run_dosxyz(ora_ctx *c, const omc_source_dosxyz *src, int nperbatch, int nbatch,
                   double *sum, double *sum2);
/* Same problem, PHILOX mode, histories spread over `nthreads` OpenMP
 * threads (results independent of nthreads up to FP64 summation order). */
int ora_run_dosxyz_parallel(const omc_tables *t, const omc_regions *r, const omc_phantom *g,
                            const omc_source_dosxyz *src, uint32_t s0, uint32_t s1,
                            int64_t hist0, int nperbatch, int nbatch, int nthreads,
                            double *sum, double *sum2, ora_counters *cnt);

/* matRad beamlet loop for one beamlet (omc_matrad.c:1389-1414): fills
 * per-batch sums like ora_run_dosxyz. */
int ora_run_beamlet(ora_ctx *c, const omc_source_beamlets *src, int ibeamlet,
                    int64_t hist0, int nperbatch, int nbatch, double *sum, double *sum2);

/* Primary particle of history `hid` of a beamlet (initHistory,
 * omc_matrad.c:1084-1254). */
int ora_beamlet_primary(ora_ctx *c, const omc_source_beamlets *src, int ibeamlet, int64_t hid,
                        ora_particle *out);

/* ---- per-function entry points for parity vectors ------------------------ */
/* howfar/hownear on one particle state. */
void   ora_howfar(const ora_ctx *c, const ora_particle *p, int *idisc, int *irnew, double *ustep);
double ora_hownear(const ora_ctx *c, const ora_particle *p);
double ora_compute_drange(const ora_ctx *c, int imed, int iq, int lelke, double ekei,
                          double ekef, double elkei, double elkef);
double ora_compute_eloss(const ora_ctx *c, int imed, int iq, double rhof, double tustep,
                         double range, double eke, double elke, int lelke);
/* msdist on particle p (position/direction updated in *out); returns ustep. */
double ora_msdist(ora_ctx *c, const ora_particle *p, int imed, double rhof, double de,
                  double tustep, double eke, ora_particle *out);
/* Runs one interaction on `p` (kind: 0 compton, 1 pair, 2 photo, 3 rayleigh,
 * 4 brems, 5 moller, 6 bhabha, 7 annih, 8 rannih) and returns the resulting
 * particles (up to 3) in out[]; returns their count. */
int    ora_interaction(ora_ctx *c, int kind, const ora_particle *p, ora_particle *out);
/* Full shower of one prepared primary; deposits go to `endep` ([nreg]). */
int    ora_shower_one(ora_ctx *c, const ora_particle *primary, double *endep);

/* Philox4x32-10 block (for RNG known-answer tests). */
void   ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
