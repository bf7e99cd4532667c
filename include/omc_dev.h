/*
 * omc_dev.h - C ABI of libomc_hip.so, the MI355X (gfx950) transport engine.
 *
 * This is the drop-in boundary for ompMC's shower() path.  In ompMC the user
 * code drives transport one history at a time:
 *
 *     for each batch:                                   omc_dosxyz.c:1237-1263
 *         #pragma omp parallel for                      omc_matrad.c:1389-1414
 *         for each history: initHistory(); shower();    ompmc.c:5436-5447
 *         accumEndep();                                 omc_dosxyz.c:696-717
 *
 * with shower() calling back into howfar/hownear/ausgab (ompmc.h:37-39).
 * A GPU cannot call host callbacks, so the boundary moves up one level: the
 * engine takes the whole batch loop -- source sampling, transport in egsphant
 * voxel geometry, scoring -- and returns the batch sums accumEndep would have
 * accumulated.  Host code (omc_host.h) keeps building tables, phantoms and
 * sources exactly as ompMC does, and keeps the statistics and output files.
 *
 * Conventions: every call returns OMC_OK (0) or a negative OMC_E* code from
 * omc_host.h and never exits; omc_dev_last_error() describes the failure.
 * Inputs are copied (the caller keeps ownership); outputs go to caller
 * allocated host arrays.  One host thread per context.
 */
#ifndef OMC_DEV_H
#define OMC_DEV_H

#include "omc_host.h"
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct omc_dev_ctx omc_dev_ctx;

#define OMC_MAX_NSPLIT 1023   /* photon splitting factor the device supports */

/* Entries per history of the optional per-history ledger: energy deposited
 * in the phantom, energy deposited in region 0 (outside), energy carried
 * out by photons, and a path checksum (the sum over deposits in the phantom
 * of energy times region index), which differs as soon as a history puts
 * any energy into a different voxel. */
#define OMC_LEDGER 4

enum { OMC_PREC_FP64 = 0,   /* all transport in FP64 (plain GPU kernel) */
       OMC_PREC_FP32 = 1 }; /* FP32 transport, FP64 tallies and CSDA inversion */

typedef struct {
    int     device;          /* HIP device ordinal */
    int     precision;       /* OMC_PREC_FP64 / OMC_PREC_FP32 */
    int64_t hbm_budget;      /* bytes the context may hold; 0 = 90% of free HBM */
    int64_t chunk;           /* live particles per pipeline before new primaries
                                are held back; 0 = automatic */
    int     pipelines;       /* concurrent transport pipelines (streams),
                                at most 16; 0 = automatic (6; the autotuner
                                tries up to 12).  Each stream, and the context's
                                helper stream, needs a hardware queue of its
                                own: set GPU_MAX_HW_QUEUES >= pipelines + 1
                                before HIP starts */
    int     yield_trips;     /* transport loop trips a particle may take in
                                one launch before it continues in the next
                                (its lane state travels with it, so results
                                do not depend on this); 0 = automatic */
    int     drain_lanes;     /* once a launch's input is used up, a wavefront
                                with at most this many live lanes hands its
                                particles to the next launch (exactly, as a
                                yield); 0 = automatic, < 0 = never */
    int     drain_trips;     /* ...after they have taken this many trips in
                                the launch; 0 = automatic */
    int     event_div;       /* a photon wavefront runs its waiting
                                interactions once 1/event_div of its live
                                lanes wait for one; 0 = automatic */
    int     fetch_min;       /* free lanes of a wavefront take new particles
                                together once this many are free (or none
                                is live); 0 = automatic (4) */
    int64_t drain_min;       /* launches whose input holds fewer particles
                                than this do not drain: on a device they
                                cannot fill, repacking stragglers gains
                                nothing and costs a launch per repack;
                                0 = automatic (the lanes the device holds
                                resident), < 0 = every launch drains */
    int     woodcock;        /* photon flights by delta (Woodcock) tracking
                                where a phantom allows it (uniform voxel
                                grid, one photon energy grid for its media):
                                interaction sites come from the same
                                distribution as with voxel stepping, but
                                histories are not the reference's step for
                                step.  0 = automatic (on for the FP32
                                engine, off for FP64, which keeps its
                                history-by-history parity), > 0 on, < 0 off */
    int64_t tally_bytes;     /* largest batch tally ([nbatch][nreg] FP64) a
                                job holds whole.  A larger one is admitted
                                batch by batch: each batch gets one grid,
                                folded into the job's sum and sum of squares
                                grids as soon as its last particle is gone
                                (the batch method of accumEndep, batches
                                summed in the order they end).  Batches in
                                flight: as many as the limit leaves room
                                for, but at least two and enough histories
                                for half of pipelines x chunk, and at most
                                16; no split where that saves no grid.  Results equal the whole-tally ones to
                                FP64 summation order.  0 = 512 MiB, < 0 =
                                never split */
    int64_t sort_min;        /* charged-particle launches whose input holds at
                                least this many particles take them in order
                                of falling energy (a counting sort of the
                                queue's records into 64 log-energy buckets
                                before the launch): lanes of a wavefront
                                start on similar particles and the longest
                                tracks start first.  The histories do not
                                depend on it.  0 = automatic (launches
                                larger than the lanes the device holds
                                resident), < 0 = never */
} omc_dev_opts;

/* Device counters of one run (all sums over the run). */
typedef struct {
    int64_t histories;
    int64_t photon_steps;       /* photon voxel steps (howfar calls) */
    int64_t electron_steps;     /* electron ustep iterations */
    int64_t deposits;           /* ausgab calls */
    int64_t photons_pushed, electrons_pushed;
    int64_t generations;        /* photon+electron launch pairs, all pipelines */
    int64_t injections;         /* source launches (primaries handed to a pipeline) */
    double  transport_ms;       /* wall time of the run on the engine clock */
    double  photon_kernel_ms, electron_kernel_ms, source_kernel_ms;   /* sums of
                                   launch durations; pipelines overlap, so these
                                   may add up to more than transport_ms */
    int64_t photon_launches, electron_launches;
    int64_t killed;             /* particles stopped by the step-limit watchdog */
    int64_t photon_crossings, electron_crossings;   /* voxel boundaries crossed */
    int64_t photon_deposits, electron_deposits;     /* ausgab calls per kernel */
    int64_t yields;             /* particles handed on to a later launch */
    double  host_ms;            /* scheduler thread busy (admission, launches,
                                   generation ends), not waiting */
    double  pipe_idle_ms;       /* summed over pipelines: device time between a
                                   generation's end and the next one's start,
                                   within the run */
    int64_t woodcock_over;      /* delta-tracking sites whose attenuation
                                   exceeded the majorant (0 unless the
                                   majorant is wrong) */
    /* Physics events of the run, counted where the reference calls the
     * routine named: with the same histories (FP64 engine, voxel-stepped
     * photons) each equals the count of the CPU restatement exactly. */
    int64_t compton, photo, pair, rayleigh;          /* ompmc.c:2046-2065 */
    int64_t brems, moller, bhabha;                   /* ompmc.c:5375-5429 */
    int64_t annih_flight;       /* annih(), positron in flight */
    int64_t annih_rest;         /* rannih(): positron cut off or leaving */
    int64_t fict_rejections;    /* discrete events rejected by sigf/sig0
                                   (ompmc.c:5354-5372) */
    int64_t msdist;             /* condensed-history steps (msdist calls) */
    int64_t sscat;              /* single elastic scatterings (sscat calls) */
    int64_t howfar_charged;     /* howfar calls for charged particles */
    /* Loop trips of the longest photon and charged-particle track of the
     * run (one trip: a voxel step or flight, an interaction, one
     * condensed-history step).  A launch lasts as long as its longest
     * track, so these bound a small job's latency. */
    int64_t longest_photon_trips, longest_electron_trips;
} omc_dev_stats;

/* Context: device, pipelines (each a stream with its own particle queues),
 * HBM budget, precision.  Replaces the per-thread initStack/initRandom
 * set-up (ompmc.c:47-64, omc_random.c:58-119). */
int  omc_dev_create(omc_dev_ctx **ctx, const omc_dev_opts *opts);
int  omc_dev_destroy(omc_dev_ctx *ctx);
const char *omc_dev_last_error(void);
int  omc_dev_precision(const omc_dev_ctx *ctx);

/* Deep copy of every table shower() reads (photon_data, rayleigh_data,
 * pair_data, electron_data, spin_data, mscat_data, pegs_data;
 * ompmc.h:104-401) into the device layout.  Replaces the global tables that
 * initMediaData fills (ompmc.c:5450-5487). */
int  omc_dev_upload_media(omc_dev_ctx *ctx, const omc_tables *t);

/* Resident voxel phantom and its region data (omc_dosxyz.c:62-175,
 * 890-962).  Several phantoms may be resident; each gets an id. */
int  omc_dev_upload_phantom(omc_dev_ctx *ctx, const omc_phantom *ph,
                            const omc_regions *reg, int *phantom_id);
int  omc_dev_release_phantom(omc_dev_ctx *ctx, int phantom_id);

typedef struct {
    int      phantom_id;
    int      nsplit;            /* photon splitting (vrt.nsplit, ompmc.c:1903-2118,
                                   roulette 4149-4164, 4338-4353, 4627-4642):
                                   1 = off, at most OMC_MAX_NSPLIT */
    uint32_t seed[2];           /* run seed (e.g. the .inp "rng seeds") */
    int64_t  hist0;             /* index of the first history (stream id) */
    int64_t  nperbatch;         /* histories per statistical batch */
    int      nbatch;            /* number of batches */
    double  *ledger;            /* optional [OMC_LEDGER*nperbatch*nbatch]:
                                   per history the entries OMC_LEDGER names */
    omc_dev_stats *stats;       /* optional */
} omc_dev_run_params;

/* The omc_dosxyz batch loop on the device: nbatch x nperbatch histories of
 * initHistory (omc_dosxyz.c:964-1068) + shower().  On return sum[i] and
 * sum2[i] ([nreg] host arrays, region 0 included) have been increased by
 * the per-batch energy and its square, exactly what accumEndep adds
 * (omc_dosxyz.c:696-717), so omc_batch_results() gives the .3ddose values. */
int  omc_dev_run_dosxyz(omc_dev_ctx *ctx, const omc_source_dosxyz *src,
                        const omc_dev_run_params *rp, double *sum, double *sum2);

/* The same batch loop for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): no counterpart in the reference.  The job's batch tallies and
 * spectrum live in a static device array of OMC_NOALLOC_SCRATCH_BYTES per
 * GPU instead of memory the call allocates; the context, its particle
 * queues, tables and phantoms are set up beforehand as usual, and this call
 * makes no hipMalloc, hipFree or other allocating device call.  The tallies
 * must fit: 8 * nreg * nbatch bytes plus 16 KiB, else OMC_ERANGE (the C2
 * slab at 10 batches takes 21 MB).  No ledger (rp->ledger must be NULL), and
 * no other work may be queued in the context.  Every call on one GPU uses
 * that GPU's one array, so calls that overlap on a GPU (several contexts on
 * it, from several threads) race; each call clears the tallies it uses. */
#define OMC_NOALLOC_SCRATCH_BYTES (64u << 20)
int  omc_dev_run_dosxyz_noalloc(omc_dev_ctx *ctx, const omc_source_dosxyz *src,
                                const omc_dev_run_params *rp, double *sum, double *sum2);

/* matRad dose influence: every beamlet of `src` in one device pass instead
 * of the serial beamlet loop (omc_matrad.c:1389-1414).  Beamlet b runs
 * nperbatch_b * rp->nbatch histories with nperbatch_b = nhist[b] / nbatch
 * (1 if that is 0, as omc_matrad.c:1332-1337 does for its single count);
 * its histories take stream indices after those of beamlets 0..b-1,
 * starting at rp->hist0.  rp->nperbatch is not used.
 * sum [b*nreg + i] and sum2[b*nreg + i] are increased by the batch sums,
 * as omc_dev_run_dosxyz does for one field.  rp->ledger, if set, receives
 * OMC_LEDGER entries per history of the run, beamlet after beamlet. */
int  omc_dev_run_beamlets(omc_dev_ctx *ctx, const omc_source_beamlets *src,
                          const int64_t *nhist, const omc_dev_run_params *rp,
                          double *sum, double *sum2);

/* Same beamlet pass, with the matRad Dij assembled on the device: for each
 * beamlet the dose per incident history (accumulateResults with iout = 1,
 * omc_matrad.c:839-918, normalised by the beamlet's own history count)
 * above rel_thr * beamlet max becomes one CSC column, rows in voxel order
 * (omc_matrad.c:1416-1479).  jcs has nbeamlets+1 entries; irs/vals have
 * `cap` entries; *nnz receives the total.  When it exceeds cap the call
 * returns OMC_ERANGE, writes nothing and keeps the matrix in the context:
 * a call with src == NULL (nhist and rp unused) then copies the kept
 * matrix into larger arrays without running the beamlets again. */
int  omc_dev_beamlets_sparse(omc_dev_ctx *ctx, const omc_source_beamlets *src,
                             const int64_t *nhist, const omc_dev_run_params *rp,
                             double rel_thr, int64_t *jcs, int64_t *irs,
                             double *vals, int64_t cap, int64_t *nnz);

/* Autotuner for one workload (no counterpart in the reference, whose only
 * schedule knob is the OpenMP thread count).  Times the omc_dosxyz batch loop
 * of `src` on `probe` histories (rp->nbatch batches; rp->nperbatch, ledger
 * and stats unused) under each setting of one option at a time, from the
 * context's current settings (coordinate descent, one pass):
 *   pipelines  2, 4, 6, 8, 10, 12 (at most GPU_MAX_HW_QUEUES - 1 when that is set)
 *   chunk      1M, 2M, 4M, 8M live particles per pipeline (within the budget)
 *   yield_trips 1024, 4096, 16384;  drain_lanes 16, 32, 48;  drain_trips 8, 16, 32
 *   event_div  1, 2, 4;  fetch_min 1, 4, 8, 16
 * Each setting is the better of two runs.  The histories, and so the
 * results, do not depend on any of these.  The context keeps the fastest
 * settings; *best receives them (device, precision and budget as they were)
 * and *hps the histories/s measured for them.  No work may be queued. */
int  omc_dev_tune(omc_dev_ctx *ctx, const omc_source_dosxyz *src, const omc_dev_run_params *rp,
                  int64_t probe, omc_dev_opts *best, double *hps);

/* The same autotuner on a matRad beamlet batch: times omc_dev_run_beamlets of
 * `src` with nhist[b] histories per beamlet (the caller's probe, e.g. the
 * batch itself or a subset) under each setting; no dose is copied back. */
int  omc_dev_tune_beamlets(omc_dev_ctx *ctx, const omc_source_beamlets *src, const int64_t *nhist,
                           const omc_dev_run_params *rp, omc_dev_opts *best, double *hps);

/* Diagnostic, no counterpart in the reference: evaluates the elementary
 * functions the transport kernels use -- fn 0 log, 1 exp, 2 sqrt, 3 sin,
 * 4 cos, 5 pow(x, y) -- on the device in the given precision, exactly as
 * the kernels of that precision compile them, for n arguments (y only for
 * pow; inputs are rounded to FP32 for OMC_PREC_FP32).  Comparing the results
 * with the host C library names the functions through which a device
 * history can depart from the CPU one.  Allocates its own device memory. */
int  omc_dev_math_eval(int device, int precision, int fn, int64_t n, const double *x, const double *y,
                       double *out);

/* initPhotonData on the GPU (src/ompmc.c:216-665, the part omc_build_photon
 * restates on the host): the photon mean free path, branching ratios and
 * Rayleigh ratio of every medium of t (gmfp0/1, gbr10/11, gbr20/21,
 * cohe0/1, and ge0/ge1) from the XCOM element data in data_folder, written
 * into t's arrays in omc_build_photon's layout.  t's pegs4 fields (ap, up,
 * rho, ne, elements; omc_read_pegs) must be filled and its photon arrays
 * allocated.  The device evaluates the host builder's arithmetic in the
 * same order; an entry differs from the host's only where the device's exp
 * or log rounds differently from glibc's (a few ulps at most). */
int  omc_dev_build_photon(int device, omc_tables *t, const char *data_folder);

/* ---- serving: continuous batching of dose requests ----------------------- */
typedef struct {
    int      phantom_id;
    int      kind;              /* 0 = dosxyz field (src_dosxyz), 1 = beamlet */
    const omc_source_dosxyz   *src_dosxyz;
    const omc_source_beamlets *src_beamlets;
    int      beamlet;           /* for kind 1 */
    int64_t  nhist;             /* histories (split into nbatch groups) */
    int      nbatch;
    uint32_t seed[2];
} omc_dev_request;

typedef struct {
    int     ticket;
    int     done;               /* 1 when finished (or failed) */
    int     status;             /* OMC_OK, or the error that ended it */
    double  submit_ms, admit_ms, finish_ms;   /* engine clock */
    int64_t nhist;              /* histories run: nperbatch * nbatch */
} omc_dev_result;

/* Queues a request (its source is copied); it is admitted once its dose
 * grid fits the HBM budget next to the resident tables, phantoms and
 * particle queues.  Histories are nhist / nbatch per batch (at least 1).
 * Returns a ticket >= 0. */
int  omc_dev_submit(omc_dev_ctx *ctx, const omc_dev_request *req);
/* Advances the engine: admits waiting requests whose memory fits, starts a
 * transport generation on every idle pipeline that has particles or
 * admitted histories (new primaries of all admitted requests share each
 * generation), waits for one generation to end and finishes the requests
 * whose last particle is gone.  Returns the number of submitted requests not
 * yet finished (0 when idle), or a negative error. */
int  omc_dev_step(omc_dev_ctx *ctx);
/* Result of a ticket.  When done, sum/sum2 ([nreg] each, may be NULL) get
 * the request's batch sums (set, not added) and the ticket is released. */
int  omc_dev_poll(omc_dev_ctx *ctx, int ticket, omc_dev_result *res,
                  double *sum, double *sum2);

#ifdef __cplusplus
}
#endif
#endif /* OMC_DEV_H */
