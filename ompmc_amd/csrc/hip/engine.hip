// Host side of the device engine: the C ABI declared in include/omc_dev.h.
//
// A context owns the uploaded tables, resident phantoms and a few transport
// pipelines.  A pipeline is one HIP stream with four particle queues
// (photons and charged particles, current and next generation); it runs one
// generation at a time -- source (new primaries), photon kernel, charged
// particle kernel -- and the pipelines run concurrently, so the tail of one
// generation (its slowest particle) overlaps the bulk of the others.
//
// Work arrives as jobs: one dosxyz field, one beamlet or one serving
// request, each with its own FP64 tally [nbatch][nreg] in HBM.  Jobs are
// admitted in order when their memory fits the HBM budget; their histories
// are handed to the pipelines as they drain (round robin over the admitted
// jobs), and a job is finished once every history has been injected and no
// pipeline holds one of its particles.  Every particle carries its own
// Philox stream, so the schedule never changes a history.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "omc_dev.h"
#include "launch.h"
#include "transport.h"

using namespace omc;

// Device memory of omc_dev_run_dosxyz_noalloc: the batch tallies and the
// spectrum of its one job live here instead of in memory the call
// allocates.  Every GPU the library runs on holds its own copy for the life
// of the process.  In host code `scratch` names a host-side stand-in, not
// the device address: the address comes from hipGetSymbolAddress, once per
// device (scratch_on), and only that address is handed to kernels and
// copies.
static __device__ __align__(256) unsigned char scratch[OMC_NOALLOC_SCRATCH_BYTES];

namespace {

char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define OMC_TRY_DEV(x) do { int r__ = (x); if (r__ < 0) return r__; } while (0)

#define HIPCHK(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(OMC_EDEVICE, "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

constexpr int MAX_JOBS = 4096;
constexpr int MAX_SEGS = 512;
constexpr int MAX_PIPES = 16;
// Pipelines x chunk on the 64^3 slab phantom with the 6 MV spectrum (steps
// of 1e8 histories, 16 hardware queues): the FP32 engine gains from
// overlapping pipelines up to 10 (histories/s at 4M live particles each: 6
// pipelines 2.15e8, 7 2.18e8, 8 2.28e8, 10 2.38e8, 12 2.38e8), while the
// FP64 engine, whose charged-particle kernel holds 2 waves per SIMD, is flat
// from 6 to 12 (1.07e8-1.09e8).  The default stays at 6 for both: the
// serving loop has been exercised at 6, and the autotuner reaches 10-12
// where a workload gains.  Each pipeline stream and the helper stream need
// a hardware queue of their own (GPU_MAX_HW_QUEUES=16, which the drivers
// set); pipelines that share a queue serialise.
constexpr int DEFAULT_PIPES = 6;
constexpr int64_t DEFAULT_CHUNK = 1 << 22;
// Batches of one work admitted batch by batch that may be in flight at once
// (each is a job with its own grid).  Runs have held up to 13; a serving
// request of about 2000 small batches in flight faulted the device with a
// cause not yet found, so the count is held near the range that has run.
constexpr int64_t MAX_SPLIT_INFLIGHT = 16;
constexpr uint32_t DEFAULT_YIELD_TRIPS = 4096;
// Drain compaction: once a launch's input is used up, a wavefront with at
// most 32 live lanes hands its particles (after 16 trips in the launch) to
// the next launch, which packs them into full wavefronts (slab sweep: FP32
// 38 -> 55M histories/s, FP64 37 -> 42M).
constexpr uint32_t DEFAULT_DRAIN_LANES = 32;
constexpr uint32_t DEFAULT_DRAIN_TRIPS = 16;
// Repacking only pays while the device is full: a launch whose input is
// smaller than the lanes the device holds resident (CUs x 4 SIMDs x the
// electron kernel's waves per SIMD x 64) runs without draining.  A lone
// 2e4-history job otherwise takes ~40 short generations, its stragglers
// handed on every 16 trips.
constexpr int64_t drain_min_default(int cus, int precision)
{
    return (int64_t)cus * 4 * (precision == OMC_PREC_FP64 ? 2 : 4) * 64;
}
// A photon wavefront runs its waiting events (interaction, re-entry) once
// half its live lanes wait for one (slab sweep of 2/4/8/16: 2 first, 22%
// more histories/s than running each event as it comes).  Free lanes take
// new particles as soon as they are free: with delta-tracked photon flights
// the FP32 slab ran 133 / 140 / 144 / 145 M histories/s at fetch thresholds
// of 16 / 4 / 8 / 1 lanes (FP64 and the beamlet batch within noise).
constexpr uint32_t DEFAULT_EVENT_DIV = 2;
// Whole batch tallies up to 512 MiB per job (the C3 beamlet's 10 grids of the
// 128x128x96 phantom take 126 MB); a C5-sized job (84 MB per grid, 100
// batches) is admitted batch by batch, its batches in flight beside its sum
// grids.
constexpr int64_t DEFAULT_TALLY_CAP = (int64_t)512 << 20;
// Free lanes of a wavefront take new particles once this many are free (or
// none is live).  Slab, steps of 1e8 histories, FP32 / FP64 histories/s:
// 1 2.20e8 / 1.07e8, 2 2.23e8 / 1.08e8, 4 2.29e8 / 1.10e8, 8 2.26e8 / 1.10e8.
constexpr uint32_t DEFAULT_FETCH_MIN = 4;

struct PhantomSlot {
    bool used = false;
    int nreg = 0;
    int isize = 0, jsize = 0, ksize = 0;
    DevBuf xb, yb, zb, med, rhof, rho;
    // delta tracking: largest relative density of each medium (-1 = absent)
    // and the grid, if uniform
    double rho_max[MAXMED];
    bool uniform = false;
    double x0 = 0, y0 = 0, z0 = 0, dx = 0, dy = 0, dz = 0;
    DevBuf wc;                    // majorant [MXGE]
};

// Host mirror of a pipeline's small device state, in pinned memory so the
// end-of-generation copies are asynchronous.
struct PipeMirror {
    uint32_t counts[8];
    DevStatus status;
    uint32_t flags[MAX_JOBS];
    SourceSeg segs[MAX_SEGS];
};

struct Pipe {
    hipStream_t stream = nullptr;
    void *q[4] = {};              // particle queues: photon cur/next, charged cur/next
    uint32_t *d_counts = nullptr; // 4 tails + 2 fetch counters
    DevStatus *d_status = nullptr;
    uint32_t *d_flags = nullptr;  // [MAX_JOBS]
    SourceSeg *d_segs = nullptr;  // [MAX_SEGS]
    void *d_sorted = nullptr;     // [qcap] records: a charged-particle launch's input by falling energy
    uint32_t *d_sort = nullptr;   // [SORT_BUCKETS * SORT_BLOCKS] its histograms / offsets
    PipeMirror *h = nullptr;
    hipEvent_t ev_src0 = nullptr, ev_src1 = nullptr, ev_ph0 = nullptr, ev_ph1 = nullptr;
    hipEvent_t ev_el0 = nullptr, ev_el1 = nullptr, ev_begin = nullptr;
    hipEvent_t ev_end[2] = { nullptr, nullptr };   // end of generation gen (slot gen & 1)
    uint64_t gen = 0;             // generations launched on this pipeline
    uint64_t idle_from = 0;       // first generation of the current run (idle counts after it)
    int cur = 0;
    uint32_t nP = 0, nE = 0;      // particles waiting in the current queues
    bool inflight = false;
    bool had_src = false, had_ph = false, had_el = false;
    uint32_t injected = 0;        // primaries injected by the generation in flight
    uint32_t in_e = 0;            // charged particles in its queue at launch (known to the host)
    std::vector<int> open;        // jobs that may hold particles here
    DevStatus base;               // status at the start of the current run
};

enum WorkState { W_WAITING = 0, W_RUNNING = 1, W_DONE = 2, W_FAILED = 3 };

struct RunSink;

// An admission unit: a source plus its histories.  Serving requests and the
// beamlets / field of a synchronous run are all Works.
struct Work {
    int ticket = -1;              // serving ticket, or -1 for a synchronous run
    RunSink *sink = nullptr;      // synchronous run receiving the result
    int index = 0;                // beamlet index within the run
    int phantom = 0;
    DevJob tmpl;                  // source and stream fields; memory filled on admission
    std::vector<double> cdf;      // [2*OMC_INVDIM] inverse CDF
    int64_t nhist = 0, nperbatch = 0;
    int nbatch = 1;
    int slots = 1;                // tally grids: nbatch, or 1 when only the mean is wanted
    int64_t nreg = 0;
    bool ledger = false;
    int64_t ledger_off = 0;       // first history of this work in the run's ledger
    bool scratch = false;         // tally and spectrum go to the device's scratch array
    // Admitted batch by batch (the whole tally would exceed the context's
    // tally_bytes): at most split_k batches in flight, each its own job with
    // one grid, folded into acc = [sum | sum of squares] when it ends.
    int split_k = 0;
    int next_batch = 0, batches_live = 0, batches_done = 0;
    double *acc = nullptr;
    int state = W_WAITING;
    int status = OMC_OK;
    int job = -1;
    double submit_ms = 0, admit_ms = 0, finish_ms = 0;
    std::vector<double> sum, sum2;
};

// Where a synchronous run wants its results.
struct RunSink {
    bool beamlets = false;
    double *sum = nullptr, *sum2 = nullptr;   // [nreg] or [nb*nreg]
    double *ledger = nullptr;                 // [3 * histories of the run]
    bool sparse = false;
    double rel_thr = 0.0;
    std::vector<std::vector<int64_t>> rows;
    std::vector<std::vector<double>> vals;
    int remaining = 0;
    int rc = OMC_OK;
};

// Stages of collecting a finished job's results on the helper stream.
enum { FIN_NONE = 0, FIN_QUEUED, FIN_FOLD, FIN_COLUMN };

struct JobRec {
    bool used = false;
    bool borrowed = false;        // tally and spectrum in the device's scratch array
    Work *w = nullptr;
    int batch = -1;               // batch of a work admitted batch by batch (-1: the whole work)
    int64_t nhist = 0;            // histories of this job
    int64_t injected = 0;
    uint32_t pipes = 0;           // pipelines that may hold its particles
    int64_t quantum = 0;          // histories per turn in the round robin
    double *tally = nullptr;
    size_t tally_bytes = 0;
    double *cdf = nullptr;
    double *ledger = nullptr;
    size_t ledger_bytes = 0;
    int fin = FIN_NONE;           // result collection, see pump_finish
    uint32_t nnz = 0;             // sparse column length, once known
};

}  // namespace

struct omc_dev_ctx {
    int device = 0;
    int precision = OMC_PREC_FP64;
    hipStream_t aux = nullptr;    // admission and finishing work
    int64_t budget = 0;
    int64_t used = 0;
    void *scratch = nullptr;          // this device's scratch array (device address)
    int64_t tally_cap = 0;            // largest whole batch tally of a job (bytes; 0 = no limit)
    int64_t sort_min = 0;             // charged-particle launches of at least this many are energy-ordered (0 = never)
    int64_t chunk = 0;
    uint32_t qcap = 0;
    uint32_t yield_trips = 0;
    uint32_t drain_lanes = 0, drain_trips = 0;
    int64_t drain_min = 0;            // smallest launch input that drains
    uint32_t event_div = 0, fetch_min = 0;
    int cus = 256;
    std::chrono::steady_clock::time_point t0;

    // tables
    bool media = false;
    int nmed = 0;
    bool woodcock = false;            // photon flights by delta tracking
    // host copy of the photon attenuation tables, for the majorants:
    // {gmfp1, gmfp0, cohe1, cohe0} per [imed*MXGE + bin]
    std::vector<double> h_mfp;
    double h_ge0[MAXMED] = {}, h_ge1[MAXMED] = {};
    void *small = nullptr, *ph_mfp = nullptr, *ph_br = nullptr, *ph_pmax = nullptr;
    void *el = nullptr, *csda = nullptr, *spin = nullptr, *ms = nullptr;
    size_t tab_bytes[8] = {};         // sizes of the eight table buffers above, in that order
    // phantoms
    PhantomSlot ph[MAX_PHANTOMS];
    void *d_phantoms = nullptr;       // DevPhantom<R>[MAX_PHANTOMS]
    // finishing scratch (fold output + sparse column), sized for the largest phantom
    int64_t fin_nvox = -1;
    double *fin_sums = nullptr;       // [2*nreg]
    SparseScratch sp = {};
    size_t fin_bytes = 0;
    // pinned host side of the finishing copies, same sizes
    double *h_sums = nullptr;         // [2*nreg]
    int64_t *h_rows = nullptr;        // [nvox]
    double *h_vals = nullptr;         // [nvox]
    uint32_t *h_nnz = nullptr;
    std::deque<int> finishing;        // jobs whose results are being collected, oldest first
    hipEvent_t ev_fin = nullptr;      // end of the head job's current finishing stage
    // jobs
    DevJob *d_jobs = nullptr;
    JobRec jobs[MAX_JOBS];
    int jobs_hi = 0;                  // one past the highest used slot
    std::deque<int> runnable;         // admitted jobs with histories left to inject
    std::deque<Work *> waiting;       // works not yet admitted
    std::multimap<size_t, void *> pool;   // freed tallies kept for reuse
    // pipelines
    std::vector<Pipe> pipes;
    omc_dev_stats acc;                // accumulated over the current run
    // serving
    std::map<int, Work *> tickets;
    int next_ticket = 0;
    int unfinished = 0;               // submitted tickets not yet done
    // last sparse Dij that did not fit the caller's arrays
    std::vector<std::vector<int64_t>> kept_rows;
    std::vector<std::vector<double>> kept_vals;
};

namespace {

double now_ms(const omc_dev_ctx *c)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c->t0).count();
}

int dalloc(omc_dev_ctx *c, void **p, size_t bytes)
{
    if (c->budget > 0 && c->used + (int64_t)bytes > c->budget)
        return fail(OMC_EBUDGET, "HBM budget exceeded: %lld + %zu > %lld", (long long)c->used, bytes,
                    (long long)c->budget);
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(OMC_EDEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    c->used += (int64_t)bytes;
    return OMC_OK;
}

void dfree(omc_dev_ctx *c, void *p, size_t bytes)
{
    if (!p) return;
    (void)hipFree(p);
    c->used -= (int64_t)bytes;
}

// Tallies come and go with jobs; buffers of a size seen before are reused
// instead of returned, and returned only when the budget needs the room.
int pool_alloc(omc_dev_ctx *c, void **p, size_t bytes)
{
    auto it = c->pool.find(bytes);
    if (it != c->pool.end()) {
        *p = it->second;
        c->pool.erase(it);
        return OMC_OK;
    }
    while (c->budget > 0 && c->used + (int64_t)bytes > c->budget && !c->pool.empty()) {
        auto big = std::prev(c->pool.end());
        dfree(c, big->second, big->first);
        c->pool.erase(big);
    }
    return dalloc(c, p, bytes);
}

void pool_free(omc_dev_ctx *c, void *p, size_t bytes)
{
    if (p) c->pool.emplace(bytes, p);
}

// per queue slot: 8 reals, ir/meta/hist/k0, k1+ctr, 6 resume reals, sflags, trips
template <typename R> size_t particle_bytes() { return QR_N * sizeof(R) + QU_N * sizeof(uint32_t); }

template <typename R> Queue<R> queue_view(const omc_dev_ctx *c, const Pipe &P, int k)
{
    Queue<R> q;
    q.base = (char *)P.q[k];
    q.count = P.d_counts + k;
    q.cap = c->qcap;
    return q;
}

template <typename R> int alloc_pipe(omc_dev_ctx *c, Pipe &P)
{
    if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess) return fail(OMC_EDEVICE, "stream");
    hipEvent_t *evs[] = { &P.ev_src0, &P.ev_src1, &P.ev_ph0, &P.ev_ph1, &P.ev_el0, &P.ev_el1, &P.ev_begin,
                          &P.ev_end[0], &P.ev_end[1] };
    for (auto *e : evs) HIPCHK(hipEventCreate(e));
    for (int k = 0; k < 4; ++k) OMC_TRY_DEV(dalloc(c, &P.q[k], particle_bytes<R>() * (size_t)c->qcap));
    OMC_TRY_DEV(dalloc(c, (void **)&P.d_counts, 8 * sizeof(uint32_t)));
    OMC_TRY_DEV(dalloc(c, (void **)&P.d_status, sizeof(DevStatus)));
    OMC_TRY_DEV(dalloc(c, (void **)&P.d_flags, MAX_JOBS * sizeof(uint32_t)));
    OMC_TRY_DEV(dalloc(c, (void **)&P.d_segs, MAX_SEGS * sizeof(SourceSeg)));
    OMC_TRY_DEV(dalloc(c, &P.d_sorted, particle_bytes<R>() * (size_t)c->qcap));
    OMC_TRY_DEV(dalloc(c, (void **)&P.d_sort, sizeof(uint32_t) * SORT_BUCKETS * SORT_BLOCKS));
    HIPCHK(hipHostMalloc((void **)&P.h, sizeof(PipeMirror), hipHostMallocDefault));
    memset(P.h, 0, sizeof(PipeMirror));
    HIPCHK(hipMemset(P.d_counts, 0, 8 * sizeof(uint32_t)));
    HIPCHK(hipMemset(P.d_status, 0, sizeof(DevStatus)));
    memset(&P.base, 0, sizeof P.base);
    return OMC_OK;
}

void free_pipe(Pipe &P)
{
    if (P.stream) (void)hipStreamSynchronize(P.stream);
    for (auto &p : P.q) if (p) { (void)hipFree(p); p = nullptr; }
    void *ptrs[] = { P.d_counts, P.d_status, P.d_flags, P.d_segs, P.d_sorted, P.d_sort };
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (P.h) (void)hipHostFree(P.h);
    hipEvent_t evs[] = { P.ev_src0, P.ev_src1, P.ev_ph0, P.ev_ph1, P.ev_el0, P.ev_el1, P.ev_begin,
                         P.ev_end[0], P.ev_end[1] };
    for (auto e : evs) if (e) (void)hipEventDestroy(e);
    if (P.stream) (void)hipStreamDestroy(P.stream);
}

}  // namespace

// ---------------------------------------------------------------- tables --
// Majorant of the photon attenuation coefficient over a phantom, per energy
// bin of the energy grid its media share (delta tracking, see photon_flight):
// within a bin the tables are linear in log(E), so gmfp and cohfac take their
// smallest values at an edge of the bin, and mu = rhof / (gmfp * cohfac) is at
// most rho_max / (min gmfp * min cohfac).  The end bins also cover a bin's
// width beyond the grid; a 0.1% margin covers FP32 rounding.  A phantom with
// a non-uniform grid, media on different energy grids or a table that is not
// positive keeps voxel stepping (null majorant).
template <typename R> static int build_woodcock(omc_dev_ctx *c, int slot)
{
    PhantomSlot &P = c->ph[slot];
    if (P.wc.p) { dfree(c, P.wc.p, P.wc.bytes); P.wc.p = nullptr; P.wc.bytes = 0; }
    Woodcock<R> w;
    memset(&w, 0, sizeof w);
    int ref = -1;
    bool ok = P.uniform && c->media;
    for (int m = 0; ok && m < c->nmed; ++m) {
        if (P.rho_max[m] < 0) continue;
        if (ref < 0) ref = m;
        else if (c->h_ge0[m] != c->h_ge0[ref] || c->h_ge1[m] != c->h_ge1[ref]) ok = false;
    }
    std::vector<R> mu(MXGE, (R)1.0E-30);
    if (ok && ref >= 0) {
        const double ge0 = c->h_ge0[ref], ge1 = c->h_ge1[ref];
        if (!(ge1 > 0.0)) ok = false;
        for (int k = 0; ok && k < MXGE; ++k) {
            // bin k holds ge1 * log(E) + ge0 in [k + 1, k + 2) (pwlf_bin)
            double lo = (k + 1 - ge0) / ge1, hi = (k + 2 - ge0) / ge1;
            const double wbin = hi - lo;
            if (k == 0) lo -= wbin;
            if (k == MXGE - 1) hi += wbin;
            double top = 0.0;
            for (int m = 0; m < c->nmed; ++m) {
                if (P.rho_max[m] <= 0) continue;
                const double *t = &c->h_mfp[4 * ((size_t)m * MXGE + k)];
                const double amin = std::min(lo * t[0] + t[1], hi * t[0] + t[1]);
                const double bmin = std::min(lo * t[2] + t[3], hi * t[2] + t[3]);
                if (!(amin > 0.0 && bmin > 0.0)) { ok = false; break; }
                top = std::max(top, P.rho_max[m] / (amin * bmin));
            }
            mu[k] = (R)std::max(1.001 * top, 1.0E-30);
        }
    }
    if (ok) {
        P.wc.bytes = sizeof(R) * MXGE;
        OMC_TRY_DEV(dalloc(c, &P.wc.p, P.wc.bytes));
        HIPCHK(hipMemcpy(P.wc.p, mu.data(), P.wc.bytes, hipMemcpyHostToDevice));
        w.mu = (const R *)P.wc.p;
        w.ge0 = c->h_ge0[ref >= 0 ? ref : 0];
        w.ge1 = c->h_ge1[ref >= 0 ? ref : 0];
        w.x0 = P.x0; w.y0 = P.y0; w.z0 = P.z0;
        w.idx = (R)(1.0 / P.dx); w.idy = (R)(1.0 / P.dy); w.idz = (R)(1.0 / P.dz);
    }
    HIPCHK(hipMemcpy((char *)c->d_phantoms + slot * sizeof(DevPhantom<R>) + offsetof(DevPhantom<R>, wc), &w,
                     sizeof w, hipMemcpyHostToDevice));
    return OMC_OK;
}

template <typename R> static int upload_tables(omc_dev_ctx *c, const omc_tables *t)
{
    const int nmed = t->nmed;
    SmallTables S;
    memset(&S, 0, sizeof S);
    S.nmed = nmed;
    S.b2spin_min = t->b2spin_min; S.dbeta2i = t->dbeta2i; S.espml = t->espml;
    S.dleneri = t->dleneri; S.dqq1i = t->dqq1i; S.dllambi = t->dllambi; S.dqmsi = t->dqmsi;
    for (int j = 0; j < OMC_MXRAYFF; ++j) {
        S.xgrid[j] = t->xgrid[j]; S.fcum[j] = t->fcum[j];
        S.b_array[j] = t->b_array[j]; S.c_array[j] = t->c_array[j];
    }
    for (int j = 0; j < OMC_RAYCDFSIZE; ++j) S.i_array[j] = t->i_array[j];
    for (int m = 0; m < nmed; ++m) {
        MediumScalars &M = S.med[m];
        M.ge0 = t->ge0[m]; M.ge1 = t->ge1[m]; M.eke0 = t->eke0[m]; M.eke1 = t->eke1[m];
        M.blcc = t->blcc[m]; M.xcc = t->xcc[m]; M.esig_e = t->esig_e[m]; M.psig_e = t->psig_e[m];
        M.te = t->te[m]; M.thmoll = t->thmoll[m]; M.ap = t->ap[m]; M.delcm = t->delcm[m];
        M.zbrang = t->zbrang[m]; M.bpar0 = t->bpar0[m]; M.bpar1 = t->bpar1[m];
        M.mono_e = t->sig_ismonotone[0][m]; M.mono_p = t->sig_ismonotone[1][m]; M.meke = t->meke[m];
        for (int l = 0; l < 8; ++l) {
            M.dl1[l] = t->dl1[m * 8 + l]; M.dl2[l] = t->dl2[m * 8 + l]; M.dl3[l] = t->dl3[m * 8 + l];
            M.dl4[l] = t->dl4[m * 8 + l]; M.dl5[l] = t->dl5[m * 8 + l]; M.dl6[l] = t->dl6[m * 8 + l];
        }
    }
    const size_t G = (size_t)nmed * MXGE, K = (size_t)nmed * 2 * MXEKE;
    std::vector<R4<R>> mfp(G), br(G);
    std::vector<R2<R>> pm(G);
    for (size_t i = 0; i < G; ++i) {
        mfp[i] = { (R)t->gmfp1[i], (R)t->gmfp0[i], (R)t->cohe1[i], (R)t->cohe0[i] };
        br[i] = { (R)t->gbr11[i], (R)t->gbr10[i], (R)t->gbr21[i], (R)t->gbr20[i] };
        pm[i] = { (R)t->pmax1[i], (R)t->pmax0[i] };
    }
    std::vector<ElecBin<R>> el(K);
    std::vector<ElecCsda> cs(K);
    for (int m = 0; m < nmed; ++m)
        for (int q = 0; q < 2; ++q)
            for (int k = 0; k < MXEKE; ++k) {
                size_t s = (size_t)m * MXEKE + k;
                ElecBin<R> &b = el[((size_t)m * 2 + q) * MXEKE + k];
                b.sig1 = (R)(q ? t->psig1[s] : t->esig1[s]); b.sig0 = (R)(q ? t->psig0[s] : t->esig0[s]);
                b.tmxs1 = (R)t->tmxs1[s]; b.tmxs0 = (R)t->tmxs0[s];
                b.eta1 = (R)(q ? t->etap_ms1[s] : t->etae_ms1[s]); b.eta0 = (R)(q ? t->etap_ms0[s] : t->etae_ms0[s]);
                b.blc1 = (R)t->blcce1[s]; b.blc0 = (R)t->blcce0[s];
                b.q1c1 = (R)(q ? t->q1cp_ms1[s] : t->q1ce_ms1[s]); b.q1c0 = (R)(q ? t->q1cp_ms0[s] : t->q1ce_ms0[s]);
                b.q2c1 = (R)(q ? t->q2cp_ms1[s] : t->q2ce_ms1[s]); b.q2c0 = (R)(q ? t->q2cp_ms0[s] : t->q2ce_ms0[s]);
                b.br11 = (R)(q ? t->pbr11[s] : t->ebr11[s]); b.br10 = (R)(q ? t->pbr10[s] : t->ebr10[s]);
                b.br21 = (R)(q ? t->pbr21[s] : 0.0); b.br20 = (R)(q ? t->pbr20[s] : 0.0);
                ElecCsda &d = cs[((size_t)m * 2 + q) * MXEKE + k];
                d.dedx1 = q ? t->pdedx1[s] : t->ededx1[s];
                d.dedx0 = q ? t->pdedx0[s] : t->ededx0[s];
                d.range = t->range_ep[(size_t)q * nmed * MXEKE + s];
                d.earr = t->e_array[s];
            }
    std::vector<R> spin((size_t)nmed * OMC_SPIN_PER_MED);
    for (size_t i = 0; i < spin.size(); ++i) spin[i] = (R)t->spin_rej[i];
    std::vector<MsAlias<R>> ms(OMC_MS_SIZE);
    for (int i = 0; i < OMC_MS_SIZE; ++i) {
        const int row = i - i % 32, a = t->ims[i];
        if (a < 0 || a >= 32) return fail(OMC_EFORMAT, "multiple-scattering alias index %d out of its row", a);
        ms[i] = { (R)t->wms[i], (R)t->ums[i], (R)t->ums[row + a], (R)0 };
    }

    struct { void **dst; const void *src; size_t bytes; } up[] = {
        { &c->small, &S, sizeof S },
        { &c->ph_mfp, mfp.data(), mfp.size() * sizeof(mfp[0]) },
        { &c->ph_br, br.data(), br.size() * sizeof(br[0]) },
        { &c->ph_pmax, pm.data(), pm.size() * sizeof(pm[0]) },
        { &c->el, el.data(), el.size() * sizeof(el[0]) },
        { &c->csda, cs.data(), cs.size() * sizeof(cs[0]) },
        { &c->spin, spin.data(), spin.size() * sizeof(spin[0]) },
        { &c->ms, ms.data(), ms.size() * sizeof(ms[0]) },
    };
    for (int b = 0; b < 8; ++b) {
        auto &u = up[b];
        if (*u.dst) { dfree(c, *u.dst, c->tab_bytes[b]); *u.dst = nullptr; c->tab_bytes[b] = 0; }
        OMC_TRY_DEV(dalloc(c, u.dst, u.bytes));
        c->tab_bytes[b] = u.bytes;
        HIPCHK(hipMemcpy(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice));
    }
    c->media = true;
    c->nmed = nmed;
    c->h_mfp.resize(G * 4);
    for (size_t i = 0; i < G; ++i) {
        c->h_mfp[4 * i] = t->gmfp1[i]; c->h_mfp[4 * i + 1] = t->gmfp0[i];
        c->h_mfp[4 * i + 2] = t->cohe1[i]; c->h_mfp[4 * i + 3] = t->cohe0[i];
    }
    for (int m = 0; m < nmed; ++m) { c->h_ge0[m] = t->ge0[m]; c->h_ge1[m] = t->ge1[m]; }
    for (int i = 0; i < MAX_PHANTOMS; ++i)
        if (c->ph[i].used) OMC_TRY_DEV(build_woodcock<R>(c, i));
    return OMC_OK;
}

template <typename R> static DevTables<R> tables_view(omc_dev_ctx *c)
{
    DevTables<R> T;
    T.small = (const SmallTables *)c->small;
    T.ph_mfp = (const R4<R> *)c->ph_mfp;
    T.ph_br = (const R4<R> *)c->ph_br;
    T.ph_pmax = (const R2<R> *)c->ph_pmax;
    T.el = (const ElecBin<R> *)c->el;
    T.csda = (const ElecCsda *)c->csda;
    T.spin = (const R *)c->spin;
    T.ms = (const MsAlias<R> *)c->ms;
    return T;
}

// --------------------------------------------------------------- phantoms --
// Scratch for finishing a job on the device: the folded sums and the sparse
// column of a beamlet, sized for the largest resident phantom.
static int drain_finish(omc_dev_ctx *c);

static int grow_finish_scratch(omc_dev_ctx *c, int64_t nvox)
{
    if (nvox <= c->fin_nvox) return OMC_OK;
    OMC_TRY_DEV(drain_finish(c));
    if (c->fin_sums) dfree(c, c->fin_sums, c->fin_bytes);
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    c->fin_sums = nullptr;
    c->h_sums = nullptr;
    c->fin_nvox = -1;
    const int nblk = sparse_blocks(nvox);
    const size_t sums = 2 * sizeof(double) * (size_t)(nvox + 1);
    const size_t dose = sizeof(double) * (size_t)nvox;
    const size_t blk = sizeof(uint32_t) * ((size_t)nblk + 2);
    const size_t rows = sizeof(int64_t) * (size_t)nvox, vals = sizeof(double) * (size_t)nvox;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    c->fin_bytes = up(sums) + up(dose) + up(8) + up(blk) + up(rows) + up(vals);
    char *p = nullptr;
    OMC_TRY_DEV(dalloc(c, (void **)&p, c->fin_bytes));
    c->fin_sums = (double *)p;              p += up(sums);
    c->sp.dose = (double *)p;               p += up(dose);
    c->sp.maxkey = (unsigned long long *)p; p += up(8);
    c->sp.blk = (uint32_t *)p;              p += up(blk);
    c->sp.rows = (int64_t *)p;              p += up(rows);
    c->sp.vals = (double *)p;
    char *h = nullptr;
    if (hipHostMalloc((void **)&h, up(sums) + up(rows) + up(vals) + up(4), hipHostMallocDefault) != hipSuccess)
        return fail(OMC_EDEVICE, "pinned finishing buffers (%zu bytes)", up(sums) + up(rows) + up(vals) + up(4));
    c->h_sums = (double *)h;                h += up(sums);
    c->h_rows = (int64_t *)h;               h += up(rows);
    c->h_vals = (double *)h;                h += up(vals);
    c->h_nnz = (uint32_t *)h;
    c->fin_nvox = nvox;
    return OMC_OK;
}

template <typename R> static int upload_phantom(omc_dev_ctx *c, const omc_phantom *ph, const omc_regions *reg, int *id)
{
    int slot = -1;
    for (int i = 0; i < MAX_PHANTOMS; ++i) if (!c->ph[i].used) { slot = i; break; }
    if (slot < 0) return fail(OMC_ERANGE, "more than %d resident phantoms", MAX_PHANTOMS);
    if (!c->media) return fail(OMC_EARG, "upload media before phantoms");
    const int nreg = reg->nreg;
    const int64_t nvox = (int64_t)ph->isize * ph->jsize * ph->ksize;
    if ((int64_t)nreg != nvox + 1)
        return fail(OMC_EARG, "regions (%d) do not match phantom %dx%dx%d", nreg, ph->isize, ph->jsize, ph->ksize);
    PhantomSlot &P = c->ph[slot];
    std::vector<uint8_t> med(nreg);
    std::vector<R> rhof(nreg);
    double ecut[MAXMED + 1], pcut[MAXMED + 1];
    for (int m = 0; m <= MAXMED; ++m) { ecut[m] = -1.0; pcut[m] = -1.0; }
    ecut[0] = 0.0; pcut[0] = 0.0;
    for (int i = 0; i < nreg; ++i) {
        int m = reg->med[i];
        if (m >= c->nmed) return fail(OMC_EARG, "region %d uses medium %d, tables have %d", i, m, c->nmed);
        med[i] = m < 0 ? 0xff : (uint8_t)m;
        rhof[i] = (R)reg->rhof[i];
        if (m >= 0) {
            // transport cuts are per medium (region cut = max(global, PEGS))
            if (ecut[m + 1] < 0) { ecut[m + 1] = reg->ecut[i]; pcut[m + 1] = reg->pcut[i]; }
            else if (ecut[m + 1] != reg->ecut[i] || pcut[m + 1] != reg->pcut[i])
                return fail(OMC_ENOTSUP, "region %d: cut-offs differ within medium %d", i, m);
        }
    }
    for (int m = 1; m <= MAXMED; ++m) { if (ecut[m] < 0) ecut[m] = 0.0; if (pcut[m] < 0) pcut[m] = 0.0; }
    for (int m = 0; m < MAXMED; ++m) P.rho_max[m] = -1.0;
    for (int i = 1; i < nreg; ++i)
        if (med[i] != 0xff) P.rho_max[med[i]] = std::max(P.rho_max[med[i]], (double)rhof[i]);
    // a uniform grid lets delta tracking find a voxel from a position
    auto uniform_axis = [](const double *b, int n, double &lo, double &d) {
        lo = b[0];
        d = (b[n] - b[0]) / n;
        if (!(d > 0.0)) return false;
        for (int i = 0; i < n; ++i)
            if (std::fabs((b[i + 1] - b[i]) - d) > 1e-6 * d) return false;
        return true;
    };
    P.uniform = uniform_axis(ph->xbounds, ph->isize, P.x0, P.dx) && uniform_axis(ph->ybounds, ph->jsize, P.y0, P.dy) &&
                uniform_axis(ph->zbounds, ph->ksize, P.z0, P.dz);
    OMC_TRY_DEV(grow_finish_scratch(c, nvox));
    P.nreg = nreg;
    P.isize = ph->isize; P.jsize = ph->jsize; P.ksize = ph->ksize;
    P.xb.bytes = sizeof(double) * (ph->isize + 1);
    P.yb.bytes = sizeof(double) * (ph->jsize + 1);
    P.zb.bytes = sizeof(double) * (ph->ksize + 1);
    P.med.bytes = nreg;
    P.rhof.bytes = sizeof(R) * (size_t)nreg;
    P.rho.bytes = sizeof(double) * (size_t)nvox;
    DevBuf *bufs[] = { &P.xb, &P.yb, &P.zb, &P.med, &P.rhof, &P.rho };
    const void *src[] = { ph->xbounds, ph->ybounds, ph->zbounds, med.data(), rhof.data(), ph->rho };
    for (int b = 0; b < 6; ++b) {
        OMC_TRY_DEV(dalloc(c, &bufs[b]->p, bufs[b]->bytes));
        HIPCHK(hipMemcpy(bufs[b]->p, src[b], bufs[b]->bytes, hipMemcpyHostToDevice));
    }
    DevPhantom<R> d;
    memset(&d, 0, sizeof d);
    d.isize = ph->isize; d.jsize = ph->jsize; d.ksize = ph->ksize; d.nreg = nreg;
    d.xb = (const double *)P.xb.p; d.yb = (const double *)P.yb.p; d.zb = (const double *)P.zb.p;
    d.med = (const uint8_t *)P.med.p; d.rhof = (const R *)P.rhof.p;
    for (int m = 0; m <= MAXMED; ++m) { d.ecut[m] = ecut[m]; d.pcut[m] = pcut[m]; }
    HIPCHK(hipMemcpy((char *)c->d_phantoms + slot * sizeof(DevPhantom<R>), &d, sizeof d, hipMemcpyHostToDevice));
    P.used = true;
    OMC_TRY_DEV(build_woodcock<R>(c, slot));
    *id = slot;
    return OMC_OK;
}

// -------------------------------------------------------------- scheduler --
static void release_job(omc_dev_ctx *c, int j)
{
    JobRec &J = c->jobs[j];
    if (!J.borrowed) {
        pool_free(c, J.tally, J.tally_bytes);
        if (J.cdf) dfree(c, J.cdf, 2 * OMC_INVDIM * sizeof(double));
    }
    if (J.ledger) dfree(c, J.ledger, J.ledger_bytes);
    J = JobRec();
    while (c->jobs_hi > 0 && !c->jobs[c->jobs_hi - 1].used) c->jobs_hi--;
}

static void end_work(omc_dev_ctx *c, Work *w, int status)
{
    // a work admitted batch by batch has several jobs but ends once
    if (w->state == W_DONE || w->state == W_FAILED) return;
    if (w->acc) {
        dfree(c, w->acc, 2 * sizeof(double) * (size_t)w->nreg);
        w->acc = nullptr;
    }
    w->state = status == OMC_OK ? W_DONE : W_FAILED;
    w->status = status;
    w->finish_ms = now_ms(c);
    if (w->sink) {
        if (status != OMC_OK && w->sink->rc == OMC_OK) w->sink->rc = status;
        w->sink->remaining--;
    } else if (w->ticket >= 0) {
        c->unfinished--;
    }
}

// Bytes a work needs on admission.
static size_t work_bytes(const Work *w)
{
    if (w->scratch) return 0;
    if (w->split_k)   // the next batch's grid, spectrum and ledger, and the sums with the first batch
        return sizeof(double) * (size_t)w->nreg * (w->acc ? 1 : 3) + 2 * OMC_INVDIM * sizeof(double) +
               (w->ledger ? sizeof(double) * OMC_LEDGER * (size_t)w->nperbatch : 0);
    return sizeof(double) * (size_t)w->nreg * w->slots + 2 * OMC_INVDIM * sizeof(double) +
           (w->ledger ? sizeof(double) * OMC_LEDGER * (size_t)w->nhist : 0);
}

static int any_running(const omc_dev_ctx *c)
{
    for (int j = 0; j < c->jobs_hi; ++j) if (c->jobs[j].used) return 1;
    return 0;
}

// Admit waiting works in order while their memory fits.  A work that cannot
// fit even with nothing else admitted fails with OMC_EBUDGET.  A work
// admitted batch by batch stays in line until its last batch is admitted;
// while its batches in flight are at its limit, the works behind it go on.
static int admit(omc_dev_ctx *c)
{
    size_t at = 0;
    while (at < c->waiting.size()) {
        Work *w = c->waiting[at];
        if (w->split_k && w->batches_live >= w->split_k) { ++at; continue; }
        int slot = -1;
        for (int j = 0; j < MAX_JOBS; ++j) if (!c->jobs[j].used) { slot = j; break; }
        if (slot < 0) return OMC_OK;
        int64_t pooled = 0;
        for (auto &kv : c->pool) pooled += (int64_t)kv.first;
        if (c->budget > 0 && c->used - pooled + (int64_t)work_bytes(w) > c->budget) {
            if (any_running(c)) return OMC_OK;          // wait for memory to free up
            c->waiting.erase(c->waiting.begin() + at);
            fail(OMC_EBUDGET, "request needs %zu bytes, budget leaves %lld", work_bytes(w),
                 (long long)(c->budget - (c->used - pooled)));
            end_work(c, w, OMC_EBUDGET);
            continue;
        }
        JobRec &J = c->jobs[slot];
        J = JobRec();
        const bool split = w->split_k > 0;
        const int64_t jn = split ? w->nperbatch : w->nhist;      // histories of this job
        const int slots = split ? 1 : w->slots;
        J.tally_bytes = sizeof(double) * (size_t)w->nreg * slots;
        int rc = OMC_OK;
        if (split && !w->acc) {
            rc = dalloc(c, (void **)&w->acc, 2 * sizeof(double) * (size_t)w->nreg);
            if (!rc) launch_zero_f64(w->acc, 2 * w->nreg, c->aux);
        }
        if (!rc && w->scratch) {
            // the caller checked that both fit (omc_dev_run_dosxyz_noalloc)
            J.borrowed = true;
            J.tally = (double *)c->scratch;
            J.cdf = (double *)((char *)c->scratch + ((J.tally_bytes + 255) & ~(size_t)255));
        } else if (!rc) {
            rc = pool_alloc(c, (void **)&J.tally, J.tally_bytes);
            if (!rc) rc = dalloc(c, (void **)&J.cdf, 2 * OMC_INVDIM * sizeof(double));
        }
        if (!rc && w->ledger) {
            J.ledger_bytes = sizeof(double) * OMC_LEDGER * (size_t)jn;
            rc = dalloc(c, (void **)&J.ledger, J.ledger_bytes);
        }
        if (rc) {
            J.used = true;
            release_job(c, slot);
            if (any_running(c)) return OMC_OK;
            c->waiting.erase(c->waiting.begin() + at);
            end_work(c, w, rc);
            continue;
        }
        J.used = true;
        J.w = w;
        J.nhist = jn;
        J.batch = split ? w->next_batch : -1;
        // serving requests share each refill in slices of chunk/8 (spreading
        // a 6.5e5-history request over all six pipelines instead raised the
        // FP32 p50 of the latency stream from 57 to 69-77 ms); a synchronous
        // run spreads its histories over all pipelines
        const int64_t npipes = (int64_t)c->pipes.size();
        J.quantum = w->ticket >= 0 ? std::max<int64_t>(4096, c->chunk / 8)
                                   : std::max<int64_t>(65536, (jn + npipes - 1) / npipes);
        launch_zero_f64(J.tally, w->nreg * slots, c->aux);
        if (J.ledger) launch_zero_f64(J.ledger, OMC_LEDGER * jn, c->aux);
        HIPCHK(hipMemcpyAsync(J.cdf, w->cdf.data(), 2 * OMC_INVDIM * sizeof(double), hipMemcpyHostToDevice, c->aux));
        DevJob d = w->tmpl;
        d.phantom = w->phantom;
        d.cdf1 = J.cdf;
        d.cdf2 = J.cdf + OMC_INVDIM;
        // one grid per batch; a job that keeps one grid puts every history in
        // it.  A batch admitted on its own is a job of its histories: their
        // streams start at the batch's first history, as in the whole work.
        d.nperbatch = slots == w->nbatch ? w->nperbatch : jn;
        if (split) d.hist0 = w->tmpl.hist0 + (uint64_t)J.batch * (uint64_t)w->nperbatch;
        d.tally = J.tally;
        d.ledger = J.ledger;
        HIPCHK(hipMemcpyAsync(c->d_jobs + slot, &d, sizeof d, hipMemcpyHostToDevice, c->aux));
        HIPCHK(hipStreamSynchronize(c->aux));
        c->jobs_hi = std::max(c->jobs_hi, slot + 1);
        w->job = slot;
        if (w->state == W_WAITING) {
            w->state = W_RUNNING;
            w->admit_ms = now_ms(c);
        }
        c->runnable.push_back(slot);
        if (split) {
            w->batches_live++;
            if (++w->next_batch < w->nbatch) continue;   // the work stays in line
        }
        c->waiting.erase(c->waiting.begin() + at);
    }
    return OMC_OK;
}

// Finishing a job: its last particle is gone, so its tally is final.  The
// results are collected on the helper stream while the pipelines go on:
//   FIN_FOLD    fold the batches (accumEndep, omc_dosxyz.c:696-717) and/or
//               build the beamlet's sparse column; copy the sums and the
//               column length to pinned memory
//   FIN_COLUMN  copy the column's rows and values to pinned memory
// then the host delivers the result.  The device scratch is shared, so one
// job is collected at a time, in the order the jobs ended.
static bool wants_sums(const Work *w)
{
    const RunSink *k = w->sink;
    return !k || k->sum || !k->beamlets;
}

static int finish_fold(omc_dev_ctx *c, int j)
{
    JobRec &J = c->jobs[j];
    const Work *w = J.w;
    const RunSink *k = w->sink;
    const int64_t nreg = w->nreg;
    hipStream_t s = c->aux;
    if (J.batch >= 0) {
        // one batch of a work admitted batch by batch: into the work's sums;
        // after its last batch the sums go to pinned memory
        if (w->acc) {
            launch_accum(J.tally, nreg, w->acc, w->acc + nreg, s);
            if (w->batches_done + 1 == w->nbatch &&
                hipMemcpyAsync(c->h_sums, w->acc, 2 * (size_t)nreg * sizeof(double), hipMemcpyDeviceToHost, s) !=
                    hipSuccess)
                return fail(OMC_EDEVICE, "copy tallies");
        }
    } else if (wants_sums(w)) {
        launch_fold(J.tally, w->nbatch, nreg, c->fin_sums, c->fin_sums + nreg, s);
        if (hipMemcpyAsync(c->h_sums, c->fin_sums, 2 * (size_t)nreg * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess)
            return fail(OMC_EDEVICE, "copy tallies");
    }
    if (k && k->sparse) {
        const PhantomSlot &P = c->ph[w->phantom];
        const double inc = (double)(w->nperbatch * (int64_t)w->nbatch);
        launch_beamlet_column(J.tally, w->slots, w->nbatch, P.isize, P.jsize, P.ksize, (const double *)P.xb.p,
                              (const double *)P.yb.p, (const double *)P.zb.p, (const double *)P.rho.p, inc,
                              k->rel_thr, c->sp, s);
        const int nblk = sparse_blocks((int64_t)P.isize * P.jsize * P.ksize);
        if (hipMemcpyAsync(c->h_nnz, c->sp.blk + nblk, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
            return fail(OMC_EDEVICE, "sparse count");
    }
    if (hipEventRecord(c->ev_fin, s) != hipSuccess) return fail(OMC_EDEVICE, "finishing event");
    J.fin = FIN_FOLD;
    return OMC_OK;
}

static int finish_column(omc_dev_ctx *c, int j)
{
    JobRec &J = c->jobs[j];
    hipStream_t s = c->aux;
    J.nnz = *c->h_nnz;
    if (J.nnz > (uint32_t)c->fin_nvox) return fail(OMC_EDEVICE, "sparse column of %u entries", J.nnz);
    if (J.nnz && (hipMemcpyAsync(c->h_rows, c->sp.rows, J.nnz * sizeof(int64_t), hipMemcpyDeviceToHost, s) !=
                      hipSuccess ||
                  hipMemcpyAsync(c->h_vals, c->sp.vals, J.nnz * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess))
        return fail(OMC_EDEVICE, "copy sparse column");
    if (hipEventRecord(c->ev_fin, s) != hipSuccess) return fail(OMC_EDEVICE, "finishing event");
    J.fin = FIN_COLUMN;
    return OMC_OK;
}

// Host side: hand the collected results to the work's owner.
static int finish_deliver(omc_dev_ctx *c, int j)
{
    JobRec &J = c->jobs[j];
    Work *w = J.w;
    RunSink *k = w->sink;
    const int64_t nreg = w->nreg;
    const bool last = J.batch < 0 || w->batches_done + 1 == w->nbatch;
    if (last && wants_sums(w)) {
        const double *h = c->h_sums;
        if (!k) {
            w->sum.assign(h, h + nreg);
            w->sum2.assign(h + nreg, h + 2 * nreg);
        } else if (k->sum) {
            double *s1 = k->sum + (k->beamlets ? (size_t)w->index * nreg : 0);
            double *s2 = k->sum2 + (k->beamlets ? (size_t)w->index * nreg : 0);
            for (int64_t i = 0; i < nreg; ++i) { s1[i] += h[i]; s2[i] += h[nreg + i]; }
        }
    }
    const int64_t first = w->ledger_off + (J.batch >= 0 ? (int64_t)J.batch * w->nperbatch : 0);
    if (k && k->ledger && J.ledger &&
        hipMemcpy(k->ledger + OMC_LEDGER * first, J.ledger, J.ledger_bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(OMC_EDEVICE, "copy ledger");
    if (k && k->sparse) {
        k->rows[w->index].assign(c->h_rows, c->h_rows + J.nnz);
        k->vals[w->index].assign(c->h_vals, c->h_vals + J.nnz);
    }
    return OMC_OK;
}

// Advances the collection of finished jobs without blocking.  Returns the
// number of jobs delivered (a batch of a work admitted batch by batch counts), or an error (the failing job is ended with it);
// *worked tells whether any stage advanced (otherwise it only polled).
static int pump_finish(omc_dev_ctx *c, bool *worked)
{
    int delivered = 0;
    *worked = false;
    while (!c->finishing.empty()) {
        const int j = c->finishing.front();
        JobRec &J = c->jobs[j];
        int rc = OMC_OK;
        if (J.fin == FIN_QUEUED) {
            rc = finish_fold(c, j);
            *worked = true;
        } else {
            const hipError_t e = hipEventQuery(c->ev_fin);
            if (e == hipErrorNotReady) return delivered;
            *worked = true;
            if (e != hipSuccess) rc = fail(OMC_EDEVICE, "finishing: %s", hipGetErrorString(e));
            else if (J.fin == FIN_FOLD && J.w->sink && J.w->sink->sparse) rc = finish_column(c, j);
            else {
                rc = finish_deliver(c, j);
                if (!rc) {
                    Work *w = J.w;
                    const bool batch = J.batch >= 0;
                    c->finishing.pop_front();
                    release_job(c, j);
                    if (batch) {
                        w->batches_live--;
                        w->batches_done++;
                    }
                    if (!batch || w->batches_done == w->nbatch) end_work(c, w, OMC_OK);
                    // a finished batch frees memory for the work's next one:
                    // the scheduler admits before it waits again
                    delivered++;
                }
            }
        }
        if (rc) {
            Work *w = J.w;
            c->finishing.pop_front();
            release_job(c, j);
            end_work(c, w, rc);
            return rc;
        }
    }
    return delivered;
}

// Collects every finished job, waiting for the helper stream.
static int drain_finish(omc_dev_ctx *c)
{
    while (!c->finishing.empty()) {
        HIPCHK(hipStreamSynchronize(c->aux));
        bool worked;
        const int r = pump_finish(c, &worked);
        if (r < 0) return r;
    }
    return OMC_OK;
}

// Hand up to `room` new primaries to pipeline p (round robin over the
// admitted jobs); fills the pinned segment table.
static int pull_segments(omc_dev_ctx *c, int pi, int64_t room, uint32_t &injP, uint32_t &injE)
{
    Pipe &P = c->pipes[pi];
    int nseg = 0;
    uint32_t off = 0;
    injP = injE = 0;
    size_t turns = c->runnable.size();
    while (room > 0 && nseg < MAX_SEGS && turns-- > 0 && !c->runnable.empty()) {
        int j = c->runnable.front();
        c->runnable.pop_front();
        JobRec &J = c->jobs[j];
        int64_t left = J.nhist - J.injected;
        // A primary photon of a splitting job is followed as nsplit
        // sub-photons in one lane, each setting up to two charged particles
        // in motion, all pushed before the lane lets go of it: it counts as
        // nsplit particles against the room, so one generation's secondaries
        // stay within the queues (four times the chunk).  One primary at a
        // time still fits the smallest queue (4096 >= 2 * OMC_MAX_NSPLIT).
        const int64_t weight = J.w->tmpl.nsplit > 1 ? J.w->tmpl.nsplit : 1;
        int64_t take = std::min<int64_t>(std::min<int64_t>(left, std::max<int64_t>(room / weight, 1)), J.quantum);
        SourceSeg &sg = P.h->segs[nseg++];
        sg.first = (uint64_t)J.injected;
        sg.job = (uint32_t)j;
        sg.n = (uint32_t)take;
        sg.off = off;
        sg.pad = 0;
        off += (uint32_t)take;
        room -= take * weight;
        J.injected += take;
        if (J.w->tmpl.charge == 0) injP += (uint32_t)take; else injE += (uint32_t)take;
        if (!(J.pipes & (1u << pi))) {
            J.pipes |= 1u << pi;
            P.open.push_back(j);
        }
        if (J.injected < J.nhist) c->runnable.push_back(j);
    }
    return nseg;
}

template <typename R> static int launch_generation(omc_dev_ctx *c, int pi)
{
    Pipe &P = c->pipes[pi];
    hipStream_t s = P.stream;
    const int nxt = P.cur ^ 1;
    const int64_t live = (int64_t)P.nP + P.nE;
    uint32_t injP = 0, injE = 0;
    const int nseg = live < c->chunk ? pull_segments(c, pi, c->chunk - live, injP, injE) : 0;
    const uint32_t ntot = injP + injE;
    const uint32_t inP = P.nP + injP, inE = P.nE + injE;
    if (!inP && !inE) return OMC_OK;
    HIPCHK(hipEventRecord(P.ev_begin, s));
    launch_gen_reset(P.d_counts, nxt, s);
    const Queue<R> qpc = queue_view<R>(c, P, P.cur), qec = queue_view<R>(c, P, 2 + P.cur);
    P.had_src = nseg > 0;
    if (P.had_src) {
        HIPCHK(hipMemcpyAsync(P.d_segs, P.h->segs, nseg * sizeof(SourceSeg), hipMemcpyHostToDevice, s));
        HIPCHK(hipEventRecord(P.ev_src0, s));
        Launch<R>::source(c->d_jobs, (const DevPhantom<R> *)c->d_phantoms, P.d_segs, nseg, ntot, qpc, qec,
                          P.d_status, s);
        HIPCHK(hipEventRecord(P.ev_src1, s));
    }
    // The photon kernel appends the electrons it sets in motion to the
    // current charged-particle queue, whose count the electron kernel reads
    // when it starts, after the photon kernel on the same stream: a
    // photon-to-electron hand-over costs no generation.  Everything the
    // electron kernel makes, and every photon, goes to the next queues.
    KArgs<R> A;
    A.T = tables_view<R>(c);
    A.phantoms = (const DevPhantom<R> *)c->d_phantoms;
    A.jobs = c->d_jobs;
    A.qp = queue_view<R>(c, P, nxt);
    A.status = P.d_status;
    A.yield_trips = c->yield_trips;
    A.drain_lanes = c->drain_lanes;
    A.drain_trips = c->drain_trips;
    A.drain_min = (uint32_t)std::min<int64_t>(c->drain_min, UINT32_MAX);
    A.event_div = c->event_div;
    A.fetch_min = c->fetch_min;
    A.woodcock = c->woodcock ? 1u : 0u;
    const int max_grid = c->cus * 8;
    P.had_ph = inP > 0;
    P.in_e = inE;
    if (P.had_ph) {
        A.qin = qpc;
        A.qe = qec;
        A.in_count = P.d_counts + P.cur;
        A.fetch = P.d_counts + 4;
        A.longest = P.d_counts + 7;
        HIPCHK(hipEventRecord(P.ev_ph0, s));
        Launch<R>::photon(A, (int)std::min<int64_t>(max_grid, ((int64_t)inP + 255) / 256), s);
        HIPCHK(hipEventRecord(P.ev_ph1, s));
    }
    P.had_el = inE > 0 || inP > 0;
    if (P.had_el) {
        // the persistent grid covers the known electrons and about two per
        // photon; lanes fetch until the queue's final count either way
        A.qin = qec;
        A.qe = queue_view<R>(c, P, 2 + nxt);
        A.in_count = P.d_counts + 2 + P.cur;
        A.fetch = P.d_counts + 5;
        A.longest = P.d_counts + 6;
        HIPCHK(hipEventRecord(P.ev_el0, s));
        // the input's final count is on the device once the photon kernel is
        // done; the order is sorted there, inside the timed launch
        if (c->sort_min > 0 && (int64_t)inE + 2 * (int64_t)inP >= c->sort_min) {
            Queue<R> sorted = qec;
            sorted.base = (char *)P.d_sorted;
            Launch<R>::sort_charged(qec, P.d_sort, sorted, s);
            A.qin = sorted;
        }
        Launch<R>::electron(A, (int)std::min<int64_t>(max_grid, ((int64_t)inE + 2 * (int64_t)inP + 255) / 256), s);
        HIPCHK(hipEventRecord(P.ev_el1, s));
    }
    HIPCHK(hipGetLastError());
    const int njobs = std::max(c->jobs_hi, 1);
    launch_set_u32(P.d_flags, njobs, 0u, s);
    const Queue<R> qpn = queue_view<R>(c, P, nxt), qen = queue_view<R>(c, P, 2 + nxt);
    launch_job_flags(qpn.meta(), qpn.count, qen.meta(), qen.count, c->qcap, P.d_flags, njobs, s);
    HIPCHK(hipMemcpyAsync(P.h->counts, P.d_counts, 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&P.h->status, P.d_status, sizeof(DevStatus), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(P.h->flags, P.d_flags, njobs * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(P.ev_end[P.gen & 1], s));
    P.injected = ntot;
    P.inflight = true;
    return OMC_OK;
}

static int complete_generation(omc_dev_ctx *c, int pi)
{
    Pipe &P = c->pipes[pi];
    P.inflight = false;
    P.cur ^= 1;
    P.nP = std::min(P.h->counts[P.cur], c->qcap);
    P.nE = std::min(P.h->counts[2 + P.cur], c->qcap);
    omc_dev_stats &S = c->acc;
    float ms = 0.f;
    if (P.gen > P.idle_from) {
        // stream idle between the previous generation's end and this one's
        // start (not before a run's first generation: that gap is the
        // caller's time between runs)
        HIPCHK(hipEventElapsedTime(&ms, P.ev_end[(P.gen - 1) & 1], P.ev_begin));
        S.pipe_idle_ms += ms;
    }
    P.gen++;
    if (P.had_src) {
        HIPCHK(hipEventElapsedTime(&ms, P.ev_src0, P.ev_src1));
        S.source_kernel_ms += ms;
        S.injections++;
    }
    if (P.had_ph) {
        HIPCHK(hipEventElapsedTime(&ms, P.ev_ph0, P.ev_ph1));
        S.photon_kernel_ms += ms;
        S.photon_launches++;
    }
    if (P.had_el) {
        HIPCHK(hipEventElapsedTime(&ms, P.ev_el0, P.ev_el1));
        S.electron_kernel_ms += ms;
        S.electron_launches++;
    }
    S.generations++;
    S.longest_electron_trips = std::max<int64_t>(S.longest_electron_trips, P.h->counts[6]);
    S.longest_photon_trips = std::max<int64_t>(S.longest_photon_trips, P.h->counts[7]);
    S.photons_pushed += P.h->counts[P.cur];
    // electrons into the next queue, and those the photon kernel appended
    // to the queue this generation consumed
    S.electrons_pushed += P.h->counts[2 + P.cur] + (std::min(P.h->counts[2 + (P.cur ^ 1)], c->qcap) - P.in_e);
    if (P.h->status.overflow)
        return fail(OMC_EOVERFLOW, "particle queue overflow (capacity %u per queue)", c->qcap);
    // jobs with no particle left in this pipeline
    std::vector<int> still;
    std::vector<int> check;
    for (int j : P.open) {
        if (P.h->flags[j]) { still.push_back(j); continue; }
        c->jobs[j].pipes &= ~(1u << pi);
        check.push_back(j);
    }
    P.open.swap(still);
    for (int j : check) {
        JobRec &J = c->jobs[j];
        if (J.used && J.fin == FIN_NONE && J.pipes == 0 && J.injected == J.nhist) {
            J.fin = FIN_QUEUED;
            c->finishing.push_back(j);
        }
    }
    return OMC_OK;
}

// Stops everything after a fatal error: waits for the pipelines, empties
// their queues and fails every admitted or waiting work.
static void reset_engine(omc_dev_ctx *c, int rc)
{
    char keep[sizeof g_err];
    memcpy(keep, g_err, sizeof keep);
    (void)hipStreamSynchronize(c->aux);
    c->finishing.clear();
    for (auto &P : c->pipes) {
        (void)hipStreamSynchronize(P.stream);
        (void)hipMemsetAsync(P.d_counts, 0, 8 * sizeof(uint32_t), P.stream);
        (void)hipMemsetAsync(P.d_status, 0, sizeof(DevStatus), P.stream);
        (void)hipStreamSynchronize(P.stream);
        P.cur = 0;
        P.nP = P.nE = 0;
        P.inflight = false;
        P.open.clear();
        memset(&P.base, 0, sizeof P.base);
        memset(&P.h->status, 0, sizeof P.h->status);
    }
    c->runnable.clear();
    for (int j = 0; j < c->jobs_hi; ++j) {
        if (!c->jobs[j].used) continue;
        Work *w = c->jobs[j].w;
        release_job(c, j);
        end_work(c, w, rc);
    }
    while (!c->waiting.empty()) {
        Work *w = c->waiting.front();
        c->waiting.pop_front();
        end_work(c, w, rc);
    }
    memcpy(g_err, keep, sizeof keep);
}

// One scheduler turn: admit, start a generation on every idle pipeline that
// has work, then wait until a generation ends (and process it) or a finished
// job's results are delivered.  Returns 1 after either, 0 if the engine is
// idle.
template <typename R> static int turn(omc_dev_ctx *c)
{
    auto h0 = std::chrono::steady_clock::now();
    int rc = admit(c);
    for (size_t p = 0; !rc && p < c->pipes.size(); ++p)
        if (!c->pipes[p].inflight) rc = launch_generation<R>(c, (int)p);
    c->acc.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    if (rc) { reset_engine(c, rc); return rc; }
    bool any = !c->finishing.empty();
    for (auto &P : c->pipes) any |= P.inflight;
    if (!any) return 0;
    int spins = 0;
    for (;;) {
        if (!c->finishing.empty()) {
            auto h1 = std::chrono::steady_clock::now();
            bool worked;
            const int d = pump_finish(c, &worked);
            if (worked)      // polling a stage still in flight is waiting, not work
                c->acc.host_ms +=
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h1).count();
            if (d < 0) { reset_engine(c, d); return d; }
            if (d > 0) return 1;
        }
        int oldest = -1;
        for (size_t p = 0; p < c->pipes.size(); ++p) {
            Pipe &P = c->pipes[p];
            if (!P.inflight) continue;
            hipError_t e = hipEventQuery(P.ev_end[P.gen & 1]);
            if (e == hipSuccess) { oldest = (int)p; break; }
            if (e != hipErrorNotReady) {
                rc = fail(OMC_EDEVICE, "pipeline %zu: %s", p, hipGetErrorString(e));
                reset_engine(c, rc);
                return rc;
            }
        }
        if (oldest >= 0) {
            auto h1 = std::chrono::steady_clock::now();
            rc = complete_generation(c, oldest);
            c->acc.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h1).count();
            if (rc) { reset_engine(c, rc); return rc; }
            return 1;
        }
        // generations end every few milliseconds: keep the host thread
        // polling so the next one starts at once, and back off only when
        // nothing has ended for a long time
        if (++spins < 200000) std::this_thread::yield();
        else std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

static int turn_any(omc_dev_ctx *c)
{
    return c->precision == OMC_PREC_FP64 ? turn<double>(c) : turn<float>(c);
}

static void stats_begin(omc_dev_ctx *c)
{
    memset(&c->acc, 0, sizeof c->acc);
    for (auto &P : c->pipes) {
        P.base = P.h->status;
        P.idle_from = P.gen;
    }
}

static void stats_end(omc_dev_ctx *c, omc_dev_stats *out, int64_t histories, double t_start)
{
    if (!out) return;
    omc_dev_stats S = c->acc;
    for (auto &P : c->pipes) {
        // the pipelines' mirrors hold the status after their last generation
        S.photon_steps += (int64_t)(P.h->status.photon_steps - P.base.photon_steps);
        S.electron_steps += (int64_t)(P.h->status.electron_steps - P.base.electron_steps);
        S.photon_crossings += (int64_t)(P.h->status.photon_cross - P.base.photon_cross);
        S.electron_crossings += (int64_t)(P.h->status.electron_cross - P.base.electron_cross);
        S.photon_deposits += (int64_t)(P.h->status.photon_dep - P.base.photon_dep);
        S.electron_deposits += (int64_t)(P.h->status.electron_dep - P.base.electron_dep);
        S.killed += (int64_t)(P.h->status.killed - P.base.killed);
        S.woodcock_over += (int64_t)(P.h->status.wc_over - P.base.wc_over);
        S.yields += (int64_t)(P.h->status.yields - P.base.yields);
        int64_t *ev[EV_N] = { &S.compton, &S.photo, &S.pair, &S.rayleigh, &S.brems, &S.moller, &S.bhabha,
                              &S.annih_flight, &S.annih_rest, &S.fict_rejections, &S.msdist, &S.sscat,
                              &S.howfar_charged };
        for (int k = 0; k < EV_N; ++k) *ev[k] += (int64_t)(P.h->status.ev[k] - P.base.ev[k]);
    }
    S.deposits = S.photon_deposits + S.electron_deposits;
    S.histories = histories;
    S.transport_ms = now_ms(c) - t_start;
    *out = S;
}

// Runs the works of one synchronous call to completion (serving requests
// already admitted keep running alongside).
// A work whose whole batch tally would exceed the context's tally_bytes is
// admitted batch by batch, with as many batches in flight as the limit
// leaves room for beside its two sum grids, but at least two and at least
// enough histories to keep half of the pipelines' refill room (pipelines x
// chunk) busy: each batch holds its grid until its slowest particle is
// gone, and with small batches in flight the pipelines would run mostly
// such tails (C5 plan, 6e5 histories per batch, four batches per beam in
// flight: 6.1e7 histories/s against 7.8e7 with whole tallies, 1036
// generations against 105).  No split where it would not save a grid.
static void plan_split(const omc_dev_ctx *c, Work &w)
{
    w.split_k = 0;
    const int64_t grid = (int64_t)sizeof(double) * w.nreg;
    if (c->tally_cap <= 0 || w.scratch || w.nbatch < 2 || w.slots != w.nbatch) return;
    if (grid * w.nbatch <= c->tally_cap) return;
    const int64_t inflight = (int64_t)c->pipes.size() * c->chunk / 2;
    const int64_t k = std::min<int64_t>(std::max<int64_t>({ c->tally_cap / grid - 2,
                                                            (inflight + w.nperbatch - 1) / w.nperbatch, (int64_t)2 }),
                                        MAX_SPLIT_INFLIGHT);
    if (k + 2 >= w.nbatch) return;
    w.split_k = (int)k;
}

static int run_sink(omc_dev_ctx *c, RunSink &k, std::vector<Work> &works, omc_dev_stats *stats)
{
    k.remaining = (int)works.size();
    for (auto &w : works) {
        plan_split(c, w);
        w.sink = &k;
        w.submit_ms = now_ms(c);
        c->waiting.push_back(&w);
    }
    stats_begin(c);
    const double t = now_ms(c);
    int64_t hist = 0;
    for (auto &w : works) hist += w.nhist;
    while (k.remaining > 0) {
        int r = turn_any(c);
        if (r < 0) break;                // reset_engine ended every work
        if (r == 0 && k.remaining > 0) {
            // admission either starts a work or fails it, so an idle engine
            // with works left is inconsistent: stop rather than spin
            reset_engine(c, fail(OMC_EDEVICE, "scheduler stalled with %d works left", k.remaining));
            break;
        }
    }
    stats_end(c, stats, hist, t);
    return k.rc;
}

// Device address of the scratch array on `device` (current), looked up once
// per device and kept.
static int scratch_on(int device, void **out)
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(OMC_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        it = addr.emplace(device, p).first;
    }
    *out = it->second;
    return OMC_OK;
}

// ------------------------------------------------------------------ C ABI --
extern "C" {

const char *omc_dev_last_error(void) { return g_err; }

// Device bytes one pipeline holds (mirrors alloc_pipe).
static int64_t pipe_bytes(const omc_dev_ctx *c)
{
    const size_t pb = c->precision == OMC_PREC_FP64 ? particle_bytes<double>() : particle_bytes<float>();
    return 5 * (int64_t)pb * c->qcap + 8 * sizeof(uint32_t) + sizeof(DevStatus) + MAX_JOBS * sizeof(uint32_t) +
           MAX_SEGS * sizeof(SourceSeg) + sizeof(uint32_t) * SORT_BUCKETS * SORT_BLOCKS;
}

// (Re)builds the pipelines: npipes streams, each with queues four times
// `chunk` deep (a generation's secondaries) and the sorted copy of a
// charged-particle input; the five take at most half the budget, which can
// lower the chunk.  No work may be queued.
static int configure_pipes(omc_dev_ctx *c, int npipes, int64_t chunk)
{
    for (auto &P : c->pipes) free_pipe(P);
    if (!c->pipes.empty()) c->used -= (int64_t)c->pipes.size() * pipe_bytes(c);
    c->pipes.clear();
    npipes = std::max(1, std::min(npipes, MAX_PIPES));
    const size_t pb = c->precision == OMC_PREC_FP64 ? particle_bytes<double>() : particle_bytes<float>();
    c->chunk = chunk;
    int64_t cap = std::max<int64_t>(4 * c->chunk, 1 << 12);
    const int64_t cap_mem = c->budget / 2 / npipes / 5 / (int64_t)pb;
    if (cap > cap_mem) {
        cap = cap_mem;
        c->chunk = std::max<int64_t>(cap / 4, 1);
    }
    cap = std::min<int64_t>(cap, 1 << 26);
    c->chunk = std::min<int64_t>(c->chunk, cap / 4 > 0 ? cap / 4 : 1);
    if (cap < 64) return fail(OMC_EBUDGET, "HBM budget %lld too small for the particle queues", (long long)c->budget);
    c->qcap = (uint32_t)cap;
    c->pipes.resize(npipes);
    for (auto &P : c->pipes) {
        const int rc = c->precision == OMC_PREC_FP64 ? alloc_pipe<double>(c, P) : alloc_pipe<float>(c, P);
        if (rc) return rc;
    }
    return OMC_OK;
}

int omc_dev_create(omc_dev_ctx **out, const omc_dev_opts *opts)
{
    if (!out) return fail(OMC_EARG, "null context pointer");
    *out = nullptr;
    omc_dev_opts o = { 0, OMC_PREC_FP64, 0, 0, 0, 0, 0, 0 };
    if (opts) o = *opts;
    if (o.precision != OMC_PREC_FP64 && o.precision != OMC_PREC_FP32)
        return fail(OMC_EARG, "unknown precision %d", o.precision);
    if (o.pipelines < 0 || o.pipelines > MAX_PIPES) return fail(OMC_ERANGE, "pipelines %d (1..%d)", o.pipelines, MAX_PIPES);
    if (o.chunk < 0) return fail(OMC_EARG, "chunk %lld", (long long)o.chunk);
    if (o.yield_trips < 0) return fail(OMC_EARG, "yield_trips %d", o.yield_trips);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (o.device < 0 || o.device >= ndev) return fail(OMC_EDEVICE, "device %d not present (%d devices)", o.device, ndev);
    HIPCHK(hipSetDevice(o.device));
    void *scratch_dev = nullptr;
    OMC_TRY_DEV(scratch_on(o.device, &scratch_dev));
    omc_dev_ctx *c = new omc_dev_ctx();
    c->scratch = scratch_dev;
    c->t0 = std::chrono::steady_clock::now();
    c->device = o.device;
    c->precision = o.precision;
    c->yield_trips = o.yield_trips > 0 ? (uint32_t)o.yield_trips : DEFAULT_YIELD_TRIPS;
    c->drain_lanes = o.drain_lanes > 0 ? (uint32_t)std::min(o.drain_lanes, 64) : o.drain_lanes < 0 ? 0u : DEFAULT_DRAIN_LANES;
    c->drain_trips = o.drain_trips > 0 ? (uint32_t)o.drain_trips : DEFAULT_DRAIN_TRIPS;
    c->event_div = o.event_div > 0 ? (uint32_t)o.event_div : DEFAULT_EVENT_DIV;
    c->fetch_min = o.fetch_min > 0 ? (uint32_t)std::min(o.fetch_min, 64) : DEFAULT_FETCH_MIN;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, o.device) == hipSuccess) c->cus = prop.multiProcessorCount;
    c->drain_min = o.drain_min > 0 ? o.drain_min : o.drain_min < 0 ? 0 : drain_min_default(c->cus, c->precision);
    c->woodcock = o.woodcock > 0 || (o.woodcock == 0 && o.precision == OMC_PREC_FP32);
    c->tally_cap = o.tally_bytes > 0 ? o.tally_bytes : o.tally_bytes < 0 ? 0 : DEFAULT_TALLY_CAP;
    // energy-ordered charged-particle launches once their input exceeds the
    // lanes the device holds resident: only then does the order in which
    // particles start matter (slab: FP32 +10%, FP64 +8%; same histories)
    c->sort_min = o.sort_min > 0 ? o.sort_min : o.sort_min < 0 ? 0 : drain_min_default(c->cus, c->precision);
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    c->budget = o.hbm_budget > 0 ? o.hbm_budget : (int64_t)(0.9 * (double)freeb);
    int rc = OMC_OK;
    if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess) rc = fail(OMC_EDEVICE, "stream");
    if (!rc && hipEventCreateWithFlags(&c->ev_fin, hipEventDisableTiming) != hipSuccess) rc = fail(OMC_EDEVICE, "event");
    if (!rc) rc = dalloc(c, (void **)&c->d_jobs, MAX_JOBS * sizeof(DevJob));
    const size_t phb = o.precision == OMC_PREC_FP64 ? sizeof(DevPhantom<double>) : sizeof(DevPhantom<float>);
    if (!rc) rc = dalloc(c, &c->d_phantoms, MAX_PHANTOMS * phb);
    if (!rc) rc = configure_pipes(c, o.pipelines > 0 ? o.pipelines : DEFAULT_PIPES, o.chunk > 0 ? o.chunk : DEFAULT_CHUNK);
    if (rc) { omc_dev_destroy(c); return rc; }
    *out = c;
    return OMC_OK;
}

int omc_dev_destroy(omc_dev_ctx *c)
{
    if (!c) return OMC_OK;
    (void)hipSetDevice(c->device);
    for (auto &P : c->pipes) free_pipe(P);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    c->finishing.clear();
    for (int j = 0; j < c->jobs_hi; ++j) if (c->jobs[j].used) release_job(c, j);
    if (c->h_sums) (void)hipHostFree(c->h_sums);
    if (c->ev_fin) (void)hipEventDestroy(c->ev_fin);
    for (auto &kv : c->pool) (void)hipFree(kv.second);
    void *ptrs[] = { c->small, c->ph_mfp, c->ph_br, c->ph_pmax, c->el, c->csda, c->spin, c->ms,
                     c->d_jobs, c->d_phantoms, c->fin_sums };
    for (void *p : ptrs) if (p) (void)hipFree(p);
    for (auto &P : c->ph) {
        DevBuf *bufs[] = { &P.xb, &P.yb, &P.zb, &P.med, &P.rhof, &P.rho, &P.wc };
        for (auto *b : bufs) if (b->p) (void)hipFree(b->p);
    }
    for (auto &kv : c->tickets) {
        if (kv.second->acc) (void)hipFree(kv.second->acc);
        delete kv.second;
    }
    if (c->aux) (void)hipStreamDestroy(c->aux);
    delete c;
    return OMC_OK;
}

int omc_dev_precision(const omc_dev_ctx *c) { return c ? c->precision : -1; }

int omc_dev_upload_media(omc_dev_ctx *c, const omc_tables *t)
{
    if (!c || !t) return fail(OMC_EARG, "null argument");
    if (t->nmed < 1 || t->nmed > OMC_MXMED) return fail(OMC_ERANGE, "nmed %d", t->nmed);
    if (any_running(c) || !c->waiting.empty()) return fail(OMC_EARG, "media cannot change while work is queued");
    // the tables must still cover every medium a resident phantom uses
    for (int i = 0; i < MAX_PHANTOMS; ++i) {
        if (!c->ph[i].used) continue;
        for (int m = t->nmed; m < MAXMED; ++m)
            if (c->ph[i].rho_max[m] >= 0)
                return fail(OMC_EARG, "resident phantom %d uses medium %d, new tables have %d", i, m, t->nmed);
    }
    HIPCHK(hipSetDevice(c->device));
    return c->precision == OMC_PREC_FP64 ? upload_tables<double>(c, t) : upload_tables<float>(c, t);
}

int omc_dev_upload_phantom(omc_dev_ctx *c, const omc_phantom *ph, const omc_regions *reg, int *id)
{
    if (!c || !ph || !reg || !id) return fail(OMC_EARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    return c->precision == OMC_PREC_FP64 ? upload_phantom<double>(c, ph, reg, id) : upload_phantom<float>(c, ph, reg, id);
}

int omc_dev_release_phantom(omc_dev_ctx *c, int id)
{
    if (!c || id < 0 || id >= MAX_PHANTOMS || !c->ph[id].used) return fail(OMC_EARG, "bad phantom id %d", id);
    for (int j = 0; j < c->jobs_hi; ++j)
        if (c->jobs[j].used && c->jobs[j].w->phantom == id) return fail(OMC_EARG, "phantom %d is in use", id);
    for (auto *w : c->waiting)
        if (w->phantom == id) return fail(OMC_EARG, "phantom %d is in use", id);
    PhantomSlot &P = c->ph[id];
    DevBuf *bufs[] = { &P.xb, &P.yb, &P.zb, &P.med, &P.rhof, &P.rho, &P.wc };
    for (auto *b : bufs) { if (b->p) dfree(c, b->p, b->bytes); b->p = nullptr; b->bytes = 0; }
    P.used = false;
    return OMC_OK;
}

static int check_phantom(omc_dev_ctx *c, int id)
{
    if (!c->media) return fail(OMC_EARG, "media tables not uploaded");
    if (id < 0 || id >= MAX_PHANTOMS || !c->ph[id].used) return fail(OMC_EARG, "bad phantom id %d", id);
    return OMC_OK;
}

static int check_run(omc_dev_ctx *c, const omc_dev_run_params *rp)
{
    if (!c || !rp) return fail(OMC_EARG, "null argument");
    OMC_TRY_DEV(check_phantom(c, rp->phantom_id));
    if (rp->nsplit < 1 || rp->nsplit > OMC_MAX_NSPLIT)
        return fail(OMC_EARG, "nsplit %d outside 1..%d", rp->nsplit, OMC_MAX_NSPLIT);
    if (rp->nbatch < 1) return fail(OMC_EARG, "nbatch must be >= 1");
    return OMC_OK;
}

static void source_fields(DevJob &j, const omc_spectrum &spec, int charge)
{
    j.charge = charge;
    j.spectrum = spec.spectrum;
    j.energy = spec.energy;
    j.deltak = spec.deltak;
}

static void dosxyz_work(Work &w, const omc_source_dosxyz *src)
{
    memset(&w.tmpl, 0, sizeof w.tmpl);
    w.tmpl.kind = 0;
    source_fields(w.tmpl, src->spec, src->charge);
    w.tmpl.ssd = src->ssd; w.tmpl.xinl = src->xinl; w.tmpl.yinl = src->yinl;
    w.tmpl.xsize = src->xsize; w.tmpl.ysize = src->ysize;
    w.tmpl.ixinl = src->ixinl; w.tmpl.iyinl = src->iyinl;
    w.cdf.assign(src->spec.cdfinv1, src->spec.cdfinv1 + OMC_INVDIM);
    w.cdf.insert(w.cdf.end(), src->spec.cdfinv2, src->spec.cdfinv2 + OMC_INVDIM);
}

static int beamlet_work(Work &w, const omc_source_beamlets *src, int b)
{
    if (b < 0 || b >= src->nbeamlets) return fail(OMC_ERANGE, "beamlet %d of %d", b, src->nbeamlets);
    const int ib = src->ibeam[b];
    if (ib < 0 || ib >= src->nbeams) return fail(OMC_ERANGE, "beamlet %d: beam %d of %d", b, ib, src->nbeams);
    memset(&w.tmpl, 0, sizeof w.tmpl);
    w.tmpl.kind = 1;
    source_fields(w.tmpl, src->spec, src->charge);
    for (int d = 0; d < 3; ++d) {
        w.tmpl.src[d] = src->src[3 * ib + d];
        w.tmpl.corner[d] = src->corner[3 * b + d];
        w.tmpl.side1[d] = src->side1[3 * b + d];
        w.tmpl.side2[d] = src->side2[3 * b + d];
    }
    w.cdf.assign(src->spec.cdfinv1, src->spec.cdfinv1 + OMC_INVDIM);
    w.cdf.insert(w.cdf.end(), src->spec.cdfinv2, src->spec.cdfinv2 + OMC_INVDIM);
    return OMC_OK;
}

static void stream_fields(Work &w, const uint32_t seed[2], int64_t hist0)
{
    w.tmpl.hist0 = (uint64_t)hist0;
    w.tmpl.s0 = seed[0];
    w.tmpl.s1 = seed[1];
}

static int set_histories(Work &w, int64_t nperbatch, int nbatch)
{
    if (nperbatch < 1) nperbatch = 1;
    if (nperbatch * (int64_t)nbatch > 0xffffffffLL) return fail(OMC_ERANGE, "more than 2^32 histories in one job");
    w.nperbatch = nperbatch;
    w.nbatch = nbatch;
    w.slots = nbatch;
    w.nhist = nperbatch * nbatch;
    return OMC_OK;
}

int omc_dev_run_dosxyz(omc_dev_ctx *c, const omc_source_dosxyz *src, const omc_dev_run_params *rp,
                       double *sum, double *sum2)
{
    OMC_TRY_DEV(check_run(c, rp));
    if (!src || !sum || !sum2) return fail(OMC_EARG, "null argument");
    if (src->charge < -1 || src->charge > 1) return fail(OMC_EARG, "charge %d", src->charge);
    if (rp->nperbatch < 1) return fail(OMC_EARG, "nperbatch must be >= 1");
    HIPCHK(hipSetDevice(c->device));
    std::vector<Work> works(1);
    Work &w = works[0];
    dosxyz_work(w, src);
    stream_fields(w, rp->seed, rp->hist0);
    w.tmpl.nsplit = rp->nsplit;
    OMC_TRY_DEV(set_histories(w, rp->nperbatch, rp->nbatch));
    w.phantom = rp->phantom_id;
    w.nreg = c->ph[rp->phantom_id].nreg;
    w.ledger = rp->ledger != nullptr;
    RunSink k;
    k.sum = sum;
    k.sum2 = sum2;
    k.ledger = rp->ledger;
    return run_sink(c, k, works, rp->stats);
}

int omc_dev_run_dosxyz_noalloc(omc_dev_ctx *c, const omc_source_dosxyz *src, const omc_dev_run_params *rp,
                               double *sum, double *sum2)
{
    OMC_TRY_DEV(check_run(c, rp));
    if (!src || !sum || !sum2) return fail(OMC_EARG, "null argument");
    if (src->charge < -1 || src->charge > 1) return fail(OMC_EARG, "charge %d", src->charge);
    if (rp->nperbatch < 1) return fail(OMC_EARG, "nperbatch must be >= 1");
    if (rp->ledger) return fail(OMC_EARG, "the allocation-free run keeps no per-history ledger");
    if (!c->scratch) return fail(OMC_EDEVICE, "no scratch array on device %d", c->device);
    const int64_t nreg = c->ph[rp->phantom_id].nreg;
    const size_t tally = sizeof(double) * (size_t)nreg * (size_t)rp->nbatch;
    const size_t need = ((tally + 255) & ~(size_t)255) + 2 * OMC_INVDIM * sizeof(double);
    if (need > OMC_NOALLOC_SCRATCH_BYTES)
        return fail(OMC_ERANGE, "%d batches of %lld regions need %zu bytes of scratch, at most %zu", rp->nbatch,
                    (long long)nreg, need, (size_t)OMC_NOALLOC_SCRATCH_BYTES);
    if (any_running(c) || !c->waiting.empty())
        return fail(OMC_EARG, "the allocation-free run needs a context with no queued work");
    HIPCHK(hipSetDevice(c->device));
    std::vector<Work> works(1);
    Work &w = works[0];
    dosxyz_work(w, src);
    stream_fields(w, rp->seed, rp->hist0);
    w.tmpl.nsplit = rp->nsplit;
    OMC_TRY_DEV(set_histories(w, rp->nperbatch, rp->nbatch));
    w.phantom = rp->phantom_id;
    w.nreg = nreg;
    w.scratch = true;
    RunSink k;
    k.sum = sum;
    k.sum2 = sum2;
    return run_sink(c, k, works, rp->stats);
}

// ------------------------------------------------------------- autotuner --
namespace {
struct TuneSetting {
    int pipes;
    int64_t chunk;
    int yield_trips, drain_lanes, drain_trips, event_div, fetch_min;
    int64_t sort_min;
};

TuneSetting current_setting(const omc_dev_ctx *c)
{
    return { (int)c->pipes.size(), c->chunk, (int)c->yield_trips, (int)c->drain_lanes, (int)c->drain_trips,
             (int)c->event_div, (int)c->fetch_min, c->sort_min };
}

int apply_setting(omc_dev_ctx *c, const TuneSetting &t)
{
    if (t.pipes != (int)c->pipes.size() || t.chunk != c->chunk) OMC_TRY_DEV(configure_pipes(c, t.pipes, t.chunk));
    c->yield_trips = (uint32_t)t.yield_trips;
    c->drain_lanes = (uint32_t)t.drain_lanes;
    c->drain_trips = (uint32_t)t.drain_trips;
    c->event_div = (uint32_t)t.event_div;
    c->fetch_min = (uint32_t)t.fetch_min;
    c->sort_min = t.sort_min;
    return OMC_OK;
}

}  // namespace

// Coordinate descent over the engine settings (one option at a time, from
// the current settings, one pass); each setting is timed as the better of
// two calls of `run` (which runs `hist` histories).  Leaves the context with
// the fastest settings and reports them.
static int tune_impl(omc_dev_ctx *c, const std::function<int()> &run, int64_t hist, omc_dev_opts *best,
                     double *hps)
{
    int max_pipes = MAX_PIPES;
    if (const char *q = getenv("GPU_MAX_HW_QUEUES")) {
        const int n = atoi(q);
        if (n > 1) max_pipes = std::min(max_pipes, n - 1);
    }
    auto measure = [&](const TuneSetting &t, double &v) -> int {
        OMC_TRY_DEV(apply_setting(c, t));
        v = 0;
        for (int k = 0; k < 2; ++k) {
            const auto t0 = std::chrono::steady_clock::now();
            OMC_TRY_DEV(run());
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            v = std::max(v, (double)hist / std::max(sec, 1e-9));
        }
        return OMC_OK;
    };
    const std::vector<int64_t> axes[8] = {
        { 2, 4, 6, 8, 10, 12 },                      // pipelines
        { 1 << 20, 1 << 21, 1 << 22, 1 << 23 },      // live particles per pipeline
        { 1024, 4096, 16384 },                       // yield budget (trips)
        { 16, 32, 48 },                              // drain lanes
        { 8, 16, 32 },                               // drain trips
        { 1, 2, 4 },                                 // photon event threshold divisor
        { 1, 4, 8, 16 },                             // fetch threshold (lanes)
        { 0, drain_min_default(c->cus, c->precision) },   // energy-ordered charged launches (0 = never)
    };
    TuneSetting bst = current_setting(c);
    double bv = 0;
    int rc = measure(bst, bv);
    for (int a = 0; a < 8 && !rc; ++a) {
        for (int64_t val : axes[a]) {
            TuneSetting t = bst;
            int *fields[8] = { &t.pipes, nullptr, &t.yield_trips, &t.drain_lanes, &t.drain_trips, &t.event_div,
                               &t.fetch_min, nullptr };
            if (a == 1) {
                if (t.chunk == val) continue;
                t.chunk = val;
            } else if (a == 7) {
                if (t.sort_min == val) continue;
                t.sort_min = val;
            } else {
                if (*fields[a] == (int)val) continue;
                *fields[a] = (int)val;
            }
            if (t.pipes > max_pipes) continue;
            double v = 0;
            rc = measure(t, v);
            if (rc) break;
            if (v > bv) {
                bv = v;
                bst = t;
            }
        }
    }
    const int rc2 = apply_setting(c, bst);
    if (rc) return rc;
    OMC_TRY_DEV(rc2);
    memset(best, 0, sizeof *best);
    best->device = c->device;
    best->precision = c->precision;
    best->hbm_budget = c->budget;
    best->chunk = c->chunk;
    best->pipelines = (int)c->pipes.size();
    best->yield_trips = (int)c->yield_trips;
    best->drain_lanes = (int)c->drain_lanes;
    best->drain_trips = (int)c->drain_trips;
    best->drain_min = c->drain_min > 0 ? c->drain_min : -1;
    best->woodcock = c->woodcock ? 1 : -1;
    best->event_div = (int)c->event_div;
    best->fetch_min = (int)c->fetch_min;
    best->tally_bytes = c->tally_cap > 0 ? c->tally_cap : -1;
    best->sort_min = c->sort_min > 0 ? c->sort_min : -1;
    *hps = bv;
    return OMC_OK;
}

int omc_dev_tune(omc_dev_ctx *c, const omc_source_dosxyz *src, const omc_dev_run_params *rp, int64_t probe,
                 omc_dev_opts *best, double *hps)
{
    OMC_TRY_DEV(check_run(c, rp));
    if (!src || !best || !hps) return fail(OMC_EARG, "null argument");
    if (probe < rp->nbatch) return fail(OMC_EARG, "probe of %lld histories for %d batches", (long long)probe, rp->nbatch);
    if (any_running(c) || !c->waiting.empty()) return fail(OMC_EARG, "the autotuner needs a context with no queued work");
    HIPCHK(hipSetDevice(c->device));
    omc_dev_run_params r = *rp;
    r.nperbatch = probe / rp->nbatch;
    r.ledger = nullptr;
    r.stats = nullptr;
    const int nreg = c->ph[rp->phantom_id].nreg;
    std::vector<double> s1(nreg), s2(nreg);
    return tune_impl(c, [&] { return omc_dev_run_dosxyz(c, src, &r, s1.data(), s2.data()); },
                     r.nperbatch * r.nbatch, best, hps);
}

static int beamlet_works(omc_dev_ctx *c, const omc_source_beamlets *src, const int64_t *nhist,
                         const omc_dev_run_params *rp, std::vector<Work> &works)
{
    OMC_TRY_DEV(check_run(c, rp));
    if (!src || !nhist) return fail(OMC_EARG, "null argument");
    if (src->nbeamlets < 1 || !src->ibeam || !src->src || !src->corner || !src->side1 || !src->side2)
        return fail(OMC_EARG, "empty beamlet source");
    if (src->charge < -1 || src->charge > 1) return fail(OMC_EARG, "charge %d", src->charge);
    works.resize(src->nbeamlets);
    int64_t h0 = rp->hist0;
    for (int b = 0; b < src->nbeamlets; ++b) {
        Work &w = works[b];
        if (nhist[b] < 1) return fail(OMC_EARG, "beamlet %d: %lld histories", b, (long long)nhist[b]);
        OMC_TRY_DEV(beamlet_work(w, src, b));
        OMC_TRY_DEV(set_histories(w, nhist[b] / rp->nbatch, rp->nbatch));
        stream_fields(w, rp->seed, h0);
        w.tmpl.nsplit = rp->nsplit;
        w.ledger = rp->ledger != nullptr;
        w.ledger_off = h0 - rp->hist0;
        h0 += w.nhist;
        w.index = b;
        w.phantom = rp->phantom_id;
        w.nreg = c->ph[rp->phantom_id].nreg;
    }
    return OMC_OK;
}

int omc_dev_run_beamlets(omc_dev_ctx *c, const omc_source_beamlets *src, const int64_t *nhist,
                         const omc_dev_run_params *rp, double *sum, double *sum2)
{
    if (!sum || !sum2) return fail(OMC_EARG, "null argument");
    std::vector<Work> works;
    OMC_TRY_DEV(beamlet_works(c, src, nhist, rp, works));
    HIPCHK(hipSetDevice(c->device));
    RunSink k;
    k.beamlets = true;
    k.sum = sum;
    k.sum2 = sum2;
    k.ledger = rp->ledger;
    return run_sink(c, k, works, rp->stats);
}

int omc_dev_beamlets_sparse(omc_dev_ctx *c, const omc_source_beamlets *src, const int64_t *nhist,
                            const omc_dev_run_params *rp, double rel_thr, int64_t *jcs, int64_t *irs,
                            double *vals, int64_t cap, int64_t *nnz)
{
    if (!c || !jcs || !nnz || (cap > 0 && (!irs || !vals))) return fail(OMC_EARG, "null argument");
    if (src) {
        std::vector<Work> works;
        OMC_TRY_DEV(beamlet_works(c, src, nhist, rp, works));
        // the Dij uses the mean dose only (accumulateResults' uncertainty is
        // not part of it, omc_matrad.c:1416-1479): one FP64 grid per beamlet
        // instead of one per batch, and no ledger
        for (auto &w : works) { w.ledger = false; w.slots = 1; }
        HIPCHK(hipSetDevice(c->device));
        RunSink k;
        k.beamlets = true;
        k.sparse = true;
        k.rel_thr = rel_thr;
        k.rows.resize(works.size());
        k.vals.resize(works.size());
        c->kept_rows.clear();
        c->kept_vals.clear();
        OMC_TRY_DEV(run_sink(c, k, works, rp->stats));
        c->kept_rows.swap(k.rows);
        c->kept_vals.swap(k.vals);
    } else if (c->kept_rows.empty()) {
        return fail(OMC_EARG, "no kept sparse matrix to copy");
    }
    int64_t total = 0;
    for (auto &r : c->kept_rows) total += (int64_t)r.size();
    *nnz = total;
    if (total > cap) return fail(OMC_ERANGE, "sparse Dij has %lld entries, capacity %lld", (long long)total, (long long)cap);
    int64_t at = 0;
    jcs[0] = 0;
    for (size_t b = 0; b < c->kept_rows.size(); ++b) {
        std::copy(c->kept_rows[b].begin(), c->kept_rows[b].end(), irs + at);
        std::copy(c->kept_vals[b].begin(), c->kept_vals[b].end(), vals + at);
        at += (int64_t)c->kept_rows[b].size();
        jcs[b + 1] = at;
    }
    c->kept_rows.clear();
    c->kept_vals.clear();
    return OMC_OK;
}

int omc_dev_tune_beamlets(omc_dev_ctx *c, const omc_source_beamlets *src, const int64_t *nhist,
                          const omc_dev_run_params *rp, omc_dev_opts *best, double *hps)
{
    std::vector<Work> probe;
    OMC_TRY_DEV(beamlet_works(c, src, nhist, rp, probe));      // validates the arguments
    if (!best || !hps) return fail(OMC_EARG, "null argument");
    if (any_running(c) || !c->waiting.empty()) return fail(OMC_EARG, "the autotuner needs a context with no queued work");
    HIPCHK(hipSetDevice(c->device));
    int64_t hist = 0;
    for (auto &w : probe) hist += w.nhist;
    omc_dev_run_params r = *rp;
    r.stats = nullptr;
    return tune_impl(c, [&] {
        std::vector<Work> works;
        OMC_TRY_DEV(beamlet_works(c, src, nhist, &r, works));
        for (auto &w : works) w.ledger = false;
        RunSink k;
        k.beamlets = true;     // no sums wanted: only the transport is timed
        return run_sink(c, k, works, nullptr);
    }, hist, best, hps);
}

int omc_dev_build_photon(int device, omc_tables *t, const char *data_folder)
{
    if (!t || !data_folder) return fail(OMC_EARG, "null argument");
    const int nmed = t->nmed;
    if (nmed < 1 || nmed > OMC_MXMED) return fail(OMC_ERANGE, "nmed %d", nmed);
    double *outs[8] = { t->gmfp1, t->gmfp0, t->gbr11, t->gbr10, t->gbr21, t->gbr20, t->cohe1, t->cohe0 };
    for (double *p : outs) if (!p) return fail(OMC_EARG, "photon arrays not allocated");
    // per-medium scalars and the element order, as the host builder forms them
    std::vector<PhotonBuildMedium> meds((size_t)nmed);
    for (int i = 0; i < nmed; ++i) {
        PhotonBuildMedium &m = meds[(size_t)i];
        memset(&m, 0, sizeof m);
        const int ne = t->ne[i];
        if (ne < 1 || ne > OMC_MXELEMENT) return fail(OMC_ERANGE, "medium %d: %d elements", i, ne);
        t->ge1[i] = (double)(OMC_MXGE - 1) / log(t->up[i] / t->ap[i]);
        t->ge0[i] = 1.0 - t->ge1[i] * log(t->ap[i]);
        double sum_a = 0.0, sum_z = 0.0;
        for (int j = 0; j < ne; ++j) {
            sum_a += t->elements[i][j].pz * t->elements[i][j].wa;
            sum_z += t->elements[i][j].pz * t->elements[i][j].z;
        }
        int idx[OMC_MXELEMENT];
        for (int j = 0; j < ne; ++j) idx[j] = j;
        for (int j = 1; j < ne; ++j) {           // ascending Z (distinct Z: one order)
            const int k = idx[j];
            int q = j - 1;
            while (q >= 0 && t->elements[i][idx[q]].z > t->elements[i][k].z) { idx[q + 1] = idx[q]; --q; }
            idx[q + 1] = k;
        }
        for (int j = 0; j < ne; ++j) {
            m.z[j] = t->elements[i][idx[j]].z;
            m.pz[j] = t->elements[i][idx[j]].pz;
            const int zi = (int)(m.z[j] + 0.5);
            if (zi < 1 || zi > OMC_MXELEMENT)
                return fail(OMC_ERANGE, "medium %d: element Z = %d outside the XCOM data (1..%d)", i, zi, OMC_MXELEMENT);
        }
        m.ne = ne;
        m.ge0 = t->ge0[i];
        m.ge1 = t->ge1[i];
        m.sum_z = sum_z;
        m.con2 = t->rho[i] / (sum_a * 1.6605655);
    }
    omc_xcom_data x;
    if (omc_xcom_load(&x, data_folder) != OMC_OK) return fail(OMC_EIO, "%s", omc_last_error());
    HIPCHK(hipSetDevice(device));
    const size_t G = (size_t)nmed * OMC_MXGE;
    const size_t bytes_meds = sizeof(PhotonBuildMedium) * (size_t)nmed;
    const size_t bytes_n = sizeof(int) * 4 * OMC_MXELEMENT, bytes_off = sizeof(int64_t) * 4 * OMC_MXELEMENT;
    const size_t bytes_pts = sizeof(double) * (size_t)x.total;
    char *d = nullptr;
    const size_t total = bytes_meds + bytes_n + bytes_off + 2 * bytes_pts + sizeof(double) * (4 * G + 8 * G) + 6 * 256;
    int rc = OMC_OK;
    if (hipMalloc((void **)&d, total) != hipSuccess) {
        omc_xcom_free(&x);
        return fail(OMC_EDEVICE, "hipMalloc(%zu) for the photon tables", total);
    }
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char *p = d;
    PhotonBuildMedium *d_meds = (PhotonBuildMedium *)p; p += up(bytes_meds);
    int *d_n = (int *)p;                                 p += up(bytes_n);
    int64_t *d_off = (int64_t *)p;                       p += up(bytes_off);
    double *d_le = (double *)p;                          p += up(bytes_pts);
    double *d_ls = (double *)p;                          p += up(bytes_pts);
    double *d_nodes = (double *)p;                       p += up(sizeof(double) * 4 * G);
    double *d_tab = (double *)p;
    if (hipMemcpy(d_meds, meds.data(), bytes_meds, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_n, x.n, bytes_n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_off, x.off, bytes_off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_le, x.le, bytes_pts, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_ls, x.ls, bytes_pts, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(OMC_EDEVICE, "copy the XCOM data");
    if (!rc) {
        XcomView xv{ d_n, d_off, d_le, d_ls };
        launch_photon_build(d_meds, nmed, xv, d_nodes, d_tab, nullptr);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail(OMC_EDEVICE, "photon table kernels");
    }
    for (int q = 0; !rc && q < 8; ++q)
        if (hipMemcpy(outs[q], d_tab + (size_t)q * G, sizeof(double) * G, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(OMC_EDEVICE, "copy the photon tables");
    (void)hipFree(d);
    omc_xcom_free(&x);
    return rc;
}

int omc_dev_math_eval(int device, int precision, int fn, int64_t n, const double *x, const double *y, double *out)
{
    if (!x || !out || n < 0 || fn < 0 || fn > 5) return fail(OMC_EARG, "bad argument");
    if (precision != OMC_PREC_FP64 && precision != OMC_PREC_FP32) return fail(OMC_EARG, "precision %d", precision);
    if (n == 0) return OMC_OK;
    HIPCHK(hipSetDevice(device));
    double *d = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n;
    HIPCHK(hipMalloc((void **)&d, 3 * bytes));
    int rc = OMC_OK;
    if (hipMemcpy(d, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, y ? y : x, bytes, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(OMC_EDEVICE, "copy arguments");
    if (!rc) {
        if (precision == OMC_PREC_FP64) Launch<double>::math_eval(fn, n, d, d + n, d + 2 * n, nullptr);
        else Launch<float>::math_eval(fn, n, d, d + n, d + 2 * n, nullptr);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(OMC_EDEVICE, "math evaluation");
    }
    (void)hipFree(d);
    return rc;
}

int omc_dev_submit(omc_dev_ctx *c, const omc_dev_request *req)
{
    if (!c || !req) return fail(OMC_EARG, "null argument");
    OMC_TRY_DEV(check_phantom(c, req->phantom_id));
    if (req->nhist < 1 || req->nbatch < 1) return fail(OMC_EARG, "nhist and nbatch must be >= 1");
    Work *w = new Work();
    int rc = OMC_OK;
    if (req->kind == 0) {
        if (!req->src_dosxyz) rc = fail(OMC_EARG, "dosxyz request without a source");
        else if (req->src_dosxyz->charge < -1 || req->src_dosxyz->charge > 1) rc = fail(OMC_EARG, "charge");
        else dosxyz_work(*w, req->src_dosxyz);
    } else if (req->kind == 1) {
        if (!req->src_beamlets) rc = fail(OMC_EARG, "beamlet request without a source");
        else rc = beamlet_work(*w, req->src_beamlets, req->beamlet);
    } else {
        rc = fail(OMC_EARG, "request kind %d", req->kind);
    }
    if (!rc) rc = set_histories(*w, req->nhist / req->nbatch, req->nbatch);
    if (rc) { delete w; return rc; }
    stream_fields(*w, req->seed, 0);
    w->phantom = req->phantom_id;
    w->nreg = c->ph[req->phantom_id].nreg;
    w->ticket = c->next_ticket++;
    w->submit_ms = now_ms(c);
    plan_split(c, *w);
    c->tickets[w->ticket] = w;
    c->waiting.push_back(w);
    c->unfinished++;
    return w->ticket;
}

int omc_dev_step(omc_dev_ctx *c)
{
    if (!c) return fail(OMC_EARG, "null argument");
    HIPCHK(hipSetDevice(c->device));
    int r = turn_any(c);
    if (r < 0) return r;
    if (r == 0 && c->unfinished > 0) {
        // requests can only wait while something runs (see admit)
        int rc = fail(OMC_EDEVICE, "scheduler stalled with %d requests left", c->unfinished);
        reset_engine(c, rc);
        return rc;
    }
    return c->unfinished;
}

int omc_dev_poll(omc_dev_ctx *c, int ticket, omc_dev_result *res, double *sum, double *sum2)
{
    if (!c || !res) return fail(OMC_EARG, "null argument");
    auto it = c->tickets.find(ticket);
    if (it == c->tickets.end()) return fail(OMC_EARG, "unknown ticket %d", ticket);
    Work *w = it->second;
    memset(res, 0, sizeof *res);
    res->ticket = ticket;
    res->done = w->state == W_DONE || w->state == W_FAILED;
    res->status = w->status;
    res->submit_ms = w->submit_ms;
    res->admit_ms = w->admit_ms;
    res->finish_ms = w->finish_ms;
    res->nhist = w->nhist;
    if (!res->done) return OMC_OK;
    if (w->state == W_DONE) {
        if (sum) std::copy(w->sum.begin(), w->sum.end(), sum);
        if (sum2) std::copy(w->sum2.begin(), w->sum2.end(), sum2);
    }
    c->tickets.erase(it);
    delete w;
    return OMC_OK;
}

}  // extern "C"
