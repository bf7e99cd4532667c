// Host side of the device engine: the C ABI declared in include/omc_dev.h.
//
// A context owns one HIP stream, the uploaded tables, resident phantoms and
// four particle queues (photons and charged particles, current and next
// generation).  A run is processed in chunks of primaries; each chunk is
// sampled by the source kernel and then transported generation by
// generation until both queues are empty.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

#include "omc_dev.h"
#include "launch.h"
#include "transport.h"

using namespace omc;

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

constexpr int MAX_PHANTOMS = 16;
constexpr int MAX_JOBS = 4096;

struct PhantomSlot {
    bool used = false;
    int nreg = 0;
    int nmed = 0;
    DevBuf xb, yb, zb, med, rhof;
};

}  // namespace

struct omc_dev_ctx {
    int device = 0;
    int precision = OMC_PREC_FP64;
    hipStream_t stream = nullptr;
    int64_t budget = 0;
    int64_t used = 0;
    int64_t chunk = 0;
    uint32_t qcap = 0;
    int cus = 256;

    std::vector<void *> allocs;
    // tables
    bool media = false;
    int nmed = 0;
    void *small = nullptr, *ph_mfp = nullptr, *ph_br = nullptr, *ph_pmax = nullptr;
    void *el = nullptr, *csda = nullptr, *spin = nullptr, *ms = nullptr;
    // phantoms
    PhantomSlot ph[MAX_PHANTOMS];
    void *d_phantoms = nullptr;       // DevPhantom<R>[MAX_PHANTOMS]
    std::vector<unsigned char> h_phantoms;
    // jobs
    DevJob *d_jobs = nullptr;
    // queues: [0] photon cur, [1] photon next, [2] electron cur, [3] electron next
    void *q[4][13] = {};
    uint32_t *d_counts = nullptr;     // 4 queue tails + 2 fetch counters
    DevStatus *d_status = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // serving
    struct Req;
    std::map<int, Req *> reqs;
    int next_ticket = 0;
};

namespace {

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

template <typename R> size_t particle_bytes() { return 8 * sizeof(R) + 6 * sizeof(uint32_t); }

template <typename R> Queue<R> queue_view(omc_dev_ctx *c, int k)
{
    Queue<R> q;
    void **a = c->q[k];
    q.e = (R *)a[0]; q.x = (R *)a[1]; q.y = (R *)a[2]; q.z = (R *)a[3];
    q.u = (R *)a[4]; q.v = (R *)a[5]; q.w = (R *)a[6]; q.wt = (R *)a[7];
    q.ir = (int32_t *)a[8]; q.meta = (uint32_t *)a[9]; q.hist = (uint32_t *)a[10];
    q.k0 = (uint32_t *)a[11]; q.k1 = (uint32_t *)a[12];
    q.ctr = (uint32_t *)a[12] + c->qcap;     // k1 and ctr share one allocation
    q.count = c->d_counts + k;
    q.cap = c->qcap;
    return q;
}

template <typename R> int alloc_queues(omc_dev_ctx *c)
{
    for (int k = 0; k < 4; ++k) {
        for (int f = 0; f < 8; ++f) OMC_TRY_DEV(dalloc(c, &c->q[k][f], sizeof(R) * (size_t)c->qcap));
        for (int f = 8; f < 12; ++f) OMC_TRY_DEV(dalloc(c, &c->q[k][f], 4 * (size_t)c->qcap));
        OMC_TRY_DEV(dalloc(c, &c->q[k][12], 8 * (size_t)c->qcap));
    }
    return OMC_OK;
}

}  // namespace

// ---------------------------------------------------------------- tables --
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
    std::vector<MsEntry> ms(OMC_MS_SIZE);
    for (int i = 0; i < OMC_MS_SIZE; ++i) ms[i] = { t->ums[i], t->wms[i], t->ims[i], 0 };

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
    for (auto &u : up) {
        if (*u.dst) { (void)hipFree(*u.dst); *u.dst = nullptr; }
        OMC_TRY_DEV(dalloc(c, u.dst, u.bytes));
        HIPCHK(hipMemcpy(*u.dst, u.src, u.bytes, hipMemcpyHostToDevice));
    }
    c->media = true;
    c->nmed = nmed;
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
    T.ms = (const MsEntry *)c->ms;
    return T;
}

// --------------------------------------------------------------- phantoms --
template <typename R> static int upload_phantom(omc_dev_ctx *c, const omc_phantom *ph, const omc_regions *reg, int *id)
{
    int slot = -1;
    for (int i = 0; i < MAX_PHANTOMS; ++i) if (!c->ph[i].used) { slot = i; break; }
    if (slot < 0) return fail(OMC_ERANGE, "more than %d resident phantoms", MAX_PHANTOMS);
    if (!c->media) return fail(OMC_EARG, "upload media before phantoms");
    const int nreg = reg->nreg;
    if ((int64_t)nreg != (int64_t)ph->isize * ph->jsize * ph->ksize + 1)
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
    P.nreg = nreg;
    P.xb.bytes = sizeof(double) * (ph->isize + 1);
    P.yb.bytes = sizeof(double) * (ph->jsize + 1);
    P.zb.bytes = sizeof(double) * (ph->ksize + 1);
    P.med.bytes = nreg;
    P.rhof.bytes = sizeof(R) * (size_t)nreg;
    DevBuf *bufs[] = { &P.xb, &P.yb, &P.zb, &P.med, &P.rhof };
    const void *src[] = { ph->xbounds, ph->ybounds, ph->zbounds, med.data(), rhof.data() };
    for (int b = 0; b < 5; ++b) {
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
    *id = slot;
    return OMC_OK;
}

// ---------------------------------------------------------- transport -----
struct RunJob {
    DevJob j;
    int64_t nhist;
};

template <typename R>
static int transport_chunk(omc_dev_ctx *c, const std::vector<std::pair<int, std::pair<uint64_t, uint32_t>>> &src,
                           omc_dev_stats *stats)
{
    // src: (job id, (first history, count)) segments sampled into the queues
    hipStream_t s = c->stream;
    launch_set_u32(c->d_counts, 8, 0u, s);
    Queue<R> qp = queue_view<R>(c, 0), qe = queue_view<R>(c, 2);
    auto t0 = std::chrono::steady_clock::now();
    for (auto &seg : src)
        Launch<R>::source(c->d_jobs, (const DevPhantom<R> *)c->d_phantoms, seg.first, seg.second.first,
                          seg.second.second, qp, qe, c->d_status, s);
    int cur = 0;
    uint32_t counts[4];
    HIPCHK(hipMemcpyAsync(counts, c->d_counts, sizeof counts, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (stats) stats->source_kernel_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    {
        DevStatus st0;
        HIPCHK(hipMemcpy(&st0, c->d_status, sizeof st0, hipMemcpyDeviceToHost));
        if (st0.overflow) return fail(OMC_EOVERFLOW, "source overflow (capacity %u)", c->qcap);
    }
    uint32_t nP = std::min(counts[0], c->qcap), nE = std::min(counts[2], c->qcap);
    const int max_grid = c->cus * 8;
    while (nP || nE) {
        int nxt = cur ^ 1;
        // reset next-generation tails and the fetch counters
        launch_set_u32(c->d_counts + nxt, 1, 0u, s);
        launch_set_u32(c->d_counts + 2 + nxt, 1, 0u, s);
        launch_set_u32(c->d_counts + 4, 2, 0u, s);
        KArgs<R> A;
        A.T = tables_view<R>(c);
        A.phantoms = (const DevPhantom<R> *)c->d_phantoms;
        A.jobs = c->d_jobs;
        A.qp = queue_view<R>(c, nxt);
        A.qe = queue_view<R>(c, 2 + nxt);
        A.status = c->d_status;
        float ms_p = 0, ms_e = 0;
        if (nP) {
            A.qin = queue_view<R>(c, cur);
            A.in_begin = 0;
            A.in_end = nP;
            A.fetch = c->d_counts + 4;
            int grid = std::min<int64_t>(max_grid, ((int64_t)nP + 255) / 256);
            HIPCHK(hipEventRecord(c->ev0, s));
            Launch<R>::photon(A, grid, s);
            HIPCHK(hipEventRecord(c->ev1, s));
            HIPCHK(hipEventSynchronize(c->ev1));
            HIPCHK(hipEventElapsedTime(&ms_p, c->ev0, c->ev1));
            if (stats) { stats->photon_kernel_ms += ms_p; stats->photon_launches++; }
        }
        if (nE) {
            A.qin = queue_view<R>(c, 2 + cur);
            A.in_begin = 0;
            A.in_end = nE;
            A.fetch = c->d_counts + 5;
            int grid = std::min<int64_t>(max_grid, ((int64_t)nE + 255) / 256);
            HIPCHK(hipEventRecord(c->ev0, s));
            Launch<R>::electron(A, grid, s);
            HIPCHK(hipEventRecord(c->ev1, s));
            HIPCHK(hipEventSynchronize(c->ev1));
            HIPCHK(hipEventElapsedTime(&ms_e, c->ev0, c->ev1));
            if (stats) { stats->electron_kernel_ms += ms_e; stats->electron_launches++; }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(counts, c->d_counts, sizeof counts, hipMemcpyDeviceToHost, s));
        DevStatus st;
        HIPCHK(hipMemcpyAsync(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (st.overflow) return fail(OMC_EOVERFLOW, "particle queue overflow (capacity %u)", c->qcap);
        if (stats) {
            stats->generations++;
            stats->transport_ms += ms_p + ms_e;
            stats->photons_pushed += counts[nxt];
            stats->electrons_pushed += counts[2 + nxt];
        }
        nP = std::min(counts[nxt], c->qcap);
        nE = std::min(counts[2 + nxt], c->qcap);
        cur = nxt;
    }
    return OMC_OK;
}

static void fill_source_job(DevJob &j, const omc_spectrum &spec, int charge, double *cdf_dev)
{
    j.charge = charge;
    j.spectrum = spec.spectrum;
    j.energy = spec.energy;
    j.deltak = spec.deltak;
    j.cdf1 = cdf_dev;
    j.cdf2 = cdf_dev + OMC_INVDIM;
}

template <typename R>
static int run_jobs(omc_dev_ctx *c, std::vector<RunJob> &jobs, omc_dev_stats *stats)
{
    if ((int)jobs.size() > MAX_JOBS) return fail(OMC_ERANGE, "more than %d jobs in one run", MAX_JOBS);
    std::vector<DevJob> dj(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) dj[i] = jobs[i].j;
    HIPCHK(hipMemcpy(c->d_jobs, dj.data(), dj.size() * sizeof(DevJob), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->d_status, 0, sizeof(DevStatus)));
    // chunks: consecutive history ranges over the jobs
    std::vector<std::pair<int, std::pair<uint64_t, uint32_t>>> segs;
    int64_t in_chunk = 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        int64_t done = 0;
        while (done < jobs[i].nhist) {
            int64_t take = std::min<int64_t>(jobs[i].nhist - done, c->chunk - in_chunk);
            segs.push_back({ (int)i, { (uint64_t)done, (uint32_t)take } });
            done += take;
            in_chunk += take;
            if (in_chunk == c->chunk) {
                OMC_TRY_DEV(transport_chunk<R>(c, segs, stats));
                if (stats) stats->chunks++;
                segs.clear();
                in_chunk = 0;
            }
        }
    }
    if (!segs.empty()) {
        OMC_TRY_DEV(transport_chunk<R>(c, segs, stats));
        if (stats) stats->chunks++;
    }
    if (stats) {
        DevStatus st;
        HIPCHK(hipMemcpy(&st, c->d_status, sizeof st, hipMemcpyDeviceToHost));
        stats->photon_steps += (int64_t)st.photon_steps;
        stats->electron_steps += (int64_t)st.electron_steps;
        stats->deposits += (int64_t)st.deposits;
        stats->killed += (int64_t)st.killed;
        for (auto &j : jobs) stats->histories += j.nhist;
    }
    return OMC_OK;
}

// ------------------------------------------------------------------ C ABI --
extern "C" {

const char *omc_dev_last_error(void) { return g_err; }

int omc_dev_create(omc_dev_ctx **out, const omc_dev_opts *opts)
{
    if (!out) return fail(OMC_EARG, "null context pointer");
    *out = nullptr;
    omc_dev_opts o = { 0, OMC_PREC_FP64, 0, 0 };
    if (opts) o = *opts;
    if (o.precision != OMC_PREC_FP64 && o.precision != OMC_PREC_FP32)
        return fail(OMC_EARG, "unknown precision %d", o.precision);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (o.device < 0 || o.device >= ndev) return fail(OMC_EDEVICE, "device %d not present (%d devices)", o.device, ndev);
    HIPCHK(hipSetDevice(o.device));
    omc_dev_ctx *c = new omc_dev_ctx();
    c->device = o.device;
    c->precision = o.precision;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, o.device) == hipSuccess) c->cus = prop.multiProcessorCount;
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    c->budget = o.hbm_budget > 0 ? o.hbm_budget : (int64_t)(0.9 * (double)freeb);
    size_t pb = o.precision == OMC_PREC_FP64 ? particle_bytes<double>() : particle_bytes<float>();
    // four queues take at most a quarter of the budget, capped at 16M slots each
    int64_t cap = c->budget / 4 / 4 / (int64_t)pb;
    cap = std::min<int64_t>(cap, 16 << 20);
    cap = std::max<int64_t>(cap, 1 << 16);
    c->qcap = (uint32_t)cap;
    c->chunk = o.chunk > 0 ? std::min<int64_t>(o.chunk, cap / 2) : cap / 4;
    int rc = OMC_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = fail(OMC_EDEVICE, "stream");
    if (!rc) rc = hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess ? OMC_OK : fail(OMC_EDEVICE, "events");
    if (!rc) rc = o.precision == OMC_PREC_FP64 ? alloc_queues<double>(c) : alloc_queues<float>(c);
    if (!rc) rc = dalloc(c, (void **)&c->d_counts, 16 * sizeof(uint32_t));
    if (!rc) rc = dalloc(c, (void **)&c->d_status, sizeof(DevStatus));
    if (!rc) rc = dalloc(c, (void **)&c->d_jobs, MAX_JOBS * sizeof(DevJob));
    size_t phb = o.precision == OMC_PREC_FP64 ? sizeof(DevPhantom<double>) : sizeof(DevPhantom<float>);
    if (!rc) rc = dalloc(c, &c->d_phantoms, MAX_PHANTOMS * phb);
    if (rc) { omc_dev_destroy(c); return rc; }
    *out = c;
    return OMC_OK;
}

int omc_dev_destroy(omc_dev_ctx *c)
{
    if (!c) return OMC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &qq : c->q) for (auto &p : qq) if (p) { (void)hipFree(p); p = nullptr; }
    void *ptrs[] = { c->small, c->ph_mfp, c->ph_br, c->ph_pmax, c->el, c->csda, c->spin, c->ms,
                     c->d_counts, c->d_status, c->d_jobs, c->d_phantoms };
    for (void *p : ptrs) if (p) (void)hipFree(p);
    for (auto &P : c->ph) {
        DevBuf *bufs[] = { &P.xb, &P.yb, &P.zb, &P.med, &P.rhof };
        for (auto *b : bufs) if (b->p) (void)hipFree(b->p);
    }
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OMC_OK;
}

int omc_dev_precision(const omc_dev_ctx *c) { return c ? c->precision : -1; }

int omc_dev_upload_media(omc_dev_ctx *c, const omc_tables *t)
{
    if (!c || !t) return fail(OMC_EARG, "null argument");
    if (t->nmed < 1 || t->nmed > OMC_MXMED) return fail(OMC_ERANGE, "nmed %d", t->nmed);
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
    PhantomSlot &P = c->ph[id];
    DevBuf *bufs[] = { &P.xb, &P.yb, &P.zb, &P.med, &P.rhof };
    for (auto *b : bufs) { dfree(c, b->p, b->bytes); b->p = nullptr; b->bytes = 0; }
    P.used = false;
    return OMC_OK;
}

static int check_run(omc_dev_ctx *c, const omc_dev_run_params *rp)
{
    if (!c || !rp) return fail(OMC_EARG, "null argument");
    if (!c->media) return fail(OMC_EARG, "media tables not uploaded");
    if (rp->phantom_id < 0 || rp->phantom_id >= MAX_PHANTOMS || !c->ph[rp->phantom_id].used)
        return fail(OMC_EARG, "bad phantom id %d", rp->phantom_id);
    if (rp->nsplit != 1)
        return fail(OMC_ENOTSUP, "photon splitting (nsplit=%d) is not supported by the device path", rp->nsplit);
    if (rp->nbatch < 1 || rp->nperbatch < 1) return fail(OMC_EARG, "nbatch and nperbatch must be >= 1");
    if ((int64_t)rp->nbatch * rp->nperbatch > 0xffffffffLL)
        return fail(OMC_ERANGE, "more than 2^32 histories in one run");
    return OMC_OK;
}

int omc_dev_run_dosxyz(omc_dev_ctx *c, const omc_source_dosxyz *src, const omc_dev_run_params *rp,
                       double *sum, double *sum2)
{
    OMC_TRY_DEV(check_run(c, rp));
    if (!src || !sum || !sum2) return fail(OMC_EARG, "null argument");
    if (src->charge < -1 || src->charge > 1) return fail(OMC_EARG, "charge %d", src->charge);
    HIPCHK(hipSetDevice(c->device));
    const PhantomSlot &P = c->ph[rp->phantom_id];
    const int64_t nreg = P.nreg;
    const int64_t nhist = rp->nperbatch * (int64_t)rp->nbatch;
    double *tally = nullptr, *cdf = nullptr, *led = nullptr, *fs = nullptr;
    size_t tb = sizeof(double) * (size_t)nreg * rp->nbatch;
    size_t lb = rp->ledger ? sizeof(double) * 3 * (size_t)nhist : 0;
    int rc = dalloc(c, (void **)&tally, tb);
    if (!rc) rc = dalloc(c, (void **)&cdf, 2 * OMC_INVDIM * sizeof(double));
    if (!rc && lb) rc = dalloc(c, (void **)&led, lb);
    if (!rc) rc = dalloc(c, (void **)&fs, 2 * sizeof(double) * (size_t)nreg);
    if (!rc) {
        launch_zero_f64(tally, (int64_t)nreg * rp->nbatch, c->stream);
        if (led) launch_zero_f64(led, 3 * nhist, c->stream);
        std::vector<double> cd(2 * OMC_INVDIM);
        memcpy(cd.data(), src->spec.cdfinv1, OMC_INVDIM * sizeof(double));
        memcpy(cd.data() + OMC_INVDIM, src->spec.cdfinv2, OMC_INVDIM * sizeof(double));
        if (hipMemcpy(cdf, cd.data(), cd.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(OMC_EDEVICE, "copy spectrum");
    }
    if (!rc) {
        RunJob j;
        memset(&j, 0, sizeof j);
        j.j.phantom = rp->phantom_id;
        j.j.kind = 0;
        fill_source_job(j.j, src->spec, src->charge, cdf);
        j.j.ssd = src->ssd; j.j.xinl = src->xinl; j.j.yinl = src->yinl;
        j.j.xsize = src->xsize; j.j.ysize = src->ysize;
        j.j.ixinl = src->ixinl; j.j.iyinl = src->iyinl;
        j.j.hist0 = (uint64_t)rp->hist0;
        j.j.nperbatch = rp->nperbatch;
        j.j.s0 = rp->seed[0]; j.j.s1 = rp->seed[1];
        j.j.tally = tally;
        j.j.ledger = led;
        j.nhist = nhist;
        std::vector<RunJob> jobs(1, j);
        rc = c->precision == OMC_PREC_FP64 ? run_jobs<double>(c, jobs, rp->stats) : run_jobs<float>(c, jobs, rp->stats);
    }
    if (!rc) {
        launch_fold(tally, rp->nbatch, nreg, fs, fs + nreg, c->stream);
        std::vector<double> h(2 * (size_t)nreg);
        if (hipMemcpyAsync(h.data(), fs, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(OMC_EDEVICE, "copy tallies");
        for (int64_t i = 0; !rc && i < nreg; ++i) { sum[i] += h[i]; sum2[i] += h[nreg + i]; }
        if (!rc && led && hipMemcpy(rp->ledger, led, lb, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(OMC_EDEVICE, "copy ledger");
    }
    dfree(c, tally, tb);
    dfree(c, cdf, 2 * OMC_INVDIM * sizeof(double));
    dfree(c, led, lb);
    dfree(c, fs, 2 * sizeof(double) * (size_t)nreg);
    return rc;
}

}  // extern "C"
