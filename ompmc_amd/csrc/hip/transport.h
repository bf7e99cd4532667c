// Event-based transport: per-lane state machines for photons and charged
// particles, and the kernels that drive them over structure-of-arrays
// particle queues.
//
// ompMC runs each history depth-first through a per-thread stack
// (shower(), ompmc.c:5436-5447).  Here a wavefront's 64 lanes each own one
// particle of a single type; one loop trip advances every lane by one photon
// voxel step or one electron ustep iteration, and a lane whose particle ends
// refills itself from the input queue.  Secondaries go to the next
// generation's queues.  Because every particle draws from its own Philox
// stream, the order in which particles are processed does not change any
// history.
#pragma once

#include "physics_e.h"

namespace omc {

constexpr double SGMFP = 1.0E-05;
// Loop trips one particle may take over all launches before it is given up
// (counted in DevStatus::killed); far above any physical history.
constexpr uint32_t MAX_ITERS = 1u << 22;

template <typename R> struct KArgs {
    DevTables<R> T;
    const DevPhantom<R> *phantoms;
    const DevJob *jobs;
    Queue<R> qin;            // particles consumed by this launch
    const uint32_t *in_count;    // their number, final before the launch starts
    uint32_t *fetch;         // fetch counter (zero at launch)
    Queue<R> qp;             // photon output queue
    Queue<R> qe;             // charged-particle output queue
    DevStatus *status;
    uint32_t yield_trips;    // loop trips a particle may take per launch
    uint32_t drain_lanes;    // once the input is used up, a wave with at most
                             // this many live lanes hands its particles on...
    uint32_t drain_trips;    // ...if they have taken this many trips here...
    uint32_t drain_min;      // ...and the launch's input holds at least this many
    uint32_t event_div;      // photon events run once 1/event_div of the live lanes wait
    uint32_t fetch_min;      // free lanes take new particles once this many are free
    uint32_t woodcock;       // photon flights by delta tracking where the phantom allows
    uint32_t *longest;       // trips of the longest track that ended in this launch (atomicMax)
};

// Per-lane geometry / scoring context for the particle the lane holds.  The
// phantom records are copied to LDS at kernel start, so a lane keeps one
// pointer to its record instead of the record's fields.
template <typename R> struct LaneGeo {
    const DevPhantom<R> *g;  // LDS copy of the particle's phantom record
    gmem<double> *tally;     // job tally row of this history's batch
    gmem<double> *ledger;    // job ledger entry of this history (or null)

    __device__ __forceinline__ void bind(const KArgs<R> &A, const DevPhantom<R> *lds, const Particle<R> &p)
    {
        const DevJob &J = A.jobs[p.job];
        g = &lds[J.phantom];
        uint32_t batch = (uint32_t)((uint64_t)p.hist / (uint64_t)J.nperbatch);
        tally = glob(J.tally) + (size_t)batch * (size_t)g->nreg;
        ledger = J.ledger ? glob(J.ledger) + OMC_LEDGER * (size_t)p.hist : nullptr;
    }
    __device__ __forceinline__ int medium(int ir) const { int m = glob(g->med)[ir]; return m == 0xff ? -1 : m; }
    __device__ __forceinline__ R rho(int ir) const { return glob(g->rhof)[ir]; }
};

// Copies the resident phantom records to LDS (all threads of the block).
template <typename R>
__device__ __forceinline__ void stage_phantoms(const DevPhantom<R> *src, DevPhantom<R> *lds)
{
    constexpr int words = (int)(sizeof(DevPhantom<R>) * MAX_PHANTOMS / 4);
    static_assert(sizeof(DevPhantom<R>) % 4 == 0, "phantom record must be word sized");
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(lds);
    for (int i = threadIdx.x; i < words; i += blockDim.x) d[i] = s[i];
    __syncthreads();
}

// Deposit accumulator: consecutive deposits into the same region are summed
// in a register and written with one FP64 atomic when the region changes.
template <typename R> struct Scorer {
    int ir = -1;
    double acc = 0.0;
    uint32_t ndep = 0;
    __device__ __forceinline__ void flush(const LaneGeo<R> &G)
    {
        if (ir >= 0) {
            atomicAdd((double *)(G.tally + ir), acc);
            if (G.ledger) {
                atomicAdd((double *)(G.ledger + (ir > 0 ? 0 : 1)), acc);
                // path checksum: energy times the region it went to
                if (ir > 0) atomicAdd((double *)(G.ledger + 3), acc * (double)ir);
            }
        }
        ir = -1;
        acc = 0.0;
    }
    // ausgab (omc_dosxyz.c:683-694): wt * edep into region ir
    __device__ __forceinline__ void add(const LaneGeo<R> &G, int region, double e)
    {
        ndep++;
        if (region != ir) {
            flush(G);
            ir = region;
        }
        acc += e;
    }
};

template <typename R>
__device__ __forceinline__ void push(const Queue<R> &q, const Particle<R> &p, DevStatus *st)
{
    uint32_t i = wave_reserve(q.count, 1);
    if (i < q.cap) q_store(q, i, p);
    else atomicOr(&st->overflow, 1u);
}

// Packing of a lane's integer variables into one word of the resume state:
// bits 0-1 phase, 2 compute_tstep, 3-6 medium + 1, 7-10 previous medium + 1,
// 11-22 table bin + 1.
__device__ __forceinline__ uint32_t pack_lane(int phase, int flag, int imed, int medold, int bin)
{
    return (uint32_t)phase | ((uint32_t)flag << 2) | ((uint32_t)(imed + 1) << 3) | ((uint32_t)(medold + 1) << 7) |
           ((uint32_t)(bin + 1) << 11);
}
__device__ __forceinline__ void unpack_lane(uint32_t w, int &phase, int &flag, int &imed, int &medold, int &bin)
{
    phase = (int)(w & 3u);
    flag = (int)((w >> 2) & 1u);
    imed = (int)((w >> 3) & 15u) - 1;
    medold = (int)((w >> 7) & 15u) - 1;
    bin = (int)((w >> 11) & 4095u) - 1;
}

// Hands a particle that is still in flight to the next launch together with
// the state its lane needs to continue (see PhotonLane / ElectronLane).
template <typename R, typename Lane>
__device__ __forceinline__ void yield_push(const Queue<R> &q, const Particle<R> &p, const Lane &s, uint32_t trips,
                                           DevStatus *st)
{
    uint32_t i = wave_reserve(q.count, 1);
    if (i < q.cap) {
        q_store(q, i, p, META_RESUME);
        s.save(q, i);
        q.trips()[i] = trips;
    } else {
        atomicOr(&st->overflow, 1u);
    }
}

// ============================================================ photons ======
template <typename R> struct PhotonLane {
    int phase;               // 0: (re)enter photon(), 1: transport step, 2: interaction
    int imed, lgle;
    R gle, dpmfp, gmfp, cohfac, eig;
    // Not part of the saved state: the density of the current voxel and the
    // table values of medium cmed at energy gle (gmfpr0, cohfac, lgle),
    // reused while the photon keeps its energy and medium (-2 = none).
    R rhof, gmfpr0;
    int cmed;
    // Photon splitting in progress (0 = none): nsplit | i_survive << 10 |
    // (index of the next sub-photon) << 20, the stratification variable
    // eta' and the mean free paths used so far.  A lane does not yield while
    // it splits, so none of this is saved.
    uint32_t split;
    R eta, dpmfp_old;

    // After the particle's voxel or energy was set from outside the step.
    __device__ __forceinline__ void attach(const LaneGeo<R> &G, const Particle<R> &p)
    {
        rhof = G.rho(p.ir);
        cmed = -2;
    }

    __device__ __forceinline__ void save(const Queue<R> &q, uint32_t i) const
    {
        q.st(0)[i] = gle; q.st(1)[i] = dpmfp; q.st(2)[i] = gmfp; q.st(3)[i] = cohfac; q.st(4)[i] = eig;
        q.sflags()[i] = pack_lane(phase, 0, imed, -1, lgle);
    }
    __device__ __forceinline__ void load(const Queue<R> &q, uint32_t i)
    {
        gle = q.st(0)[i]; dpmfp = q.st(1)[i]; gmfp = q.st(2)[i]; cohfac = q.st(3)[i]; eig = q.st(4)[i];
        int flag, medold;
        unpack_lane(q.sflags()[i], phase, flag, imed, medold, lgle);
        split = 0u;
    }
};

// A copy of the lane's particle that leaves the lane while the lane's stream
// goes on (photon splitting) takes a child stream of its own.
template <typename R>
__device__ __forceinline__ void detach_stream(const Particle<R> &p, Particle<R> &q)
{
    p.rng.child(q.rng.k0, q.rng.k1, 2u);
    q.rng.ctr = 0u;
    q.rng.reset_cache();
}

// Photon splitting (ompmc.c:1903-2118): the photon that enters photon() is
// followed as nsplit sub-photons of weight wt/nsplit along its direction,
// each starting where the previous one interacted, over the stratified mean
// free paths -log(eta') (eta' stepping down by 1/nsplit).  A charged
// particle a sub-photon sets in motion goes on with weight wt/nsplit; a
// scattered photon only from the one sub-photon drawn to survive, with its
// weight restored, and it is split again when it enters photon().  The lane
// keeps the sub-photon; each interaction works on a copy.
//
// Next sub-photon, or false when all are done (ompmc.c:1924-1935).
template <typename R>
__device__ __forceinline__ bool next_split(Particle<R> &p, PhotonLane<R> &s, LaneGeo<R> &G)
{
    const uint32_t nsplit = s.split & 1023u, isplit = s.split >> 20;
    if (isplit >= nsplit) return false;
    s.eta -= (R)1.0 / (R)nsplit;
    if (s.eta <= (R)0.0) return false;
    const R dpmfp = -m_log(s.eta) - s.dpmfp_old;
    s.dpmfp_old += dpmfp;
    s.dpmfp = dpmfp;
    s.split += 1u << 20;
    s.imed = G.medium(p.ir);
    s.phase = 1;
    return true;
}

// Interaction of a sub-photon (ompmc.c:2022-2093), then the next one.
template <typename R>
__device__ __forceinline__ bool split_interaction(const KArgs<R> &A, Particle<R> &p, PhotonLane<R> &s,
                                                  LaneGeo<R> &G)
{
    const DevTables<R> &T = A.T;
    const uint32_t nsplit = s.split & 1023u;
    const bool survivor = ((s.split >> 10) & 1023u) == (s.split >> 20) - 1u;
    const int imed = s.imed;
    Particle<R> q = p;
    R rnno = urand<R>(q.rng);
    if (rnno <= (R)1.0 - s.cohfac) {
        if (survivor) {
            count_event_global(A.status, EV_RAYLEIGH);
            q.wt *= (R)nsplit;
            rayleigh<R>(T, q, imed, s.eig, s.gle, s.lgle);
            p.rng = q.rng;
            detach_stream<R>(p, q);
            push<R>(A.qp, q, A.status);
        } else {
            p.rng = q.rng;
        }
        return next_split<R>(p, s, G);
    }
    rnno = urand<R>(q.rng);
    R4<R> br = T.ph_br[imed * MXGE + s.lgle];
    R gbr1 = s.gle * br.a1 + br.a0;
    Particle<R> b;
    if (rnno <= gbr1 && s.eig > (R)2.0 * (R)RM) {
        count_event_global(A.status, EV_PAIR);
        pair<R>(T, q, b, imed);
        p.rng = q.rng;
        detach_stream<R>(p, q);
        push<R>(A.qe, q, A.status);
        push<R>(A.qe, b, A.status);
    } else if (rnno < s.gle * br.b1 + br.b0) {
        count_event_global(A.status, EV_COMPTON);
        compton<R>(q, b);
        p.rng = q.rng;
        push<R>(A.qe, b, A.status);
        if (survivor) {
            q.wt *= (R)nsplit;
            detach_stream<R>(p, q);
            push<R>(A.qp, q, A.status);
        }
    } else {
        count_event_global(A.status, EV_PHOTO);
        photo<R>(q, (R)G.g->ecut[imed + 1]);
        p.rng = q.rng;
        detach_stream<R>(p, q);
        push<R>(A.qe, q, A.status);
    }
    return next_split<R>(p, s, G);
}

// A photon lane alternates transport steps (phase 1) with events: the
// interaction at the end of a step (phase 2) and the (re-)entry into
// photon() (phase 0) that follows it or starts a particle.  Events touch
// few lanes per step, so the kernel runs them for all waiting lanes of a
// wavefront together; the operations of each photon and their order are
// those of the reference's loop.

// Event: the pending interaction (ompmc.c:2026-2068), then the entry into
// photon() (ompmc.c:1884-1935, nsplit = 1) of the photon that goes on.
// Returns true while the photon is still owned by the lane.  Electrons it
// sets in motion are pushed to A.qe.
template <typename R>
__device__ __forceinline__ bool photon_event(const KArgs<R> &A, Particle<R> &p, PhotonLane<R> &s, LaneGeo<R> &G,
                                             Scorer<R> &sc)
{
    const DevTables<R> &T = A.T;
    if (s.phase == 2) {
        if (s.split) return split_interaction<R>(A, p, s, G);
        const int imed = s.imed;
        R rnno = urand<R>(p.rng);
        if (rnno <= (R)1.0 - s.cohfac) {
            count_event_global(A.status, EV_RAYLEIGH);
            rayleigh<R>(T, p, imed, s.eig, s.gle, s.lgle);
        } else {
            rnno = urand<R>(p.rng);
            R4<R> br = T.ph_br[imed * MXGE + s.lgle];
            R gbr1 = s.gle * br.a1 + br.a0;
            Particle<R> b;
            if (rnno <= gbr1 && s.eig > (R)2.0 * (R)RM) {
                count_event_global(A.status, EV_PAIR);
                pair<R>(T, p, b, imed);
                push<R>(A.qe, p, A.status);
                push<R>(A.qe, b, A.status);
                return false;
            }
            R gbr2 = s.gle * br.b1 + br.b0;
            if (rnno < gbr2) {
                count_event_global(A.status, EV_COMPTON);
                compton<R>(p, b);
                push<R>(A.qe, b, A.status);
            } else {
                count_event_global(A.status, EV_PHOTO);
                photo<R>(p, (R)G.g->ecut[imed + 1]);
                push<R>(A.qe, p, A.status);
                return false;
            }
        }
    }
    // photon() entry
    int irl = p.ir;
    s.imed = G.medium(irl);
    s.attach(G, p);
    s.eig = p.e;
    if (p.e <= (R)G.g->pcut[s.imed + 1] || p.wt == (R)0) {
        sc.add(G, irl, (double)p.wt * (double)p.e);
        return false;
    }
    s.gle = m_log(p.e);
    const int nsplit = A.jobs[p.job].nsplit;
    if (nsplit > 1) {
        R rnno = urand<R>(p.rng);
        rnno /= (R)nsplit;
        s.eta = (R)1.0 - rnno + (R)1.0 / (R)nsplit;
        const uint32_t isurv = (uint32_t)(urand<R>(p.rng) * (R)nsplit);
        s.split = (uint32_t)nsplit | isurv << 10;
        s.dpmfp_old = (R)0.0;
        p.wt /= (R)nsplit;
        return next_split<R>(p, s, G);
    }
    R rnno = urand<R>(p.rng);
    rnno /= (R)1.0;
    R d_eta = (R)1.0;
    R eta_prime = (R)1.0 - rnno + d_eta;
    (void)urand<R>(p.rng);          // survivor selection, trivial for nsplit = 1
    eta_prime -= d_eta;
    s.dpmfp = -m_log(eta_prime) - (R)0.0;
    s.phase = 1;
    return true;
}

// Voxel crossings a photon trip traces ahead and runs.  On the slab phantom
// FP32 trips of 1 / 4 / 8 / 12 / 16 crossings gave 61 / 74 / 80 / 81 / 81 M
// histories/s (16 spills registers); FP64 fits 12 without spills.
#ifndef OMC_PHOTON_STEPS_32
#define OMC_PHOTON_STEPS_32 12
#endif
#ifndef OMC_PHOTON_STEPS_64
#define OMC_PHOTON_STEPS_64 12
#endif
template <typename R> struct PhotonSteps;
template <> struct PhotonSteps<float> { static constexpr int value = OMC_PHOTON_STEPS_32; };
template <> struct PhotonSteps<double> { static constexpr int value = OMC_PHOTON_STEPS_64; };

// One trip of photon transport: up to PhotonSteps<R> iterations of the
// reference's transport loop (ompmc.c:1951-2019).  A straight run of voxel
// crossings needs the medium and density of each voxel entered, and fetching
// them one crossing at a time puts a memory round trip into every
// iteration.  So the trip first traces the straight path ahead through up to
// PhotonSteps<R> voxel boundaries with howfar (with no step limit the axis
// it picks is the one the step picks whenever the step reaches a boundary)
// and moves a copy of the position as the steps will, then loads all media
// and densities at once, then runs the iterations on them.  It stops at the
// first iteration that ends inside its voxel: an interaction (phase 2), or a
// step cut short by rounding, after which the traced path no longer applies.
// Returns true while the photon is still owned by the lane.
template <typename R>
__device__ __forceinline__ void advance(R &x, R &y, R &z, R u, R v, R w, R ustep)
{
    x += ustep * u;
    y += ustep * v;
    z += ustep * w;
}

template <typename R>
__device__ __forceinline__ bool photon_step(const KArgs<R> &A, Particle<R> &p, PhotonLane<R> &s, LaneGeo<R> &G,
                                            uint32_t &nsteps, uint32_t &ncross)
{
    const DevTables<R> &T = A.T;
    constexpr int K = PhotonSteps<R>::value;
    int irk[K], medk[K];
    R dk[K], rhok[K];
    int n = 0;
    {
        int irc = p.ir;
        R x = p.x, y = p.y, z = p.z;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (irc != 0 && n == k) {
                int irn = irc;
                R us = (R)3.0E38;
                howfar<R>(*G.g, irc, x, y, z, p.u, p.v, p.w, irn, us);
                advance<R>(x, y, z, p.u, p.v, p.w, us);
                irk[k] = irn;
                dk[k] = us;
                irc = irn;
                n = k + 1;
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < n) {
                medk[k] = G.medium(irk[k]);
                rhok[k] = G.rho(irk[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (k > 0 && k >= n) break;
        nsteps++;
        const int irl = p.ir;
        R tstep;
        if (s.imed != -1) {
            if (s.imed != s.cmed) {
                const MediumScalars &M = T.small->med[s.imed];
                s.lgle = pwlf_bin(s.gle, M.ge1, M.ge0, MXGE);
                R4<R> c = T.ph_mfp[s.imed * MXGE + s.lgle];
                s.gmfpr0 = s.gle * c.a1 + c.a0;
                s.cohfac = s.gle * c.b1 + c.b0;
                s.cmed = s.imed;
            }
            s.gmfp = s.gmfpr0 / s.rhof;
            s.gmfp *= s.cohfac;
            tstep = s.gmfp * s.dpmfp;
        } else {
            tstep = (R)1.0E8;
        }
        if (irl == 0) {                    // left the phantom: discard
            if (G.ledger) atomicAdd((double *)(G.ledger + 2), (double)p.wt * (double)p.e);
            return false;
        }
        // howfar with the step limit tstep: the traced boundary if nearer
        R ustep = tstep;
        int irnew = irl;
        if (dk[k] < ustep) {
            ustep = dk[k];
            irnew = irk[k];
        }
        advance<R>(p.x, p.y, p.z, p.u, p.v, p.w, ustep);
        if (s.imed != -1) s.dpmfp = m_fmax((R)0.0, s.dpmfp - ustep / s.gmfp);
        if (irnew == irl) {
            if (s.imed != -1 && s.dpmfp <= (R)SGMFP) s.phase = 2;
            return true;
        }
        ncross++;
        p.ir = irnew;
        s.imed = medk[k];
        s.rhof = rhok[k];
        if (s.imed != -1 && s.dpmfp <= (R)SGMFP) {
            s.phase = 2;
            return true;
        }
    }
    return true;
}

// Photon flight by delta (Woodcock) tracking, in place of photon_step for a
// photon that is not splitting, in a phantom with a uniform grid whose media
// share one energy grid.  Photons deposit nothing in flight, so only the
// interaction sites matter: the flight runs dpmfp mean free paths of the
// majorant mu_max(E) >= mu(E, voxel), and at the site a real interaction
// happens with probability mu/mu_max (otherwise a fresh exponential flight
// follows).  The sites have the distribution voxel stepping gives, without a
// stop at every voxel boundary: one medium and density lookup per site.
// At a real site the lane holds the voxel's medium, density and table values
// as photon_step would, and the interaction event follows.  A site outside
// the grid ends the photon (it left the phantom).  A trip visits up to
// OMC_WOODCOCK_SITES sites (FP32 slab / beamlets, histories/s: 1 site
// 1.31e8 / 1.00e8, 2 1.48e8 / 1.00e8, 4 1.44e8 / 1.08e8, 8 1.48e8 / 1.11e8).
#ifndef OMC_WOODCOCK_SITES
#define OMC_WOODCOCK_SITES 8
#endif
template <typename R>
__device__ __forceinline__ bool photon_flight(const KArgs<R> &A, Particle<R> &p, PhotonLane<R> &s, LaneGeo<R> &G,
                                              uint32_t &nsteps, uint32_t &ncross)
{
    const DevTables<R> &T = A.T;
    const DevPhantom<R> &g = *G.g;
    const Woodcock<R> &W = g.wc;
    const R mu_max = glob(W.mu)[pwlf_bin(s.gle, (R)W.ge1, (R)W.ge0, MXGE)];
    for (int k = 0; k < OMC_WOODCOCK_SITES; ++k) {
        nsteps++;
        advance<R>(p.x, p.y, p.z, p.u, p.v, p.w, s.dpmfp / mu_max);
        const R fx = (p.x - (R)W.x0) * W.idx, fy = (p.y - (R)W.y0) * W.idy, fz = (p.z - (R)W.z0) * W.idz;
        if (!(fx >= (R)0 && fy >= (R)0 && fz >= (R)0 && fx <= (R)g.isize && fy <= (R)g.jsize &&
              fz <= (R)g.ksize)) {                                     // left the phantom: discard
            if (G.ledger) atomicAdd((double *)(G.ledger + 2), (double)p.wt * (double)p.e);
            return false;
        }
        const int ix = min((int)fx, g.isize - 1), iy = min((int)fy, g.jsize - 1), iz = min((int)fz, g.ksize - 1);
        const int ir = 1 + ix + iy * g.isize + iz * g.isize * g.jsize;
        ncross++;
        p.ir = ir;
        const int m = G.medium(ir);
        s.imed = m;
        if (m >= 0) {
            const R rho = G.rho(ir);
            if (m != s.cmed) {
                const MediumScalars &M = T.small->med[m];
                s.lgle = pwlf_bin(s.gle, M.ge1, M.ge0, MXGE);
                R4<R> c = T.ph_mfp[m * MXGE + s.lgle];
                s.gmfpr0 = s.gle * c.a1 + c.a0;
                s.cohfac = s.gle * c.b1 + c.b0;
                s.cmed = m;
            }
            const R mu = rho / (s.gmfpr0 * s.cohfac);
            if (mu > mu_max) atomicAdd(&A.status->wc_over, 1ull);
            if (urand<R>(p.rng) * mu_max <= mu) {                     // real interaction
                s.rhof = rho;
                s.gmfp = s.gmfpr0 / rho * s.cohfac;
                s.dpmfp = (R)0.0;
                s.phase = 2;
                return true;
            }
        }
        s.dpmfp = -m_log((R)1.0 - urand<R>(p.rng));
    }
    return true;
}

// ================================================== electrons/positrons ====
template <typename R> struct ElectronLane {
    int phase;               // 0 entry, 1 head of tstep loop, 2 ustep iteration
    int imed, medold, lelke, compute_tstep;
    R rhof, eke, elke, demfp, sig0, ustep;

    __device__ __forceinline__ void save(const Queue<R> &q, uint32_t i) const
    {
        q.st(0)[i] = rhof; q.st(1)[i] = eke; q.st(2)[i] = elke;
        q.st(3)[i] = demfp; q.st(4)[i] = sig0; q.st(5)[i] = ustep;
        q.sflags()[i] = pack_lane(phase, compute_tstep, imed, medold, lelke);
    }
    __device__ __forceinline__ void load(const Queue<R> &q, uint32_t i)
    {
        rhof = q.st(0)[i]; eke = q.st(1)[i]; elke = q.st(2)[i];
        demfp = q.st(3)[i]; sig0 = q.st(4)[i]; ustep = q.st(5)[i];
        unpack_lane(q.sflags()[i], phase, compute_tstep, imed, medold, lelke);
    }
};

enum { E_ALIVE = 0, E_DONE = 1 };

// Photons a charged particle makes go to the photon queue; while photon
// splitting is on each first plays Russian roulette (ompmc.c:4149-4164,
// 4338-4353, 4627-4642): it survives with probability 1/nsplit, its weight
// multiplied by nsplit.  The draws come from the stream of the particle the
// lane holds.
template <typename R>
__device__ __forceinline__ bool roulette(const KArgs<R> &A, Rng &rng, int job, Particle<R> &g)
{
    const int ns = A.jobs[job].nsplit;
    if (ns <= 1) return true;
    if (urand<R>(rng) * (R)ns > (R)1.0) return false;
    g.wt *= (R)ns;
    return true;
}

template <typename R>
__device__ __forceinline__ void push_made_photon(const KArgs<R> &A, Particle<R> &p, Particle<R> &b)
{
    if (roulette<R>(A, p.rng, p.job, b)) push<R>(A.qp, b, A.status);
}

// Both annihilation photons: a is the lane's own particle, whose stream makes
// both roulette draws before it leaves with a.
template <typename R>
__device__ __forceinline__ void push_photon_pair(const KArgs<R> &A, Particle<R> &a, Particle<R> &b)
{
    const bool ka = roulette<R>(A, a.rng, a.job, a);
    const bool kb = roulette<R>(A, a.rng, a.job, b);
    if (ka) push<R>(A.qp, a, A.status);
    if (kb) push<R>(A.qp, b, A.status);
}

// electron_cutoff: kinetic energy deposited; a positron then annihilates at
// rest and the lane hands both photons to the photon queue.
template <typename R>
__device__ __forceinline__ void electron_cutoff(const KArgs<R> &A, Particle<R> &p, LaneGeo<R> &G,
                                                Scorer<R> &sc, R eie)
{
    R edep = p.e - (R)RM;
    sc.add(G, p.ir, (double)p.wt * (double)edep);
    if (p.iq > 0 && edep < eie) {
        count_event(EV_RANNIH);
        Particle<R> b;
        rannih<R>(p, b);
        push_photon_pair<R>(A, p, b);
    }
}

// One ustep iteration of electron() (ompmc.c:4765-5341) plus, where the
// reference reaches them, the tstep-loop head (4694-4763), the fictitious
// rejection (5347-5372) and the discrete interaction (5375-5429).
// End of the ustep loop (ompmc.c:5347-5433): fictitious-cross-section
// rejection, then the discrete interaction.
template <typename R>
__device__ __forceinline__ int electron_interaction(const KArgs<R> &A, Particle<R> &p, ElectronLane<R> &s)
{
    const DevTables<R> &T = A.T;
    const int iq = p.iq;
    const int qel = (1 + iq) / 2;
    // ---- end of the ustep loop: fictitious-cross-section rejection ----
    {
        const ElecBin<R> &eb = T.el[(s.imed * 2 + qel) * MXEKE + s.lelke];
        const ElecCsda &cs = T.csda[(s.imed * 2 + qel) * MXEKE + s.lelke];
        R sigf = s.elke * eb.sig1 + eb.sig0;
        R dedx0 = s.elke * (R)cs.dedx1 + (R)cs.dedx0;
        sigf /= dedx0;
        R sigratio = sigf / s.sig0;
        R rfict = urand<R>(p.rng);
        if (rfict >= sigratio) {
            count_event(EV_FICT);
            s.phase = 1;
            return E_ALIVE;
        }
    }

    // ---- discrete interaction ----
    const int m = s.imed;
    const ElecBin<R> &eb = T.el[(m * 2 + qel) * MXEKE + s.lelke];
    Particle<R> b;
    s.phase = 0;
    if (iq < 0) {
        R ebr1 = s.elke * eb.br11 + eb.br10;
        R rnno = urand<R>(p.rng);
        if (rnno <= ebr1) {
            count_event(EV_BREMS);
            brems<R>(T, p, b, m);
            push_made_photon<R>(A, p, b);
        } else if (p.e <= (R)T.small->med[m].thmoll) {
            if (ebr1 <= (R)0) return E_ALIVE;   // neither process possible: re-enter
            count_event(EV_BREMS);
            brems<R>(T, p, b, m);
            push_made_photon<R>(A, p, b);
        } else {
            count_event(EV_MOLLER);
            if (moller<R>(T, p, b, m)) push<R>(A.qe, b, A.status);
        }
        return E_ALIVE;
    }
    R pbr1 = s.elke * eb.br11 + eb.br10;
    R rnno = urand<R>(p.rng);
    if (rnno < pbr1) {
        count_event(EV_BREMS);
        brems<R>(T, p, b, m);
        push_made_photon<R>(A, p, b);
        return E_ALIVE;
    }
    R pbr2 = s.elke * eb.br21 + eb.br20;
    if (rnno < pbr2) {
        count_event(EV_BHABHA);
        bhabha<R>(T, p, b, m);
        push<R>(A.qe, b, A.status);
        return E_ALIVE;
    }
    count_event(EV_ANNIH);
    annih<R>(p, b);
    push_photon_pair<R>(A, p, b);
    return E_DONE;
}

// The ustep iteration that follows a step in matter while the tstep loop
// goes on (ompmc.c:4780-5345).  total_tstep starts again at 0 in every
// iteration (ompmc.c:4787), so here tstep = 0 and the step length is
// min(0, tmxs): the msdist branch cannot be taken, one random number is
// drawn for the single-scattering test, the particle does not move, and the
// tstep loop restarts.  Run in the same trip as the step before it; when an
// unusual table value would make the general iteration behave otherwise,
// the general iteration runs instead in the next trip.
template <typename R>
__device__ __forceinline__ int zero_ustep(const KArgs<R> &A, Particle<R> &p, ElectronLane<R> &s, LaneGeo<R> &G,
                                          Scorer<R> &sc, uint32_t &nsteps, uint32_t &ncross)
{
    if (s.compute_tstep || !(s.sig0 > (R)0.0) || s.imed == -1) return E_ALIVE;
    const DevTables<R> &T = A.T;
    const R rm = (R)RM;
    const int iq = p.iq;
    const int qel = (1 + iq) / 2;
    const R eie = p.e;
    int irl = p.ir;
    const int imed = s.imed;
    const MediumScalars *M = &T.small->med[imed];
    const R rhof = s.rhof;
    const R tstep = (R)(0.0 / (double)rhof);
    const ElecBin<R> &eb = T.el[(imed * 2 + qel) * MXEKE + s.lelke];
    const ElecCsda &cs = T.csda[(imed * 2 + qel) * MXEKE + s.lelke];
    R dedx0 = s.elke * (R)cs.dedx1 + (R)cs.dedx0;
    R dedx = rhof * dedx0;
    R tmxs = s.elke * eb.tmxs1 + eb.tmxs0;
    tmxs /= rhof;
    // the range term of min(tstep, tmxs, range) is >= 0 and cannot lower it
    R tustep = m_fmin(tstep, tmxs);
    R tperp = hownear<R>(*G.g, irl, p.x, p.y, p.z);
    R blccl = rhof * (R)M->blcc;
    R xccl = rhof * (R)M->xcc;
    R p2 = s.eke * (s.eke + (R)2.0 * rm);
    R beta2 = p2 / (p2 + sq(rm));
    R etap = s.elke * eb.eta1 + eb.eta0;
    R ms_corr = s.elke * eb.blc1 + eb.blc0;
    blccl = blccl / etap / ((R)1.0 + (R)0.25 * etap * xccl / blccl / p2) * ms_corr;
    R ssmfp = beta2 / blccl;
    R skindepth = (R)3 * ssmfp;
    if (!(skindepth > (R)0.0)) return E_ALIVE;
    tustep = m_fmin(tustep, m_fmax(tperp, skindepth));
    // from here on the general iteration's path is fixed
    nsteps++;
    R rnno = urand<R>(p.rng);
    if (rnno < (R)1.0E-30) rnno = (R)1.0E-30;
    R lambda = (R)(-1.0) * m_log((R)1.0 - rnno);
    R lambda_max = (R)0.5 * blccl * rm / dedx;
    lambda_max *= (s.eke / rm + (R)1.0) * (s.eke / rm + (R)1.0) * (s.eke / rm + (R)1.0);
    if (!(lambda >= (R)0.0 && lambda_max > (R)0.0)) return E_DONE;   // dropped unscored
    // tuss >= 0 >= tustep: single scattering stays off
    R ustep = tustep;
    int irnew = irl;
    const int irold = irl;
    if (!(ustep < tperp)) {
        count_lane_event(EV_HOWFAR_E);
        howfar<R>(*G.g, irl, p.x, p.y, p.z, p.u, p.v, p.w, irnew, ustep);
    }
    if (ustep < (R)0) ustep = (R)0.0;
    s.ustep = ustep;
    if (irnew != irold) {
        ncross++;
        p.ir = irnew;
        irl = irnew;
        s.imed = G.medium(irl);
        s.rhof = G.rho(irl);
    }
    if (eie <= (R)G.g->ecut[s.imed + 1]) { electron_cutoff<R>(A, p, G, sc, eie); return E_DONE; }
    s.phase = 1;
    return E_ALIVE;
}

template <typename R>
__device__ __forceinline__ int electron_iteration(const KArgs<R> &A, Particle<R> &p, ElectronLane<R> &s, LaneGeo<R> &G,
                                  Scorer<R> &sc, uint32_t &nsteps, uint32_t &ncross)
{
    const DevTables<R> &T = A.T;
    const R rm = (R)RM;
    const int iq = p.iq;
    const int qel = (1 + iq) / 2;
    R eie = p.e;

    if (s.phase == 0) {
        int irl = p.ir;
        s.imed = G.medium(irl);
        s.rhof = G.rho(irl);
        s.medold = s.imed;
        if (eie <= (R)G.g->ecut[s.imed + 1]) { electron_cutoff<R>(A, p, G, sc, eie); return E_DONE; }
        s.phase = 1;
    }
    const MediumScalars *M = s.imed >= 0 ? &T.small->med[s.imed] : nullptr;
    if (s.phase == 1) {
        s.compute_tstep = 1;
        s.eke = eie - rm;
        if (s.imed != -1) {
            R rnno = urand<R>(p.rng);
            if (rnno == (R)0.0) rnno = (R)1.0E-30;
            s.demfp = m_fmax(-m_log(rnno), (R)EPSEMFP);
            s.elke = m_log(s.eke);
            s.lelke = pwlf_bin(s.elke, M->eke1, M->eke0, MXEKE);
            if (qel == 0 ? M->mono_e : M->mono_p) {
                const ElecBin<R> &eb = T.el[(s.imed * 2 + qel) * MXEKE + s.lelke];
                const ElecCsda &cs = T.csda[(s.imed * 2 + qel) * MXEKE + s.lelke];
                s.sig0 = s.elke * eb.sig1 + eb.sig0;
                R dedx0 = s.elke * (R)cs.dedx1 + (R)cs.dedx0;
                s.sig0 /= dedx0;
            } else {
                s.sig0 = (R)(qel == 0 ? M->esig_e : M->psig_e);
            }
        }
        s.phase = 2;
    }

    // ---- one ustep iteration ----
    nsteps++;
    int irl = p.ir;
    int imed = s.imed;
    bool call_howfar, do_single = false, called_msdist = false;
    R tstep, tustep, range = 0, tvstep = 0, de = 0, tperp = 0;
    double total_tstep = 0.0;             // fresh every iteration (ompmc.c:4787)
    const ElecCsda *csb = nullptr;
    R ustep;
    R xf = 0, yf = 0, zf = 0, uf = 0, vf = 0, wf = 0;
    if (imed == -1) {
        tstep = (R)10.0E8;
        ustep = tstep;
        tustep = ustep;
        call_howfar = true;
    } else {
        csb = T.csda + (imed * 2 + qel) * MXEKE;
        const double eke0 = M->eke0, eke1 = M->eke1;
        const R rhof = s.rhof;          // density of the current voxel, loaded on entering it
        if (s.sig0 <= (R)0.0) {
            tstep = (R)10.0E8;
            s.sig0 = (R)1.0E-15;
        } else {
            if (s.compute_tstep) {
                double total_de = (double)s.demfp / (double)s.sig0;
                double eke = (double)s.eke, elke = (double)s.elke;
                double ekef = eke - total_de;
                double ts;
                if (ekef <= csb[0].earr) {
                    ts = 10.0E8;
                } else {
                    double elkef = log(ekef);
                    int lelkef = pwlf_bin(elkef, eke1, eke0, MXEKE);
                    if (lelkef == s.lelke) {
                        ts = drange(&csb[s.lelke], iq, eke, ekef, elke, elkef);
                    } else {
                        double ekei = csb[s.lelke].earr;
                        double elkei = (s.lelke + 1 - eke0) / eke1;
                        double tuss = drange(&csb[s.lelke], iq, eke, ekei, elke, elkei);
                        ekei = csb[lelkef + 1].earr;
                        elkei = ((lelkef + 2) - eke0) / eke1;
                        ts = drange(&csb[lelkef], iq, ekei, ekef, elkei, elkef);
                        ts += tuss + csb[s.lelke].range - csb[lelkef + 1].range;
                    }
                }
                total_tstep = ts;
                s.compute_tstep = 0;
            }
            tstep = (R)(total_tstep / (double)rhof);
        }
        const ElecBin<R> &eb = T.el[(imed * 2 + qel) * MXEKE + s.lelke];
        const ElecCsda &cs = csb[s.lelke];
        R dedx0 = s.elke * (R)cs.dedx1 + (R)cs.dedx0;
        R dedx = rhof * dedx0;
        R tmxs = s.elke * eb.tmxs1 + eb.tmxs0;
        tmxs /= rhof;
        {
            double ekei = cs.earr;
            double elkei = (s.lelke + 1 - eke0) / eke1;
            double rg = drange(&cs, iq, (double)s.eke, ekei, (double)s.elke, elkei);
            rg += cs.range;
            rg /= (double)rhof;
            range = (R)rg;
        }
        tustep = m_fmin(m_fmin(tstep, tmxs), range);
        tperp = hownear<R>(*G.g, irl, p.x, p.y, p.z);
        R blccl = rhof * (R)M->blcc;
        R xccl = rhof * (R)M->xcc;
        R p2 = s.eke * (s.eke + (R)2.0 * rm);
        R beta2 = p2 / (p2 + sq(rm));
        R etap = s.elke * eb.eta1 + eb.eta0;
        R ms_corr = s.elke * eb.blc1 + eb.blc0;
        blccl = blccl / etap / ((R)1.0 + (R)0.25 * etap * xccl / blccl / p2) * ms_corr;
        R ssmfp = beta2 / blccl;
        R skindepth = (R)3 * ssmfp;
        tustep = m_fmin(tustep, m_fmax(tperp, skindepth));
        if ((tustep <= tperp) && (tustep > skindepth)) {
            call_howfar = false;
            do_single = false;
            called_msdist = true;
            de = (R)eloss(csb, M->te, eke0, eke1, (double)rhof, (double)tustep, (double)range,
                          (double)s.eke, (double)s.elke, s.lelke);
            tvstep = tustep;
            // msdist moves a copy; the original position is kept until
            // after ausgab, which scores in the pre-step region
            Particle<R> q = p;
            count_lane_event(EV_MSDIST);
            ustep = msdist<R>(T, q, imed, rhof, de, tustep, s.eke);
            p.rng = q.rng;
            xf = q.x; yf = q.y; zf = q.z; uf = q.u; vf = q.v; wf = q.w;
        } else {
            called_msdist = false;
            R rnno = urand<R>(p.rng);
            if (rnno < (R)1.0E-30) rnno = (R)1.0E-30;
            R lambda = (R)(-1.0) * m_log((R)1.0 - rnno);
            R lambda_max = (R)0.5 * blccl * rm / dedx;
            lambda_max *= (s.eke / rm + (R)1.0) * (s.eke / rm + (R)1.0) * (s.eke / rm + (R)1.0);
            if (lambda >= (R)0.0 && lambda_max > (R)0.0) {
                R tuss;
                if (lambda < lambda_max) tuss = lambda * ssmfp * ((R)1.0 - (R)0.5 * lambda / lambda_max);
                else tuss = (R)0.5 * lambda * ssmfp;
                if (tuss < tustep) { tustep = tuss; do_single = true; }
                else do_single = false;
            } else {
                return E_DONE;               // reference drops the particle unscored
            }
            ustep = tustep;
            call_howfar = !(ustep < tperp);
        }
    }

    int irold = p.ir, irnew = p.ir;
    if (call_howfar) {
        count_lane_event(EV_HOWFAR_E);
        if (irl == 0) {                      // howfar discard (outside)
            R edep = (iq > 0) ? p.e + rm : p.e - rm;
            sc.add(G, p.ir, (double)p.wt * (double)edep);
            if (iq > 0 && edep < eie) {
                count_event(EV_RANNIH);
                Particle<R> b;
                rannih<R>(p, b);
                push_photon_pair<R>(A, p, b);
            }
            return E_DONE;
        }
        howfar<R>(*G.g, irl, p.x, p.y, p.z, p.u, p.v, p.w, irnew, ustep);
    }
    if (ustep < (R)0) ustep = (R)0.0;
    s.ustep = ustep;
    if (ustep == (R)0.0 || imed == -1) {
        if (ustep != (R)0.0) {
            R vstep = ustep;
            p.x += p.u * vstep;
            p.y += p.v * vstep;
            p.z += p.w * vstep;
        }
        if (irnew != irold) {
            ncross++;
            p.ir = irnew;
            irl = irnew;
            s.imed = G.medium(irl);
            s.rhof = G.rho(irl);
        }
        if (eie <= (R)G.g->ecut[s.imed + 1]) { electron_cutoff<R>(A, p, G, sc, eie); return E_DONE; }
        // leave the ustep loop; the tstep loop restarts unless a step through
        // vacuum brought the particle back into the medium it last lost energy
        // in, where the reference goes on to the fictitious-sigma test
        // (ompmc.c:5347; medold is only updated by steps in matter)
        if (s.imed != s.medold || ustep == (R)0.0 || s.imed == -1) {
            s.phase = 1;
            return E_ALIVE;
        }
        return electron_interaction<R>(A, p, s);
    }
    R vstep = ustep;
    const R rhof = s.rhof;
    if (call_howfar) {
        tvstep = vstep;
        if (tvstep != tustep) do_single = false;
        de = (R)eloss(csb, M->te, M->eke0, M->eke1, (double)rhof, (double)tvstep, (double)range,
                      (double)s.eke, (double)s.elke, s.lelke);
    } else {
        tvstep = tustep;
        if (!called_msdist)
            de = (R)eloss(csb, M->te, M->eke0, M->eke1, (double)rhof, (double)tvstep, (double)range,
                          (double)s.eke, (double)s.elke, s.lelke);
    }
    R edep = de;
    R ekef = s.eke - de;
    if (!called_msdist) {
        R sinthe, costhe;
        if (do_single) {
            R ekems = m_fmax(ekef, (R)G.g->ecut[imed + 1] - rm);
            R p2 = ekems * (ekems + (R)2.0 * rm);
            R beta2 = p2 / (p2 + sq(rm));
            R chia2 = (R)M->xcc / ((R)4.0 * (R)M->blcc * p2);
            R elkems = m_log(ekems);
            int lelkems = pwlf_bin(elkems, M->eke1, M->eke0, MXEKE);
            const ElecBin<R> &ebs = T.el[(imed * 2 + qel) * MXEKE + lelkems];
            R etap = elkems * ebs.eta1 + ebs.eta0;
            chia2 *= etap;
            count_lane_event(EV_SSCAT);
            sscat<R>(T, p.rng, imed, qel, chia2, elkems, beta2, costhe, sinthe);
        } else {
            sinthe = (R)0.0;
            costhe = (R)1.0;
        }
        xf = p.x + p.u * vstep;
        yf = p.y + p.v * vstep;
        zf = p.z + p.w * vstep;
        if (do_single) {
            Uphi<R> h;
            uphi21<R>(p, h, costhe, sinthe);
        }
        uf = p.u; vf = p.v; wf = p.w;
    }
    sc.add(G, p.ir, (double)p.wt * (double)edep);
    p.x = xf; p.y = yf; p.z = zf;
    p.u = uf; p.v = vf; p.w = wf;
    irold = p.ir;
    eie -= edep;
    p.e = eie;
    if (irnew == irl && eie <= (R)G.g->ecut[imed + 1]) { electron_cutoff<R>(A, p, G, sc, eie); return E_DONE; }
    s.medold = imed;
    if (imed != -1) {
        s.eke = eie - rm;
        s.elke = m_log(s.eke);
        s.lelke = pwlf_bin(s.elke, M->eke1, M->eke0, MXEKE);
    }
    if (irnew != irold) {
        ncross++;
        p.ir = irnew;
        irl = irnew;
        s.imed = G.medium(irl);
        s.rhof = G.rho(irl);
    }
    if (eie <= (R)G.g->ecut[s.imed + 1]) { electron_cutoff<R>(A, p, G, sc, eie); return E_DONE; }
    if (s.imed != s.medold) { s.phase = 1; return E_ALIVE; }
    s.demfp -= de * s.sig0;
    total_tstep -= (double)tvstep * (double)rhof;
    if (total_tstep < 1.0E-9) s.demfp = (R)0.0;
    if (s.demfp >= (R)EPSEMFP) return zero_ustep<R>(A, p, s, G, sc, nsteps, ncross);

    return electron_interaction<R>(A, p, s);
}

// ============================================================ kernels =====
// Minimum waves per SIMD the register allocation of each transport kernel
// must allow (the second __launch_bounds__ argument); values above what the
// kernel needs make the compiler keep fewer values in VGPRs.  Chosen by
// sweeps on the slab phantom: FP32 photon 4 (no spills) / electron 4 (128
// VGPRs and a few spills; 3-4% more histories/s than 3 waves without
// spills), FP64 photon 2 / electron 2 (a few spills, still faster than one
// wave).
#ifndef OMC_WAVES_P32
#define OMC_WAVES_P32 4
#endif
#ifndef OMC_WAVES_E32
#define OMC_WAVES_E32 4
#endif
#ifndef OMC_WAVES_P64
#define OMC_WAVES_P64 2
#endif
#ifndef OMC_WAVES_E64
#define OMC_WAVES_E64 2
#endif
// Philox blocks each lane keeps ready in its LDS ring at the head of a trip
// (at most RNG_RING; 0 = no ring).  An electron trip draws about 15
// numbers; a photon trip draws a few and its kernel waits on voxel loads,
// so it keeps the occupancy the ring's LDS would cost.
#ifndef OMC_RNG_DEPTH_E
#define OMC_RNG_DEPTH_E 8
#endif
#ifndef OMC_RNG_DEPTH_P
#define OMC_RNG_DEPTH_P 0
#endif
static_assert(OMC_RNG_DEPTH_E <= RNG_RING && OMC_RNG_DEPTH_P <= RNG_RING, "ring depth");

template <typename R> struct Waves;
template <> struct Waves<float> { static constexpr int photon = OMC_WAVES_P32, electron = OMC_WAVES_E32; };
template <> struct Waves<double> { static constexpr int photon = OMC_WAVES_P64, electron = OMC_WAVES_E64; };

template <typename R>
__global__ void __launch_bounds__(256, Waves<R>::photon) k_photon(KArgs<R> A)
{
    __shared__ DevPhantom<R> lds_ph[MAX_PHANTOMS];
    stage_phantoms<R>(A.phantoms, lds_ph);
    Particle<R> p;
    PhotonLane<R> s;
    LaneGeo<R> G;
    Scorer<R> sc;
    bool alive = false, exhausted = false;
    uint32_t nsteps = 0, ncross = 0, trips = 0, here = 0, nyield = 0, longest = 0;
    const uint32_t in_end = min(*A.in_count, A.qin.cap);
    for (;;) {
        // free lanes take new particles together (one counter atomic and one
        // round of queue loads for all of them), once enough are free or
        // none is live
        // (both ballots are taken by the whole wavefront: inside the branch
        // only free lanes are active, and a ballot of `alive` there is 0)
        const unsigned long long want = __ballot(!alive && !exhausted);
        const bool none_live = __ballot(alive) == 0;
        if (!alive && !exhausted && ((uint32_t)__popcll(want) >= A.fetch_min || none_live)) {
            uint32_t i = wave_reserve(A.fetch, 1);
            if (i < in_end) {
                if (q_load(A.qin, i, p)) {
                    s.load(A.qin, i);
                    trips = A.qin.trips()[i];
                } else {
                    s.phase = 0;
                    s.split = 0u;
                    trips = 0;
                }
                G.bind(A, lds_ph, p);
                s.attach(G, p);
                here = 0;
                alive = true;
            } else {
                exhausted = true;
            }
        }
        const unsigned long long live = __ballot(alive);
        if (live == 0) break;
        // input used up and the wave mostly empty: hand the rest on, the
        // next launch packs them into full waves
        const uint32_t nlive = (uint32_t)__popcll(live);
        const bool drain = in_end >= A.drain_min && __ballot(exhausted) != 0 && nlive <= A.drain_lanes;
        // events (interaction, entry) run for all waiting lanes of the
        // wavefront together, once enough of the live lanes wait for one or
        // none can step; until then the waiting lanes sit out the steps
        const uint32_t nev = (uint32_t)__popcll(__ballot(alive && s.phase != 1));
        const bool events = nev * A.event_div >= nlive || nev == nlive;
        if (alive && (s.phase == 1 || events)) {
            if (OMC_RNG_DEPTH_P) p.rng.top_up(OMC_RNG_DEPTH_P);
            if (s.phase != 1)
                alive = photon_event<R>(A, p, s, G, sc);
            else if (A.woodcock && G.g->wc.mu && s.split == 0u)
                alive = photon_flight<R>(A, p, s, G, nsteps, ncross);
            else
                alive = photon_step<R>(A, p, s, G, nsteps, ncross);
            ++trips;
            ++here;
            if (alive && trips > MAX_ITERS) {           // watchdog: no particle runs forever
                atomicAdd(&A.status->killed, 1u);
                alive = false;
            } else if (alive && s.split == 0u && (here >= A.yield_trips || (drain && here >= A.drain_trips))) {
                sc.flush(G);
                yield_push<R>(A.qp, p, s, trips, A.status);
                nyield++;
                alive = false;
                trips = 0;
            }
            if (!alive) {
                sc.flush(G);
                longest = max(longest, trips);
            }
        }
    }
    wave_max_to(A.longest, longest);
    sc.flush(G);
    if (nyield) atomicAdd(&A.status->yields, (unsigned long long)nyield);
    atomicAdd(&A.status->photon_steps, (unsigned long long)nsteps);
    atomicAdd(&A.status->photon_cross, (unsigned long long)ncross);
    atomicAdd(&A.status->photon_dep, (unsigned long long)sc.ndep);
}

template <typename R>
__global__ void __launch_bounds__(256, Waves<R>::electron) k_electron(KArgs<R> A)
{
    __shared__ DevPhantom<R> lds_ph[MAX_PHANTOMS];
    events_begin<true>();
    stage_phantoms<R>(A.phantoms, lds_ph);
    Particle<R> p;
    ElectronLane<R> s;
    LaneGeo<R> G;
    Scorer<R> sc;
    bool alive = false, exhausted = false;
    uint32_t nsteps = 0, ncross = 0, trips = 0, here = 0, nyield = 0, longest = 0;
    const uint32_t in_end = min(*A.in_count, A.qin.cap);
    for (;;) {
        // free lanes take new particles together (one counter atomic and one
        // round of queue loads for all of them), once enough are free or
        // none is live
        // (both ballots are taken by the whole wavefront: inside the branch
        // only free lanes are active, and a ballot of `alive` there is 0)
        const unsigned long long want = __ballot(!alive && !exhausted);
        const bool none_live = __ballot(alive) == 0;
        if (!alive && !exhausted && ((uint32_t)__popcll(want) >= A.fetch_min || none_live)) {
            uint32_t i = wave_reserve(A.fetch, 1);
            if (i < in_end) {
                if (q_load(A.qin, i, p)) {
                    s.load(A.qin, i);
                    trips = A.qin.trips()[i];
                } else {
                    s.phase = 0;
                    trips = 0;
                }
                G.bind(A, lds_ph, p);
                here = 0;
                alive = true;
            } else {
                exhausted = true;
            }
        }
        const unsigned long long live = __ballot(alive);
        if (live == 0) break;
        const bool drain =
            in_end >= A.drain_min && __ballot(exhausted) != 0 && (uint32_t)__popcll(live) <= A.drain_lanes;
        if (alive) {
            p.rng.top_up(OMC_RNG_DEPTH_E);
            int r = electron_iteration<R>(A, p, s, G, sc, nsteps, ncross);
            ++trips;
            ++here;
            if (r != E_DONE && trips > MAX_ITERS) {     // watchdog: no particle runs forever
                atomicAdd(&A.status->killed, 1u);
                r = E_DONE;
            } else if (r != E_DONE && (here >= A.yield_trips || (drain && here >= A.drain_trips))) {
                sc.flush(G);
                yield_push<R>(A.qe, p, s, trips, A.status);
                nyield++;
                r = E_DONE;
                trips = 0;
            }
            if (r == E_DONE) {
                alive = false;
                sc.flush(G);
                longest = max(longest, trips);
            }
        }
    }
    wave_max_to(A.longest, longest);
    sc.flush(G);
    if (nyield) atomicAdd(&A.status->yields, (unsigned long long)nyield);
    atomicAdd(&A.status->electron_steps, (unsigned long long)nsteps);
    atomicAdd(&A.status->electron_cross, (unsigned long long)ncross);
    atomicAdd(&A.status->electron_dep, (unsigned long long)sc.ndep);
    events_end<true>(A.status);
}

}  // namespace omc
