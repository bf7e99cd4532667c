// Device-side building blocks for the gfx950 transport engine: the
// per-particle Philox4x32-10 stream, 64-wide wavefront helpers, and the
// layouts of tables, phantoms, jobs and particle queues in HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "omc_dev.h"

namespace omc {

// Global-memory view of a pointer.  A pointer read from LDS or from memory
// carries no address space, so the compiler reaches it with FLAT
// instructions, which also wait on the LDS counter; the cast makes the loads
// and atomics through it global ones.
template <typename T> using gmem = __attribute__((address_space(1))) T;
template <typename T> __device__ __forceinline__ gmem<T> *glob(T *p) { return (gmem<T> *)p; }

constexpr double RM = OMC_RM;
constexpr int MXGE = OMC_MXGE;
constexpr int MXEKE = OMC_MXEKE;
constexpr int MAXMED = OMC_MXMED;
constexpr int MAX_PHANTOMS = 16;     // resident phantoms per context

// ---------------------------------------------------------------- RNG -----
// Philox4x32-10 (Salmon et al., SC'11).  Each particle owns a stream: key
// (k0,k1) and a draw counter `ctr`; draw d uses word d&3 of block d>>2 and
// becomes the 24-bit uniform (w >> 8) * 2^-24, the same granularity as the
// RANMAR numbers ompMC uses.  A secondary's key is the Philox block
// (parent ctr, 1, 0, 0) under the parent key, taken after the interaction's
// last draw, so every history is independent of processing order.
__device__ __forceinline__ void philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// The same function with its rounds kept as a loop: for the rare paths
// inlined at every draw site (a block the ring does not hold, a secondary's
// key).  Unrolled there, Philox was half of the charged-particle kernel's
// instructions (about a hundred inlined copies), far more code than the
// instruction cache a CU pair shares holds.
__device__ __forceinline__ void philox10_rolled(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll 1
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Look-ahead of a lane's stream in LDS: RNG_RING Philox blocks per lane,
// block b in slot b % RNG_RING of the lane's column.  A transport kernel
// that uses it tops the ring up where the wavefront's live lanes are
// converged (the head of each trip), so the draw sites read 16 bytes from
// LDS instead of running Philox in whichever few lanes crossed a block
// boundary there.  The words and their order are those of the stream
// itself.  The ring is the launch's dynamic LDS (RNG_RING_BYTES), so only
// the launches that top it up pay for it in occupancy.
constexpr uint32_t RNG_RING = 8;
constexpr int RNG_LANES = 256;           // threads per transport block
constexpr size_t RNG_RING_BYTES = (size_t)RNG_RING * RNG_LANES * sizeof(uint4);
extern __shared__ uint4 rng_ring[];      // [RNG_RING][RNG_LANES]

struct Rng {
    uint32_t k0, k1, ctr;
    uint32_t blk;            // block held in w[] (0xffffffff = none)
    uint32_t w0, w1, w2, w3;
    uint32_t tail;           // blocks in [ctr >> 2, tail) are in this thread's LDS ring

    // No cached block, nothing in the ring: every block is computed at the
    // draw (secondaries, fresh primaries, particles read from a queue).
    // Only the stream of the particle a transport lane owns is topped up,
    // so only that stream ever finds blocks in the thread's ring column.
    __device__ __forceinline__ void reset_cache() { blk = 0xffffffffu; tail = 0u; }

    // Fill the thread's ring up to `depth` blocks ahead of the next draw
    // (transport kernels, for the stream of the lane's own particle).
    __device__ __forceinline__ void top_up(uint32_t depth)
    {
        const uint32_t cur = ctr >> 2;
        if (tail < cur) tail = cur;
        while (tail < cur + depth) {
            uint32_t o[4];
            philox10(tail, 0u, 0u, 0u, k0, k1, o);
            rng_ring[(tail % RNG_RING) * RNG_LANES + threadIdx.x] = make_uint4(o[0], o[1], o[2], o[3]);
            tail++;
        }
    }

    __device__ __forceinline__ uint32_t next_u24()
    {
        uint32_t b = ctr >> 2;
        if (b != blk) {
            if (b < tail) {
                const uint4 v = rng_ring[(b % RNG_RING) * RNG_LANES + threadIdx.x];
                w0 = v.x; w1 = v.y; w2 = v.z; w3 = v.w;
            } else {
                uint32_t o[4];
                philox10_rolled(b, 0u, 0u, 0u, k0, k1, o);
                w0 = o[0]; w1 = o[1]; w2 = o[2]; w3 = o[3];
                tail = b + 1;
            }
            blk = b;
        }
        // word (ctr & 3) of the block by two selects on its bits (a select
        // chain on the whole value becomes an indexed private array)
        const uint32_t s = ctr & 3u;
        const uint32_t even = (s & 2u) ? w2 : w0;
        const uint32_t odd = (s & 2u) ? w3 : w1;
        const uint32_t w = (s & 1u) ? odd : even;
        ctr++;
        return w >> 8;
    }
    // Key of a secondary: the Philox block (ctr, tag, 0, 0) under this key.
    // Interactions use tag 1 for the particles they create; tag 2 is for a
    // copy of the stream's own particle handed on while the stream goes on
    // (photon splitting), so the two never share a key.
    __device__ __forceinline__ void child(uint32_t &c0, uint32_t &c1, uint32_t tag = 1u) const
    {
        uint32_t o[4];
        philox10_rolled(ctr, tag, 0u, 0u, k0, k1, o);
        c0 = o[0]; c1 = o[1];
    }
};

// next uniform of a particle stream, in [0,1) with 24-bit granularity
template <typename R> __device__ __forceinline__ R urand(Rng &r)
{
    return (R)r.next_u24() * (R)(1.0 / 16777216.0);
}

__device__ __forceinline__ void primary_key(uint32_t s0, uint32_t s1, uint64_t hid, uint32_t &k0, uint32_t &k1)
{
    uint32_t o[4];
    philox10((uint32_t)hid, (uint32_t)(hid >> 32), 2u, 0u, s0, s1, o);
    k0 = o[0]; k1 = o[1];
}

// ------------------------------------------------------------ wavefront ----
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// One atomic per wavefront for all active lanes that call it together; every
// calling lane gets a distinct index.  Works in divergent code (only the
// currently active lanes take part).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *counter, uint32_t n_per_lane)
{
    uint64_t m = __ballot(1);
    int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t total = (uint32_t)__popcll(m) * n_per_lane;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    return base + lanes_below(m) * n_per_lane;
}

// Largest v over the lanes of a converged wavefront, into *p with one
// atomicMax per wavefront.
__device__ __forceinline__ void wave_max_to(uint32_t *p, uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
    if (lane_id() == 0 && v) atomicMax(p, v);
}

// ------------------------------------------------------------- layouts ----
template <typename R> struct R2 { R c1, c0; };           // pwlf slope / intercept
template <typename R> struct R4 { R a1, a0, b1, b0; };   // two pwlf pairs

// Per (medium, charge, bin) electron step data used at every CH step
// (sig, tmxs, eta, blcce) and by msdist / the interaction choice.
template <typename R> struct ElecBin {
    R sig1, sig0, tmxs1, tmxs0;
    R eta1, eta0, blc1, blc0;
    R q1c1, q1c0, q2c1, q2c0;
    R br11, br10, br21, br20;
};

// CSDA data, FP64 in every precision mode: restricted stopping power,
// range to the first table energy, and that energy.
struct ElecCsda {
    double dedx1, dedx0, range, earr;
};

// multiple-scattering surface entry: alias threshold, alias index, u node
// One node of a multiple-scattering alias table (mscatSample tables): the
// alias weight, the node's value and the value of the node it aliases
// (same row), so that a draw is one load.
template <typename R> struct MsAlias {
    R wms, ums, ualias, pad;
};

// Per-medium scalars and the small tables.
struct MediumScalars {
    double ge0, ge1, eke0, eke1, blcc, xcc, esig_e, psig_e;
    double te, thmoll, ap, delcm, zbrang, bpar0, bpar1, pad;
    int mono_e, mono_p, meke, pad2;
    double dl1[8], dl2[8], dl3[8], dl4[8], dl5[8], dl6[8];
};

struct SmallTables {
    int nmed;
    int pad;
    double b2spin_min, dbeta2i, espml, dleneri, dqq1i, dllambi, dqmsi;
    // Rayleigh: ompMC samples with medium-0 form factors (ompmc.c:1120-1131)
    double xgrid[OMC_MXRAYFF], fcum[OMC_MXRAYFF], b_array[OMC_MXRAYFF], c_array[OMC_MXRAYFF];
    int i_array[OMC_RAYCDFSIZE];
    MediumScalars med[MAXMED];
};

// Device tables for precision R.  Photon bins are interleaved so one voxel
// step reads one 32-byte record.
template <typename R> struct DevTables {
    const SmallTables *small;
    const R4<R> *ph_mfp;     // {gmfp1, gmfp0, cohe1, cohe0} [imed*MXGE + bin]
    const R4<R> *ph_br;      // {gbr11, gbr10, gbr21, gbr20}
    const R2<R> *ph_pmax;    // {pmax1, pmax0}
    const ElecBin<R> *el;    // [(imed*2 + qel)*MXEKE + bin]
    const ElecCsda *csda;    // [(imed*2 + qel)*MXEKE + bin]
    const R *spin;           // [imed*32768 + qel*16384 + i*512 + j*32 + k]
    const MsAlias<R> *ms;    // [i*256 + j*32 + k]
};

// Resident phantom: region r = voxel + 1, region 0 = outside.
// Delta tracking of photon flights in a phantom with a uniform voxel grid
// whose media share one photon energy grid: mu[bin] bounds the attenuation
// coefficient (1/cm) of every voxel at every energy of that bin.
template <typename R> struct Woodcock {
    const R *mu;             // [MXGE] majorant; null = voxel stepping only
    double ge0, ge1;         // lgle = ge1 * log(E) + ge0 for every medium
    double x0, y0, z0;       // lower corner of the grid
    R idx, idy, idz;         // inverse voxel sizes
    int pad;
};

template <typename R> struct DevPhantom {
    int isize, jsize, ksize, nreg;
    const double *xb, *yb, *zb;
    const uint8_t *med;      // [nreg] medium index, 0xff = vacuum / outside
    const R *rhof;           // [nreg]
    double ecut[MAXMED + 1]; // [imed + 1]; entry 0 is vacuum / outside
    double pcut[MAXMED + 1];
    Woodcock<R> wc;
};

// One source job (a dosxyz field, one beamlet, one serving request).
struct DevJob {
    int phantom;             // index into the phantom table
    int kind;                // 0 dosxyz field, 1 beamlet
    int charge;
    int spectrum;
    double energy;           // mono energy
    const double *cdf1, *cdf2;   // [1000] inverse CDF (device)
    double deltak;
    // dosxyz field
    double ssd, xinl, yinl, xsize, ysize;
    int ixinl, iyinl;
    // beamlet
    double src[3], corner[3], side1[3], side2[3];
    // histories
    uint64_t hist0;          // stream index of the first history
    int64_t nperbatch;
    uint32_t s0, s1;         // seed
    int nsplit;              // photon splitting (vrt.nsplit); 0 or 1 = off
    int pad_nsplit;
    double *tally;           // [slot*nreg + region], slot = batch
    double *ledger;          // optional [3*nhist]
};

// A run of consecutive histories [first, first + n) of one job, placed at
// thread offset `off` of a source launch.
struct SourceSeg {
    uint64_t first;
    uint32_t job, n, off, pad;
};

// Structure-of-arrays particle queue.  A particle that gave up its lane in
// the middle of its transport (a yield) also carries the lane's state, so
// the next launch resumes it exactly where it stopped.
//
// One allocation per queue: QR_N arrays of R (e x y z u v w wt, then the six
// real-valued resume variables), then QU_N arrays of uint32 (ir, meta =
// job << 3 | resume << 2 | (iq + 1), history index within the job, Philox
// key and counter, packed integer resume variables, loop trips taken so
// far), each `cap` long.  A queue is then three scalars -- base, tail
// counter, capacity -- and a kernel's three queues fit its scalar
// registers; with one pointer per array they took 135 of them and the
// compiler spilled pointers into vector-register lanes, reading them back
// before every table or queue access.
enum { QR_E, QR_X, QR_Y, QR_Z, QR_U, QR_V, QR_W, QR_WT, QR_ST, QR_N = QR_ST + 6 };
enum { QU_IR, QU_META, QU_HIST, QU_K0, QU_K1, QU_CTR, QU_SFLAGS, QU_TRIPS, QU_N };

template <typename R> struct Queue {
    char *base;
    uint32_t *count;         // device tail counter
    uint32_t cap;

    __host__ __device__ R *r(int f) const { return reinterpret_cast<R *>(base) + (size_t)f * cap; }
    __host__ __device__ uint32_t *u(int f) const
    {
        return reinterpret_cast<uint32_t *>(base + sizeof(R) * (size_t)QR_N * cap) + (size_t)f * cap;
    }
    __host__ __device__ R *e() const { return r(QR_E); }
    __host__ __device__ R *x() const { return r(QR_X); }
    __host__ __device__ R *y() const { return r(QR_Y); }
    __host__ __device__ R *z() const { return r(QR_Z); }
    __host__ __device__ R *uu() const { return r(QR_U); }
    __host__ __device__ R *v() const { return r(QR_V); }
    __host__ __device__ R *w() const { return r(QR_W); }
    __host__ __device__ R *wt() const { return r(QR_WT); }
    __host__ __device__ R *st(int k) const { return r(QR_ST + k); }
    __host__ __device__ int32_t *ir() const { return reinterpret_cast<int32_t *>(u(QU_IR)); }
    __host__ __device__ uint32_t *meta() const { return u(QU_META); }
    __host__ __device__ uint32_t *hist() const { return u(QU_HIST); }
    __host__ __device__ uint32_t *k0() const { return u(QU_K0); }
    __host__ __device__ uint32_t *k1() const { return u(QU_K1); }
    __host__ __device__ uint32_t *ctr() const { return u(QU_CTR); }
    __host__ __device__ uint32_t *sflags() const { return u(QU_SFLAGS); }
    __host__ __device__ uint32_t *trips() const { return u(QU_TRIPS); }
};

constexpr uint32_t META_RESUME = 4u;

// energy-ordered fetch of charged-particle launches (kernels_impl.h)
constexpr int SORT_BUCKETS = 64;
constexpr int SORT_BLOCKS = 128;

template <typename R> struct Particle {
    R e, x, y, z, u, v, w, wt;
    int ir, iq;
    uint32_t job, hist;
    Rng rng;
};

// Returns true if the entry carries a resume state.
template <typename R>
__device__ __forceinline__ bool q_load(const Queue<R> &q, uint32_t i, Particle<R> &p)
{
    p.e = q.e()[i]; p.x = q.x()[i]; p.y = q.y()[i]; p.z = q.z()[i];
    p.u = q.uu()[i]; p.v = q.v()[i]; p.w = q.w()[i]; p.wt = q.wt()[i];
    p.ir = q.ir()[i];
    uint32_t m = q.meta()[i];
    p.job = m >> 3;
    p.iq = (int)(m & 3u) - 1;
    p.hist = q.hist()[i];
    p.rng.k0 = q.k0()[i]; p.rng.k1 = q.k1()[i]; p.rng.ctr = q.ctr()[i];
    p.rng.reset_cache();
    return (m & META_RESUME) != 0u;
}

template <typename R>
__device__ __forceinline__ void q_store(const Queue<R> &q, uint32_t i, const Particle<R> &p, uint32_t resume = 0u)
{
    q.e()[i] = p.e; q.x()[i] = p.x; q.y()[i] = p.y; q.z()[i] = p.z;
    q.uu()[i] = p.u; q.v()[i] = p.v; q.w()[i] = p.w; q.wt()[i] = p.wt;
    q.ir()[i] = p.ir;
    q.meta()[i] = (p.job << 3) | resume | (uint32_t)(p.iq + 1);
    q.hist()[i] = p.hist;
    q.k0()[i] = p.rng.k0; q.k1()[i] = p.rng.k1; q.ctr()[i] = p.rng.ctr;
}

// Physics events counted by the transport kernels, in the order of the
// omc_dev_stats fields.  The rare ones (interactions) are counted per block
// with LDS atomics; the frequent ones (every condensed history step) in a
// per-thread LDS slot, so neither holds a register.
enum EvKind {
    EV_COMPTON, EV_PHOTO, EV_PAIR, EV_RAYLEIGH, EV_BREMS, EV_MOLLER, EV_BHABHA, EV_ANNIH, EV_RANNIH, EV_FICT,
    EV_NRARE,
    EV_MSDIST = EV_NRARE, EV_SSCAT, EV_HOWFAR_E,
    EV_N
};

// Counters and error flags shared by one launch.
struct DevStatus {
    uint32_t overflow;       // a push found its queue full
    uint32_t killed;         // particles stopped by the per-particle step limit
    unsigned long long yields;                         // particles handed to the next launch
    unsigned long long photon_steps, electron_steps;   // transport loop trips
    unsigned long long photon_cross, electron_cross;   // voxel boundaries crossed
    unsigned long long photon_dep, electron_dep;       // ausgab calls
    unsigned long long wc_over;                        // delta-tracking sites above the majorant
    unsigned long long ev[EV_N];                       // physics events (EvKind)
};

static __shared__ uint32_t ev_block[EV_NRARE];
static __shared__ uint32_t ev_lane[EV_N - EV_NRARE][RNG_LANES];

// Zeroes the block's event counters (the per-thread slots only where the
// kernel counts frequent events); the caller synchronises the block before
// the first count.
template <bool LANES> __device__ __forceinline__ void events_begin()
{
    if (threadIdx.x < EV_NRARE) ev_block[threadIdx.x] = 0u;
    if (LANES)
        for (int k = 0; k < EV_N - EV_NRARE; ++k) ev_lane[k][threadIdx.x] = 0u;
}

// A rare event in this lane: an LDS atomic (same-address lanes of a
// wavefront are serialised by the LDS, which costs little for events that
// come a few times per history).
__device__ __forceinline__ void count_event(int k) { atomicAdd(&ev_block[k], 1u); }

// A rare event in every active lane, straight into the launch status: one
// global atomic per wavefront (kernels that keep no block counters).
__device__ __forceinline__ void count_event_global(DevStatus *st, int k)
{
    const uint64_t m = __ballot(1);
    if (lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(&st->ev[k], (unsigned long long)__popcll(m));
}

// A frequent event of this thread's lane.
__device__ __forceinline__ void count_lane_event(int k) { ev_lane[k - EV_NRARE][threadIdx.x] += 1u; }

// Adds the block's counts to the launch status (every thread of the block
// calls it once, after its transport loop).
template <bool LANES> __device__ __forceinline__ void events_end(DevStatus *st)
{
    __syncthreads();
    const int k = threadIdx.x;
    if (k < EV_NRARE) {
        if (ev_block[k]) atomicAdd(&st->ev[k], (unsigned long long)ev_block[k]);
    } else if (LANES && k < EV_N) {
        unsigned long long n = 0;
        for (int t = 0; t < RNG_LANES; ++t) n += ev_lane[k - EV_NRARE][t];
        if (n) atomicAdd(&st->ev[k], n);
    }
}

}  // namespace omc
