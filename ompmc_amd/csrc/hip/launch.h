// Host-side launch interface of the transport kernels, one instantiation per
// transport precision (k64.hip: double, k32.hip: float), plus the
// precision-independent helper kernels of the engine.
#pragma once

#include <hip/hip_runtime.h>
#include "common.h"

namespace omc {

template <typename R> struct KArgs;

template <typename R> struct Launch {
    static void source(const DevJob *jobs, const DevPhantom<R> *ph, const SourceSeg *segs, int nseg,
                       uint32_t ntot, const Queue<R> &qp, const Queue<R> &qe, DevStatus *st, hipStream_t s);
    static void photon(const KArgs<R> &a, int grid, hipStream_t s);
    static void electron(const KArgs<R> &a, int grid, hipStream_t s);
    static void math_eval(int fn, int64_t n, const double *x, const double *y, double *out, hipStream_t s);
    // out[0, count) = the queue's records, highest total energy first
    // (hist: SORT_BUCKETS * SORT_BLOCKS words of scratch)
    static void sort_charged(const Queue<R> &q, uint32_t *hist, const Queue<R> &out, hipStream_t s);
};

// One medium for the photon table builder (initPhotonData): its elements in
// ascending Z with their atom fractions, and the per-medium scalars the
// host computes in the host builder's order (log-energy grid ge0/ge1, the
// sum of Z and the conversion to 1/cm).
struct PhotonBuildMedium {
    int ne;
    int pad;
    double z[OMC_MXELEMENT], pz[OMC_MXELEMENT];
    double ge0, ge1, sum_z, con2;
};

// XCOM element data on the device: set s (photo, Rayleigh, pair, triplet),
// element Z = z + 1 has n[s*OMC_MXELEMENT + z] points at off[...] in le, ls.
struct XcomView {
    const int *n;
    const int64_t *off;
    const double *le, *ls;
};

// Photon tables of nmed media: node values [nmed][4][MXGE] (gmfp, gbr1,
// gbr2, cohe at the grid energies), then the per-bin fits into
// tab [8][nmed*MXGE] (gmfp1, gmfp0, gbr11, gbr10, gbr21, gbr20, cohe1, cohe0).
void launch_photon_build(const PhotonBuildMedium *meds, int nmed, const XcomView &x, double *nodes, double *tab,
                         hipStream_t st);

// precision-independent helpers (defined in k64.hip)
void launch_fold(const double *tally, int nbatch, int64_t nreg, double *s, double *s2, hipStream_t st);
void launch_accum(const double *grid, int64_t nreg, double *s, double *s2, hipStream_t st);
void launch_set_u32(uint32_t *p, int n, uint32_t v, hipStream_t st);
void launch_zero_f64(double *p, int64_t n, hipStream_t st);

// Start of a generation: zero the next generation's queue tails and both
// fetch counters.  counts = [photon cur/next, charged cur/next, fetch x2].
void launch_gen_reset(uint32_t *counts, int nxt, hipStream_t st);

// flags[j] = 1 for every job j that owns a particle in either queue
// (meta[] >> 3), over the first min(*count, cap) entries; flags must be
// cleared beforehand.
void launch_job_flags(const uint32_t *meta_p, const uint32_t *count_p, const uint32_t *meta_e,
                      const uint32_t *count_e, uint32_t cap, uint32_t *flags, int njobs, hipStream_t st);

// matRad beamlet column on the device (omc_matrad.c:839-918 with iout = 1,
// then 1420-1479): dose[v] from the tally grids (nslots of them: one per
// batch, or a single one holding every batch) summed and divided by nbatch,
// its maximum, and the compaction of the voxels above rel_thr * max into
// rows/vals in voxel order.
struct SparseScratch {
    double *dose;                 // [nvox]
    unsigned long long *maxkey;   // [1] order-preserving key of the maximum
    uint32_t *blk;                // [nblk + 1] per-block counts, then offsets
    int64_t *rows;                // [nvox]
    double *vals;                 // [nvox]
};
int sparse_blocks(int64_t nvox);
void launch_beamlet_column(const double *tally, int nslots, int nbatch, int isize, int jsize, int ksize, const double *xb,
                           const double *yb, const double *zb, const double *rho, double inc_fluence,
                           double rel_thr, const SparseScratch &S, hipStream_t st);

}  // namespace omc
