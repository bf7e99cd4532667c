// Kernel bodies shared by the FP64 (k64.hip) and FP32 (k32.hip) builds:
// source sampling, batch folding and the launch wrappers the engine calls.
#pragma once

#include "launch.h"
#include "transport.h"

namespace omc {

// The voxel index the reference's linear scans find, `i = lo; while (i < hi
// && b[i + 1] < x) ++i` (omc_dosxyz.c:1040-1055, omc_matrad.c:1236-1248),
// found by bisection: the bounds ascend, so `b[i + 1] < x` holds for a
// prefix of [lo, hi) and the scan stops at the first i where it fails, or
// at hi.  The comparisons are the scan's own (bound rounded to R first), so
// the index is the same; a thread reads log2(hi - lo) bounds instead of up
// to hi - lo, one after the other.
template <typename R>
__device__ __forceinline__ int scan_bounds(const double *b, int lo, int hi, R x)
{
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((R)b[mid + 1] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// initHistory for a dosxyz field (omc_dosxyz.c:964-1068) or a matRad beamlet
// (omc_matrad.c:1084-1254): one thread per history, primary pushed to the
// photon or charged-particle queue.  One launch serves `nseg` segments of
// consecutive histories, possibly of different jobs; segs[].off is the
// exclusive prefix of the segment sizes.
template <typename R>
__global__ void __launch_bounds__(256) k_source(const DevJob *jobs, const DevPhantom<R> *phs,
                                                const SourceSeg *segs, int nseg, uint32_t ntot,
                                                Queue<R> qp, Queue<R> qe, DevStatus *st)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntot) return;
    int lo = 0, hi = nseg - 1;             // last segment with off <= t
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].off <= t) lo = mid; else hi = mid - 1;
    }
    const SourceSeg sg = segs[lo];
    const int job = (int)sg.job;
    const DevJob &J = jobs[job];
    const DevPhantom<R> &g = phs[J.phantom];
    const uint64_t h = sg.first + (t - sg.off);
    Particle<R> p;
    p.job = (uint32_t)job;
    p.hist = (uint32_t)h;
    primary_key(J.s0, J.s1, J.hist0 + h, p.rng.k0, p.rng.k1);
    p.rng.ctr = 0u;
    p.rng.reset_cache();
    p.iq = J.charge;
    R ein;
    if (J.spectrum) {
        R r1 = urand<R>(p.rng), r2 = urand<R>(p.rng);
        int k = (int)fmin(J.deltak * (double)r1, J.deltak - 1.0);
        ein = (R)J.cdf1[k] + r2 * (R)J.cdf2[k];
    } else {
        ein = (R)J.energy;
    }
    p.e = (p.iq != 0) ? ein + (R)RM : ein;
    p.wt = (R)1.0;
    if (J.kind == 0) {
        R rxyz;
        const R ssd = (R)J.ssd;
        if (J.xsize == 0.0 || J.ysize == 0.0) {
            p.x = (R)J.xinl;
            p.y = (R)J.yinl;
            rxyz = m_sqrt(sq(ssd) + sq(p.x) + sq(p.y));
            p.w = ssd / rxyz;
        } else {
            R fw, r3;
            do {
                r3 = urand<R>(p.rng);
                p.x = r3 * (R)J.xsize + (R)J.xinl;
                r3 = urand<R>(p.rng);
                p.y = r3 * (R)J.ysize + (R)J.yinl;
                r3 = urand<R>(p.rng);
                rxyz = m_sqrt(ssd * ssd + p.x * p.x + p.y * p.y);
                p.w = ssd / rxyz;
                fw = p.w * p.w * p.w;
            } while (r3 >= fw);
        }
        p.z = (R)g.zb[0];
        p.u = p.x / rxyz;
        p.v = p.y / rxyz;
        int ix, iy;
        ix = J.xsize == 0.0 ? J.ixinl : scan_bounds<R>(g.xb, J.ixinl - 1, g.isize - 1, p.x);
        iy = J.ysize == 0.0 ? J.iyinl : scan_bounds<R>(g.yb, J.iyinl - 1, g.jsize - 1, p.y);
        p.ir = 1 + ix + iy * g.isize;
    } else {
        R r1 = urand<R>(p.rng), r2 = urand<R>(p.rng);
        R xiso = r1 * (R)J.side1[0] + r2 * (R)J.side2[0] + (R)J.corner[0];
        R yiso = r1 * (R)J.side1[1] + r2 * (R)J.side2[1] + (R)J.corner[1];
        R ziso = r1 * (R)J.side1[2] + r2 * (R)J.side2[2] + (R)J.corner[2];
        R sx = (R)J.src[0], sy = (R)J.src[1], sz = (R)J.src[2];
        R vnorm = m_sqrt(sq(xiso - sx) + sq(yiso - sy) + sq(ziso - sz));
        R u = -(xiso - sx) / vnorm, v = -(yiso - sy) / vnorm, w = -(ziso - sz) / vnorm;
        R ustep = (R)1.0E5, dist;
        if (u > (R)0) { dist = ((R)g.xb[g.isize] - xiso) / u; if (dist < ustep) ustep = dist; }
        if (u < (R)0) { dist = -(xiso - (R)g.xb[0]) / u; if (dist < ustep) ustep = dist; }
        if (v > (R)0) { dist = ((R)g.yb[g.jsize] - yiso) / v; if (dist < ustep) ustep = dist; }
        if (v < (R)0) { dist = -(yiso - (R)g.yb[0]) / v; if (dist < ustep) ustep = dist; }
        if (w > (R)0) { dist = ((R)g.zb[g.ksize] - ziso) / w; if (dist < ustep) ustep = dist; }
        if (w < (R)0) { dist = -(ziso - (R)g.zb[0]) / w; if (dist < ustep) ustep = dist; }
        p.x = xiso + ustep * u;
        p.y = yiso + ustep * v;
        p.z = ziso + ustep * w;
        p.u = -u; p.v = -v; p.w = -w;
        const R tiny = (R)(2.0 * 2.2250738585072014e-308);
        if (p.x < (R)g.xb[0]) p.x = (R)g.xb[0] + tiny;
        if (p.x > (R)g.xb[g.isize]) p.x = (R)g.xb[g.isize] - tiny;
        if (p.y < (R)g.yb[0]) p.y = (R)g.yb[0] + tiny;
        if (p.y > (R)g.yb[g.jsize]) p.y = (R)g.yb[g.jsize] - tiny;
        if (p.z < (R)g.zb[0]) p.z = (R)g.yb[0] + tiny;      // reference's ybounds (omc_matrad.c:1225)
        if (p.z > (R)g.zb[g.ksize]) p.z = (R)g.zb[g.ksize] - tiny;
        const int ix = scan_bounds<R>(g.xb, 0, g.isize - 1, p.x);
        const int iy = scan_bounds<R>(g.yb, 0, g.jsize - 1, p.y);
        const int iz = scan_bounds<R>(g.zb, 0, g.ksize - 1, p.z);
        p.ir = 1 + ix + iy * g.isize + iz * g.isize * g.jsize;
    }
    push<R>(p.iq == 0 ? qp : qe, p, st);
}

template <typename R>
void Launch<R>::source(const DevJob *jobs, const DevPhantom<R> *ph, const SourceSeg *segs, int nseg,
                       uint32_t ntot, const Queue<R> &qp, const Queue<R> &qe, DevStatus *st, hipStream_t s)
{
    if (!ntot || nseg < 1) return;
    hipLaunchKernelGGL(k_source<R>, dim3((ntot + 255) / 256), dim3(256), 0, s, jobs, ph, segs, nseg, ntot, qp, qe,
                       st);
}

template <typename R>
void Launch<R>::photon(const KArgs<R> &a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_photon<R>, dim3(grid), dim3(RNG_LANES), OMC_RNG_DEPTH_P ? RNG_RING_BYTES : 0, s, a);
}

template <typename R>
void Launch<R>::electron(const KArgs<R> &a, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_electron<R>, dim3(grid), dim3(RNG_LANES), OMC_RNG_DEPTH_E ? RNG_RING_BYTES : 0, s, a);
}

// The transport's math functions (m_log ... m_pow in the precision R, as
// the kernels of this build compile them) on n arguments: a diagnostic of
// where the device's arithmetic can differ from the host C library's.
template <typename R>
__global__ void k_math_eval(int fn, int64_t n, const double *x, const double *y, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const R a = (R)x[i], b = (R)y[i];
    R r;
    switch (fn) {
    case 0: r = m_log(a); break;
    case 1: r = m_exp(a); break;
    case 2: r = m_sqrt(a); break;
    case 3: r = m_sin(a); break;
    case 4: r = m_cos(a); break;
    default: r = m_pow(a, b); break;
    }
    out[i] = (double)r;
}

template <typename R>
void Launch<R>::math_eval(int fn, int64_t n, const double *x, const double *y, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_math_eval<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fn, n, x, y, out);
}

// ---- energy-ordered fetch of a charged-particle launch ---------------------
// A counting sort of the queue's records into SORT_BUCKETS buckets of total
// energy, highest first, so that the lanes of a wavefront start on particles
// of similar energy (their table lookups fall on few cache lines, their
// steps are alike) and the longest tracks start first.  The order within a
// bucket follows the LDS atomics and is not fixed; the histories do not
// depend on it.  The queue's count is read on the device, after the photon
// kernel on the same stream has appended to the queue.

// Bucket of total energy e: 12 per octave from 0.5 MeV (below every
// charged-particle cut-off) up, highest energy in bucket 0.
template <typename R> __device__ __forceinline__ int sort_bucket(R e)
{
    const float l = (log2f((float)e) + 1.0f) * 12.0f;
    int b = l > 0.0f ? (int)l : 0;
    b = b > SORT_BUCKETS - 1 ? SORT_BUCKETS - 1 : b;
    return SORT_BUCKETS - 1 - b;
}

// Share of the queue each of the SORT_BLOCKS blocks takes.
__device__ __forceinline__ void sort_share(const uint32_t *count, uint32_t cap, uint32_t &lo, uint32_t &hi)
{
    const uint32_t n = min(*count, cap);
    const uint32_t per = (n + SORT_BLOCKS - 1) / SORT_BLOCKS;
    lo = min(blockIdx.x * per, n);
    hi = min(lo + per, n);
}

// Per-block histograms, bucket-major: hist[b * SORT_BLOCKS + block].
template <typename R>
__global__ void __launch_bounds__(256) k_sort_count(Queue<R> q, uint32_t *hist)
{
    __shared__ uint32_t h[SORT_BUCKETS];
    for (int b = threadIdx.x; b < SORT_BUCKETS; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    uint32_t lo, hi;
    sort_share(q.count, q.cap, lo, hi);
    const R *e = q.e();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) atomicAdd(&h[sort_bucket<R>(e[i])], 1u);
    __syncthreads();
    for (int b = threadIdx.x; b < SORT_BUCKETS; b += blockDim.x) hist[b * SORT_BLOCKS + blockIdx.x] = h[b];
}

// Exclusive prefix sum of the SORT_BUCKETS * SORT_BLOCKS counts, in place,
// by one wavefront (each lane sums a run of the array).  One wavefront fits
// beside the transport kernels of the other pipelines; a block of 1024
// threads waited for a whole CU to empty, 6 ms per launch on average.
template <typename R> __global__ void __launch_bounds__(64) k_sort_scan(uint32_t *hist)
{
    constexpr int N = SORT_BUCKETS * SORT_BLOCKS, RUN = N / 64;
    static_assert(N % 64 == 0, "sort table size");
    const int t = threadIdx.x;
    uint32_t s = 0;
    for (int k = 0; k < RUN; ++k) s += hist[t * RUN + k];
    uint32_t incl = s;                              // inclusive scan of the run sums
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, d);
        if (t >= d) incl += v;
    }
    uint32_t run = incl - s;
    for (int k = 0; k < RUN; ++k) {
        const uint32_t c = hist[t * RUN + k];
        hist[t * RUN + k] = run;
        run += c;
    }
}

// Each block copies its share's records to its offsets in the sorted queue
// (the resume state only for records that carry one, a particle that
// yielded).  The transport kernel then reads the sorted queue in order: its
// fetches stay coalesced (reading through an index instead scatters every
// lane's loads of a wavefront over as many cache lines, one or two records
// used in each).
template <typename R>
__global__ void __launch_bounds__(256) k_sort_scatter(Queue<R> q, const uint32_t *offsets, Queue<R> out)
{
    __shared__ uint32_t base[SORT_BUCKETS];
    for (int b = threadIdx.x; b < SORT_BUCKETS; b += blockDim.x) base[b] = offsets[b * SORT_BLOCKS + blockIdx.x];
    __syncthreads();
    uint32_t lo, hi;
    sort_share(q.count, q.cap, lo, hi);
    const R *e = q.e();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint32_t j = atomicAdd(&base[sort_bucket<R>(e[i])], 1u);
#pragma unroll
        for (int f = 0; f < QR_ST; ++f) out.r(f)[j] = q.r(f)[i];
#pragma unroll
        for (int f = 0; f < QU_SFLAGS; ++f) out.u(f)[j] = q.u(f)[i];
        if (q.meta()[i] & META_RESUME) {
#pragma unroll
            for (int f = QR_ST; f < QR_N; ++f) out.r(f)[j] = q.r(f)[i];
#pragma unroll
            for (int f = QU_SFLAGS; f < QU_N; ++f) out.u(f)[j] = q.u(f)[i];
        }
    }
}

template <typename R>
void Launch<R>::sort_charged(const Queue<R> &q, uint32_t *hist, const Queue<R> &out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sort_count<R>, dim3(SORT_BLOCKS), dim3(256), 0, s, q, hist);
    hipLaunchKernelGGL(k_sort_scan<R>, dim3(1), dim3(64), 0, s, hist);
    hipLaunchKernelGGL(k_sort_scatter<R>, dim3(SORT_BLOCKS), dim3(256), 0, s, q, (const uint32_t *)hist, out);
}

}  // namespace omc
