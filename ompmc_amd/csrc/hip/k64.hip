// FP64 transport kernels (built with FMA contraction off so that every
// operation rounds as in the CPU reference).
#include "kernels_impl.h"

namespace omc {

template struct Launch<double>;

// accumEndep over all batches (omc_dosxyz.c:696-717): per region the sum of
// the batch energies and of their squares, batches in order.
__global__ void __launch_bounds__(256) k_fold(const double *tally, int nbatch, int64_t nreg,
                                              double *s, double *s2)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nreg) return;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < nbatch; ++k) {
        double e = tally[(size_t)k * nreg + i];
        a += e;
        b += e * e;
    }
    s[i] = a;
    s2[i] = b;
}

// ---- photon tables on the device (initPhotonData, src/ompmc.c:216-665) ----
// The host builder's arithmetic in its order (FP64, no contraction): a
// table entry differs from the host's only where the device's exp or log
// rounds a value differently from glibc's.

// Klein-Nishina total cross section per electron (src/ompmc.c:451-469).
__device__ double kn_sigma0(double e)
{
    const double con = 0.1274783851;
    const double ko = e / RM;
    if (ko < 0.01) return 8.0 * con / 3.0 * (1.0 - ko * (2.0 - ko * (5.2 - 13.3 * ko))) / RM;
    const double c1 = 1.0 / (ko * ko);
    const double c2 = 1.0 - 2.0 * (1.0 + ko) * c1;
    const double c3 = (1.0 + 2.0 * ko) * c1;
    const double eps2 = 1.0;
    const double eps1 = 1.0 / (1.0 + 2.0 * ko);
    return (c1 * (1.0 / eps1 - 1.0 / eps2) + c2 * log(eps2 / eps1) + eps2 * (c3 + 0.5 * eps2) -
            eps1 * (c3 + 0.5 * eps1)) / e * con;
}

// Partial cross section of a medium at log energy gle (get_data,
// src/ompmc.c:347-449): each element's data interpolated log-log, summed in
// ascending Z weighted by the atom fractions.  Pair (kind 1) and triplet
// (kind 2) data get a threshold point at eth = 2m / 4m, log sigma corrected
// by -3 log(1 - eth/E) at every point, and the factor (1 - eth/E)^3.
__device__ double medium_sigma(int kind, const PhotonBuildMedium &m, const XcomView &x, int set, double gle,
                               double e)
{
    double res = 0.0;
    for (int i = 0; i < m.ne; ++i) {
        const int z = (int)(m.z[i] + 0.5) - 1;
        const int n0 = x.n[set * OMC_MXELEMENT + z];
        const double *le = x.le + x.off[set * OMC_MXELEMENT + z];
        const double *ls = x.ls + x.off[set * OMC_MXELEMENT + z];
        double sig = 0.0;
        if (kind == 0) {
            if (!((gle < le[0]) || (gle >= le[n0 - 1]))) {
                int k;
                for (k = 0; k < n0 - 1; ++k)
                    if ((gle >= le[k]) && (gle < le[k + 1])) break;
                const double p = (gle - le[k]) / (le[k + 1] - le[k]);
                sig = exp(p * ls[k + 1] + (1.0 - p) * ls[k]);
            }
        } else {
            const double eth = kind == 1 ? 2.0 * RM : 4.0 * RM;
            const int n = n0 + 1;
            // point j of the extended set: j = 0 at the threshold (with the
            // corrected log sigma of the first data point), j >= 1 data point j - 1
            const double le_th = log(eth);
            auto lej = [&](int j) { return j == 0 ? le_th : le[j - 1]; };
            auto lsj = [&](int j) {
                const int q = j == 0 ? 0 : j - 1;
                return ls[q] - 3.0 * log(1.0 - eth / exp(le[q]));
            };
            if ((gle < lej(0)) || (gle >= lej(n - 1))) {
                sig = (gle < lej(0)) ? 0.0 : exp(lsj(n - 1));
            } else {
                int k;
                for (k = 0; k < n - 1; ++k)
                    if ((gle >= lej(k)) && (gle < lej(k + 1))) break;
                const double p = (gle - lej(k)) / (lej(k + 1) - lej(k));
                sig = exp(p * lsj(k + 1) + (1.0 - p) * lsj(k));
            }
            if (e > eth) sig *= (1.0 - eth / e) * (1.0 - eth / e) * (1.0 - eth / e);
        }
        res += m.pz[i] * sig;
    }
    return res;
}

// Node values at the grid energies, one block per medium.
__global__ void __launch_bounds__(256) k_photon_nodes(const PhotonBuildMedium *meds, XcomView x, double *nodes)
{
    const PhotonBuildMedium &m = meds[blockIdx.x];
    double *out = nodes + (size_t)blockIdx.x * 4 * MXGE;
    for (int j = threadIdx.x; j < MXGE; j += blockDim.x) {
        const double gle = ((double)(j + 1) - m.ge0) / m.ge1;
        const double e = exp(gle);
        const double s_photo = medium_sigma(0, m, x, 0, gle, e);
        const double s_rayl = medium_sigma(0, m, x, 1, gle, e);
        const double s_pair = medium_sigma(1, m, x, 2, gle, e);
        const double s_trip = medium_sigma(2, m, x, 3, gle, e);
        const double sig_kn = m.sum_z * kn_sigma0(e);
        const double sig_p = s_pair + s_trip;
        const double sigma = sig_kn + sig_p + s_photo;
        const double gbr1 = sig_p / sigma;
        out[j] = 1.0 / (sigma * m.con2);
        out[MXGE + j] = gbr1;
        out[2 * MXGE + j] = gbr1 + sig_kn / sigma;
        out[3 * MXGE + j] = sigma / (s_rayl + sigma);
    }
}

// Linear fits in log energy: bin b from nodes b and b + 1 at node b + 1's
// log energy; the last bin keeps the slope of the one before it.
__global__ void __launch_bounds__(256) k_photon_bins(const PhotonBuildMedium *meds, int nmed, const double *nodes,
                                                     double *tab)
{
    const int m = blockIdx.x;
    const double ge0 = meds[m].ge0, ge1 = meds[m].ge1;
    const double *nd = nodes + (size_t)m * 4 * MXGE;
    const size_t plane = (size_t)nmed * MXGE;
    for (int b = threadIdx.x; b < MXGE; b += blockDim.x) {
        const int j = b < MXGE - 1 ? b + 1 : MXGE - 1;
        const double gle = ((double)(j + 1) - ge0) / ge1;
        const size_t ix = (size_t)m * MXGE + b;
        for (int q = 0; q < 4; ++q) {
            const double v = nd[q * MXGE + j];
            const double slope = (v - nd[q * MXGE + j - 1]) * ge1;
            tab[(2 * q) * plane + ix] = slope;
            tab[(2 * q + 1) * plane + ix] = v - slope * gle;
        }
    }
}

void launch_photon_build(const PhotonBuildMedium *meds, int nmed, const XcomView &x, double *nodes, double *tab,
                         hipStream_t st)
{
    hipLaunchKernelGGL(k_photon_nodes, dim3(nmed), dim3(256), 0, st, meds, x, nodes);
    hipLaunchKernelGGL(k_photon_bins, dim3(nmed), dim3(256), 0, st, meds, nmed, (const double *)nodes, tab);
}

// One finished batch of a job admitted batch by batch: its energies added to
// the job's sums and their squares to the sums of squares (the same per-batch
// terms as k_fold, in the order the batches end).
__global__ void __launch_bounds__(256) k_accum(const double *grid, int64_t nreg, double *s, double *s2)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nreg) return;
    const double e = grid[i];
    s[i] += e;
    s2[i] += e * e;
}

__global__ void k_set_u32(uint32_t *p, int n, uint32_t v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_f64(double *p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0.0;
}

__global__ void k_gen_reset(uint32_t *counts, int nxt)
{
    if (threadIdx.x == 0) {
        counts[nxt] = 0u;
        counts[2 + nxt] = 0u;
        counts[4] = 0u;
        counts[5] = 0u;
        counts[6] = 0u;         // longest charged-particle track of the generation
        counts[7] = 0u;         // longest photon track
    }
}

// One flag store per distinct job in a wavefront: lanes holding the same job
// elect one writer.
__device__ __forceinline__ void flag_jobs(const uint32_t *meta, uint32_t n, uint32_t *flags, int njobs)
{
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
        uint32_t i = base + threadIdx.x;
        bool have = i < n;
        uint32_t job = have ? (meta[i] >> 3) : 0u;
        while (__ballot(have)) {
            uint32_t lead = __shfl(job, __ffsll((unsigned long long)__ballot(have)) - 1);
            bool mine = have && job == lead;
            if (mine && __ffsll((unsigned long long)__ballot(mine)) - 1 == (int)__lane_id() && lead < (uint32_t)njobs)
                flags[lead] = 1u;
            if (mine) have = false;
        }
    }
}

__global__ void __launch_bounds__(256) k_job_flags(const uint32_t *meta_p, const uint32_t *count_p,
                                                   const uint32_t *meta_e, const uint32_t *count_e, uint32_t cap,
                                                   uint32_t *flags, int njobs)
{
    flag_jobs(meta_p, min(*count_p, cap), flags, njobs);
    flag_jobs(meta_e, min(*count_e, cap), flags, njobs);
}

// ---- matRad beamlet column ------------------------------------------------
// Order-preserving map of doubles to unsigned keys (negative values reverse).
__device__ __forceinline__ unsigned long long dkey(double d)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dunkey(unsigned long long k)
{
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

__global__ void k_bl_init(unsigned long long *maxkey)
{
    if (threadIdx.x == 0) *maxkey = dkey(0.0);   // doseMax starts at 0 (omc_matrad.c:1421)
}

// accumulateResults with iout = 1 (omc_matrad.c:839-918) for one beamlet:
// batch sums in batch order, mean over batches, conversion to Gy per
// incident history, zero in voxels with density < 0.044.
__global__ void __launch_bounds__(256) k_bl_dose(const double *tally, int nslots, int nbatch, int isize, int jsize,
                                                 int ksize,
                                                 const double *xb, const double *yb, const double *zb,
                                                 const double *rho, double inc_fluence, double *dose,
                                                 unsigned long long *maxkey)
{
    __shared__ unsigned long long wmax[4];
    const int64_t nvox = (int64_t)isize * jsize * ksize, nreg = nvox + 1;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long key = dkey(0.0);
    if (v < nvox) {
        double e = 0.0;
        for (int b = 0; b < nslots; ++b) e += tally[(size_t)b * nreg + v + 1];
        e /= (double)nbatch;
        const int ix = (int)(v % isize), iy = (int)((v / isize) % jsize), iz = (int)(v / ((int64_t)isize * jsize));
        double mass = (xb[ix + 1] - xb[ix]) * (yb[iy + 1] - yb[iy]) * (zb[iz + 1] - zb[iz]);
        mass *= rho[v];
        e *= 1.602E-10 / (mass * inc_fluence);
        if (rho[v] < 0.044) e = 0.0;
        dose[v] = e;
        key = dkey(e);
    }
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long k2 = __shfl_xor(key, o);
        key = k2 > key ? k2 : key;
    }
    if (__lane_id() == 0) wmax[threadIdx.x >> 6] = key;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmax[0];
        for (int w = 1; w < 4; ++w) m = wmax[w] > m ? wmax[w] : m;
        atomicMax(maxkey, m);
    }
}

__device__ __forceinline__ bool bl_keep(const double *dose, int64_t v, int64_t nvox, double thresh)
{
    return v < nvox && dose[v] > thresh;
}

__global__ void __launch_bounds__(256) k_bl_count(const double *dose, int64_t nvox, const unsigned long long *maxkey,
                                                  double rel_thr, uint32_t *blk)
{
    __shared__ uint32_t wc[4];
    const double thresh = dunkey(*maxkey) * rel_thr;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t c = (uint32_t)__popcll(__ballot(bl_keep(dose, v, nvox, thresh)));
    if (__lane_id() == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Exclusive scan of blk[0..n) in place, total in blk[n]; one block.
__global__ void __launch_bounds__(1024) k_bl_scan(uint32_t *blk, int n)
{
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        uint32_t x = i < n ? blk[i] : 0u, incl = x;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t before = carry;
        for (int k = 0; k < w; ++k) before += wsum[k];
        if (i < n) blk[i] = before + incl - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) blk[n] = carry;
}

__global__ void __launch_bounds__(256) k_bl_emit(const double *dose, int64_t nvox, const unsigned long long *maxkey,
                                                 double rel_thr, const uint32_t *blk, int64_t *rows, double *vals)
{
    __shared__ uint32_t wc[4];
    const double thresh = dunkey(*maxkey) * rel_thr;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool keep = bl_keep(dose, v, nvox, thresh);
    const unsigned long long m = __ballot(keep);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (__lane_id() == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    if (keep) {
        uint32_t off = blk[blockIdx.x] + below;
        for (int k = 0; k < (int)(threadIdx.x >> 6); ++k) off += wc[k];
        rows[off] = v;
        vals[off] = dose[v];
    }
}

int sparse_blocks(int64_t nvox) { return (int)((nvox + 255) / 256); }

void launch_beamlet_column(const double *tally, int nslots, int nbatch, int isize, int jsize, int ksize, const double *xb,
                           const double *yb, const double *zb, const double *rho, double inc_fluence,
                           double rel_thr, const SparseScratch &S, hipStream_t st)
{
    const int64_t nvox = (int64_t)isize * jsize * ksize;
    const int nblk = sparse_blocks(nvox);
    hipLaunchKernelGGL(k_bl_init, dim3(1), dim3(64), 0, st, S.maxkey);
    hipLaunchKernelGGL(k_bl_dose, dim3(nblk), dim3(256), 0, st, tally, nslots, nbatch, isize, jsize, ksize, xb, yb, zb,
                       rho, inc_fluence, S.dose, S.maxkey);
    hipLaunchKernelGGL(k_bl_count, dim3(nblk), dim3(256), 0, st, S.dose, nvox, S.maxkey, rel_thr, S.blk);
    hipLaunchKernelGGL(k_bl_scan, dim3(1), dim3(1024), 0, st, S.blk, nblk);
    hipLaunchKernelGGL(k_bl_emit, dim3(nblk), dim3(256), 0, st, S.dose, nvox, S.maxkey, rel_thr, S.blk, S.rows,
                       S.vals);
}

void launch_gen_reset(uint32_t *counts, int nxt, hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_reset, dim3(1), dim3(64), 0, st, counts, nxt);
}

void launch_job_flags(const uint32_t *meta_p, const uint32_t *count_p, const uint32_t *meta_e,
                      const uint32_t *count_e, uint32_t cap, uint32_t *flags, int njobs, hipStream_t st)
{
    hipLaunchKernelGGL(k_job_flags, dim3(512), dim3(256), 0, st, meta_p, count_p, meta_e, count_e, cap, flags,
                       njobs);
}

void launch_fold(const double *tally, int nbatch, int64_t nreg, double *s, double *s2, hipStream_t st)
{
    if (nreg <= 0) return;
    hipLaunchKernelGGL(k_fold, dim3((unsigned)((nreg + 255) / 256)), dim3(256), 0, st, tally, nbatch, nreg, s, s2);
}

void launch_accum(const double *grid, int64_t nreg, double *s, double *s2, hipStream_t st)
{
    if (nreg <= 0) return;
    hipLaunchKernelGGL(k_accum, dim3((unsigned)((nreg + 255) / 256)), dim3(256), 0, st, grid, nreg, s, s2);
}

void launch_set_u32(uint32_t *p, int n, uint32_t v, hipStream_t st)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_set_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

void launch_zero_f64(double *p, int64_t n, hipStream_t st)
{
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}

}  // namespace omc
