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
