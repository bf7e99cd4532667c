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


void launch_fold(const double *tally, int nbatch, int64_t nreg, double *s, double *s2, hipStream_t st)
{
    if (nreg <= 0) return;
    hipLaunchKernelGGL(k_fold, dim3((unsigned)((nreg + 255) / 256)), dim3(256), 0, st, tally, nbatch, nreg, s, s2);
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
