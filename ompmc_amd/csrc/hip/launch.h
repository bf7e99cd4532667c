// Host-side launch interface of the transport kernels, one instantiation per
// transport precision (k64.hip: double, k32.hip: float).
#pragma once

#include <hip/hip_runtime.h>
#include "common.h"

namespace omc {

template <typename R> struct KArgs;

template <typename R> struct Launch {
    static void source(const DevJob *jobs, const DevPhantom<R> *ph, int job, uint64_t first, uint32_t n,
                       const Queue<R> &qp, const Queue<R> &qe, DevStatus *st, hipStream_t s);
    static void photon(const KArgs<R> &a, int grid, hipStream_t s);
    static void electron(const KArgs<R> &a, int grid, hipStream_t s);
};

// precision-independent helpers (defined in k64.hip)
void launch_fold(const double *tally, int nbatch, int64_t nreg, double *s, double *s2, hipStream_t st);
void launch_set_u32(uint32_t *p, int n, uint32_t v, hipStream_t st);
void launch_zero_f64(double *p, int64_t n, hipStream_t st);

}  // namespace omc
