/*
 * omc_dosxyz - drop-in for ompMC's omc_dosxyz executable with the transport
 * on the GPU.
 *
 *     omc_dosxyz -i <input stem> -o <output name> [--verbose] [--fp32] [--device N]
 *
 * Same .inp keys, loaders and .3ddose output as the reference
 * (omc_dosxyz.c:1072-1307); the batch loop `initHistory(); shower();` +
 * `accumEndep()` (1237-1263) becomes one omc_dev_run_dosxyz() call.
 * Histories draw from per-particle Philox streams keyed by "rng seeds"
 * instead of RANMAR, so results agree with the reference statistically, and
 * with the oracle's PHILOX mode history for history.  Photon splitting
 * (nsplit > 1) runs on the device as in the reference (ompmc.c:1903-2118).
 */
#include "driver_common.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    const double tbegin = drv_time();
    drv_args a;
    drv_parse_args(&a, argc, argv);

    omc_dosxyz_problem *p = calloc(1, sizeof *p);
    if (!p) return EXIT_FAILURE;
    int rc = omc_dosxyz_problem_load(p, a.input);
    if (rc) drv_die("reading the input", rc, 0);
    if (p->nsplit > 1) printf("Photon splitting enabled, nsplit = %d\n", p->nsplit);

    omc_dev_ctx *ctx = drv_open_device(&a, &p->tab);
    int pid = -1;
    rc = omc_dev_upload_phantom(ctx, &p->ph, &p->reg, &pid);
    if (rc) drv_die("phantom upload", rc, 1);

    printf("Total number of particle histories: %d\n", p->nhist);
    printf("Number of statistical batches: %d\n", p->nbatch);
    printf("Histories per batch: %d\n", p->nperbatch);
    printf("Execution time up to this point : %8.2f seconds\n", drv_time() - tbegin);

    const int nreg = p->reg.nreg;
    double *sum = calloc((size_t)nreg, sizeof(double));
    double *sum2 = calloc((size_t)nreg, sizeof(double));
    if (!sum || !sum2) return EXIT_FAILURE;
    omc_dev_stats st;
    omc_dev_run_params rp;
    memset(&rp, 0, sizeof rp);
    rp.phantom_id = pid;
    rp.nsplit = p->nsplit < 1 ? 1 : p->nsplit;
    rp.seed[0] = (uint32_t)p->seeds[0];
    rp.seed[1] = (uint32_t)p->seeds[1];
    rp.nperbatch = p->nperbatch;
    rp.nbatch = p->nbatch;
    rp.stats = &st;
    rc = omc_dev_run_dosxyz(ctx, &p->src, &rp, sum, sum2);
    if (rc) drv_die("transport", rc, 1);
    printf("Simulation finished\n");
    printf("Execution time up to this point : %8.2f seconds\n", drv_time() - tbegin);
    printf("Transport: %.1f ms, %.4g histories/s\n", st.transport_ms,
           st.transport_ms > 0 ? 1e3 * (double)p->nhist / st.transport_ms : 0.0);

    if (a.verbose) {
        double etot = 0.0;
        for (int irl = 1; irl < nreg; ++irl) etot += sum[irl];
        printf("Energy deposited in the phantom per batch: %g MeV\n", etot / p->nbatch);
        printf("Device counters: %lld photon steps, %lld electron steps, %lld deposits, %lld generations\n",
               (long long)st.photon_steps, (long long)st.electron_steps, (long long)st.deposits,
               (long long)st.generations);
    }

    /* outputResults (omc_dosxyz.c:801-886) with iout = 1, fluence nperbatch */
    const int nvox = nreg - 1;
    double *dose = malloc(sizeof(double) * (size_t)nvox), *unc = malloc(sizeof(double) * (size_t)nvox);
    if (!dose || !unc) return EXIT_FAILURE;
    rc = omc_batch_results(&p->ph, sum, sum2, p->nbatch, (double)p->nperbatch, 1, dose, unc);
    if (rc) drv_die("results", rc, 0);
    char path[2 * OMC_INP_BUF];
    rc = drv_output_path(&p->inp, a.output, ".3ddose", path, sizeof path);
    if (rc) return EXIT_FAILURE;
    rc = omc_write_3ddose(path, &p->ph, dose, unc);
    if (rc) drv_die("writing the dose", rc, 0);

    omc_dev_destroy(ctx);
    omc_dosxyz_problem_free(p);
    free(p); free(sum); free(sum2); free(dose); free(unc);
    printf("Total execution time : %8.5f seconds\n", drv_time() - tbegin);
    return EXIT_SUCCESS;
}
