/*
 * Test-only driver for the sanitizer build (make -C oracle sanitize): the
 * host library (loaders, table builders, phantom, sources, scoring) and the
 * CPU oracle compiled together with AddressSanitizer and
 * UndefinedBehaviorSanitizer, exercised end to end on one .inp problem:
 *
 *   omc_sanitize <inp-stem> <out.3ddose>
 *
 * loads the problem, runs the oracle in RANMAR mode (serial) and in PHILOX
 * mode (2 threads), writes the .3ddose, reads it back, and frees
 * everything, so that LeakSanitizer sees every allocation returned.
 */
#include "omc_host.h"
#include "omc_oracle.h"

#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: omc_sanitize <inp-stem> <out.3ddose>\n");
        return 2;
    }
    omc_dosxyz_problem p;
    int rc = omc_dosxyz_problem_load(&p, argv[1]);
    if (rc) { fprintf(stderr, "load: %s\n", omc_last_error()); return 1; }
    const int nreg = p.reg.nreg, nvox = nreg - 1;
    double *s = calloc(nreg, sizeof(double)), *s2 = calloc(nreg, sizeof(double));
    double *dose = calloc(nvox, sizeof(double)), *unc = calloc(nvox, sizeof(double));
    ora_ctx *c = ora_create(&p.tab, &p.reg, &p.ph, p.nsplit, ORA_RNG_RANMAR);
    ora_seed_ranmar(c, p.seeds[0], p.seeds[1]);
    rc = ora_run_dosxyz(c, &p.src, p.nperbatch, p.nbatch, s, s2);
    ora_destroy(c);
    if (!rc) rc = omc_batch_results(&p.ph, s, s2, p.nbatch, (double)p.nperbatch, 1, dose, unc);
    if (!rc) rc = omc_write_3ddose(argv[2], &p.ph, dose, unc);
    int ni = 0, nj = 0, nk = 0;
    double *d2 = NULL, *u2 = NULL;
    if (!rc) rc = omc_read_3ddose(argv[2], &ni, &nj, &nk, &d2, &u2);
    if (!rc && (ni != p.ph.isize || nj != p.ph.jsize || nk != p.ph.ksize)) rc = 1;
    free(d2);
    free(u2);
    ora_counters cnt;
    if (!rc && p.nsplit == 1)
        rc = ora_run_dosxyz_parallel(&p.tab, &p.reg, &p.ph, &p.src, 5, 6, 0, p.nperbatch, p.nbatch, 2, s, s2,
                                     &cnt);
    free(s);
    free(s2);
    free(dose);
    free(unc);
    omc_dosxyz_problem_free(&p);
    if (rc) fprintf(stderr, "failed: %d %s\n", rc, omc_last_error());
    else printf("ok\n");
    return rc ? 1 : 0;
}
