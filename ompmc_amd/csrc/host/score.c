/*
 * Batch statistics and output files.
 *   omc_batch_results   accumulateResults  (ucodes/omc_dosxyz/omc_dosxyz.c:719-799)
 *   omc_write_3ddose    outputResults body (omc_dosxyz.c:841-879)
 *   omc_sparse_column   matRad Dij column  (ucodes/omc_matrad/omc_matrad.c:1420-1479)
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int omc_batch_results(const omc_phantom *p, const double *sum, const double *sum2,
                      int nbatch, double inc_fluence, int iout, double *dose, double *unc)
{
    if (nbatch < 1) return omc_set_error(OMC_EARG, "nbatch must be >= 1");
    const int is = p->isize, js = p->jsize, ks = p->ksize;
    const int imax = is, ijmax = is * js;
    for (int iz = 0; iz < ks; ++iz)
        for (int iy = 0; iy < js; ++iy)
            for (int ix = 0; ix < is; ++ix) {
                int irl = 1 + ix + iy * imax + iz * ijmax;
                double endep = sum[irl], endep2 = sum2[irl], u;
                endep /= (double)nbatch;
                endep2 /= (double)nbatch;
                if (endep != 0.0) {
                    u = endep2 - endep * endep;
                    u /= (double)(nbatch - 1);
                    u = sqrt(u) / endep;
                } else {
                    endep = 0.0;
                    u = 0.9999999;
                }
                if (iout) {
                    double mass = (p->xbounds[ix + 1] - p->xbounds[ix]) *
                                  (p->ybounds[iy + 1] - p->ybounds[iy]) *
                                  (p->zbounds[iz + 1] - p->zbounds[iz]);
                    mass *= p->rho[irl - 1];
                    endep *= 1.602E-10 / (mass * inc_fluence);
                } else {
                    endep /= inc_fluence;
                }
                /* dose in air is not reported */
                if (p->rho[irl - 1] < 0.044) { endep = 0.0; u = 0.9999999; }
                dose[irl - 1] = endep;
                unc[irl - 1] = u;
            }
    return OMC_OK;
}

int omc_write_3ddose(const char *path, const omc_phantom *p, const double *dose, const double *unc)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    fprintf(fp, "%5d%5d%5d\n", p->isize, p->jsize, p->ksize);
    for (int i = 0; i <= p->isize; ++i) fprintf(fp, "%f ", p->xbounds[i]);
    fprintf(fp, "\n");
    for (int i = 0; i <= p->jsize; ++i) fprintf(fp, "%f ", p->ybounds[i]);
    fprintf(fp, "\n");
    for (int i = 0; i <= p->ksize; ++i) fprintf(fp, "%f ", p->zbounds[i]);
    fprintf(fp, "\n");
    size_t nvox = (size_t)p->isize * p->jsize * p->ksize;
    for (size_t v = 0; v < nvox; ++v) fprintf(fp, "%e ", dose[v]);
    fprintf(fp, "\n");
    for (size_t v = 0; v < nvox; ++v) fprintf(fp, "%f ", unc[v]);
    fprintf(fp, "\n");
    if (fclose(fp)) return omc_set_error(OMC_EIO, "error writing %s", path);
    return OMC_OK;
}

int omc_read_3ddose(const char *path, int *isize, int *jsize, int *ksize,
                    double **dose, double **unc)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    int n[3];
    if (fscanf(fp, "%d %d %d", &n[0], &n[1], &n[2]) != 3) goto bad;
    double tmp;
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i <= n[d]; ++i) if (fscanf(fp, "%lf", &tmp) != 1) goto bad;
    size_t nvox = (size_t)n[0] * n[1] * n[2];
    *dose = malloc(nvox * sizeof(double));
    *unc = malloc(nvox * sizeof(double));
    if (!*dose || !*unc) { free(*dose); free(*unc); fclose(fp); return omc_set_error(OMC_ENOMEM, "oom"); }
    for (size_t v = 0; v < nvox; ++v) if (fscanf(fp, "%lf", &(*dose)[v]) != 1) goto bad2;
    for (size_t v = 0; v < nvox; ++v) if (fscanf(fp, "%lf", &(*unc)[v]) != 1) goto bad2;
    fclose(fp);
    *isize = n[0]; *jsize = n[1]; *ksize = n[2];
    return OMC_OK;
bad2:
    free(*dose); free(*unc); *dose = *unc = NULL;
bad:
    fclose(fp);
    return omc_set_error(OMC_EFORMAT, "%s: malformed 3ddose file", path);
}

int64_t omc_sparse_column(const double *dose, int64_t nvox, double rel_thr,
                          int64_t *rows, double *vals, int64_t cap)
{
    double dmax = 0.0;
    for (int64_t v = 0; v < nvox; ++v) if (dose[v] > dmax) dmax = dose[v];
    double thr = dmax * rel_thr;
    int64_t n = 0;
    for (int64_t v = 0; v < nvox; ++v) {
        if (dose[v] > thr) {
            if (n >= cap) return omc_set_error(OMC_ERANGE, "sparse column capacity %lld exceeded", (long long)cap);
            rows[n] = v;
            vals[n] = dose[v];
            ++n;
        }
    }
    return n;
}
