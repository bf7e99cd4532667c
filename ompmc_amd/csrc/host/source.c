/*
 * Source descriptions: energy spectrum inverse CDF
 * (ucodes/omc_dosxyz/omc_dosxyz.c:368-519) and the omc_dosxyz collimated
 * point source (omc_dosxyz.c:521-632).  Sampling itself happens per history
 * in the transport (device kernels / oracle).
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void omc_spectrum_mono(omc_spectrum *s, double energy)
{
    memset(s, 0, sizeof *s);
    s->spectrum = 0;
    s->energy = energy;
    s->deltak = OMC_INVDIM;
}

int omc_spectrum_from_bins(omc_spectrum *s, int nensrc, double enmin,
                           const double *ensrcd, const double *weight, int imode)
{
    if (nensrc < 1 || nensrc > OMC_MXEBIN)
        return omc_set_error(OMC_ERANGE, "spectrum with %d bins (1..%d)", nensrc, OMC_MXEBIN);
    if (imode != 0 && imode != 1)
        return omc_set_error(OMC_EFORMAT, "Invalid mode number in spectrum file.");
    memset(s, 0, sizeof *s);
    s->spectrum = 1;
    double pdf[OMC_MXEBIN], cdf[OMC_MXEBIN];
    for (int i = 0; i < nensrc; ++i) pdf[i] = weight[i];
    if (imode == 1) {            /* counts per MeV -> counts per bin */
        pdf[0] *= (ensrcd[0] - enmin);
        for (int i = 1; i < nensrc; ++i) pdf[i] *= (ensrcd[i] - ensrcd[i - 1]);
    }
    cdf[0] = pdf[0];
    for (int i = 1; i < nensrc; ++i) cdf[i] = cdf[i - 1] + pdf[i];
    double fnorm = 1.0 / cdf[nensrc - 1];
    s->deltak = OMC_INVDIM;
    double gridsz = 1.0f / s->deltak;
    for (int i = 0; i < nensrc; ++i) cdf[i] *= fnorm;
    for (int k = 0; k < s->deltak; ++k) {
        double ak = (double)k * gridsz;
        int i;
        for (i = 0; i < nensrc; ++i)
            if (ak <= cdf[i]) break;
        if (i >= nensrc) i = nensrc - 1;   /* only reachable through round-off */
        s->cdfinv1[k] = (i != 0) ? ensrcd[i - 1] : enmin;
        s->cdfinv2[k] = ensrcd[i] - s->cdfinv1[k];
    }
    return OMC_OK;
}

int omc_spectrum_read(omc_spectrum *s, const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    char buf[OMC_INP_BUF];
    int nensrc = 0, imode = 0;
    double enmin = 0.0;
    double e[OMC_MXEBIN], w[OMC_MXEBIN];
    if (!fgets(buf, sizeof buf, fp) || !fgets(buf, sizeof buf, fp) ||
        sscanf(buf, "%d %lf %d", &nensrc, &enmin, &imode) != 3) {
        fclose(fp);
        return omc_set_error(OMC_EFORMAT, "%s: malformed spectrum header", path);
    }
    if (nensrc < 1 || nensrc > OMC_MXEBIN) {
        fclose(fp);
        return omc_set_error(OMC_ERANGE, "Number of energy bins = %d is greater than max allowed = %d", nensrc, OMC_MXEBIN);
    }
    for (int i = 0; i < nensrc; ++i) {
        if (!fgets(buf, sizeof buf, fp) || sscanf(buf, "%lf %lf", &e[i], &w[i]) != 2) {
            fclose(fp);
            return omc_set_error(OMC_EFORMAT, "%s: bad spectrum bin %d", path, i);
        }
    }
    fclose(fp);
    return omc_spectrum_from_bins(s, nensrc, enmin, e, w, imode);
}

int omc_source_dosxyz_setup(omc_source_dosxyz *s, const omc_phantom *p,
                            double xinl, double xinu, double yinl, double yinu,
                            double ssd, int charge)
{
    if (charge < -1 || charge > 1) return omc_set_error(OMC_EARG, "Particle kind not recognized.");
    if (ssd < 0) return omc_set_error(OMC_EARG, "SSD must be greater than zero.");
    s->charge = charge;
    s->ssd = ssd;
    const double *xb = p->xbounds, *yb = p->ybounds;
    int is = p->isize, js = p->jsize;

    if (xinl < xb[0]) xinl = xb[0];
    if (xinu <= xinl) xinu = xinl;        /* pencil beam */
    if (xinu > xb[is]) xinu = xb[is];
    if (xinl > xb[is]) xinl = xb[is];
    int ixl = 0;
    while (ixl < is && (xb[ixl] <= xinl) && (xb[ixl + 1] < xinl)) ixl++;
    int ixu = ixl - 1;
    while (ixu + 1 <= is && (ixu < 0 || xb[ixu] <= xinu) && (xb[ixu + 1] < xinu)) ixu++;

    if (yinl < yb[0]) yinl = yb[0];
    if (yinu <= yinl) yinu = yinl;
    if (yinu > yb[js]) yinu = yb[js];
    if (yinl > yb[js]) yinl = yb[js];
    int iyl = 0;
    while (iyl < js && (yb[iyl] <= yinl) && (yb[iyl + 1] < yinl)) iyl++;
    int iyu = iyl - 1;
    while (iyu + 1 <= js && (iyu < 0 || yb[iyu] <= yinu) && (yb[iyu + 1] < yinu)) iyu++;

    s->xinl = xinl; s->xinu = xinu; s->yinl = yinl; s->yinu = yinu;
    s->ixinl = ixl; s->ixinu = ixu; s->iyinl = iyl; s->iyinu = iyu;
    s->xsize = xinu - xinl;
    s->ysize = yinu - yinl;
    return OMC_OK;
}
