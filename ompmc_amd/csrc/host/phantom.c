/*
 * egsphant voxel phantoms (ucodes/omc_dosxyz/omc_dosxyz.c:62-175) and the
 * region table the transport reads (omc_dosxyz.c:890-962).
 *
 * Voxel order is i + j*isize + k*isize*jsize; region = voxel + 1, region 0
 * is the outside world.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void omc_phantom_free(omc_phantom *p)
{
    free(p->xbounds); free(p->ybounds); free(p->zbounds);
    free(p->med); free(p->rho);
    memset(p, 0, sizeof *p);
}

int omc_phantom_alloc(omc_phantom *p, int nmed, int isize, int jsize, int ksize)
{
    memset(p, 0, sizeof *p);
    if (nmed < 1 || nmed > OMC_MXMED)
        return omc_set_error(OMC_ERANGE, "phantom with %d media (1..%d supported)", nmed, OMC_MXMED);
    if (isize < 1 || jsize < 1 || ksize < 1)
        return omc_set_error(OMC_EARG, "phantom dimensions %d %d %d", isize, jsize, ksize);
    size_t nvox = (size_t)isize * jsize * ksize;
    if (nvox >= (size_t)1 << 31)
        return omc_set_error(OMC_ERANGE, "phantom with %zu voxels exceeds 2^31", nvox);
    p->nmed = nmed;
    p->isize = isize; p->jsize = jsize; p->ksize = ksize;
    p->xbounds = calloc(isize + 1, sizeof(double));
    p->ybounds = calloc(jsize + 1, sizeof(double));
    p->zbounds = calloc(ksize + 1, sizeof(double));
    p->med = calloc(nvox, sizeof(int));
    p->rho = calloc(nvox, sizeof(double));
    if (!p->xbounds || !p->ybounds || !p->zbounds || !p->med || !p->rho) {
        omc_phantom_free(p);
        return omc_set_error(OMC_ENOMEM, "out of memory for phantom");
    }
    return OMC_OK;
}

static void strip_spaces(char *dst, const char *src, size_t cap)
{
    size_t o = 0;
    for (; *src && o + 1 < cap; ++src)
        if (!isspace((unsigned char)*src)) dst[o++] = *src;
    dst[o] = '\0';
}

int omc_phantom_read(omc_phantom *p, const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    char buf[OMC_INP_BUF];
    int nmed, is, js, ks;
    char names[OMC_MXMED][64];
    if (!fgets(buf, sizeof buf, fp)) goto bad;
    nmed = atoi(buf);
    if (nmed < 1 || nmed > OMC_MXMED) { fclose(fp); return omc_set_error(OMC_ERANGE, "%s: %d media", path, nmed); }
    for (int i = 0; i < nmed; ++i) {
        if (!fgets(buf, sizeof buf, fp)) goto bad;
        strip_spaces(names[i], buf, sizeof names[i]);
    }
    if (!fgets(buf, sizeof buf, fp)) goto bad;           /* ESTEPE line, unused */
    if (!fgets(buf, sizeof buf, fp)) goto bad;
    if (sscanf(buf, "%d %d %d", &is, &js, &ks) != 3) goto bad;
    int rc = omc_phantom_alloc(p, nmed, is, js, ks);
    if (rc) { fclose(fp); return rc; }
    for (int i = 0; i < nmed; ++i) memcpy(p->names[i], names[i], 64);
    for (int i = 0; i <= is; ++i) if (fscanf(fp, "%lf", &p->xbounds[i]) != 1) goto bad2;
    for (int i = 0; i <= js; ++i) if (fscanf(fp, "%lf", &p->ybounds[i]) != 1) goto bad2;
    for (int i = 0; i <= ks; ++i) if (fscanf(fp, "%lf", &p->zbounds[i]) != 1) goto bad2;
    if (!fgets(buf, sizeof buf, fp)) goto bad2;
    /* medium digits: isize characters per (j,k) row, a blank line per slice */
    for (int k = 0; k < ks; ++k) {
        for (int j = 0; j < js; ++j) {
            for (int i = 0; i < is; ++i) {
                int c = fgetc(fp);
                if (c == EOF) goto bad2;
                p->med[i + j * is + (size_t)k * js * is] = c - '0';
            }
            if (!fgets(buf, sizeof buf, fp)) goto bad2;
        }
        if (!fgets(buf, sizeof buf, fp)) goto bad2;
    }
    for (int k = 0; k < ks; ++k) {
        for (int j = 0; j < js; ++j)
            for (int i = 0; i < is; ++i)
                if (fscanf(fp, "%lf", &p->rho[i + j * is + (size_t)k * js * is]) != 1) goto bad2;
        if (!fgets(buf, sizeof buf, fp) && k + 1 < ks) goto bad2;
    }
    fclose(fp);
    size_t nvox = (size_t)is * js * ks;
    for (size_t v = 0; v < nvox; ++v)
        if (p->med[v] < 0 || p->med[v] > nmed) {
            int m = p->med[v];
            omc_phantom_free(p);
            return omc_set_error(OMC_EFORMAT, "%s: voxel %zu has medium %d (1..%d)", path, v, m, nmed);
        }
    return OMC_OK;
bad2:
    omc_phantom_free(p);
bad:
    fclose(fp);
    return omc_set_error(OMC_EFORMAT, "%s: malformed egsphant file", path);
}

int omc_phantom_write(const omc_phantom *p, const char *path)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return omc_set_error(OMC_EIO, "cannot write %s", path);
    fprintf(fp, "%d\n", p->nmed);
    for (int i = 0; i < p->nmed; ++i) fprintf(fp, "%s\n", p->names[i]);
    for (int i = 0; i < p->nmed; ++i) fprintf(fp, "  0.25");
    fprintf(fp, "\n%5d%5d%5d\n", p->isize, p->jsize, p->ksize);
    const double *b[3] = { p->xbounds, p->ybounds, p->zbounds };
    int n[3] = { p->isize, p->jsize, p->ksize };
    for (int d = 0; d < 3; ++d) {
        for (int i = 0; i <= n[d]; ++i)
            fprintf(fp, "%.17g%s", b[d][i], (i % 8 == 7 || i == n[d]) ? "\n" : " ");
    }
    const int is = p->isize, js = p->jsize, ks = p->ksize;
    for (int k = 0; k < ks; ++k) {
        for (int j = 0; j < js; ++j) {
            for (int i = 0; i < is; ++i) fputc('0' + p->med[i + j * is + (size_t)k * js * is], fp);
            fputc('\n', fp);
        }
        fputc('\n', fp);
    }
    for (int k = 0; k < ks; ++k) {
        for (int j = 0; j < js; ++j)
            for (int i = 0; i < is; ++i)
                fprintf(fp, "%.17g%s", p->rho[i + j * is + (size_t)k * js * is], (i % 8 == 7 || i + 1 == is) ? "\n" : " ");
        fputc('\n', fp);
    }
    if (fclose(fp)) return omc_set_error(OMC_EIO, "error writing %s", path);
    return OMC_OK;
}

void omc_regions_free(omc_regions *r)
{
    free(r->med); free(r->rhof); free(r->pcut); free(r->ecut);
    memset(r, 0, sizeof *r);
}

int omc_regions_build(omc_regions *r, const omc_phantom *p, const omc_tables *t,
                      double ecut, double pcut)
{
    memset(r, 0, sizeof *r);
    size_t nvox = (size_t)p->isize * p->jsize * p->ksize;
    int nreg = (int)nvox + 1;
    r->nreg = nreg;
    r->med = calloc(nreg, sizeof(int));
    r->rhof = calloc(nreg, sizeof(double));
    r->pcut = calloc(nreg, sizeof(double));
    r->ecut = calloc(nreg, sizeof(double));
    if (!r->med || !r->rhof || !r->pcut || !r->ecut) {
        omc_regions_free(r);
        return omc_set_error(OMC_ENOMEM, "out of memory for regions");
    }
    r->med[0] = -1;
    for (int i = 1; i < nreg; ++i) {
        int imed = p->med[i - 1] - 1;     /* egsphant counts media from 1 */
        r->med[i] = imed;
        if (imed == -1) continue;         /* vacuum: zero density and cuts */
        if (imed >= t->nmed) {
            omc_regions_free(r);
            return omc_set_error(OMC_EARG, "region %d uses medium %d but only %d built", i, imed + 1, t->nmed);
        }
        r->rhof[i] = (p->rho[i - 1] == 0.0F) ? 1.0 : p->rho[i - 1] / t->rho[imed];
        r->pcut[i] = (t->ap[imed] <= pcut) ? pcut : t->ap[imed];
        r->ecut[i] = (t->ae[imed] <= ecut) ? ecut : t->ae[imed];
    }
    return OMC_OK;
}
