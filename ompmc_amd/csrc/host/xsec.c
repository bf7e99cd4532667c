/*
 * Photon interaction tables: total mean free path and branching ratios from
 * the XCOM element data (src/ompmc.c:216-665), Rayleigh form-factor sampling
 * tables (src/ompmc.c:757-1007) and the pair/brems screening rejection
 * parameters (src/ompmc.c:1148-1318).
 *
 * The arithmetic is written in the same operation order as the reference so
 * that every table entry is bit-identical under IEEE double with no FMA
 * contraction (x86-64 SSE2, gcc -O2).
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define RM OMC_RM

/* One XCOM file: for each of the first 50 elements a count and that many
 * (log E, log sigma) pairs. */
typedef struct {
    int     n[OMC_MXELEMENT];
    double *le[OMC_MXELEMENT];
    double *ls[OMC_MXELEMENT];
} xcom_set;

static void xcom_free(xcom_set *x)
{
    for (int i = 0; i < OMC_MXELEMENT; ++i) { free(x->le[i]); free(x->ls[i]); }
    memset(x, 0, sizeof *x);
}

static int xcom_read(xcom_set *x, const char *folder, const char *name)
{
    char path[1024];
    snprintf(path, sizeof path, "%s%s", folder, name);
    memset(x, 0, sizeof *x);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    for (int i = 0; i < OMC_MXELEMENT; ++i) {
        unsigned n;
        if (fscanf(fp, "%u\n", &n) != 1) goto bad;
        x->n[i] = (int)n;
        x->le[i] = malloc(n * sizeof(double));
        x->ls[i] = malloc(n * sizeof(double));
        if (!x->le[i] || !x->ls[i]) goto bad;
        for (unsigned j = 0; j < n; ++j)
            if (fscanf(fp, "%lf %lf", &x->le[i][j], &x->ls[i][j]) != 2) goto bad;
    }
    fclose(fp);
    return OMC_OK;
bad:
    fclose(fp);
    xcom_free(x);
    return omc_set_error(OMC_EFORMAT, "Could not read the data file %s", path);
}

int omc_xcom_load(omc_xcom_data *x, const char *data_folder)
{
    static const char *names[4] = { "xcom_photo.data", "xcom_rayleigh.data", "xcom_pair.data",
                                    "xcom_triplet.data" };
    if (!x || !data_folder) return omc_set_error(OMC_EARG, "null argument");
    memset(x, 0, sizeof *x);
    xcom_set s[4];
    memset(s, 0, sizeof s);
    int rc = OMC_OK;
    for (int k = 0; k < 4 && !rc; ++k) rc = xcom_read(&s[k], data_folder, names[k]);
    if (!rc) {
        for (int k = 0; k < 4; ++k)
            for (int z = 0; z < OMC_MXELEMENT; ++z) {
                x->n[k][z] = s[k].n[z];
                x->off[k][z] = x->total;
                x->total += s[k].n[z];
            }
        x->le = malloc((size_t)x->total * sizeof(double));
        x->ls = malloc((size_t)x->total * sizeof(double));
        if (!x->le || !x->ls) rc = omc_set_error(OMC_ENOMEM, "XCOM data");
    }
    for (int k = 0; !rc && k < 4; ++k)
        for (int z = 0; z < OMC_MXELEMENT; ++z) {
            memcpy(x->le + x->off[k][z], s[k].le[z], (size_t)s[k].n[z] * sizeof(double));
            memcpy(x->ls + x->off[k][z], s[k].ls[z], (size_t)s[k].n[z] * sizeof(double));
        }
    for (int k = 0; k < 4; ++k) xcom_free(&s[k]);
    if (rc) omc_xcom_free(x);
    return rc;
}

void omc_xcom_free(omc_xcom_data *x)
{
    if (!x) return;
    free(x->le);
    free(x->ls);
    memset(x, 0, sizeof *x);
}

/* Element order used when summing partial cross sections: ascending Z.
 * (The reference heap-sorts; for a medium's distinct Z values every correct
 * sort gives this same order, which fixes the floating-point sum order.) */
static void sort_by_z(int ne, const omc_element *el, double *z_sorted, double *pz_sorted)
{
    int idx[OMC_MXELEMENT];
    for (int i = 0; i < ne; ++i) idx[i] = i;
    for (int i = 1; i < ne; ++i) {
        int k = idx[i], j = i - 1;
        while (j >= 0 && el[idx[j]].z > el[k].z) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = k;
    }
    for (int i = 0; i < ne; ++i) {
        z_sorted[i] = el[idx[i]].z;
        pz_sorted[i] = el[idx[i]].pz;
    }
}

/* Medium cross section on the log-energy grid from element log-log data
 * (get_data, src/ompmc.c:347-449).  kind 0: plain interpolation; kind 1/2:
 * pair/triplet with threshold 2m/4m handled by the (1-eth/E)^3 factor. */
static void medium_xsec(int kind, int ne, const xcom_set *xs, const double *z_sorted,
                        const double *pz_sorted, double ge0, double ge1, double *res)
{
    for (int j = 0; j < OMC_MXGE; ++j) res[j] = 0.0;
    for (int i = 0; i < ne; ++i) {
        int z = (int)(z_sorted[i] + 0.5) - 1;
        int n = xs->n[z];
        double eth = 0.0;
        double *le, *ls;
        if (kind == 0) {
            le = malloc(n * sizeof(double));
            ls = malloc(n * sizeof(double));
            for (int j = 0; j < n; ++j) { le[j] = xs->le[z][j]; ls[j] = xs->ls[z][j]; }
        } else {
            le = malloc((n + 1) * sizeof(double));
            ls = malloc((n + 1) * sizeof(double));
            for (int j = 0; j < n; ++j) { le[j + 1] = xs->le[z][j]; ls[j + 1] = xs->ls[z][j]; }
            eth = (kind == 1) ? 2.0 * RM : 4.0 * RM;
            n++;
            for (int j = 1; j < n; ++j) ls[j] -= 3.0 * log(1.0 - eth / exp(le[j]));
            le[0] = (double)log(eth);
            ls[0] = (double)ls[1];
        }
        for (int j = 0; j < OMC_MXGE; ++j) {
            double gle = ((double)(j + 1) - ge0) / ge1;
            double e = exp(gle);
            double sig = 0.0;
            if ((gle < le[0]) || (gle >= le[n - 1])) {
                if (kind != 0) sig = (gle < le[0]) ? 0.0 : exp(ls[n - 1]);
            } else {
                int k;
                for (k = 0; k < n - 1; ++k)
                    if ((gle >= le[k]) && (gle < le[k + 1])) break;
                double p = (gle - le[k]) / (le[k + 1] - le[k]);
                sig = exp(p * ls[k + 1] + (1.0 - p) * ls[k]);
            }
            if ((kind != 0) && (e > eth))
                sig *= (1.0 - eth / e) * (1.0 - eth / e) * (1.0 - eth / e);
            res[j] += pz_sorted[i] * sig;
        }
        free(le);
        free(ls);
    }
}

/* Klein-Nishina total cross section per electron (src/ompmc.c:451-469). */
double omc_kn_sigma0(double e)
{
    double con = 0.1274783851;
    double ko = e / RM;
    if (ko < 0.01)
        return 8.0 * con / 3.0 * (1.0 - ko * (2.0 - ko * (5.2 - 13.3 * ko))) / RM;
    double c1 = 1.0 / (ko * ko);
    double c2 = 1.0 - 2.0 * (1.0 + ko) * c1;
    double c3 = (1.0 + 2.0 * ko) * c1;
    double eps2 = 1.0;
    double eps1 = 1.0 / (1.0 + 2.0 * ko);
    return (c1 * (1.0 / eps1 - 1.0 / eps2) + c2 * log(eps2 / eps1) +
            eps2 * (c3 + 0.5 * eps2) - eps1 * (c3 + 0.5 * eps1)) / e * con;
}

int omc_build_photon(omc_tables *t, const char *data_folder)
{
    xcom_set photo, rayl, pair, trip;
    OMC_TRY(xcom_read(&photo, data_folder, "xcom_photo.data"));
    int rc = xcom_read(&rayl, data_folder, "xcom_rayleigh.data");
    if (!rc) rc = xcom_read(&pair, data_folder, "xcom_pair.data");
    if (!rc) rc = xcom_read(&trip, data_folder, "xcom_triplet.data");
    if (rc) { xcom_free(&photo); xcom_free(&rayl); xcom_free(&pair); return rc; }

    double *sp = malloc(4 * OMC_MXGE * sizeof(double));
    double *s_photo = sp, *s_rayl = sp + OMC_MXGE, *s_pair = sp + 2 * OMC_MXGE,
           *s_trip = sp + 3 * OMC_MXGE;

    for (int i = 0; i < t->nmed; ++i) {
        t->ge1[i] = (double)(OMC_MXGE - 1) / log(t->up[i] / t->ap[i]);
        t->ge0[i] = 1.0 - t->ge1[i] * log(t->ap[i]);

        double sumA = 0.0, sumZ = 0.0;
        for (int j = 0; j < t->ne[i]; ++j) {
            sumA += t->elements[i][j].pz * t->elements[i][j].wa;
            sumZ += t->elements[i][j].pz * t->elements[i][j].z;
        }
        double con2 = t->rho[i] / (sumA * 1.6605655);
        double zs[OMC_MXELEMENT], pzs[OMC_MXELEMENT];
        sort_by_z(t->ne[i], t->elements[i], zs, pzs);

        medium_xsec(0, t->ne[i], &photo, zs, pzs, t->ge0[i], t->ge1[i], s_photo);
        medium_xsec(0, t->ne[i], &rayl, zs, pzs, t->ge0[i], t->ge1[i], s_rayl);
        medium_xsec(1, t->ne[i], &pair, zs, pzs, t->ge0[i], t->ge1[i], s_pair);
        medium_xsec(2, t->ne[i], &trip, zs, pzs, t->ge0[i], t->ge1[i], s_trip);

        double gle = 0.0, gmfp = 0.0, gbr1 = 0.0, gbr2 = 0.0, cohe = 0.0;
        double gmfp_o = 0.0, gbr1_o = 0.0, gbr2_o = 0.0, cohe_o = 0.0;
        for (int j = 0; j < OMC_MXGE; ++j) {
            gle = ((double)(j + 1) - t->ge0[i]) / t->ge1[i];
            double e = exp(gle);
            double sig_kn = sumZ * omc_kn_sigma0(e);
            double sig_p = s_pair[j] + s_trip[j];
            double sigma = sig_kn + sig_p + s_photo[j];
            gmfp = 1.0 / (sigma * con2);
            gbr1 = sig_p / sigma;
            gbr2 = gbr1 + sig_kn / sigma;
            cohe = sigma / (s_rayl[j] + sigma);
            if (j > 0) {
                /* bin j-1 spans energies j-1..j: slope/intercept in log E */
                int ix = i * OMC_MXGE + (j - 1);
                t->gmfp1[ix] = (gmfp - gmfp_o) * t->ge1[i];
                t->gmfp0[ix] = gmfp - t->gmfp1[ix] * gle;
                t->gbr11[ix] = (gbr1 - gbr1_o) * t->ge1[i];
                t->gbr10[ix] = gbr1 - t->gbr11[ix] * gle;
                t->gbr21[ix] = (gbr2 - gbr2_o) * t->ge1[i];
                t->gbr20[ix] = gbr2 - t->gbr21[ix] * gle;
                t->cohe1[ix] = (cohe - cohe_o) * t->ge1[i];
                t->cohe0[ix] = cohe - t->cohe1[ix] * gle;
            }
            gmfp_o = gmfp; gbr1_o = gbr1; gbr2_o = gbr2; cohe_o = cohe;
        }
        int ix = i * OMC_MXGE + OMC_MXGE - 1;
        t->gmfp1[ix] = t->gmfp1[ix - 1];
        t->gmfp0[ix] = gmfp - t->gmfp1[ix] * gle;
        t->gbr11[ix] = t->gbr11[ix - 1];
        t->gbr10[ix] = gbr1 - t->gbr11[ix] * gle;
        t->gbr21[ix] = t->gbr21[ix - 1];
        t->gbr20[ix] = gbr2 - t->gbr21[ix] * gle;
        t->cohe1[ix] = t->cohe1[ix - 1];
        t->cohe0[ix] = cohe - t->cohe1[ix] * gle;
    }
    free(sp);
    xcom_free(&photo); xcom_free(&rayl); xcom_free(&pair); xcom_free(&trip);
    return OMC_OK;
}

/* Rayleigh tables, EGSnrc prepare_rayleigh_data as restated by
 * src/ompmc.c:820-1007. */
int omc_build_rayleigh(omc_tables *t, const char *pgs4form_file)
{
    FILE *fp = fopen(pgs4form_file, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", pgs4form_file);
    double xval[OMC_MXRAYFF];
    double *aff = malloc(OMC_MXELEMENT * OMC_MXRAYFF * sizeof(double));
    int ok = 1;
    for (int i = 0; i < OMC_MXRAYFF && ok; ++i) ok = fscanf(fp, "%lf", &xval[i]) == 1;
    for (int i = 0; i < OMC_MXELEMENT * OMC_MXRAYFF && ok; ++i) ok = fscanf(fp, "%lf", &aff[i]) == 1;
    fclose(fp);
    if (!ok) { free(aff); return omc_set_error(OMC_EFORMAT, "Could not read atomic form factors file %s", pgs4form_file); }

    double ff[OMC_MXRAYFF];
    double *pe = malloc(OMC_MXGE * sizeof(double));
    for (int i = 0; i < t->nmed; ++i) {
        double *xg = t->xgrid + i * OMC_MXRAYFF;
        double *fc = t->fcum + i * OMC_MXRAYFF;
        double *ba = t->b_array + i * OMC_MXRAYFF;
        double *ca = t->c_array + i * OMC_MXRAYFF;
        int *ia = t->i_array + i * OMC_RAYCDFSIZE;

        /* independent-atom form factor */
        for (int j = 0; j < OMC_MXRAYFF; ++j) {
            double s = 0.0;
            xg[j] = xval[j];
            for (int k = 0; k < t->ne[i]; ++k) {
                int z = (int)t->elements[i][k].z - 1;
                s += t->elements[i][k].pz * pow(aff[z * OMC_MXRAYFF + j], 2);
            }
            ff[j] = sqrt(s);
        }
        if (xg[0] < 1.0E-6) xg[0] = 0.0001;

        double emin = exp((1.0 - t->ge0[i]) / t->ge1[i]);
        double emax = exp((OMC_MXGE - t->ge0[i]) / t->ge1[i]);

        /* avoid log(0): an exact +0 becomes the smallest subnormal */
        for (int j = 0; j < OMC_MXRAYFF; ++j) {
            uint64_t bits;
            memcpy(&bits, &ff[j], sizeof bits);
            if (bits == 0) { bits = 1; memcpy(&ff[j], &bits, sizeof bits); }
        }

        /* cumulative distribution of ff^2 over x^2 with power-law pieces */
        double sum0 = 0.0;
        fc[0] = 0.0;
        for (int j = 0; j < OMC_MXRAYFF - 1; ++j) {
            double b = log(ff[j + 1] / ff[j]) / log(xg[j + 1] / xg[j]);
            ba[j] = b;
            double x1 = xg[j], x2 = xg[j + 1];
            double p1 = pow(x1, 2.0 * b), p2 = pow(x2, 2.0 * b);
            sum0 += pow(ff[j], 2) * (pow(x2, 2) * p2 - pow(x1, 2) * p1) / ((1.0 + b) * p1);
            fc[j + 1] = sum0;
        }

        /* maximum cumulative probability versus photon energy */
        double dle = log(emax / emin) / ((double)OMC_MXGE - 1.0);
        int idx = 1;
        for (int j = 1; j <= OMC_MXGE; ++j) {
            double e = emin * exp(dle * ((double)j - 1.0));
            double xmax = 20.607544 * 2.0 * e / RM;
            int k;
            for (k = 1; k <= OMC_MXRAYFF - 1; ++k)
                if ((xmax >= xg[k - 1]) && (xmax < xg[k])) break;
            idx = k;
            double b = ba[idx - 1];
            double x1 = xg[idx - 1], x2 = xmax;
            double p1 = pow(x1, 2.0 * b), p2 = pow(x2, 2.0 * b);
            pe[j - 1] = fc[idx - 1] + pow(ff[idx - 1], 2) *
                        (pow(x2, 2) * p2 - pow(x1, 2) * p1) / ((1.0 + b) * p1);
        }
        ia[OMC_RAYCDFSIZE - 1] = idx;

        double anorm = 1.0 / sqrt(pe[OMC_MXGE - 1]);
        double anorm1 = 1.005 / pe[OMC_MXGE - 1];
        double anorm2 = 1.0 / pe[OMC_MXGE - 1];
        for (int j = 0; j < OMC_MXGE; ++j) {
            pe[j] *= anorm1;
            if (pe[j] > 1.0) pe[j] = 1.0;
        }
        for (int j = 0; j < OMC_MXRAYFF; ++j) {
            ff[j] *= anorm;
            fc[j] *= anorm2;
            ca[j] = (1.0 + ba[j]) / pow(xg[j] * ff[j], 2);
        }

        /* uniform cumulative bins -> starting interval index */
        double dw = 1.0 / ((double)OMC_RAYCDFSIZE - 1.0);
        double xold = xg[0];
        int ibin = 1;
        double b = ba[0];
        double p1 = pow(xg[0], 2.0 * b);
        ia[0] = 1;
        for (int j = 2; j <= OMC_RAYCDFSIZE - 1; ++j) {
            double w = dw;
            for (;;) {
                double x1 = xold, x2 = xg[ibin];
                double tt = pow(x1, 2) * pow(x1, 2.0 * b);
                double p2 = pow(x2, 2.0 * b);
                double aux = pow(ff[ibin - 1], 2) * (pow(x2, 2) * p2 - tt) / ((1.0 + b) * p1);
                if (aux > w) {
                    xold = exp(log(tt + w * (1.0 + b) * p1 / pow(ff[ibin - 1], 2)) /
                               (2.0 + 2.0 * b));
                    ia[j - 1] = ibin;
                    break;
                }
                w -= aux;
                xold = x2;
                ibin++;
                b = ba[ibin - 1];
                p1 = pow(xold, 2.0 * b);
            }
        }

        /* run-time form of b */
        for (int j = 0; j < OMC_MXRAYFF; ++j) ba[j] = 0.5 / (1.0 + ba[j]);

        for (int j = 0; j < OMC_MXGE - 1; ++j) {
            double gle = ((j + 1) - t->ge0[i]) / t->ge1[i];
            t->pmax1[i * OMC_MXGE + j] = (pe[j + 1] - pe[j]) * t->ge1[i];
            t->pmax0[i * OMC_MXGE + j] = pe[j] - t->pmax1[i * OMC_MXGE + j] * gle;
        }
        t->pmax0[i * OMC_MXGE + OMC_MXGE - 1] = t->pmax0[i * OMC_MXGE + OMC_MXGE - 2];
        t->pmax1[i * OMC_MXGE + OMC_MXGE - 1] = t->pmax1[i * OMC_MXGE + OMC_MXGE - 2];
    }
    free(pe);
    free(aff);
    return OMC_OK;
}

/* Coulomb correction fc(Z) (src/ompmc.c:1148-1163). */
static double coulomb_fc(double zi)
{
    double a = 0.00729735255664 * zi;
    double fc = 1 + pow(a, 2);
    fc = 1.0 / fc;
    fc = fc + 0.20206 - 0.0369 * pow(a, 2);
    fc = fc + 0.0083 * pow(a, 4);
    fc = fc - 0.002 * pow(a, 6);
    fc = fc * pow(a, 2);
    return fc;
}

/* Atomic-electron screening term xi(Z) (src/ompmc.c:1165-1187). */
static double screening_xi(double zi, double fc)
{
    if (zi == 4) return 5.924 / (4.710 - fc);
    if (zi == 3) return 5.805 / (4.740 - fc);
    if (zi == 2) return 5.621 / (4.790 - fc);
    if (zi == 1) return 6.144 / (5.310 - fc);
    return log(1194.0 * pow(zi, -2.0 / 3.0)) / log(184.15 * pow(zi, -1.0 / 3.0)) - fc;
}

/* Screening rejection-function fits used by brems (l=0..3) and pair
 * (l=4..7) (src/ompmc.c:1189-1318). */
int omc_build_pair(omc_tables *t)
{
    for (int m = 0; m < t->nmed; ++m) {
        double Zt = 0.0, Zb = 0.0, Zf = 0.0;
        for (int i = 0; i < t->ne[m]; ++i) {
            double zi = t->elements[m][i].z;
            double pi = t->elements[m][i].pz;
            double fc = coulomb_fc(zi);
            double xi = screening_xi(zi, fc);
            double aux = pi * zi * (zi + xi);
            Zt = Zt + aux;
            Zb = Zb - aux * log(zi) / 3.0;
            Zf = Zf + aux * fc;
        }
        double Zv = (Zb - Zf) / Zt;
        double Zg = Zb / Zt;
        double fmax1 = 2.0 * (20.863 + 4.0 * Zg) - 2.0 * (20.029 + 4.0 * Zg) / 3.0;
        double fmax2 = 2.0 * (20.863 + 4.0 * Zv) - 2.0 * (20.029 + 4.0 * Zv) / 3.0;
        double *d1 = t->dl1 + m * 8, *d2 = t->dl2 + m * 8, *d3 = t->dl3 + m * 8;
        double *d4 = t->dl4 + m * 8, *d5 = t->dl5 + m * 8, *d6 = t->dl6 + m * 8;

        /* brems: BH (Zg) and Coulomb-corrected (Zv), phi1 / phi2 forms */
        d1[0] = (20.863 + 4.0 * Zg) / fmax1; d2[0] = -3.242 / fmax1; d3[0] = 0.625 / fmax1;
        d4[0] = (21.12 + 4.0 * Zg) / fmax1;  d5[0] = -4.184 / fmax1; d6[0] = 0.952;
        d1[1] = (20.029 + 4.0 * Zg) / fmax1; d2[1] = -1.93 / fmax1;  d3[1] = -0.086 / fmax1;
        d4[1] = (21.12 + 4.0 * Zg) / fmax1;  d5[1] = -4.184 / fmax1; d6[1] = 0.952;
        d1[2] = (20.863 + 4.0 * Zv) / fmax2; d2[2] = -3.242 / fmax2; d3[2] = 0.625 / fmax2;
        d4[2] = (21.12 + 4.0 * Zv) / fmax2;  d5[2] = -4.184 / fmax2; d6[2] = 0.952;
        d1[3] = (20.029 + 4.0 * Zv) / fmax2; d2[3] = -1.93 / fmax2;  d3[3] = -0.086 / fmax2;
        d4[3] = (21.12 + 4.0 * Zv) / fmax2;  d5[3] = -4.184 / fmax2; d6[3] = 0.952;

        /* pair: A = 3 phi1 - phi2 and B = 3 phi1 + phi2 combinations */
        d1[4] = (3.0 * (20.863 + 4.0 * Zg) - (20.029 + 4.0 * Zg));
        d2[4] = (3.0 * (-3.242) - (-1.930));
        d3[4] = (3.0 * (0.625) - (-0.086));
        d4[4] = (2.0 * 21.12 + 8.0 * Zg);
        d5[4] = (2.0 * (-4.184));
        d6[4] = 0.952;
        d1[5] = (3.0 * (20.863 + 4.0 * Zg) + (20.029 + 4.0 * Zg));
        d2[5] = (3.0 * (-3.242) + (-1.930));
        d3[5] = (3.0 * 0.625 + (-0.086));
        d4[5] = (4.0 * 21.12 + 16.0 * Zg);
        d5[5] = (4.0 * (-4.184));
        d6[5] = 0.952;
        d1[6] = (3.0 * (20.863 + 4.0 * Zv) - (20.029 + 4.0 * Zv));
        d2[6] = (3.0 * (-3.242) - (-1.930));
        d3[6] = (3.0 * (0.625) - (-0.086));
        d4[6] = (2.0 * 21.12 + 8.0 * Zv);
        d5[6] = (2.0 * (-4.184));
        d6[6] = 0.952;
        d1[7] = (3.0 * (20.863 + 4.0 * Zv) + (20.029 + 4.0 * Zv));
        d2[7] = (3.0 * (-3.242) + (-1.930));
        d3[7] = (3.0 * 0.625 + (-0.086));
        d4[7] = (4.0 * 21.12 + 16.0 * Zv);
        d5[7] = (4.0 * (-4.184));
        d6[7] = 0.952;

        t->bpar1[m] = d1[6] / (3.0 * d1[7] + d1[6]);
        t->bpar0[m] = 12.0 * d1[7] / (3.0 * d1[7] + d1[6]);

        /* ((1/111) Zeff^(1/3))^2 for the pair/brems angular distributions */
        double zb = 0.0, pzn = 0.0;
        for (int i = 0; i < t->ne[m]; ++i) {
            zb += (double)(t->elements[m][i].pz) * (t->elements[m][i].z) *
                  ((t->elements[m][i].z) + 1.0f);
            pzn += t->elements[m][i].pz;
        }
        t->zbrang[m] = (8.116224E-05) * pow(zb / pzn, 1.0 / 3.0);
        t->delcm[m] = t->pegs_delcm[m];
    }
    return OMC_OK;
}
