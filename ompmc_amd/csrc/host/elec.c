/*
 * Electron transport tables: screened-Rutherford multiple-scattering data
 * (src/ompmc.c:3201-3280), spin/PWA corrections and the spin rejection
 * table (src/ompmc.c:2376-2947), cubic splines (src/ompmc.c:3016-3095),
 * fictitious cross-section maxima, CSDA ranges and the maximum step table
 * tmxs (src/ompmc.c:3282-3511).
 *
 * Operation order follows the reference so each value is bit-identical.
 * Where the reference mixes float and double (the spinms header values are
 * floats) the same conversions are made.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define RM OMC_RM
#define NE_S  OMC_NE_SPIN          /* 32 */
#define NQ_S  OMC_NQ_SPIN          /* 16 */
#define NU_S  OMC_NU_SPIN          /* 32 */
#define MXE_S OMC_MXE_SPIN         /* 15 */

/* ---- cubic spline with natural end conditions ------------------------------ */
void omc_spline_setup(const double *x, const double *f, double *a, double *b,
                      double *c, double *d, int n)
{
    double s = 0, r;
    int m2 = n - 1;
    for (int m = 1; m <= m2; ++m) {
        d[m - 1] = x[m] - x[m - 1];
        r = (f[m] - f[m - 1]) / d[m - 1];
        c[m - 1] = r - s;
        s = r;
    }
    s = 0;
    r = 0;
    c[0] = 0;
    c[n - 1] = 0;
    for (int m = 2; m <= m2; ++m) {
        c[m - 1] = c[m - 1] + r * c[m - 2];
        b[m - 1] = 2 * (x[m - 2] - x[m]) - r * s;
        s = d[m - 1];
        r = s / b[m - 1];
    }
    int mr = m2;
    for (int m = 2; m <= m2; ++m) {
        c[mr - 1] = (d[mr - 1] * c[mr] - c[mr - 1]) / b[mr - 1];
        mr = mr - 1;
    }
    for (int m = 1; m <= m2; ++m) {
        s = d[m - 1];
        r = c[m] - c[m - 1];
        d[m - 1] = r / s;
        c[m - 1] = 3 * c[m - 1];
        b[m - 1] = (f[m] - f[m - 1]) / s - (c[m - 1] + r) * s;
        a[m - 1] = f[m - 1];
    }
}

double omc_spline_eval(double s, const double *x, const double *a, const double *b,
                       const double *c, const double *d, int n)
{
    int dir, mlo, mhi, m;
    if (x[0] > x[n - 1]) { dir = 1; mlo = n; mhi = 0; }
    else                 { dir = 0; mlo = 0; mhi = n; }
    if (s >= x[mhi + dir - 1]) {
        m = mhi + 2 * dir - 1;
    } else if (s <= x[mlo - dir]) {
        m = mlo - 2 * dir + 1;
    } else {
        int ml = mlo, mu = mhi;
        while (abs(mu - ml) > 1) {
            int mav = (ml + mu) / 2;
            if (s < x[mav - 1]) mu = mav; else ml = mav;
        }
        m = mu + dir - 1;
    }
    double q = s - x[m - 1];
    return a[m - 1] + q * (b[m - 1] + q * (c[m - 1] + q * d[m - 1]));
}

/* ---- screened Rutherford MS tables (msnew.data) ------------------------------ */
int omc_build_mscat(omc_tables *t, const char *data_folder)
{
    char path[1024];
    snprintf(path, sizeof path, "%smsnew.data", data_folder);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    memset(t->wms, 0, OMC_MS_SIZE * sizeof(double));
    memset(t->ims, 0, OMC_MS_SIZE * sizeof(int));
    int ok = 1;
    for (int i = 0; i < OMC_NL_MS && ok; ++i) {
        for (int j = 0; j < OMC_NQ_MS && ok; ++j) {
            int base = i * OMC_NQ_MS * OMC_NU_MS + j * OMC_NU_MS;
            for (int k = 0; k < OMC_NU_MS && ok; ++k) ok = fscanf(fp, "%lf ", &t->ums[base + k]) == 1;
            for (int k = 0; k < OMC_NU_MS && ok; ++k) ok = fscanf(fp, "%lf ", &t->fms[base + k]) == 1;
            for (int k = 0; k < OMC_NU_MS - 1 && ok; ++k) ok = fscanf(fp, "%lf ", &t->wms[base + k]) == 1;
            for (int k = 0; k < OMC_NU_MS - 1 && ok; ++k) ok = fscanf(fp, "%d ", &t->ims[base + k]) == 1;
            if (!ok) break;
            /* store the relative slope of the density and 0-based alias */
            for (int k = 0; k < OMC_NU_MS - 1; ++k) {
                t->fms[base + k] = t->fms[base + k + 1] / t->fms[base + k] - 1.0;
                t->ims[base + k] = t->ims[base + k] - 1;
            }
            t->fms[base + OMC_NU_MS - 1] = t->fms[base + OMC_NU_MS - 2];
        }
    }
    fclose(fp);
    if (!ok) return omc_set_error(OMC_EFORMAT, "Could not read MS data file %s", path);
    double llammin = log(1.0), llammax = log(1.0E5);
    double dllamb = (llammax - llammin) / (OMC_NL_MS - 1);
    t->dllambi = 1.0 / dllamb;
    double dqms = 0.5 / (OMC_NQ_MS - 1);
    t->dqmsi = 1.0 / dqms;
    return OMC_OK;
}

/* ---- spin / PWA corrections ----------------------------------------------- */
#define SPIN_REC_FLOATS 276   /* 1104-byte records */

/* Upper interpolation node for the positron screening table.  At the top
 * beta^2 node the reference reads one element past its 64-entry array
 * (src/ompmc.c:2697, 2725); the last positron node is used instead, so
 * positron entries above ~48 MeV differ from the reference's undefined
 * values.  (The electron lookup one past its half reads the positron's
 * first node, which is defined memory, and is kept as is.) */
static int pos_upper(int je)
{
    int k = NE_S + je + 1;
    return k < 2 * NE_S ? k : 2 * NE_S - 1;
}

static int build_spin(omc_tables *t, const char *data_folder)
{
    char path[1024];
    snprintf(path, sizeof path, "%sspinms.data", data_folder);
    FILE *fp = fopen(path, "rb");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    fseek(fp, 0, SEEK_END);
    long flen = ftell(fp);
    rewind(fp);
    unsigned char *raw = malloc((size_t)flen);
    if (!raw) { fclose(fp); return omc_set_error(OMC_ENOMEM, "out of memory"); }
    size_t got = fread(raw, 1, (size_t)flen, fp);
    fclose(fp);
    if ((long)got != flen || flen < SPIN_REC_FLOATS * 4) {
        free(raw);
        return omc_set_error(OMC_EFORMAT, "spinms file %s too short", path);
    }
    /* header: char version[32], char endian[4], 4 floats */
    float espin_min, espin_max, b2spin_min, b2spin_max;
    memcpy(&espin_min, raw + 36, 4);
    memcpy(&espin_max, raw + 40, 4);
    memcpy(&b2spin_min, raw + 44, 4);
    memcpy(&b2spin_max, raw + 48, 4);
    t->b2spin_min = (double)b2spin_min;

    /* energy nodes: 16 log-spaced in [espin_min, espin_max] keV, then 15
     * equidistant in beta^2 up to b2spin_max */
    double earray[NE_S];
    double dloge = log(espin_max / espin_min) / (double)MXE_S;
    double eloge = log(espin_min);
    earray[0] = espin_min;
    for (int i = 1; i <= MXE_S; ++i) { eloge += dloge; earray[i] = exp(eloge); }
    double dbeta2 = (b2spin_max - b2spin_min) / MXE_S;
    double beta2 = b2spin_min;
    earray[MXE_S + 1] = espin_max;
    for (int i = MXE_S + 2; i <= 2 * MXE_S + 1; ++i) {
        beta2 += dbeta2;
        earray[i] = (beta2 < 0.999) ? RM * 1000.0 * (1.0 / sqrt(1.0 - beta2) - 1) : 50585.1;
    }

    espin_min /= 1000.0;
    espin_max /= 1000.0;
    double dlener = log(espin_max / espin_min) / MXE_S;
    t->dleneri = 1.0 / dlener;
    t->espml = log(espin_min);
    dbeta2 = (b2spin_max - b2spin_min) / MXE_S;
    t->dbeta2i = 1.0 / dbeta2;
    double dqq1 = 0.5 / (OMC_NQ_SPIN - 1);
    t->dqq1i = 1.0 / dqq1;

    const long nrec_avail = flen / (SPIN_REC_FLOATS * 4);
    double eta_a[2 * NE_S], c_a[2 * NE_S], g_a[2 * NE_S];
    double elarray[NE_S], farray[NE_S], af[NE_S], bf[NE_S], cf[NE_S], df[NE_S];
    const double tf = 137.03604 * 0.88534138;   /* 1/alpha times Thomas-Fermi constant */

    for (int m = 0; m < t->nmed; ++m) {
        double sum_Z2 = 0.0, sum_A = 0.0, sum_pz = 0.0, sum_Z = 0.0;
        memset(eta_a, 0, sizeof eta_a);
        memset(c_a, 0, sizeof c_a);
        memset(g_a, 0, sizeof g_a);
        double *rej = t->spin_rej + (size_t)m * OMC_SPIN_PER_MED;
        memset(rej, 0, OMC_SPIN_PER_MED * sizeof(double));

        for (int ie = 0; ie < t->ne[m]; ++ie) {
            double z = t->elements[m][ie].z;
            int iz = (int)(z + 0.5);
            double pz = t->elements[m][ie].pz;
            double tmp = z * (z + 1.0) * pz;
            sum_Z2 += tmp;
            sum_Z += pz * z;
            sum_A += pz * t->elements[m][ie].wa;
            sum_pz += pz;
            double z23 = pow(z, 2.0 / 3.0);
            for (int iq = 0; iq < 2; ++iq) {
                for (int i = 0; i < NE_S; ++i) {
                    long irec = 1 + (long)(iz - 1) * 4 * (MXE_S + 1) + 2 * iq * (MXE_S + 1) + i + 1;
                    if (irec > nrec_avail) {
                        free(raw);
                        return omc_set_error(OMC_EFORMAT, "spinms file %s has no record %ld (Z=%d)", path, irec, iz);
                    }
                    const unsigned char *rec = raw + (size_t)(irec - 1) * SPIN_REC_FLOATS * 4;
                    float f4[20];
                    memcpy(f4, rec, sizeof f4);
                    double dum1 = f4[0], dum2 = f4[1], dum3 = f4[2], aux_o = f4[3];
                    eta_a[iq * NE_S + i] += tmp * log(z23 * aux_o);
                    double tau = earray[i] / (1000.0 * RM);
                    double eta = z23 / ((137.03604 * 0.88534138) * (137.03604 * 0.88534138)) *
                                 aux_o / 4 / tau / (tau + 2);
                    c_a[iq * NE_S + i] += tmp * (log(1.0 + 1.0 / eta) - 1.0 / (1.0 + eta)) * dum1 * dum3;
                    g_a[iq * NE_S + i] += tmp * dum2;
                    for (int j = 0; j < NQ_S; ++j) {
                        double fmaxj = f4[4 + j];
                        for (int k = 0; k < NU_S; ++k) {
                            int16_t s16;
                            memcpy(&s16, rec + 80 + 2 * (NU_S * j + k), 2);
                            int ii2 = s16;
                            if (ii2 < 0) ii2 += 65536;
                            double v = ii2;
                            v = v * fmaxj / 65535;
                            rej[iq * NE_S * NQ_S * NU_S + i * NQ_S * NU_S + j * NU_S + k] += tmp * v;
                        }
                    }
                }
            }
        }

        /* scale each angular row to a maximum of one */
        for (int iq = 0; iq < 2; ++iq)
            for (int i = 0; i < NE_S; ++i)
                for (int j = 0; j < NQ_S; ++j) {
                    double *row = rej + iq * NE_S * NQ_S * NU_S + i * NQ_S * NU_S + j * NU_S;
                    double flmax = 0.0;
                    for (int k = 0; k < NU_S; ++k) if (flmax < row[k]) flmax = row[k];
                    for (int k = 0; k < NU_S; ++k) row[k] = row[k] / flmax;
                }

        for (int i = 0; i < NE_S; ++i) {
            double tau = (earray[i] / RM) * 0.001;
            for (int iq = 0; iq < 2; ++iq) {
                double aux_o = exp(eta_a[iq * NE_S + i] / sum_Z2) / (tf * tf);
                eta_a[iq * NE_S + i] = 0.26112447 * aux_o * (t->blcc[m]) / (t->xcc[m]);
                double eta = aux_o / 4.0 / tau / (tau + 2);
                double gamma = 3.0 * (1.0 + eta) * (log(1.0 + 1.0 / eta) * (1.0 + 2.0 * eta) - 2.0) /
                               (log(1.0 + 1.0 / eta) * (1.0 + eta) - 1.0);
                g_a[iq * NE_S + i] = g_a[iq * NE_S + i] / sum_Z2 / gamma;
                c_a[iq * NE_S + i] = c_a[iq * NE_S + i] / sum_Z2 /
                                     (log(1.0 + 1.0 / eta) - 1.0 / (1.0 + eta));
            }
        }

        /* eta (screening) correction interpolation tables */
        const double eke0 = t->eke0[m], eke1 = t->eke1[m];
        const int neke = t->meke[m];
        double *E = t->etae_ms0 + m * OMC_MXEKE, *E1 = t->etae_ms1 + m * OMC_MXEKE;
        double *P = t->etap_ms0 + m * OMC_MXEKE, *P1 = t->etap_ms1 + m * OMC_MXEKE;
        double eil = (1.0 - eke0) / eke1;
        double e = exp(eil);
        double s1e, s1p, s2e, s2p, aae;
        int je;
        if (e <= espin_min) {
            s1e = eta_a[0];
            s1p = eta_a[NE_S];
        } else {
            if (e <= espin_max) {
                aae = (eil - t->espml) * t->dleneri;
                je = (int)aae;
                aae = aae - je;
            } else {
                double tau = e / RM;
                aae = (tau * (tau + 2.0) / ((tau + 1.0) * (tau + 1.0)) - t->b2spin_min) * t->dbeta2i;
                je = (int)aae;
                aae = aae - je;
                je = je + MXE_S + 1;
            }
            s1e = (1 - aae) * eta_a[je] + aae * eta_a[je + 1];
            s1p = (1 - aae) * eta_a[NE_S + je] + aae * eta_a[pos_upper(je)];
        }
        for (int i = 1; i < neke; ++i) {
            eil = (i + 1.0 - eke0) / eke1;
            e = exp(eil);
            if (e <= espin_min) {
                s2e = eta_a[0];
                s2p = eta_a[NE_S];
            } else {
                if (e <= espin_max) {
                    aae = (eil - t->espml) * t->dleneri;
                    je = (int)aae;
                    aae = aae - je;
                } else {
                    double tau = e / RM;
                    double b2 = tau * (tau + 2.0) / ((tau + 1.0) * (tau + 1.0));
                    aae = (b2 - t->b2spin_min) * t->dbeta2i;
                    je = (int)aae;
                    aae = aae - je;
                    je = je + MXE_S + 1;
                }
                s2e = (1.0 - aae) * eta_a[je] + aae * eta_a[je + 1];
                s2p = (1.0 - aae) * eta_a[NE_S + je] + aae * eta_a[pos_upper(je)];
            }
            E1[i - 1] = (s2e - s1e) * eke1;
            E[i - 1] = (s2e - E1[i - 1] * eil);
            P1[i - 1] = (s2p - s1p) * eke1;
            P[i - 1] = (s2p - P1[i - 1] * eil);
            s1e = s2e;
            s1p = s2p;
        }
        E1[neke - 1] = E1[neke - 2]; E[neke - 1] = E[neke - 2];
        P1[neke - 1] = P1[neke - 2]; P[neke - 1] = P[neke - 2];

        /* q1 (first moment) and q2 corrections by spline in log E */
        for (int i = 0; i <= MXE_S; ++i) elarray[i] = log(earray[i] / 1000.0);
        for (int i = MXE_S + 1; i <= 2 * MXE_S; ++i) elarray[i] = log(earray[i + 1] / 1000.0);
        elarray[NE_S - 1] = (t->ue[m] > 1.0E5) ? log(t->ue[m]) : log(1.0E5);
        farray[NE_S - 1] = 1.0;

        const double *src_f[4] = { c_a, c_a + NE_S, g_a, g_a + NE_S };
        double *dst0[4] = { t->q1ce_ms0, t->q1cp_ms0, t->q2ce_ms0, t->q2cp_ms0 };
        double *dst1[4] = { t->q1ce_ms1, t->q1cp_ms1, t->q2ce_ms1, t->q2cp_ms1 };
        for (int w = 0; w < 4; ++w) {
            for (int i = 0; i <= MXE_S; ++i) farray[i] = src_f[w][i];
            for (int i = MXE_S + 1; i <= 2 * MXE_S; ++i) farray[i] = src_f[w][i + 1];
            omc_spline_setup(elarray, farray, af, bf, cf, df, NE_S);
            double *T0 = dst0[w] + m * OMC_MXEKE, *T1 = dst1[w] + m * OMC_MXEKE;
            eil = (1.0 - eke0) / eke1;
            s1e = omc_spline_eval(eil, elarray, af, bf, cf, df, NE_S);
            for (int i = 1; i <= neke - 1; ++i) {
                eil = (i + 1 - eke0) / eke1;
                s2e = omc_spline_eval(eil, elarray, af, bf, cf, df, NE_S);
                T1[i - 1] = (s2e - s1e) * eke1;
                T0[i - 1] = s2e - T1[i - 1] * eil;
                s1e = s2e;
            }
            T1[neke - 1] = T1[neke - 2];
            /* the reference copies the electron q1 SLOPE into the last
             * intercept (src/ompmc.c:2784-2785); the value is kept */
            T0[neke - 1] = (w == 0) ? T1[neke - 2] : T0[neke - 2];
        }

        /* remove the scattering power already in discrete Moller/Bhabha */
        double tauc = t->te[m] / RM;
        double *B0 = t->blcce0 + m * OMC_MXEKE, *B1 = t->blcce1 + m * OMC_MXEKE;
        s1e = 1.0;
        for (int i = 1; i <= neke - 1; ++i) {
            eil = ((double)(i + 1) - eke0) / eke1;
            e = exp(eil);
            int leil = m * OMC_MXEKE + i;
            double tau = e / RM;
            if (tau > 2.0 * tauc) {
                double sig = t->esig1[leil] * eil + t->esig0[leil];
                double dedx = t->ededx1[leil] * eil + t->ededx0[leil];
                sig /= dedx;
                if (sig > 1.0E-6) {
                    double etap = t->etae_ms1[leil] * eil + t->etae_ms0[leil];
                    double eta = 0.25 * etap * (t->xcc[m]) / (t->blcc[m]) / tau / (tau + 2);
                    double g_r = (1.0 + 2.0 * eta) * log(1.0 + 1.0 / eta) - 2.0;
                    double g_m = log(0.5 * tau / tauc) +
                        (1.0 + ((tau + 2.0) / (tau + 1.0)) * ((tau + 2.0) / (tau + 1.0))) *
                            log(2.0 * (tau - tauc + 2.0) / (tau + 4.0)) -
                        0.25 * (tau + 2.0) * (tau + 2.0 + 2.0 * (2.0 * tau + 1.0) / ((tau + 1.0) * (tau + 1.0))) *
                            log((tau + 4.0) * (tau - tauc) / tau / (tau - tauc + 2.0)) +
                        0.5 * (tau - 2.0 * tauc) * (tau + 2.0) *
                            (1.0 / (tau - tauc) - 1.0 / ((tau + 1.0) * (tau + 1.0)));
                    if (g_m < g_r) g_m /= g_r; else g_m = 1.0;
                    s2e = 1.0 - g_m * sum_Z / sum_Z2;
                } else {
                    s2e = 1.0;
                }
            } else {
                s2e = 1.0;
            }
            B1[i - 1] = (s2e - s1e) * eke1;
            B0[i - 1] = s2e - B1[i - 1] * eil;
            s1e = s2e;
        }
        B1[neke - 1] = B1[neke - 2];
        B0[neke - 1] = B0[neke - 2];
        (void)sum_A; (void)sum_pz;
    }
    free(raw);
    return OMC_OK;
}

/* ---- electron step data ------------------------------------------------------ */
int omc_build_electron(omc_tables *t, const char *data_folder)
{
    OMC_TRY(omc_build_mscat(t, data_folder));
    for (int m = 0; m < t->nmed; ++m) {
        t->blcc[m] = 1.16699413758864573 * t->blcc[m];   /* absorbs Euler's constant */
        t->xcc[m] = pow(t->xcc[m], 2.0);
    }
    OMC_TRY(build_spin(t, data_folder));

    const int nmed = t->nmed;
    for (int m = 0; m < nmed; ++m) {
        double sigee = 1.0E-15, sigep = 1.0E-15;
        int emono = 1, pmono = 1;
        double eold = -1.0, pold = -1.0;
        for (int i = 1; i <= t->meke[m]; ++i) {
            double ei = exp(((double)i - t->eke0[m]) / t->eke1[m]);
            double eil = log(ei);
            int k = m * OMC_MXEKE + i - 1;
            double dedx = t->ededx1[k] * eil + t->ededx0[k];
            double sig = t->esig1[k] * eil + t->esig0[k];
            sig /= dedx;
            if (sig > sigee) sigee = sig;
            if (sig < eold) emono = 0;
            eold = sig;
            dedx = t->pdedx1[k] * eil + t->pdedx0[k];
            sig = t->psig1[k] * eil + t->psig0[k];
            sig /= dedx;
            if (sig > sigep) sigep = sig;
            if (sig < pold) pmono = 0;
            pold = sig;
        }
        t->sig_ismonotone[0][m] = emono;
        t->sig_ismonotone[1][m] = pmono;
        t->esig_e[m] = sigee;
        t->psig_e[m] = sigep;
    }

    for (int m = 0; m < nmed; ++m) {
        const double eke0 = t->eke0[m], eke1 = t->eke1[m];
        const int neke = t->meke[m];
        double *ea = t->e_array + m * OMC_MXEKE;
        double *re = t->range_ep + 0 * nmed * OMC_MXEKE + m * OMC_MXEKE;
        double *rp = t->range_ep + 1 * nmed * OMC_MXEKE + m * OMC_MXEKE;
        const double *ed1 = t->ededx1 + m * OMC_MXEKE, *ed0 = t->ededx0 + m * OMC_MXEKE;
        const double *pd1 = t->pdedx1 + m * OMC_MXEKE, *pd0 = t->pdedx0 + m * OMC_MXEKE;

        /* CSDA range to the first table energy, piecewise with the
         * log-interpolated restricted stopping power */
        double ei = exp((1.0 - eke0) / eke1);
        ea[0] = ei;
        t->expeke1[m] = exp(1.0 / eke1) - 1.0;
        re[0] = 0.0;
        rp[0] = 0.0;
        for (int i = 1; i <= neke - 1; ++i) {
            double eip1 = exp(((double)(i + 1) - eke0) / eke1);
            ea[i] = eip1;
            double eke = 0.5 * (eip1 + ei);
            double elke = log(eke);
            int lelke = (int)(eke1 * elke + eke0) - 1;
            double dedx = pd1[lelke] * elke + pd0[lelke];
            double aux = pd1[i - 1] / dedx;
            rp[i] = rp[i - 1] + (eip1 - ei) / dedx *
                    (1.0 + aux * (1.0 + 2.0 * aux) * pow((eip1 - ei) / eke, 2.0) / 24.0);
            dedx = ed1[lelke] * elke + ed0[lelke];
            aux = ed1[i - 1] / dedx;
            re[i] = re[i - 1] + (eip1 - ei) / dedx *
                    (1.0 + aux * (1.0 + 2.0 * aux) * pow(((eip1 - ei) / eke), 2.0) / 24.0);
            ei = eip1;
        }

        /* maximum step from the multiple-scattering energy-loss limit */
        double *tm0 = t->tmxs0 + m * OMC_MXEKE, *tm1 = t->tmxs1 + m * OMC_MXEKE;
        double eil = (1.0 - eke0) / eke1;
        ei = exp(eil);
        double p2 = ei * (ei + 2.0 * RM);
        double beta2 = p2 / (p2 + pow(RM, 2.0));
        double chi_a2 = t->xcc[m] / (4.0 * p2 * t->blcc[m]);
        double dedx0 = ed1[1] * eil + ed0[1];
        double estepx = 2.0 * p2 * beta2 * dedx0 / ei / t->xcc[m] /
                        (log(1.0 + 1.0 / chi_a2) * (1.0 + chi_a2) - 1.0);
        estepx *= 0.5;                       /* XIMAX */
        if (estepx > 0.25) estepx = 0.25;    /* ESTEPE */
        double si = estepx * ei / dedx0;
        for (int i = 1; i <= neke - 1; ++i) {
            double elke = ((double)(i + 1) - eke0) / eke1;
            double eke = exp(elke);
            int lelke = i;
            p2 = eke * (eke + 2.0 * RM);
            beta2 = p2 / (p2 + pow(RM, 2.0));
            chi_a2 = t->xcc[m] / (4.0 * p2 * t->blcc[m]);
            double dedx = ed1[lelke] * elke + ed0[lelke];
            estepx = 2.0 * p2 * beta2 * dedx / eke / (t->xcc[m]) /
                     (log(1.0 + 1.0 / chi_a2) * (1.0 + chi_a2) - 1.0);
            estepx = estepx * 0.5;
            if (estepx > 0.25) estepx = 0.25;
            double ekef = (1.0 - estepx) * eke;
            double sip1;
            int lelkef;
            if (ekef <= ea[0]) {
                sip1 = (ea[0] - ekef) / dedx0;
                ekef = ea[0];
                lelkef = 0;
            } else {
                double elkef = log(ekef);
                lelkef = eke1 * elkef + eke0 - 1;
                int leip1l = lelkef + 1;
                double eip1l = ((double)(leip1l + 1) - eke0) / eke1;
                double eip1 = ea[leip1l];
                double aux = (eip1 - ekef) / eip1;
                double elktmp = 0.5 * (elkef + eip1l + 0.25 * aux * aux * (1.0 + aux * (1.0 + 0.875 * aux)));
                double ektmp = 0.5 * (ekef + eip1);
                dedx = ed1[lelkef] * elktmp + ed0[lelkef];
                aux = ed1[lelkef] / dedx;
                sip1 = (eip1 - ekef) / dedx *
                       (1.0 + aux * (1.0 + 2.0 * aux) * (pow(((eip1 - ekef) / ektmp), 2.0) / 24.0));
            }
            sip1 += re[lelke] - re[lelkef + 1];
            tm1[i - 1] = (sip1 - si) * eke1;
            tm0[i - 1] = sip1 - tm1[i - 1] * elke;
            si = sip1;
        }
        tm0[neke - 1] = tm0[neke - 2];
        tm1[neke - 1] = tm1[neke - 2];
    }
    return OMC_OK;
}

/* ---- synthetic spinms.data ---------------------------------------------------
 * The EGSnrc spin data file is not distributed with ompMC.  This writes a
 * file in its exact binary layout whose content is smooth, deterministic and
 * physically plausible in scale, so that every code path that reads it
 * (screening correction eta, q1/q2 moment corrections, spin rejection) is
 * exercised with non-trivial values:
 *   header: espin 1..100 keV, beta^2 from beta^2(100 keV) to 0.9999;
 *   aux_o  = 1.13 + 3.76 (alpha Z)^2 / beta^2   (Moliere screening factor);
 *   dum1 = dum3 = 1, dum2 = the q2/q1 ratio of the screened Rutherford
 *   cross section for that element, so q-corrections stay near one;
 *   rejection table in (0.5, 1], decreasing with angle, stronger at high Z
 *   and low energy, with fmax[j] = 1.
 * The same file feeds the oracle, the reference build and the device, so
 * parity does not depend on how close it is to the real EGSnrc data. */
int omc_write_synthetic_spinms(const char *path)
{
    const int nz = 100;
    const long nrec = 1 + (long)nz * 64;
    const size_t rec_bytes = SPIN_REC_FLOATS * 4;
    unsigned char *buf = calloc((size_t)nrec, rec_bytes);
    if (!buf) return omc_set_error(OMC_ENOMEM, "out of memory");

    const char version[32] = "ompmc_amd synthetic spinms v1";
    memcpy(buf, version, 32);
    memcpy(buf + 32, "1234", 4);
    float emin = 1.0f, emax = 100.0f;
    double tmax = 100.0 / (1000.0 * RM);
    float b2min = (float)(tmax * (tmax + 2.0) / ((tmax + 1.0) * (tmax + 1.0)));
    float b2max = 0.9999f;
    memcpy(buf + 36, &emin, 4);
    memcpy(buf + 40, &emax, 4);
    memcpy(buf + 44, &b2min, 4);
    memcpy(buf + 48, &b2max, 4);

    /* the same energy nodes the reader derives (keV) */
    double earr[NE_S];
    double dloge = log((double)emax / emin) / MXE_S, el = log((double)emin);
    earr[0] = emin;
    for (int i = 1; i <= MXE_S; ++i) { el += dloge; earr[i] = exp(el); }
    earr[MXE_S + 1] = emax;
    double b2 = b2min, db2 = ((double)b2max - b2min) / MXE_S;
    for (int i = MXE_S + 2; i < NE_S; ++i) {
        b2 += db2;
        earr[i] = (b2 < 0.999) ? RM * 1000.0 * (1.0 / sqrt(1.0 - b2) - 1) : 50585.1;
    }

    const double alpha = 1.0 / 137.03604;
    for (int z = 1; z <= nz; ++z) {
        double z23 = pow((double)z, 2.0 / 3.0);
        for (int iq = 0; iq < 2; ++iq) {
            for (int i = 0; i < NE_S; ++i) {
                long irec = 1 + (long)(z - 1) * 64 + 32 * iq + i + 1;
                unsigned char *rec = buf + (size_t)(irec - 1) * rec_bytes;
                double tau = earr[i] / (1000.0 * RM);
                double beta2 = tau * (tau + 2.0) / ((tau + 1.0) * (tau + 1.0));
                double aux_o = 1.13 + 3.76 * (alpha * z) * (alpha * z) / beta2;
                double eta = z23 / (137.03604 * 0.88534138 * 137.03604 * 0.88534138) * aux_o / 4 / tau / (tau + 2);
                double L = log(1.0 + 1.0 / eta);
                double gam = 3.0 * (1.0 + eta) * (L * (1.0 + 2.0 * eta) - 2.0) / (L * (1.0 + eta) - 1.0);
                float f[20];
                f[0] = 1.0f;
                f[1] = (float)gam;
                f[2] = 1.0f;
                f[3] = (float)aux_o;
                for (int j = 0; j < NQ_S; ++j) f[4 + j] = 1.0f;
                memcpy(rec, f, sizeof f);
                /* strength of the large-angle suppression */
                double s = 0.45 * (z / (z + 20.0)) * (iq == 0 ? 1.0 : 0.6) / (1.0 + 0.5 * tau);
                for (int j = 0; j < NQ_S; ++j) {
                    double qj = (double)j / (NQ_S - 1);
                    for (int k = 0; k < NU_S; ++k) {
                        double u = (double)k / (NU_S - 1);
                        double r = 1.0 - s * u * u * (0.6 + 0.4 * qj) + 0.02 * sin(3.0 * u + 0.1 * z);
                        if (r > 1.0) r = 1.0;
                        if (r < 0.05) r = 0.05;
                        uint16_t v = (uint16_t)lrint(r * 65535.0);
                        memcpy(rec + 80 + 2 * (NU_S * j + k), &v, 2);
                    }
                }
            }
        }
    }
    FILE *fp = fopen(path, "wb");
    if (!fp) { free(buf); return omc_set_error(OMC_EIO, "cannot write %s", path); }
    size_t wr = fwrite(buf, rec_bytes, (size_t)nrec, fp);
    fclose(fp);
    free(buf);
    if (wr != (size_t)nrec) return omc_set_error(OMC_EIO, "short write to %s", path);
    return OMC_OK;
}
