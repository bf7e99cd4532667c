/*
 * CPU oracle: restatement of the ompMC shower() path -- TEST INFRASTRUCTURE.
 * See omc_oracle.h for who may use it and the random-number modes.
 *
 * Every function follows the reference function named in its comment, in
 * the same floating-point operation order, so that in RANMAR mode and the
 * reference's serial history loop the output is bit-identical to the
 * reference's (checked by tests/test_oracle_reference.py).
 */
#include "omc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RM        OMC_RM
#define MXGE      OMC_MXGE
#define MXEKE     OMC_MXEKE
#define MXSTACK   10000
#define SGMFP     1.0E-05
#define EPSEMFP   1.0E-5
#define HC_INVERSE 80.65506856998
#define TWICE_HC2 0.000307444456
#define LAMBMAX_MS 1.0E5
#define QMIN_MS   1.0E-3
#define QMAX_MS   0.5
#define MXU_MS    31
#define MXQ_MS    7
#define MXE_SPIN  15
#define MXQ_SPIN  15
#define MXU_SPIN  31

struct ora_ctx {
    const omc_tables  *t;
    const omc_regions *r;
    const omc_phantom *g;
    int nsplit;
    int mode;
    /* RANMAR state (src/omc_random.h:41-51) */
    int u97[97], crndm, cdrndm, cmrndm, ixx, jxx, buf[128], pos;
    /* Philox: stream of the particle being transported */
    uint32_t s0, s1;
    ora_stream cur;
    int cur_slot;
    /* tape */
    const int32_t *tape;
    int64_t tape_n, tape_pos;
    /* particle stack (src/ompmc.h:51-70) */
    int np, npold;
    ora_particle *S;
    /* scoring */
    double *endep;
    double *ledger;
    int64_t ledger_n, hid;
    ora_counters cnt;
    int err;
};

/* ======================================================================= RNG */

void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 24-bit uniform k * 2^-24 from draw number s->ctr of a particle stream. */
static double philox_draw(ora_stream *s)
{
    uint32_t ctr[4] = { s->ctr >> 2, 0u, 0u, 0u }, key[2] = { s->k0, s->k1 }, out[4];
    ora_philox4x32(ctr, key, out);
    uint32_t w = out[s->ctr & 3u];
    s->ctr++;
    return (double)(w >> 8) * (1.0 / 16777216.0);
}

/* Stream of a secondary created by the particle whose stream is `p`,
 * derived after the interaction's last draw. */
static ora_stream philox_child(const ora_stream *p)
{
    uint32_t ctr[4] = { p->ctr, 1u, 0u, 0u }, key[2] = { p->k0, p->k1 }, out[4];
    ora_philox4x32(ctr, key, out);
    ora_stream s = { out[0], out[1], 0u };
    return s;
}

static ora_stream philox_primary(uint32_t s0, uint32_t s1, int64_t hid)
{
    uint32_t ctr[4] = { (uint32_t)hid, (uint32_t)((uint64_t)hid >> 32), 2u, 0u };
    uint32_t key[2] = { s0, s1 }, out[4];
    ora_philox4x32(ctr, key, out);
    ora_stream s = { out[0], out[1], 0u };
    return s;
}

/* initRandom (src/omc_random.c:58-119) without the OpenMP thread offset. */
void ora_seed_ranmar(ora_ctx *c, int ixx, int jxx)
{
    if (ixx <= 0 || ixx > 31328) ixx = 1802;
    if (jxx <= 0 || jxx > 31328) jxx = 9373;
    int i = (ixx / 177 % 177) + 2;
    int j = (ixx % 177) + 2;
    int k = (jxx / 169 % 178) + 1;
    int l = (jxx % 169);
    for (int ii = 0; ii < 97; ++ii) {
        int s = 0, t = 8388608;
        for (int jj = 0; jj < 24; ++jj) {
            int m = ((i * j % 179) * k) % 179;
            i = j; j = k; k = m;
            l = (53 * l + 1) % 169;
            if (l * m % 64 >= 32) s += t;
            t /= 2;
        }
        c->u97[ii] = s;
    }
    c->crndm = 362436;
    c->cdrndm = 7654321;
    c->cmrndm = 16777213;
    c->ixx = 97;
    c->jxx = 33;
    c->pos = 128;
}

/* getRandom (src/omc_random.c:123-164): 128 numbers per refill. */
static void ranmar_refill(ora_ctx *c)
{
    for (int i = 0; i < 128; ++i) {
        int iopt = c->u97[c->ixx - 1] - c->u97[c->jxx - 1];
        if (iopt < 0) iopt += 16777216;
        c->u97[c->ixx - 1] = iopt;
        c->ixx -= 1;
        c->jxx -= 1;
        if (c->ixx == 0) c->ixx = 97;
        else if (c->jxx == 0) c->jxx = 97;
        c->crndm -= c->cdrndm;
        if (c->crndm < 0) c->crndm += c->cmrndm;
        iopt -= c->crndm;
        if (iopt < 0) iopt += 16777216;
        c->buf[i] = iopt;
    }
    c->pos = 0;
}

double ora_random(ora_ctx *c)
{
    c->cnt.n_draws++;
    if (c->mode == ORA_RNG_PHILOX) return philox_draw(&c->cur);
    if (c->mode == ORA_RNG_TAPE) {
        if (c->tape_pos >= c->tape_n) { c->err = 1; c->tape_pos = 0; }
        return c->tape[c->tape_pos++] * (1.0 / 16777216.0);
    }
    if (c->pos >= 128) ranmar_refill(c);
    double r = c->buf[c->pos] * (1.0 / 16777216.0);
    c->pos += 1;
    return r;
}
#define RND() ora_random(c)

/* stream bookkeeping around interactions (PHILOX mode only; harmless else) */
static void stream_enter(ora_ctx *c, int slot) { c->cur = c->S[slot].rs; c->cur_slot = slot; }
static void stream_leave(ora_ctx *c, int parent_slot, int nchild)
{
    for (int k = 1; k <= nchild; ++k) c->S[parent_slot + k].rs = philox_child(&c->cur);
    c->S[parent_slot].rs = c->cur;
}

/* ============================================================ geometry */

/* howfar (ucodes/omc_dosxyz/omc_dosxyz.c:187-297) */
void ora_howfar(const ora_ctx *c, const ora_particle *p, int *idisc, int *irnew, double *ustep)
{
    int irl = p->ir;
    double dist;
    if (irl == 0) { *idisc = 1; return; }
    const omc_phantom *g = c->g;
    int ijmax = g->isize * g->jsize, imax = g->isize;
    int irx = (irl - 1) % imax;
    int irz = (irl - 1 - irx) / ijmax;
    int iry = ((irl - 1 - irx) - irz * ijmax) / imax;
    if (p->w > 0.0) {
        dist = (g->zbounds[irz + 1] - p->z) / p->w;
        if (dist < *ustep) { *ustep = dist; *irnew = (irz != g->ksize - 1) ? irl + ijmax : 0; }
    } else if (p->w < 0.0) {
        dist = -(p->z - g->zbounds[irz]) / p->w;
        if (dist < *ustep) { *ustep = dist; *irnew = (irz != 0) ? irl - ijmax : 0; }
    }
    if (p->u > 0.0) {
        dist = (g->xbounds[irx + 1] - p->x) / p->u;
        if (dist < *ustep) { *ustep = dist; *irnew = (irx != g->isize - 1) ? irl + 1 : 0; }
    } else if (p->u < 0.0) {
        dist = -(p->x - g->xbounds[irx]) / p->u;
        if (dist < *ustep) { *ustep = dist; *irnew = (irx != 0) ? irl - 1 : 0; }
    }
    if (p->v > 0.0) {
        dist = (g->ybounds[iry + 1] - p->y) / p->v;
        if (dist < *ustep) { *ustep = dist; *irnew = (iry != g->jsize - 1) ? irl + imax : 0; }
    } else if (p->v < 0.0) {
        dist = -(p->y - g->ybounds[iry]) / p->v;
        if (dist < *ustep) { *ustep = dist; *irnew = (iry != 0) ? irl - imax : 0; }
    }
}

/* hownear (omc_dosxyz.c:299-334) */
double ora_hownear(const ora_ctx *c, const ora_particle *p)
{
    int irl = p->ir;
    if (irl == 0) return 0.0;
    const omc_phantom *g = c->g;
    int ijmax = g->isize * g->jsize, imax = g->isize;
    int irx = (irl - 1) % imax;
    int irz = (irl - 1 - irx) / ijmax;
    int iry = ((irl - 1 - irx) - irz * ijmax) / imax;
    double tperp = 1.0E10;
    tperp = fmin(tperp, g->xbounds[irx + 1] - p->x);
    tperp = fmin(tperp, p->x - g->xbounds[irx]);
    tperp = fmin(tperp, g->ybounds[iry + 1] - p->y);
    tperp = fmin(tperp, p->y - g->ybounds[iry]);
    tperp = fmin(tperp, g->zbounds[irz + 1] - p->z);
    tperp = fmin(tperp, p->z - g->zbounds[irz]);
    return tperp;
}

/* ausgab (omc_dosxyz.c:683-694) */
static void ausgab(ora_ctx *c, double edep)
{
    ora_particle *p = &c->S[c->np];
    double e = p->wt * edep;
    c->cnt.n_dep++;
    if (c->endep) c->endep[p->ir] += e;
    if (c->ledger && c->hid >= 0 && c->hid < c->ledger_n) {
        c->ledger[4 * c->hid + (p->ir > 0 ? 0 : 1)] += e;
        if (p->ir > 0) c->ledger[4 * c->hid + 3] += e * (double)p->ir;
    }
}

/* ========================================================= rotations */
typedef struct { double A, B, C, cosphi, sinphi; } ora_uphi;

/* selectAzimuthalAngle (src/ompmc.c:101-121) */
static void azimuth(ora_ctx *c, double *cphi, double *sphi)
{
    double xphi, xphi2, yphi, yphi2, rhophi2;
    do {
        xphi = RND();
        xphi = 2.0 * xphi - 1.0;
        xphi2 = xphi * xphi;
        yphi = RND();
        yphi2 = yphi * yphi;
        rhophi2 = xphi2 + yphi2;
    } while (rhophi2 > 1.0);
    rhophi2 = 1 / rhophi2;
    *cphi = (xphi2 - yphi2) * rhophi2;
    *sphi = 2.0 * xphi * yphi * rhophi2;
}

static void rotate(ora_particle *p, const ora_uphi *h, double costhe, double sinthe)
{
    double sinps2 = h->A * h->A + h->B * h->B;
    if (sinps2 < 1.0E-20) {
        p->u = sinthe * h->cosphi;
        p->v = sinthe * h->sinphi;
        p->w = h->C * costhe;
    } else {
        double sinpsi = sqrt(sinps2);
        double us = sinthe * h->cosphi, vs = sinthe * h->sinphi;
        double sindel = h->B / sinpsi, cosdel = h->A / sinpsi;
        p->u = h->C * cosdel * us - sindel * vs + h->A * costhe;
        p->v = h->C * sindel * us + cosdel * vs + h->B * costhe;
        p->w = -sinpsi * us + h->C * costhe;
    }
}

/* uphi21 (src/ompmc.c:127-164): new azimuth, rotate the particle at np. */
static void uphi21(ora_ctx *c, ora_uphi *h, double costhe, double sinthe)
{
    ora_particle *p = &c->S[c->np];
    azimuth(c, &h->cosphi, &h->sinphi);
    h->A = p->u; h->B = p->v; h->C = p->w;
    rotate(p, h, costhe, sinthe);
}

/* transferProperties (src/ompmc.c:83-99) */
static void transfer(ora_ctx *c, int to, int from)
{
    ora_particle *a = &c->S[to], *b = &c->S[from];
    a->x = b->x; a->y = b->y; a->z = b->z;
    a->u = b->u; a->v = b->v; a->w = b->w;
    a->ir = b->ir; a->wt = b->wt; a->dnear = b->dnear;
}

/* uphi32 (src/ompmc.c:166-201): second particle, same azimuth. */
static void uphi32(ora_ctx *c, ora_uphi *h, double costhe, double sinthe)
{
    transfer(c, c->np, c->np - 1);
    rotate(&c->S[c->np], h, costhe, sinthe);
}

static inline int pwlf_i(int idx, double l, const double *c1, const double *c0)
{
    return (int)(l * c1[idx] + c0[idx]);
}
static inline double pwlf_e(int idx, double l, const double *c1, const double *c0)
{
    return l * c1[idx] + c0[idx];
}

/* ========================================================= photon physics */

/* rayleigh (src/ompmc.c:1102-1145): medium-0 form-factor tables, start bin 0 */
static void ora_rayleigh(ora_ctx *c, int imed, double eig, double gle, int lgle)
{
    c->cnt.n_rayleigh++;
    const omc_tables *T = c->t;
    double pmax = pwlf_e(imed * MXGE + lgle, gle, T->pmax1, T->pmax0);
    double xmax = HC_INVERSE * eig;
    double dwi = (double)OMC_RAYCDFSIZE - 1.0;
    double rnno0, rnno1, xv, costhe, csqthe;
    do {
        rnno1 = RND();
        do {
            rnno0 = RND();
            rnno0 *= pmax;
            int ibin = (int)rnno0 * dwi;
            int ib = T->i_array[ibin] - 1;
            if ((T->i_array[ibin + 1] - 1) > ib)
                while (rnno0 >= T->fcum[ib + 1]) ib++;
            rnno0 = (rnno0 - T->fcum[ib]) * T->c_array[ib];
            xv = T->xgrid[ib] * exp(log(1.0 + rnno0) * T->b_array[ib]);
        } while (xv >= xmax);
        xv /= eig;
        costhe = 1.0 - TWICE_HC2 * pow(xv, 2.0);
        csqthe = pow(costhe, 2.0);
    } while (2.0 * rnno1 >= (1.0 + csqthe));
    double sinthe = sqrt(1.0 - csqthe);
    ora_uphi h;
    uphi21(c, &h, costhe, sinthe);
    stream_leave(c, c->np, 0);
}

/* setPairRejectionFunction (src/ompmc.c:1418-1427) */
static double pair_rej(const omc_tables *T, int imed, double xi, double esedei, double eseder, double tteig)
{
    return 2.0 + 3.0 * (esedei + eseder) - 4.0 * (esedei + eseder + 1.0 - 4.0 * pow((xi - 0.5), 2.0)) *
           (1.0 + 0.25 * log((pow(((1.0 + eseder) * (1.0 + esedei) / (2.0 * tteig)), 2.0)) +
                             T->zbrang[imed] * pow(xi, 2.0)));
}

/* pair (src/ompmc.c:1429-1667) */
static void ora_pair(ora_ctx *c, int imed)
{
    c->cnt.n_pair++;
    const omc_tables *T = c->t;
    int np = c->np;
    double eig = c->S[np].e;
    double ese1, ese2;
    int iq1, iq2, l, l1;
    c->npold = np;
    if (eig <= 2.1) {
        double r0 = RND(), r1 = RND();
        ese2 = RM + 0.5 * r0 * (eig - 2.0 * RM);
        ese1 = eig - ese2;
        if (r1 < 0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    } else {
        double Amax, Bmax, delta, aux;
        const double *d1 = T->dl1 + imed * 8, *d2 = T->dl2 + imed * 8, *d3 = T->dl3 + imed * 8;
        const double *d4 = T->dl4 + imed * 8, *d5 = T->dl5 + imed * 8, *d6 = T->dl6 + imed * 8;
        if (eig < 50.0) {
            l = 4;
            l1 = l + 1;
            delta = 4.0 * T->delcm[imed] / eig;
            if (delta < 1.0) {
                Amax = d1[l] + delta * (d2[l] + delta * d3[l]);
                Bmax = d1[l1] + delta * (d2[l1] + delta * d3[l1]);
            } else {
                aux = log(delta + d6[l]);
                Amax = d4[l] + d5[l] * aux;
                Bmax = d4[l1] + d5[l1] * aux;
            }
            aux = 1.0 - 2.0 * RM / eig;
            aux = aux * aux;
            aux *= Amax / 3.0;
            aux /= (Bmax + aux);
        } else {
            l = 6;
            Amax = d1[l];
            Bmax = d1[l + 1];
            aux = T->bpar1[imed] * (1.0 - T->bpar0[imed] * RM / eig);
        }
        double eavail = eig - 2.0 * RM, rejf, rejmax, r4;
        do {
            double br, r0 = RND(), r1 = RND();
            r4 = RND();
            if (r0 > aux) {
                br = 0.5 * r1;
                rejmax = Bmax;
                l1 = l + 1;
            } else {
                double r2 = RND(), r3 = RND();
                br = 0.5 * (1.0 - fmax(fmax(r1, r2), r3));
                rejmax = Amax;
                l1 = l;
            }
            ese2 = br * eavail + RM;
            ese1 = eig - ese2;
            delta = (eig * T->delcm[imed]) / (ese2 * ese1);
            if (delta < 1.0) rejf = d1[l1] + delta * (d2[l1] + delta * d3[l1]);
            else rejf = d4[l1] + d5[l1] * log(delta + d6[l1]);
        } while (r4 * rejmax > rejf);
        ese1 = eig - ese2;
        r4 = RND();
        if (r4 < 0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    }
    c->S[np].e = ese1;
    c->S[np + 1].e = ese2;

    ora_uphi h;
    for (int j = 0; j < 2; ++j) {
        double ese = (j == 0) ? ese1 : ese2;
        double tteig = eig / RM, ttese = ese / RM;
        double esedei = ttese / (tteig - ttese);
        double eseder = 1.0 / esedei;
        double ximin = 1.0 / (1.0 + pow((M_PI * ttese), 2.0));
        double rejmin = pair_rej(T, imed, ximin, esedei, eseder, tteig);
        double ya = pow((2.0 / tteig), 2.0);
        double xitry = fmax(0.01, fmax(ximin, fmin(0.5, sqrt(ya / T->zbrang[imed]))));
        double galpha = 1.0 + 0.25 * log(ya + T->zbrang[imed] * xitry * xitry);
        double gbeta = 0.5 * T->zbrang[imed] * xitry / (ya + T->zbrang[imed] * xitry * xitry);
        galpha -= gbeta * (xitry - 0.5);
        double ximid = galpha / (3.0 * gbeta);
        if (galpha >= 0.0) ximid = 0.5 - ximid + sqrt(pow(ximid, 2.0) + 0.25);
        else ximid = 0.5 - ximid - sqrt(pow(ximid, 2.0) + 0.25);
        ximid = fmax(0.01, fmax(ximin, fmin(0.5, ximid)));
        double rejmid = pair_rej(T, imed, ximid, esedei, eseder, tteig);
        double rejtop = 1.0 * fmax(rejmin, rejmid);
        double theta, rejtst, rtest;
        do {
            double xitst = RND();
            rejtst = pair_rej(T, imed, xitst, esedei, eseder, tteig);
            rtest = RND();
            theta = sqrt(1.0 / xitst - 1.0) / ttese;
        } while ((rtest > rejtst / rejtop) && (theta >= M_PI));  /* reference's && kept */
        double sinthe = sin(theta), costhe = cos(theta);
        if (j == 0) {
            uphi21(c, &h, costhe, sinthe);
        } else {
            sinthe = -sinthe;
            np += 1;
            c->np = np;
            uphi32(c, &h, costhe, sinthe);
        }
    }
    c->S[np].iq = iq2;
    c->S[np - 1].iq = iq1;
    stream_leave(c, np - 1, 1);
}

/* compton (src/ompmc.c:1670-1783), free-electron Klein-Nishina */
static void ora_compton(ora_ctx *c)
{
    c->cnt.n_compton++;
    int np = c->np;
    double eig = c->S[np].e;
    double ko = c->S[np].e / RM;
    double broi = 1.0 + 2.0 * ko, bro = 1.0 / broi;
    int first = 1;
    double sinthe = 0.0, costhe, br, r1, r2, r3, aux, rejf3, temp;
    double alph1 = 0.0, alph2 = 0.0, alpha = 0.0, rejmax = 0.0;
    c->npold = np;
    do {
        if (ko > 2.0) {
            if (first) {
                alph1 = log(broi);
                alph2 = ko * (broi + 1.0) * pow(bro, 2.0);
                alpha = alph1 + alph2;
            }
            do {
                r1 = RND();
                r2 = RND();
                if (r1 * alpha < alph1) br = exp(alph1 * r2) * bro;
                else br = sqrt(r2 * pow(broi, 2.0) + (1.0 - r2)) * bro;
                temp = (1.0 - br) / (ko * br);
                sinthe = fmax(0.0, temp * (2.0 - temp));
                aux = 1.0 + pow(br, 2.0);
                rejf3 = aux - br * sinthe;
                r3 = RND();
            } while (r3 * aux > rejf3);
        } else {
            if (first) rejmax = broi + bro;
            do {
                r1 = RND();
                r2 = RND();
                br = bro + (1.0 - bro) * r1;
                temp = (1.0 - br) / (ko * br);
                sinthe = fmax(0.0, temp * (2.0 - temp));
                rejf3 = 1.0 + br * br - br * sinthe;
            } while (r2 * br * rejmax > rejf3);
        }
        first = 0;
    } while ((br < bro) || (br > 1));
    costhe = 1.0 - temp;
    sinthe = sqrt(sinthe);
    double esg = br * eig, ese = eig - esg + RM;
    c->S[np].e = esg;
    ora_uphi h;
    uphi21(c, &h, costhe, sinthe);
    np += 1;
    c->np = np;
    aux = 1.0 + br * br - 2.0 * br * costhe;
    if (aux > 1.0E-8) {
        costhe = (1.0 - br * costhe) / sqrt(aux);
        sinthe = (1.0 - costhe) * (1.0 + costhe);
        sinthe = (sinthe > 0.0) ? -sqrt(sinthe) : 0.0;
    } else {
        costhe = 0.0;
        sinthe = -1.0;
    }
    uphi32(c, &h, costhe, sinthe);
    c->S[np].e = ese;
    c->S[np].iq = -1;
    stream_leave(c, np - 1, 1);
}

/* photo (src/ompmc.c:1786-1846): no binding energy, Sauter-type angle */
static void ora_photo(ora_ctx *c)
{
    c->cnt.n_photo++;
    int np = c->np;
    c->npold = np;
    c->S[np].e += RM;
    c->S[np].iq = -1;
    double eelec = c->S[np].e;
    if (eelec > c->r->ecut[c->S[np].ir]) {
        double beta = sqrt((eelec - RM) * (eelec + RM)) / eelec;
        double costhe, sinth2, xi, rnpht2;
        double gamma = eelec / RM;
        double alpha = 0.5 * gamma - 0.5 + 1.0 / gamma;
        double ratio = beta / alpha;
        do {
            double rnpht = RND();
            rnpht = 2.0 * rnpht - 1.0;
            if (ratio <= 0.2) {
                double fkappa = rnpht + 0.5 * ratio * (1.0 - rnpht) * (1.0 + rnpht);
                if (gamma < 100.0) {
                    costhe = (beta + fkappa) / (1.0 + beta * fkappa);
                } else {
                    if (fkappa > 0.0)
                        costhe = 1.0 - (1.0 - fkappa) * (gamma - 3.0) /
                                       (2.0 * (1.0 + fkappa) * pow((gamma - 1.0), 3.0));
                    else
                        costhe = (beta + fkappa) / (1.0 + beta * fkappa);
                }
                xi = (1.0 + beta * fkappa) * pow(gamma, 2.0);
            } else {
                xi = pow(gamma, 2.0) * (1.0 + alpha * (sqrt(1.0f + ratio * (2.0 * rnpht + ratio)) - 1.0));
                costhe = (1.0 - 1.0 / xi) / beta;
            }
            sinth2 = fmax((1.0 - costhe) * (1.0 + costhe), 0.0);
            rnpht2 = RND();
        } while (rnpht2 > 0.5 * (1.0 + gamma) * sinth2 * xi / gamma);
        double sinthe = sqrt(sinth2);
        ora_uphi h;
        uphi21(c, &h, costhe, sinthe);
    }
    stream_leave(c, np, 0);
}

/* photon (src/ompmc.c:1849-2121), including photon splitting */
static void ora_photon(ora_ctx *c)
{
    const omc_tables *T = c->t;
    const omc_regions *R = c->r;
    ora_particle *S = c->S;
    int np = c->np;
    int irl = S[np].ir, irold, irnew, imed = R->med[irl], idisc;
    double rhof, tstep, ustep, edep, eig = S[np].e;
    double dpmfp, dpmfp_old, gmfpr0 = 0.0, gmfp = 0.0, gbr1, gbr2, rnno;
    int nsplit = c->nsplit, i_survive_s, ip;
    double d_eta, eta_prime, a_survive;
    double xs, ys, zs, us, vs, ws, es, wts;
    int irs;

    stream_enter(c, np);
    if (eig <= R->pcut[irl] || S[np].wt == 0) {
        edep = eig;
        ausgab(c, edep);
        c->np -= 1;
        return;
    }
    int lgle = 0;
    double gle = log(eig);
    double cohfac = 0.0;
    int ptrans;

start_mfp_loop:
    rnno = RND();
    rnno /= (double)nsplit;
    d_eta = 1.0 / (double)nsplit;
    eta_prime = 1.0 - rnno + d_eta;
    xs = S[np].x; ys = S[np].y; zs = S[np].z;
    us = S[np].u; vs = S[np].v; ws = S[np].w;
    es = S[np].e; wts = S[np].wt / (double)nsplit;
    irs = S[np].ir;
    np -= 1;
    rnno = RND();
    a_survive = rnno * nsplit;
    i_survive_s = (int)a_survive;
    dpmfp_old = 0.0;

    for (int isplit = 0; isplit < nsplit; ++isplit) {
        ptrans = 1;
        eta_prime -= d_eta;
        if (eta_prime <= 0.0) goto end_mfp_loop;
        dpmfp = -log(eta_prime) - dpmfp_old;
        dpmfp_old += dpmfp;
        np += 1;
        c->np = np;
        if (np >= MXSTACK) { c->err = 2; c->np = -1; return; }
        S[np].x = xs; S[np].y = ys; S[np].z = zs;
        S[np].u = us; S[np].v = vs; S[np].w = ws;
        S[np].e = es; S[np].wt = wts;
        S[np].ir = irs; S[np].iq = 0;
        S[np].rs = c->cur;
        c->cur_slot = np;
        irl = S[np].ir;
        irold = irl;
        imed = R->med[irl];

        do {
            c->cnt.n_photon_steps++;
            if (imed != -1) {
                lgle = pwlf_i(imed, gle, T->ge1, T->ge0) - 1;
                gmfpr0 = pwlf_e(imed * MXGE + lgle, gle, T->gmfp1, T->gmfp0);
                rhof = R->rhof[irl];
                gmfp = gmfpr0 / rhof;
                cohfac = pwlf_e(imed * MXGE + lgle, gle, T->cohe1, T->cohe0);
                gmfp *= cohfac;
                tstep = gmfp * dpmfp;
            } else {
                tstep = 1.0E8;
            }
            irnew = irl;
            idisc = 0;
            ustep = tstep;
            ora_howfar(c, &S[np], &idisc, &irnew, &ustep);
            edep = 0.0;
            S[np].x += ustep * S[np].u;
            S[np].y += ustep * S[np].v;
            S[np].z += ustep * S[np].w;
            if (idisc > 0) {
                if (c->ledger && c->hid >= 0 && c->hid < c->ledger_n)
                    c->ledger[4 * c->hid + 2] += S[np].wt * S[np].e;
                np -= 1;
                c->np = np;
                if (np < 0) return;
                goto end_mfp_loop;
            }
            if (imed != -1) dpmfp = fmax(0.0, dpmfp - ustep / gmfp);
            if (irnew != irold) {
                S[np].ir = irnew;
                irl = irnew;
                irold = irnew;
                imed = R->med[irl];
                c->cnt.n_cross_photon++;
            }
            if (imed != -1 && dpmfp <= SGMFP) ptrans = 0;
        } while (ptrans);

        xs = S[np].x; ys = S[np].y; zs = S[np].z;
        irs = S[np].ir;

        rnno = RND();
        if (rnno <= 1.0 - cohfac) {
            if (isplit != i_survive_s) {
                np -= 1;
                c->np = np;
                continue;
            }
            S[np].wt *= nsplit;
            ora_rayleigh(c, imed, eig, gle, lgle);
            continue;
        }
        rnno = RND();
        gbr1 = pwlf_e(imed * MXGE + lgle, gle, T->gbr11, T->gbr10);
        if (rnno <= gbr1 && eig > 2.0 * RM) {
            ora_pair(c, imed);
            np = c->np;
        } else {
            gbr2 = pwlf_e(imed * MXGE + lgle, gle, T->gbr21, T->gbr20);
            if (rnno < gbr2) ora_compton(c);
            else ora_photo(c);
            np = c->np;
        }

        /* keep one scattered photon in nsplit (weight restored), kill others */
        ip = c->npold;
        do {
            if (S[ip].iq == 0) {
                if (isplit != i_survive_s) {
                    if (ip < np) {
                        S[ip].e = S[np].e; S[ip].iq = S[np].iq;
                        S[ip].u = S[np].u; S[ip].v = S[np].v;
                        S[ip].w = S[np].w; S[ip].wt = S[np].wt;
                    }
                    np -= 1;
                } else {
                    S[ip].wt *= nsplit;
                    ip += 1;
                }
            } else {
                ip += 1;
            }
        } while (ip <= np);
        c->np = np;
    }
end_mfp_loop:
    if (np < 0) return;
    if (S[np].iq == 0) {
        stream_enter(c, np);
        eig = S[np].e;
        irl = S[np].ir;
        imed = R->med[irl];
        if (eig <= R->pcut[irl]) {
            edep = eig;
            ausgab(c, edep);
            np -= 1;
            c->np = np;
            return;
        }
        gle = log(eig);
        goto start_mfp_loop;
    }
}

/* =============================================== multiple scattering */
typedef struct { int i, j; } ora_spinr;
typedef struct { int i, j; double omega2; } ora_mscats;

/* spinRejection (src/ompmc.c:3097-3168) */
static double spin_rejection(ora_ctx *c, int imed, int qel, double elke, double beta2, double q1,
                             double cost, int *spin_index, int is_single, ora_spinr *sr)
{
    const omc_tables *T = c->t;
    double ai, aj, ak, qq1, xi, rnno;
    int k;
    if (*spin_index) {
        *spin_index = 0;
        if (beta2 >= T->b2spin_min) {
            ai = (beta2 - T->b2spin_min) * T->dbeta2i;
            sr->i = (int)ai;
            ai -= (double)sr->i;
            sr->i += MXE_SPIN + 1;
        } else if (elke > T->espml) {
            ai = (elke - T->espml) * T->dleneri;
            sr->i = (int)ai;
            ai -= sr->i;
        } else {
            sr->i = 0;
            ai = -1.0f;
        }
        rnno = RND();
        if (rnno < ai) sr->i += 1;
        if (is_single) {
            sr->j = 0;
        } else {
            qq1 = 2.0 * q1;
            qq1 = qq1 / (1.0 + qq1);
            aj = qq1 * T->dqq1i;
            sr->j = (int)aj;
            if (sr->j >= MXQ_SPIN) {
                sr->j = MXQ_SPIN;
            } else {
                aj -= (double)sr->j;
                rnno = RND();
                if (rnno < aj) sr->j += 1;
            }
        }
    }
    xi = sqrt(0.5 * (1.0 - cost));
    ak = xi * MXU_SPIN;
    k = (int)ak;
    ak -= (double)k;
    const double *row = T->spin_rej + (size_t)imed * OMC_SPIN_PER_MED + qel * (OMC_SPIN_PER_MED / 2) +
                        sr->i * 512 + sr->j * 32;
    return (1.0 - ak) * row[k] + ak * row[k + 1];
}

/* sscat (src/ompmc.c:3170-3199) */
static void sscat(ora_ctx *c, int imed, int qel, double chia2, double elke, double beta2,
                  double *cost, double *sint)
{
    int spin_index = 1;
    ora_spinr sr = { 0, 0 };
    double xi, rejf, rnno;
    do {
        xi = RND();
        xi = 2.0 * chia2 * xi / (1.0 - xi + chia2);
        *cost = 1.0 - xi;
        rejf = spin_rejection(c, imed, qel, elke, beta2, 0, *cost, &spin_index, 1, &sr);
        rnno = RND();
    } while (rnno > rejf);
    *sint = sqrt(xi * (2.0 - xi));
}

/* mscat (src/ompmc.c:3606-3785) */
static void mscat(ora_ctx *c, int imed, int qel, int *spin_index, int *find_index, double elke,
                  double beta2, double q1, double lambda, double chia2, double *cost, double *sint,
                  ora_mscats *ms, ora_spinr *sr)
{
    const omc_tables *T = c->t;
    double xi, rejf, rnno;
    double explambda = exp(-lambda);
    if (lambda <= 13.8) {
        double sprob = RND();
        if (sprob < explambda) { *cost = 1.0; *sint = 0.0; return; }
        double wsum = (1.0 + lambda) * explambda;
        if (sprob < wsum) {
            do {
                xi = RND();
                xi = 2.0 * chia2 * xi / (1.0 - xi + chia2);
                *cost = 1.0 - xi;
                rejf = spin_rejection(c, imed, qel, elke, beta2, q1, *cost, spin_index, 0, sr);
                rnno = RND();
            } while (rnno > rejf);
            *sint = sqrt(xi * (2.0 - xi));
            return;
        }
        if (lambda <= 1) {
            int icount = 0;
            double wprob = explambda, sinz, cosz, phi;
            wsum = explambda;
            *cost = 1.0;
            *sint = 0.0;
            do {
                icount += 1;
                if (icount > 20) break;
                wprob = wprob * lambda / icount;
                wsum = wsum + wprob;
                do {
                    xi = RND();
                    xi = 2.0 * chia2 * xi / (1.0 - xi + chia2);
                    cosz = 1.0 - xi;
                    rejf = spin_rejection(c, imed, qel, elke, beta2, q1, cosz, spin_index, 0, sr);
                    rnno = RND();
                } while (rnno > rejf);
                sinz = xi * (2.0 - xi);
                if (sinz > 1.0E-20) {
                    sinz = sqrt(sinz);
                    xi = RND();
                    phi = xi * 6.2831853;
                    *cost = (*cost) * cosz - *sint * sinz * cos(phi);
                    *sint = sqrt(fmax(0.0, (double)((1.0 - (*cost)) * (1.0 + (*cost)))));
                }
            } while (wsum <= sprob);
            return;
        }
    }
    if (lambda <= LAMBMAX_MS) {
        double ai, aj, llmbda = log(lambda);
        if (*find_index) {
            ai = llmbda * T->dllambi;
            ms->i = (int)ai;
            ai -= (double)ms->i;
            xi = RND();
            if (xi < ai) ms->i += 1;
            if (q1 < QMIN_MS) {
                ms->j = 0;
            } else if (q1 < QMAX_MS) {
                aj = q1 * T->dqmsi;
                ms->j = (int)aj;
                aj -= (double)ms->j;
                xi = RND();
                if (xi < aj) ms->j += 1;
            } else {
                ms->j = MXQ_MS;
            }
            if (llmbda < 2.2299)
                ms->omega2 = chia2 * (lambda + 4.0) * (1.347006 + llmbda * (0.209364 - llmbda * (0.45525 - llmbda * (0.50142 - 0.081234 * llmbda))));
            else
                ms->omega2 = chia2 * (lambda + 4.0) * (-2.77164 + llmbda * (2.94874 - llmbda * (0.1535754 - llmbda * 0.00552888)));
            *find_index = 0;
        }
        int k;
        double a, ak, u, du, x1;
        const int base = ms->i * (MXQ_MS + 1) * (MXU_MS + 1) + ms->j * (MXU_MS + 1);
        do {
            xi = RND();
            ak = xi * MXU_MS;
            k = ak;
            ak -= k;
            if (ak > T->wms[base + k]) k = T->ims[base + k];
            a = T->fms[base + k];
            u = T->ums[base + k];
            du = T->ums[base + k] - u;   /* zero, as in the reference */
            xi = RND();
            if (fabs(a) < 0.2) {
                x1 = 0.5 * (1.0 - xi) * a;
                u += xi * du * (1.0 + x1 * (1.0 - xi * a));
            } else {
                u -= du / a * (1.0 - sqrt(1.0 + xi * a * (2.0 + a)));
            }
            xi = ms->omega2 * u / (1.0 + 0.5 * ms->omega2 - u);
            if (xi > 1.99999) xi = 1.99999;
            *cost = 1.0 - xi;
            rejf = spin_rejection(c, imed, qel, elke, beta2, q1, *cost, spin_index, 0, sr);
            rnno = RND();
        } while (rnno > rejf);
        *sint = sqrt(xi * (2.0 - xi));
    }
}

/* msdist (src/ompmc.c:3787-3976): PRESTA-II lateral and angular deflection */
static double msdist_core(ora_ctx *c, int imed, int iq, double rhof, double de, double tustep,
                          double eke, double x0, double y0, double z0, double u0, double v0, double w0,
                          double *xf, double *yf, double *zf, double *uf, double *vf, double *wf)
{
    const omc_tables *T = c->t;
    int qel = (1 + iq) / 2;
    ora_mscats ms = { 0, 0, 0.0 };
    ora_spinr sr = { 0, 0 };
    double blccc = T->blcc[imed], xcccc = T->xcc[imed];
    double e = eke - 0.5 * de;
    double tau = e / RM;
    double tau2 = pow(tau, 2.0);
    double epsilon = de / eke;
    double epsilonp = de / e;
    e *= (1.0 - pow(epsilonp, 2.0) * (6.0 + 10.0 * tau + 5.0 * tau2) / (24.0 * tau2 + 72.0 * tau + 48.0));
    double p2 = e * (e + 2.0 * RM);
    double beta2 = p2 / (p2 + pow(RM, 2.0));
    double chia2 = xcccc / (4.0 * p2 * blccc);
    double lambda = 0.5 * tustep * rhof * blccc / beta2;
    double temp2 = 0.166666 * (4.0 + tau * (6.0 + tau * (7.0 + tau * (4.0 + tau)))) *
                   (epsilonp / ((tau + 1.0) * (tau + 2.0))) * (epsilonp / ((tau + 1.0) * (tau + 2.0)));
    lambda *= (1.0 - temp2);
    double elke = log(e);
    int lelke = pwlf_i(imed, elke, T->eke1, T->eke0) - 1;
    if (lelke < 0) {
        lelke = 0;
        elke = (1.0 - T->eke0[imed]) / T->eke1[imed];
    }
    int ix = MXEKE * imed + lelke;
    double etap, xi_corr, gamma;
    if (qel == 0) {
        etap = pwlf_e(ix, elke, T->etae_ms1, T->etae_ms0);
        xi_corr = pwlf_e(ix, elke, T->q1ce_ms1, T->q1ce_ms0);
        gamma = pwlf_e(ix, elke, T->q2ce_ms1, T->q2ce_ms0);
    } else {
        etap = pwlf_e(ix, elke, T->etap_ms1, T->etap_ms0);
        xi_corr = pwlf_e(ix, elke, T->q1cp_ms1, T->q1cp_ms0);
        gamma = pwlf_e(ix, elke, T->q2cp_ms1, T->q2cp_ms0);
    }
    double ms_corr = pwlf_e(ix, elke, T->blcce1, T->blcce0);
    chia2 *= etap;
    lambda /= (etap * (1.0 + chia2));
    lambda *= ms_corr;
    double chilog = log(1.0 + 1.0 / chia2);
    double q1 = 2.0 * chia2 * (chilog * (1.0 + chia2) - 1.0);
    gamma = 6.0 * chia2 * (1.0 + chia2) * (chilog * (1.0 + 2.0 * chia2) - 2.0) / q1 * gamma;
    double xi = q1 * lambda;

    int find_index = 1, spin_index = 1;
    double w1, sint1, cphi1, sphi1, w2, sint2, cphi2, sphi2;
    mscat(c, imed, qel, &spin_index, &find_index, elke, beta2, xi, lambda, chia2, &w1, &sint1, &ms, &sr);
    azimuth(c, &cphi1, &sphi1);
    mscat(c, imed, qel, &spin_index, &find_index, elke, beta2, xi, lambda, chia2, &w2, &sint2, &ms, &sr);
    azimuth(c, &cphi2, &sphi2);

    double u2 = sint2 * cphi2;
    double v2 = sint2 * sphi2;
    double u2p = w1 * u2 + sint1 * w2;
    double us = u2p * cphi1 - v2 * sphi1;
    double vs = u2p * sphi1 + v2 * cphi1;
    double ws = w1 * w2 - sint1 * u2;

    xi *= 2 * xi_corr;
    double eta = RND();
    double eta1 = 0.5 * (1.0 - eta);
    double delta = 0.9082483 - (0.1020621 - 0.0263747 * gamma) * xi;
    double temp1 = 2.0 + tau;
    double temp = (2.0 + tau * temp1) / ((tau + 1.0) * temp1);
    temp -= (tau + 1.0) / ((tau + 2.0) * (chilog * (1.0 + chia2) - 1.0));
    temp *= epsilonp;
    temp1 = 1.0 - temp;
    delta += 0.40824829 * (epsilon * (tau + 1.0) / ((tau + 2.0) * (chilog * (1.0 + chia2) - 1.0) *
                                                    (chilog * (1.0 + 2.0 * chia2) - 2.0)) - 0.25 * pow(temp, 2.0));
    double b = eta * delta;
    double cc = eta * (1.0 - delta);
    double w1v2 = w1 * v2;
    double ut = b * sint1 * cphi1 + cc * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
    double vt = b * sint1 * sphi1 + cc * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
    double wt = eta1 * (1.0 + temp) + b * w1 + cc * w2 + eta1 * ws * temp1;
    double ustep = tustep * sqrt(ut * ut + vt * vt + wt * wt);

    double sint02 = u0 * u0 + v0 * v0;
    if (sint02 > 1.0E-20) {
        double sint0 = sqrt(sint02);
        double sint0i = 1.0 / sint0;
        double cphi0 = sint0i * u0, sphi0 = sint0i * v0;
        u2p = w0 * us + sint0 * ws;
        ws = w0 * ws - sint0 * us;
        us = u2p * cphi0 - vs * sphi0;
        vs = u2p * sphi0 + vs * cphi0;
        u2p = w0 * ut + sint0 * wt;
        wt = w0 * wt - sint0 * ut;
        ut = u2p * cphi0 - vt * sphi0;
        vt = u2p * sphi0 + vt * cphi0;
    } else {
        wt = w0 * wt;
        ws = w0 * ws;
    }
    *xf = x0 + tustep * ut;
    *yf = y0 + tustep * vt;
    *zf = z0 + tustep * wt;
    *uf = us;
    *vf = vs;
    *wf = ws;
    return ustep;
}

/* computeDrange (src/ompmc.c:3979-4014) */
double ora_compute_drange(const ora_ctx *c, int imed, int iq, int lelke, double ekei, double ekef,
                          double elkei, double elkef)
{
    const omc_tables *T = c->t;
    double fedep = 1.0 - ekef / ekei;
    double elktmp = 0.5 * (elkei + elkef + 0.25 * fedep * fedep * (1.0 + fedep * (1.0 + 0.875 * fedep)));
    double dedxmid, aux;
    int ix = MXEKE * imed + lelke;
    if (iq < 0) {
        dedxmid = pwlf_e(ix, elktmp, T->ededx1, T->ededx0);
        dedxmid = 1.0 / dedxmid;
        aux = T->ededx1[ix] * dedxmid;
    } else {
        dedxmid = pwlf_e(ix, elktmp, T->pdedx1, T->pdedx0);
        dedxmid = 1.0 / dedxmid;
        aux = T->pdedx1[ix] * dedxmid;
    }
    aux = aux * (1.0 + 2.0 * aux) * fedep * fedep / (6.0 * (2.0 - fedep) * (2.0 - fedep));
    return fedep * ekei * dedxmid * (1.0 + aux);
}

/* computeEloss (src/ompmc.c:4016-4108) */
double ora_compute_eloss(const ora_ctx *c, int imed, int iq, double rhof, double tustep,
                         double range, double eke, double elke, int lelke)
{
    const omc_tables *T = c->t;
    int nmed = T->nmed;
    int qel = (1 + iq) / 2;
    const double *rng = T->range_ep + qel * nmed * MXEKE + imed * MXEKE;
    double aux, dedxmid, de, fedep;
    double tuss = range - rng[lelke] / rhof;
    if (tuss >= tustep) {
        int ix = imed * MXEKE + lelke;
        if (iq < 0) {
            dedxmid = pwlf_e(ix, elke, T->ededx1, T->ededx0);
            aux = T->ededx1[ix] / dedxmid;
        } else {
            dedxmid = pwlf_e(ix, elke, T->pdedx1, T->pdedx0);
            aux = T->pdedx1[ix] / dedxmid;
        }
        de = dedxmid * tustep * rhof;
        fedep = de / eke;
        de *= (1.0 - 0.5 * fedep * aux * (1.0 - 0.333333 * fedep * (aux - 1.0 - 0.25 * fedep * (2.0 - aux * (4.0 - aux)))));
    } else {
        int lelktmp = lelke;
        tuss = (range - tustep) * rhof;
        if (tuss <= 0) {
            de = eke - T->te[imed] * 0.99;
        } else {
            while (tuss < rng[lelktmp]) lelktmp -= 1;
            double elktmp = (lelktmp + 2 - T->eke0[imed]) / T->eke1[imed];
            double eketmp = T->e_array[imed * MXEKE + lelktmp + 1];
            tuss = (rng[lelktmp + 1] - tuss) / rhof;
            int ix = MXEKE * imed + lelktmp;
            if (iq < 0) {
                dedxmid = pwlf_e(ix, elktmp, T->ededx1, T->ededx0);
                aux = T->ededx1[ix] / dedxmid;
            } else {
                dedxmid = pwlf_e(ix, elktmp, T->pdedx1, T->pdedx0);
                aux = T->pdedx1[ix] / dedxmid;
            }
            de = dedxmid * tuss * rhof;
            fedep = de / eketmp;
            de *= (1.0 - 0.5 * fedep * aux * (1.0 - 0.333333 * fedep * (aux - 1.0 - 0.25 * fedep * (2.0 - aux * (4.0 - aux)))));
            de += eke - eketmp;
        }
    }
    return de;
}

/* ===================================================== e-/e+ interactions */

/* Russian roulette of photons made by an e-/e+ when splitting is on
 * (src/ompmc.c:4149-4164, 4338-4353, 4627-4642). */
static void roulette_photons(ora_ctx *c)
{
    if (c->nsplit > 1) {
        for (int ip = c->npold; ip <= c->np; ++ip) {
            if (c->S[ip].iq == 0) {
                double r = RND();
                if (r * (double)c->nsplit > 1.0) { c->S[ip].wt = 0.0; c->S[ip].e = 0.0; }
                else c->S[ip].wt *= c->nsplit;
            }
        }
    }
}

/* rannih (src/ompmc.c:4111-4167): annihilation at rest */
static void ora_rannih(ora_ctx *c)
{
    c->cnt.n_rannih++;
    int np = c->np;
    c->npold = np;
    double rnno = RND();
    double costhe = 2.0 * rnno - 1;
    double sinthe = sqrt(fmax(0.0, (1.0 - costhe) * (1.0 + costhe)));
    rnno = RND();            /* drawn and unused in the reference */
    (void)rnno;
    double cphi, sphi;
    azimuth(c, &cphi, &sphi);
    ora_particle *S = c->S;
    S[np].e = RM;
    S[np].iq = 0;
    S[np].u = sinthe * cphi;
    S[np].v = sinthe * sphi;
    S[np].w = costhe;
    np += 1;
    S[np].e = RM;
    S[np].iq = 0;
    transfer(c, np, np - 1);
    S[np].u = -1.0 * S[np - 1].u;
    S[np].v = -1.0 * S[np - 1].v;
    S[np].w = -1.0 * S[np - 1].w;
    c->np = np;
    roulette_photons(c);
    stream_leave(c, np - 1, 1);
}

/* brems (src/ompmc.c:4170-4356): BH energy sampling, KM-2BS angle */
static void ora_brems(ora_ctx *c)
{
    c->cnt.n_brems++;
    const omc_tables *T = c->t;
    ora_particle *S = c->S;
    int np = c->np;
    int irl = S[np].ir;
    int imed = c->r->med[irl];
    double eie = S[np].e, phi1, phi2;
    c->npold = np;
    int l = (eie < 50.0) ? 1 : 3;
    int l1 = l + 1;
    double ekin = eie - RM;
    double brmin = T->ap[imed] / ekin;
    double waux = -log(brmin);
    double a = S[np].u, b = S[np].v, cc = S[np].w;
    double sinpsi = a * a + b * b, sindel = 0.0, cosdel = 0.0;
    if (sinpsi > 1.0E-20) {
        sinpsi = sqrt(sinpsi);
        sindel = b / sinpsi;
        cosdel = a / sinpsi;
    }
    double ztarg = T->zbrang[imed];
    double tteie = eie / RM;
    double beta = sqrt((tteie - 1.0) * (tteie + 1.0)) / tteie;
    double y2max = 2.0 * beta * (1.0 + beta) * tteie * tteie;
    double y2maxi = 1.0 / y2max;
    double z2max = y2max + 1.0;
    double z2maxi = sqrt(z2max);
    const double *d1 = T->dl1 + imed * 8, *d2 = T->dl2 + imed * 8, *d3 = T->dl3 + imed * 8;
    const double *d4 = T->dl4 + imed * 8, *d5 = T->dl5 + imed * 8, *d6 = T->dl6 + imed * 8;
    double aux, br, delta, r6, r7, rejf, ese, esg;
    do {
        r6 = RND();
        r7 = RND();
        br = brmin * exp(r6 * waux);
        esg = ekin * br;
        ese = eie - esg;
        delta = esg / eie / ese * T->delcm[imed];
        aux = ese / eie;
        if (delta < 1.0) {
            phi1 = d1[l - 1] + delta * (d2[l - 1] + delta * d3[l - 1]);
            phi2 = d1[l1 - 1] + delta * (d2[l1 - 1] + delta * d3[l1 - 1]);
        } else {
            phi1 = d4[l - 1] + d5[l - 1] * log(delta + d6[l - 1]);
            phi2 = phi1;
        }
        rejf = (1.0 + pow(aux, 2.0)) * phi1 - 2.0 * aux * phi2 / 3.0;
    } while (r7 >= rejf);

    np += 1;
    S[np].e = esg;
    S[np].iq = 0;
    transfer(c, np, np - 1);

    double y2tst = 0.0;
    double ttese = ese / RM;
    double esedei = ttese / tteie;
    double rjarg1 = 1.0 + esedei * esedei;
    double rjarg2 = rjarg1 + 2.0 * esedei;
    double rjarg3;
    double rtest = 1.0, rejtst = 0.0;
    aux = 2.0 * ese * tteie / esg;
    aux = aux * aux;
    double aux1 = aux * ztarg;
    if (aux1 > 10.0) rjarg3 = -log(T->zbrang[imed]) + (1.0 - aux1) / pow(aux1, 2.0);
    else rjarg3 = log(aux / (1.0 + aux1));
    double rejmax = rjarg1 * rjarg3 - rjarg2;
    while (rtest >= rejtst) {
        y2tst = RND();
        rtest = RND();
        double aux3 = z2maxi / (y2tst + (1.0 - y2tst) * z2maxi);
        rtest = rtest * aux3 * rejmax;
        y2tst = pow(aux3, 2.0) - 1.0;
        double y2tst1 = esedei * y2tst / pow(aux3, 4.0);
        double aux4 = 16.0 * y2tst1 - rjarg2;
        double aux5 = rjarg1 - 4.0 * y2tst1;
        if (rtest < aux4 + aux5 * rjarg3) break;
        double aux2 = log(aux / (1.0 + aux1 / pow(aux3, 4.0)));
        rejtst = aux4 + aux5 * aux2;
    }
    double costhe = 1.0 - 2.0 * y2tst * y2maxi;
    double sinthe = sqrt(fmax(0.0, (1.0 - pow(costhe, 2.0))));
    double cphi, sphi;
    azimuth(c, &cphi, &sphi);
    if (sinpsi >= 1.0E-10) {
        double us = sinthe * cphi, vs = sinthe * sphi;
        S[np].u = cc * cosdel * us - sindel * vs + a * costhe;
        S[np].v = cc * sindel * us + cosdel * vs + b * costhe;
        S[np].w = cc * costhe - sinpsi * us;
    } else {
        S[np].u = sinthe * cphi;
        S[np].v = sinthe * sphi;
        S[np].w = cc * costhe;
    }
    S[np - 1].e = ese;
    c->np = np;
    roulette_photons(c);
    stream_leave(c, np - 1, 1);
}

/* moller (src/ompmc.c:4359-4435) */
static void ora_moller(ora_ctx *c)
{
    c->cnt.n_moller++;
    const omc_tables *T = c->t;
    ora_particle *S = c->S;
    int np = c->np;
    int imed = c->r->med[S[np].ir];
    double eie = S[np].e, ekin = eie - RM;
    c->npold = np;
    if (ekin <= 2.0 * T->te[imed]) { stream_leave(c, np, 0); return; }
    double t0 = ekin / RM, e0 = t0 + 1.0;
    double extrae = eie - T->thmoll[imed];
    double g2 = pow(t0, 2.0) / pow(e0, 2.0);
    double g3 = (2.0 * t0 + 1.0) / pow(e0, 2.0);
    double gmax = (1.0 + 1.25 * g2), br, rejf4, r, r27, r28;
    do {
        r27 = RND();
        br = T->te[imed] / (ekin - extrae * r27);
        r = br / (1.0 - br);
        r28 = RND();
        rejf4 = (1.0 + g2 * pow(br, 2.0) + r * (r - g3));
        r28 *= gmax;
    } while (r28 > rejf4);
    double ekse2 = br * ekin, ese1 = eie - ekse2, ese2 = ekse2 + RM;
    S[np].e = ese1;
    S[np + 1].e = ese2;
    double h1 = (eie + RM) / ekin;
    double costh = h1 * (ese1 - RM) / (ese1 + RM);
    double sinthe = sqrt(1.0 - costh), costhe = sqrt(costh);
    ora_uphi h;
    uphi21(c, &h, costhe, sinthe);
    np += 1;
    c->np = np;
    S[np].iq = -1;
    costh = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrt(1.0 - costh);
    costhe = sqrt(costh);
    uphi32(c, &h, costhe, sinthe);
    stream_leave(c, np - 1, 1);
}

/* bhabha (src/ompmc.c:4438-4525) */
static void ora_bhabha(ora_ctx *c)
{
    c->cnt.n_bhabha++;
    const omc_tables *T = c->t;
    ora_particle *S = c->S;
    int np = c->np;
    int imed = c->r->med[S[np].ir];
    double eip = S[np].e, ekin = eip - RM;
    double t0 = ekin / RM, e0 = t0 + 1.0;
    double yy = 1.0 / (t0 + 2.0);
    double beta2 = (pow(e0, 2.0) - 1.0) / pow(e0, 2.0);
    double ep0 = T->te[imed] / ekin, ep0c = 1.0 - ep0;
    double yp = 1.0 - 2.0 * yy;
    c->npold = np;
    double b4 = pow(yp, 3.0);
    double b3 = b4 + pow(yp, 2.0);
    double b2 = yp * (3.0 + pow(yy, 2.0));
    double b1 = 2.0 - pow(yy, 2.0);
    double r3, r4, br, rejf2;
    do {
        r3 = RND();
        br = ep0 / (1.0 - ep0c * r3);
        r4 = RND();
        rejf2 = (1.0 - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4))));
    } while (r4 > rejf2);
    if (br < 0.5) {
        S[np + 1].iq = -1;
    } else {
        S[np].iq = -1;
        S[np + 1].iq = 1;
        br = 1.0 - br;
    }
    br = fmax(br, 0.0);
    double ekse2 = br * ekin, ese1 = eip - ekse2, ese2 = ekse2 + RM;
    S[np].e = ese1;
    S[np + 1].e = ese2;
    double h1 = (eip + RM) / ekin;
    double costh = fmin(1.0, h1 * (ese1 - RM) / (ese1 + RM));
    double sinthe = sqrt(1.0 - costh), costhe = sqrt(costh);
    ora_uphi h;
    uphi21(c, &h, costhe, sinthe);
    np += 1;
    c->np = np;
    costh = h1 * (ese2 - RM) / (ese2 + RM);
    sinthe = -sqrt(1.0 - costh);
    costhe = sqrt(costh);
    uphi32(c, &h, costhe, sinthe);
    stream_leave(c, np - 1, 1);
}

/* annih (src/ompmc.c:4528-4645): two-photon annihilation in flight */
static void ora_annih(ora_ctx *c)
{
    c->cnt.n_annih++;
    ora_particle *S = c->S;
    int np = c->np;
    double avip = S[np].e + RM;
    double a = avip / RM;
    double g = a - 1.0, t = g - 1.0, p = sqrt(a * t);
    c->npold = np;
    double pot = p / t;
    double ep0 = 1.0 / (a + p);
    double wsamp = log((1.0 - ep0) / ep0);
    double aa = S[np].u, bb = S[np].v, cc = S[np].w;
    double sinpsi = pow(aa, 2.0) + pow(bb, 2.0), sindel = 0.0, cosdel = 0.0;
    if (sinpsi > 1.0E-20) {
        sinpsi = sqrt(sinpsi);
        sindel = bb / sinpsi;
        cosdel = aa / sinpsi;
    }
    double ep, rejf, r1, r2;
    do {
        r1 = RND();
        ep = ep0 * exp(r1 * wsamp);
        r2 = RND();
        rejf = 1.0 - pow(ep * a - 1.0, 2.0) / (ep * (pow(a, 2.0) - 2.0));
    } while (r2 > rejf);
    double esg1 = avip * ep;
    S[np].e = esg1;
    S[np].iq = 0;
    double costhe = fmin(1.0, (esg1 - RM) * pot / esg1);
    double sinthe = sqrt(1.0 - pow(costhe, 2.0));
    double sphi, cphi;
    azimuth(c, &cphi, &sphi);
    double us, vs;
    if (sinpsi >= 1.0E-10) {
        us = sinthe * cphi;
        vs = sinthe * sphi;
        S[np].u = cc * cosdel * us - sindel * vs + aa * costhe;
        S[np].v = cc * sindel * us + cosdel * vs + bb * costhe;
        S[np].w = cc * costhe - sinpsi * us;
    } else {
        S[np].u = sinthe * cphi;
        S[np].v = sinthe * sphi;
        S[np].w = cc * costhe;
    }
    np += 1;
    double esg2 = avip - esg1;
    S[np].e = esg2;
    S[np].iq = 0;
    transfer(c, np, np - 1);
    costhe = fmin(1.0, (esg2 - RM) * pot / esg2);
    sinthe = -sqrt(1.0 - pow(costhe, 2.0));
    if (sinpsi >= 1.0E-10) {
        us = sinthe * cphi;
        vs = sinthe * sphi;
        S[np].u = cc * cosdel * us - sindel * vs + aa * costhe;
        S[np].v = cc * sindel * us + cosdel * vs + bb * costhe;
        S[np].w = cc * costhe - sinpsi * us;
    } else {
        S[np].u = sinthe * cphi;
        S[np].v = sinthe * sphi;
        S[np].w = cc * costhe;
    }
    c->np = np;
    roulette_photons(c);
    stream_leave(c, np - 1, 1);
}

/* Deposits the kinetic energy of an e-/e+ below cutoff; positrons then
 * annihilate at rest (the repeated block of src/ompmc.c:4665-4687). */
static void electron_cutoff(ora_ctx *c, int iq, double eie)
{
    double edep = c->S[c->np].e - RM;
    ausgab(c, edep);
    if (iq > 0 && edep < eie) { ora_rannih(c); return; }
    c->np -= 1;
}

/* electron (src/ompmc.c:4648-5433): class-II condensed history step */
static void ora_electron(ora_ctx *c)
{
    const omc_tables *T = c->t;
    const omc_regions *R = c->r;
    ora_particle *S = c->S;
    const int nmed = T->nmed;
    int np = c->np;
    int irl = S[np].ir;
    int imed = R->med[irl];
    double rhof = R->rhof[irl];
    double edep = 0.0;
    ora_uphi h;
    double eie = S[np].e;
    int iq = S[np].iq;
    int qel = (1 + iq) / 2;
    int medold = imed;
    double rnno;

    stream_enter(c, np);
    if (eie <= R->ecut[irl]) { electron_cutoff(c, iq, eie); return; }

    double elke = 0.0;
    int lelke = 0;
    double sigratio = 0.0, rfict = 0.0;
    double sig0 = 0.0, demfp = 0.0;

    do {
        int compute_tstep = 1;
        double eke = eie - RM;
        double sigf, dedx0, ustep = 0.0;
        if (imed != -1) {
            rnno = RND();
            if (rnno == 0.0) rnno = 1.0E-30;
            demfp = fmax(-log(rnno), EPSEMFP);
            elke = log(eke);
            lelke = pwlf_i(imed, elke, T->eke1, T->eke0) - 1;
            int ix = imed * MXEKE + lelke;
            if (T->sig_ismonotone[qel][imed]) {
                if (iq < 0) {
                    sig0 = pwlf_e(ix, elke, T->esig1, T->esig0);
                    dedx0 = pwlf_e(ix, elke, T->ededx1, T->ededx0);
                    sig0 /= dedx0;
                } else {
                    sig0 = pwlf_e(ix, elke, T->psig1, T->psig0);
                    dedx0 = pwlf_e(ix, elke, T->pdedx1, T->pdedx0);
                    sig0 /= dedx0;
                }
            } else {
                sig0 = (iq < 0) ? T->esig_e[imed] : T->psig_e[imed];
            }
        }
        do {
            /* re-initialised every ustep iteration, as in the reference
             * (src/ompmc.c:4780-4788): a second iteration therefore sees
             * tstep = 0, takes a zero-length step and restarts the tstep
             * loop after one boundary-crossing random number */
            double total_de = 0.0, total_tstep = 0.0;
            int call_howfar, do_single = 0, called_msdist = 0;
            c->cnt.n_electron_steps++;
            double tstep, tustep, ekef, ekei, elkei, tuss, range = 0.0, p2, beta2, etap;
            double tvstep = 0.0, de = 0.0;
            double xf = 0, yf = 0, zf = 0, uf = 0, vf = 0, wf = 0;
            const double *rng = T->range_ep + qel * nmed * MXEKE + imed * MXEKE;

            if (imed == -1) {
                tstep = 10.0E8;
                ustep = tstep;
                tustep = ustep;
                call_howfar = 1;
            } else {
                rhof = R->rhof[irl];
                if (sig0 <= 0.0) {
                    tstep = 10.0E8;
                    sig0 = 1.0E-15;
                } else {
                    if (compute_tstep) {
                        total_de = demfp / sig0;
                        ekef = eke - total_de;
                        if (ekef <= T->e_array[imed * MXEKE + 0]) {
                            tstep = 10.0E8;
                        } else {
                            double elkef = log(ekef);
                            int lelkef = pwlf_i(imed, elkef, T->eke1, T->eke0) - 1;
                            if (lelkef == lelke) {
                                tstep = ora_compute_drange(c, imed, iq, lelke, eke, ekef, elke, elkef);
                            } else {
                                ekei = T->e_array[imed * MXEKE + lelke];
                                elkei = (lelke + 1 - T->eke0[imed]) / T->eke1[imed];
                                tuss = ora_compute_drange(c, imed, iq, lelke, eke, ekei, elke, elkei);
                                ekei = T->e_array[imed * MXEKE + lelkef + 1];
                                elkei = ((lelkef + 2) - T->eke0[imed]) / T->eke1[imed];
                                tstep = ora_compute_drange(c, imed, iq, lelkef, ekei, ekef, elkei, elkef);
                                tstep += tuss + rng[lelke] - rng[lelkef + 1];
                            }
                        }
                        total_tstep = tstep;
                        compute_tstep = 0;
                    }
                    tstep = total_tstep / rhof;
                }
                int ix = imed * MXEKE + lelke;
                if (iq < 0) dedx0 = pwlf_e(ix, elke, T->ededx1, T->ededx0);
                else dedx0 = pwlf_e(ix, elke, T->pdedx1, T->pdedx0);
                double dedx = rhof * dedx0;
                double tmxs = pwlf_e(ix, elke, T->tmxs1, T->tmxs0);
                tmxs /= rhof;
                ekei = T->e_array[imed * MXEKE + lelke];
                elkei = (lelke + 1 - T->eke0[imed]) / T->eke1[imed];
                range = ora_compute_drange(c, imed, iq, lelke, eke, ekei, elke, elkei);
                range += rng[lelke];
                range /= rhof;
                tustep = fmin(fmin(tstep, tmxs), range);
                double tperp = ora_hownear(c, &S[np]);
                S[np].dnear = tperp;
                double blccl = rhof * T->blcc[imed];
                double xccl = rhof * T->xcc[imed];
                p2 = eke * (eke + 2.0 * RM);
                beta2 = p2 / (p2 + pow(RM, 2.0));
                if (iq < 0) etap = pwlf_e(MXEKE * imed + lelke, elke, T->etae_ms1, T->etae_ms0);
                else etap = pwlf_e(MXEKE * imed + lelke, elke, T->etap_ms1, T->etap_ms0);
                double ms_corr = pwlf_e(MXEKE * imed + lelke, elke, T->blcce1, T->blcce0);
                blccl = blccl / etap / (1.0 + 0.25 * etap * xccl / blccl / p2) * ms_corr;
                double ssmfp = beta2 / blccl;
                double skindepth = 3 * ssmfp;
                tustep = fmin(tustep, fmax(tperp, skindepth));
                if ((tustep <= tperp) && (tustep > skindepth)) {
                    call_howfar = 0;
                    do_single = 0;
                    called_msdist = 1;
                    de = ora_compute_eloss(c, imed, iq, rhof, tustep, range, eke, elke, lelke);
                    tvstep = tustep;
                    c->cnt.n_msdist++;
                    ustep = msdist_core(c, imed, iq, rhof, de, tustep, eke, S[np].x, S[np].y, S[np].z,
                                        S[np].u, S[np].v, S[np].w, &xf, &yf, &zf, &uf, &vf, &wf);
                } else {
                    called_msdist = 0;
                    rnno = RND();
                    if (rnno < 1.0E-30) rnno = 1.0E-30;
                    double lambda = (-1.0) * log(1.0 - rnno);
                    double lambda_max = 0.5 * blccl * RM / dedx;
                    lambda_max *= (eke / RM + 1.0) * (eke / RM + 1.0) * (eke / RM + 1.0);
                    if (lambda >= 0.0 && lambda_max > 0.0) {
                        if (lambda < lambda_max) tuss = lambda * ssmfp * (1.0 - 0.5 * lambda / lambda_max);
                        else tuss = 0.5 * lambda * ssmfp;
                        if (tuss < tustep) { tustep = tuss; do_single = 1; }
                        else do_single = 0;
                    } else {
                        /* the reference drops the particle without scoring */
                        c->np -= 1;
                        return;
                    }
                    ustep = tustep;
                    call_howfar = (ustep < tperp) ? 0 : 1;
                }
            }

            int irold = S[np].ir, irnew = S[np].ir, idisc = 0;
            if (call_howfar) {
                c->cnt.n_howfar_charged++;
                ora_howfar(c, &S[np], &idisc, &irnew, &ustep);
            }
            if (idisc > 0) {
                edep = (iq > 0) ? S[np].e + RM : S[np].e - RM;
                ausgab(c, edep);
                if (iq > 0 && edep < eie) { ora_rannih(c); return; }
                c->np -= 1;
                return;
            }
            if (ustep < 0) ustep = 0.0;
            double vstep;
            if (ustep == 0.0f || imed == -1) {
                if (ustep != 0.0f) {
                    vstep = ustep;
                    tvstep = vstep;
                    S[np].x += S[np].u * vstep;
                    S[np].y += S[np].v * vstep;
                    S[np].z += S[np].w * vstep;
                    S[np].dnear -= vstep;
                }
                if (irnew != irold) {
                    S[np].ir = irnew;
                    irl = irnew;
                    imed = R->med[irl];
                    c->cnt.n_cross_charged++;
                }
                if (eie <= R->ecut[irl]) { electron_cutoff(c, iq, eie); return; }
                break;
            }
            vstep = ustep;
            if (call_howfar) {
                tvstep = vstep;
                if (tvstep != tustep) do_single = 0;
                de = ora_compute_eloss(c, imed, iq, rhof, tvstep, range, eke, elke, lelke);
            } else {
                tvstep = tustep;
                if (called_msdist == 0)
                    de = ora_compute_eloss(c, imed, iq, rhof, tvstep, range, eke, elke, lelke);
            }
            edep = de;
            ekef = eke - de;
            double sinthe, costhe;
            if (called_msdist == 0) {
                if (do_single) {
                    double ekems = fmax(ekef, R->ecut[irl] - RM);
                    p2 = ekems * (ekems + 2.0 * RM);
                    beta2 = p2 / (p2 + pow(RM, 2.0));
                    double chia2 = T->xcc[imed] / (4.0 * T->blcc[imed] * p2);
                    double elkems = log(ekems);
                    int lelkems = pwlf_i(imed, elkems, T->eke1, T->eke0) - 1;
                    if (iq < 0) etap = pwlf_e(MXEKE * imed + lelkems, elkems, T->etae_ms1, T->etae_ms0);
                    else etap = pwlf_e(MXEKE * imed + lelkems, elkems, T->etap_ms1, T->etap_ms0);
                    chia2 *= etap;
                    c->cnt.n_sscat++;
                    sscat(c, imed, qel, chia2, elkems, beta2, &costhe, &sinthe);
                } else {
                    sinthe = 0.0f;
                    costhe = 1.0f;
                }
                xf = S[np].x + S[np].u * vstep;
                yf = S[np].y + S[np].v * vstep;
                zf = S[np].z + S[np].w * vstep;
                if (do_single) uphi21(c, &h, costhe, sinthe);
                uf = S[np].u; vf = S[np].v; wf = S[np].w;
            }
            ausgab(c, edep);
            S[np].x = xf; S[np].y = yf; S[np].z = zf;
            S[np].u = uf; S[np].v = vf; S[np].w = wf;
            S[np].dnear -= vstep;
            irold = S[np].ir;
            eie -= edep;
            S[np].e = eie;
            if (irnew == irl && eie <= R->ecut[irl]) { electron_cutoff(c, iq, eie); return; }
            medold = imed;
            if (imed != -1) {
                eke = eie - RM;
                elke = log(eke);
                lelke = pwlf_i(imed, elke, T->eke1, T->eke0) - 1;
            }
            if (irnew != irold) {
                S[np].ir = irnew;
                irl = irnew;
                imed = R->med[irl];
                c->cnt.n_cross_charged++;
            }
            if (eie <= R->ecut[irl]) { electron_cutoff(c, iq, eie); return; }
            if (imed != medold) break;
            demfp -= de * sig0;
            total_de -= de;
            total_tstep -= tvstep * rhof;
            if (total_tstep < 1.0E-9) demfp = 0.0;
        } while (demfp >= EPSEMFP);

        if ((imed != medold) || (ustep == 0.0) || (imed == -1)) continue;

        int ix = imed * MXEKE + lelke;
        if (iq < 0) {
            sigf = pwlf_e(ix, elke, T->esig1, T->esig0);
            dedx0 = pwlf_e(ix, elke, T->ededx1, T->ededx0);
            sigf /= dedx0;
        } else {
            sigf = pwlf_e(ix, elke, T->psig1, T->psig0);
            dedx0 = pwlf_e(ix, elke, T->pdedx1, T->pdedx0);
            sigf /= dedx0;
        }
        sigratio = sigf / sig0;
        rfict = RND();
        if (rfict >= sigratio) c->cnt.n_fict++;
    } while (rfict >= sigratio);

    int ix = imed * MXEKE + lelke;
    if (iq < 0) {
        double ebr1 = pwlf_e(ix, elke, T->ebr11, T->ebr10);
        rnno = RND();
        if (rnno <= ebr1) {
            ora_brems(c);
        } else if (S[np].e <= T->thmoll[imed]) {
            if (ebr1 <= 0) { stream_leave(c, np, 0); return; }
            ora_brems(c);
        } else {
            ora_moller(c);
        }
    } else {
        double pbr1 = pwlf_e(ix, elke, T->pbr11, T->pbr10);
        rnno = RND();
        if (rnno < pbr1) {
            ora_brems(c);
        } else {
            double pbr2 = pwlf_e(ix, elke, T->pbr21, T->pbr20);
            if (rnno < pbr2) ora_bhabha(c);
            else ora_annih(c);
        }
    }
}

/* shower (src/ompmc.c:5436-5447) */
static void ora_shower(ora_ctx *c)
{
    while (c->np >= 0 && !c->err) {
        if (c->S[c->np].iq == 0) ora_photon(c);
        else ora_electron(c);
    }
}

/* ============================================================== sources */

/* initHistory (omc_dosxyz.c:964-1068); `hid` selects the PHILOX stream */
static void init_history_dosxyz(ora_ctx *c, const omc_source_dosxyz *s, int64_t hid)
{
    const omc_phantom *g = c->g;
    ora_particle *p = &c->S[0];
    c->np = 0;
    c->hid = hid;
    if (c->mode == ORA_RNG_PHILOX) { c->cur = philox_primary(c->s0, c->s1, hid); c->cur_slot = 0; }
    p->iq = s->charge;
    double ein;
    if (s->spec.spectrum) {
        double r1 = RND(), r2 = RND();
        int k = (int)fmin(s->spec.deltak * r1, s->spec.deltak - 1.0);
        ein = s->spec.cdfinv1[k] + r2 * s->spec.cdfinv2[k];
    } else {
        ein = s->spec.energy;
    }
    p->e = (p->iq != 0) ? ein + RM : ein;
    double rxyz;
    if (s->xsize == 0.0 || s->ysize == 0.0) {
        p->x = s->xinl;
        p->y = s->yinl;
        rxyz = sqrt(pow(s->ssd, 2.0) + pow(p->x, 2.0) + pow(p->y, 2.0));
        p->w = s->ssd / rxyz;
    } else {
        double fw, r3;
        do {
            r3 = RND();
            p->x = r3 * s->xsize + s->xinl;
            r3 = RND();
            p->y = r3 * s->ysize + s->yinl;
            r3 = RND();
            rxyz = sqrt(s->ssd * s->ssd + p->x * p->x + p->y * p->y);
            p->w = s->ssd / rxyz;
            fw = p->w * p->w * p->w;
        } while (r3 >= fw);
    }
    p->z = g->zbounds[0];
    p->u = p->x / rxyz;
    p->v = p->y / rxyz;
    int ix, iy;
    if (s->xsize == 0.0) {
        ix = s->ixinl;
    } else {
        ix = s->ixinl - 1;
        while ((g->xbounds[ix + 1] < p->x) && ix < g->isize - 1) ix++;
    }
    if (s->ysize == 0.0) {
        iy = s->iyinl;
    } else {
        iy = s->iyinl - 1;
        while ((g->ybounds[iy + 1] < p->y) && iy < g->jsize - 1) iy++;
    }
    p->ir = 1 + ix + iy * g->isize;
    p->wt = 1.0;
    p->dnear = 0.0;
    p->rs = c->cur;
}

/* initHistory(ibeamlet) (omc_matrad.c:1084-1254) */
static void init_history_beamlet(ora_ctx *c, const omc_source_beamlets *s, int ib, int64_t hid)
{
    const omc_phantom *g = c->g;
    ora_particle *p = &c->S[0];
    int ijmax = g->isize * g->jsize, imax = g->isize;
    c->np = 0;
    c->hid = hid;
    if (c->mode == ORA_RNG_PHILOX) { c->cur = philox_primary(c->s0, c->s1, hid); c->cur_slot = 0; }
    p->iq = s->charge;
    double ein;
    if (s->spec.spectrum) {
        double r1 = RND(), r2 = RND();
        int k = (int)fmin(s->spec.deltak * r1, s->spec.deltak - 1.0);
        ein = s->spec.cdfinv1[k] + r2 * s->spec.cdfinv2[k];
    } else {
        ein = s->spec.energy;
    }
    p->e = (p->iq != 0) ? ein + RM : ein;
    double r1 = RND(), r2 = RND();
    const double *cr = s->corner + 3 * ib, *a1 = s->side1 + 3 * ib, *a2 = s->side2 + 3 * ib;
    double xiso = r1 * a1[0] + r2 * a2[0] + cr[0];
    double yiso = r1 * a1[1] + r2 * a2[1] + cr[1];
    double ziso = r1 * a1[2] + r2 * a2[2] + cr[2];
    const double *sp = s->src + 3 * s->ibeam[ib];
    double vnorm = sqrt(pow(xiso - sp[0], 2.0) + pow(yiso - sp[1], 2.0) + pow(ziso - sp[2], 2.0));
    double u = -(xiso - sp[0]) / vnorm, v = -(yiso - sp[1]) / vnorm, w = -(ziso - sp[2]) / vnorm;
    double ustep = 1.0E5, dist;
    if (u > 0.0) { dist = (g->xbounds[g->isize] - xiso) / u; if (dist < ustep) ustep = dist; }
    if (u < 0.0) { dist = -(xiso - g->xbounds[0]) / u; if (dist < ustep) ustep = dist; }
    if (v > 0.0) { dist = (g->ybounds[g->jsize] - yiso) / v; if (dist < ustep) ustep = dist; }
    if (v < 0.0) { dist = -(yiso - g->ybounds[0]) / v; if (dist < ustep) ustep = dist; }
    if (w > 0.0) { dist = (g->zbounds[g->ksize] - ziso) / w; if (dist < ustep) ustep = dist; }
    if (w < 0.0) { dist = -(ziso - g->zbounds[0]) / w; if (dist < ustep) ustep = dist; }
    p->x = xiso + ustep * u;
    p->y = yiso + ustep * v;
    p->z = ziso + ustep * w;
    p->u = -u; p->v = -v; p->w = -w;
    const double tiny = 2.0 * 2.2250738585072014e-308;
    if (p->x < g->xbounds[0]) p->x = g->xbounds[0] + tiny;
    if (p->x > g->xbounds[g->isize]) p->x = g->xbounds[g->isize] - tiny;
    if (p->y < g->ybounds[0]) p->y = g->ybounds[0] + tiny;
    if (p->y > g->ybounds[g->jsize]) p->y = g->ybounds[g->jsize] - tiny;
    /* the reference clamps z to ybounds[0] here (omc_matrad.c:1225) */
    if (p->z < g->zbounds[0]) p->z = g->ybounds[0] + tiny;
    if (p->z > g->zbounds[g->ksize]) p->z = g->zbounds[g->ksize] - tiny;
    int ix = 0, iy = 0, iz = 0;
    while (ix < g->isize - 1 && g->xbounds[ix + 1] < p->x) ix++;
    while (iy < g->jsize - 1 && g->ybounds[iy + 1] < p->y) iy++;
    while (iz < g->ksize - 1 && g->zbounds[iz + 1] < p->z) iz++;
    p->ir = 1 + ix + iy * imax + iz * ijmax;
    p->wt = 1.0;
    p->dnear = 0.0;
    p->rs = c->cur;
}

/* ================================================================ public */

ora_ctx *ora_create(const omc_tables *t, const omc_regions *regions, const omc_phantom *phantom,
                    int nsplit, int rng_mode)
{
    ora_ctx *c = calloc(1, sizeof *c);
    if (!c) return NULL;
    c->t = t;
    c->r = regions;
    c->g = phantom;
    c->nsplit = nsplit < 1 ? 1 : nsplit;
    c->mode = rng_mode;
    c->S = calloc(MXSTACK + 2, sizeof(ora_particle));
    c->hid = -1;
    c->np = -1;
    ora_seed_ranmar(c, 97, 33);
    return c;
}

void ora_destroy(ora_ctx *c)
{
    if (!c) return;
    free(c->S);
    free(c);
}

void ora_seed_philox(ora_ctx *c, uint32_t s0, uint32_t s1) { c->s0 = s0; c->s1 = s1; }
void ora_set_tape(ora_ctx *c, const int32_t *tape, int64_t n) { c->tape = tape; c->tape_n = n; c->tape_pos = 0; }
int64_t ora_tape_pos(const ora_ctx *c) { return c->tape_pos; }
void ora_get_counters(const ora_ctx *c, ora_counters *out) { *out = c->cnt; }
void ora_set_ledger(ora_ctx *c, double *ledger3, int64_t n) { c->ledger = ledger3; c->ledger_n = n; }

int ora_run_dosxyz(ora_ctx *c, const omc_source_dosxyz *src, int nperbatch, int nbatch,
                   double *sum, double *sum2)
{
    if (c->mode == ORA_RNG_PHILOX && c->nsplit != 1) return OMC_ENOTSUP;
    int nreg = c->r->nreg;
    double *endep = calloc(nreg, sizeof(double));
    if (!endep) return OMC_ENOMEM;
    c->endep = endep;
    for (int ib = 0; ib < nbatch; ++ib) {
        for (int ih = 0; ih < nperbatch; ++ih) {
            init_history_dosxyz(c, src, (int64_t)ib * nperbatch + ih);
            c->cnt.n_hist++;
            ora_shower(c);
            if (c->err) { free(endep); c->endep = NULL; return OMC_EOVERFLOW; }
        }
        for (int i = 0; i < nreg; ++i) {
            double e = endep[i];
            sum[i] += e;
            sum2[i] += e * e;
        }
        memset(endep, 0, nreg * sizeof(double));
    }
    free(endep);
    c->endep = NULL;
    return OMC_OK;
}

int ora_run_beamlet(ora_ctx *c, const omc_source_beamlets *src, int ibeamlet, int64_t hist0,
                    int nperbatch, int nbatch, double *sum, double *sum2)
{
    if (c->mode == ORA_RNG_PHILOX && c->nsplit != 1) return OMC_ENOTSUP;
    int nreg = c->r->nreg;
    double *endep = calloc(nreg, sizeof(double));
    if (!endep) return OMC_ENOMEM;
    c->endep = endep;
    for (int ib = 0; ib < nbatch; ++ib) {
        for (int ih = 0; ih < nperbatch; ++ih) {
            init_history_beamlet(c, src, ibeamlet, hist0 + (int64_t)ib * nperbatch + ih);
            c->cnt.n_hist++;
            ora_shower(c);
            if (c->err) { free(endep); c->endep = NULL; return OMC_EOVERFLOW; }
        }
        for (int i = 0; i < nreg; ++i) {
            double e = endep[i];
            sum[i] += e;
            sum2[i] += e * e;
        }
        memset(endep, 0, nreg * sizeof(double));
    }
    free(endep);
    c->endep = NULL;
    return OMC_OK;
}

int ora_beamlet_primary(ora_ctx *c, const omc_source_beamlets *src, int ibeamlet, int64_t hid, ora_particle *out)
{
    if (ibeamlet < 0 || ibeamlet >= src->nbeamlets) return OMC_ERANGE;
    init_history_beamlet(c, src, ibeamlet, hid);
    *out = c->S[0];
    return OMC_OK;
}

int ora_run_dosxyz_parallel(const omc_tables *t, const omc_regions *r, const omc_phantom *g,
                            const omc_source_dosxyz *src, uint32_t s0, uint32_t s1, int64_t hist0,
                            int nperbatch, int nbatch, int nthreads, double *sum, double *sum2,
                            ora_counters *cnt)
{
    int nreg = r->nreg;
    double *batch = calloc((size_t)nbatch * nreg, sizeof(double));
    if (!batch) return OMC_ENOMEM;
    int rc = OMC_OK;
    ora_counters tot;
    memset(&tot, 0, sizeof tot);
#ifdef _OPENMP
    if (nthreads < 1) nthreads = omp_get_num_procs();
#pragma omp parallel num_threads(nthreads)
#endif
    {
        ora_ctx *c = ora_create(t, r, g, 1, ORA_RNG_PHILOX);
        ora_seed_philox(c, s0, s1);
        double *mine = calloc((size_t)nbatch * nreg, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t h = 0; h < (int64_t)nbatch * nperbatch; ++h) {
            int b = (int)(h / nperbatch);
            c->endep = mine + (size_t)b * nreg;
            init_history_dosxyz(c, src, hist0 + h);
            c->cnt.n_hist++;
            ora_shower(c);
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (size_t i = 0; i < (size_t)nbatch * nreg; ++i) batch[i] += mine[i];
            if (c->err) rc = OMC_EOVERFLOW;
            {
                /* every field is an int64_t count */
                int64_t *a = (int64_t *)&tot;
                const int64_t *b = (const int64_t *)&c->cnt;
                for (size_t k = 0; k < sizeof tot / sizeof(int64_t); ++k) a[k] += b[k];
            }
        }
        free(mine);
        ora_destroy(c);
    }
    for (int b = 0; b < nbatch; ++b)
        for (int i = 0; i < nreg; ++i) {
            double e = batch[(size_t)b * nreg + i];
            sum[i] += e;
            sum2[i] += e * e;
        }
    free(batch);
    if (cnt) *cnt = tot;
    return rc;
}

/* ---- per-function entry points ---------------------------------------- */

double ora_msdist(ora_ctx *c, const ora_particle *p, int imed, double rhof, double de,
                  double tustep, double eke, ora_particle *out)
{
    *out = *p;
    return msdist_core(c, imed, p->iq, rhof, de, tustep, eke, p->x, p->y, p->z, p->u, p->v, p->w,
                       &out->x, &out->y, &out->z, &out->u, &out->v, &out->w);
}

int ora_interaction(ora_ctx *c, int kind, const ora_particle *p, ora_particle *out)
{
    c->S[0] = *p;
    c->np = 0;
    stream_enter(c, 0);
    int imed = c->r->med[p->ir];
    switch (kind) {
    case 0: ora_compton(c); break;
    case 1: ora_pair(c, imed); break;
    case 2: ora_photo(c); break;
    case 3: {
        double gle = log(p->e);
        int lgle = pwlf_i(imed, gle, c->t->ge1, c->t->ge0) - 1;
        ora_rayleigh(c, imed, p->e, gle, lgle);
        break;
    }
    case 4: ora_brems(c); break;
    case 5: ora_moller(c); break;
    case 6: ora_bhabha(c); break;
    case 7: ora_annih(c); break;
    case 8: ora_rannih(c); break;
    default: return -1;
    }
    int n = c->np + 1;
    for (int i = 0; i < n; ++i) out[i] = c->S[i];
    return n;
}

int ora_shower_one(ora_ctx *c, const ora_particle *primary, double *endep)
{
    c->S[0] = *primary;
    c->np = 0;
    c->endep = endep;
    ora_shower(c);
    c->endep = NULL;
    return c->err ? OMC_EOVERFLOW : OMC_OK;
}
