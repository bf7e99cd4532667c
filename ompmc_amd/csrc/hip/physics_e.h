// Device physics of charged particles: CSDA range and energy loss,
// multiple and single elastic scattering with spin rejection, and the
// discrete e-/e+ interactions.  Same operation order as the ompMC routines
// named beside each function.  Range / energy-loss inversion is FP64 in both
// precision modes.
#pragma once

#include "physics.h"

namespace omc {

constexpr double EPSEMFP = 1.0E-5;
constexpr double LAMBMAX_MS = 1.0E5;
constexpr double QMIN_MS = 1.0E-3;
constexpr double QMAX_MS = 0.5;

// computeDrange (ompmc.c:3979-4014): path length from ekei down to ekef,
// both in bin lelke of medium imed / charge qel.
__device__ __forceinline__ double drange(const ElecCsda *cs, int iq, double ekei, double ekef,
                                         double elkei, double elkef)
{
    (void)iq;
    double fedep = 1.0 - ekef / ekei;
    double elktmp = 0.5 * (elkei + elkef + 0.25 * fedep * fedep * (1.0 + fedep * (1.0 + 0.875 * fedep)));
    double dedxmid = elktmp * cs->dedx1 + cs->dedx0;
    dedxmid = 1.0 / dedxmid;
    double aux = cs->dedx1 * dedxmid;
    aux = aux * (1.0 + 2.0 * aux) * fedep * fedep / (6.0 * (2.0 - fedep) * (2.0 - fedep));
    return fedep * ekei * dedxmid * (1.0 + aux);
}

// computeEloss (ompmc.c:4016-4108); `bins` points at bin 0 of (imed, qel)
__device__ __forceinline__ double eloss(const ElecCsda *bins, double te, double eke0, double eke1,
                                        double rhof, double tustep, double range, double eke,
                                        double elke, int lelke)
{
    double aux, dedxmid, de, fedep;
    double tuss = range - bins[lelke].range / rhof;
    if (tuss >= tustep) {
        const ElecCsda &c = bins[lelke];
        dedxmid = elke * c.dedx1 + c.dedx0;
        aux = c.dedx1 / dedxmid;
        de = dedxmid * tustep * rhof;
        fedep = de / eke;
        de *= (1.0 - 0.5 * fedep * aux * (1.0 - 0.333333 * fedep * (aux - 1.0 - 0.25 * fedep * (2.0 - aux * (4.0 - aux)))));
    } else {
        int lt = lelke;
        tuss = (range - tustep) * rhof;
        if (tuss <= 0) {
            de = eke - te * 0.99;
        } else {
            while (lt > 0 && tuss < bins[lt].range) lt -= 1;
            double elktmp = (lt + 2 - eke0) / eke1;
            double eketmp = bins[lt + 1].earr;
            tuss = (bins[lt + 1].range - tuss) / rhof;
            const ElecCsda &c = bins[lt];
            dedxmid = elktmp * c.dedx1 + c.dedx0;
            aux = c.dedx1 / dedxmid;
            de = dedxmid * tuss * rhof;
            fedep = de / eketmp;
            de *= (1.0 - 0.5 * fedep * aux * (1.0 - 0.333333 * fedep * (aux - 1.0 - 0.25 * fedep * (2.0 - aux * (4.0 - aux)))));
            de += eke - eketmp;
        }
    }
    return de;
}

struct SpinIdx { int i, j; };

// spinRejection (ompmc.c:3097-3168)
template <typename R>
__device__ __forceinline__ R spin_rejection(const DevTables<R> &T, Rng &rng, int imed, int qel, R elke,
                                            R beta2, R q1, R cost, bool &spin_index, bool is_single,
                                            SpinIdx &si)
{
    const SmallTables &S = *T.small;
    if (spin_index) {
        spin_index = false;
        R ai;
        if (beta2 >= (R)S.b2spin_min) {
            ai = (beta2 - (R)S.b2spin_min) * (R)S.dbeta2i;
            si.i = (int)ai;
            ai -= (R)si.i;
            si.i += OMC_MXE_SPIN + 1;
        } else if (elke > (R)S.espml) {
            ai = (elke - (R)S.espml) * (R)S.dleneri;
            si.i = (int)ai;
            ai -= (R)si.i;
        } else {
            si.i = 0;
            ai = (R)-1.0;
        }
        R rnno = urand<R>(rng);
        if (rnno < ai) si.i += 1;
        if (is_single) {
            si.j = 0;
        } else {
            R qq1 = (R)2.0 * q1;
            qq1 = qq1 / ((R)1.0 + qq1);
            R aj = qq1 * (R)S.dqq1i;
            si.j = (int)aj;
            if (si.j >= 15) {
                si.j = 15;
            } else {
                aj -= (R)si.j;
                rnno = urand<R>(rng);
                if (rnno < aj) si.j += 1;
            }
        }
    }
    R xi = m_sqrt((R)0.5 * ((R)1.0 - cost));
    R ak = xi * (R)31;
    int k = (int)ak;
    ak -= (R)k;
    const R *row = T.spin + (size_t)imed * OMC_SPIN_PER_MED + qel * (OMC_SPIN_PER_MED / 2) + si.i * 512 + si.j * 32;
    return ((R)1.0 - ak) * row[k] + ak * row[k + 1];
}

// sscat (ompmc.c:3170-3199): one elastic scattering, spin rejection on
template <typename R>
__device__ __forceinline__ void sscat(const DevTables<R> &T, Rng &rng, int imed, int qel, R chia2, R elke, R beta2,
                      R &cost, R &sint)
{
    bool spin_index = true;
    SpinIdx si = {0, 0};
    R xi, rejf, rnno;
    do {
        xi = urand<R>(rng);
        xi = (R)2.0 * chia2 * xi / ((R)1.0 - xi + chia2);
        cost = (R)1.0 - xi;
        rejf = spin_rejection<R>(T, rng, imed, qel, elke, beta2, (R)0, cost, spin_index, true, si);
        rnno = urand<R>(rng);
    } while (rnno > rejf);
    sint = m_sqrt(xi * ((R)2.0 - xi));
}

struct MsIdx { int i, j; };

// mscat (ompmc.c:3606-3785): angle after lambda elastic mean free paths
template <typename R>
__device__ void mscat(const DevTables<R> &T, Rng &rng, int imed, int qel, bool &spin_index,
                      bool &find_index, R elke, R beta2, R q1, R lambda, R chia2, R &cost, R &sint,
                      MsIdx &mi, R &omega2, SpinIdx &si)
{
    const SmallTables &S = *T.small;
    R xi, rejf, rnno;
    R explambda = m_exp(-lambda);
    if (lambda <= (R)13.8) {
        R sprob = urand<R>(rng);
        if (sprob < explambda) { cost = (R)1.0; sint = (R)0.0; return; }
        R wsum = ((R)1.0 + lambda) * explambda;
        if (sprob < wsum) {
            do {
                xi = urand<R>(rng);
                xi = (R)2.0 * chia2 * xi / ((R)1.0 - xi + chia2);
                cost = (R)1.0 - xi;
                rejf = spin_rejection<R>(T, rng, imed, qel, elke, beta2, q1, cost, spin_index, false, si);
                rnno = urand<R>(rng);
            } while (rnno > rejf);
            sint = m_sqrt(xi * ((R)2.0 - xi));
            return;
        }
        if (lambda <= (R)1) {
            int icount = 0;
            R wprob = explambda, sinz, cosz, phi;
            wsum = explambda;
            cost = (R)1.0;
            sint = (R)0.0;
            do {
                icount += 1;
                if (icount > 20) break;
                wprob = wprob * lambda / (R)icount;
                wsum = wsum + wprob;
                do {
                    xi = urand<R>(rng);
                    xi = (R)2.0 * chia2 * xi / ((R)1.0 - xi + chia2);
                    cosz = (R)1.0 - xi;
                    rejf = spin_rejection<R>(T, rng, imed, qel, elke, beta2, q1, cosz, spin_index, false, si);
                    rnno = urand<R>(rng);
                } while (rnno > rejf);
                sinz = xi * ((R)2.0 - xi);
                if (sinz > (R)1.0E-20) {
                    sinz = m_sqrt(sinz);
                    xi = urand<R>(rng);
                    phi = xi * (R)6.2831853;
                    cost = cost * cosz - sint * sinz * m_cos(phi);
                    sint = m_sqrt(m_fmax((R)0.0, ((R)1.0 - cost) * ((R)1.0 + cost)));
                }
            } while (wsum <= sprob);
            return;
        }
    }
    if (lambda <= (R)LAMBMAX_MS) {
        R llmbda = m_log(lambda);
        if (find_index) {
            R ai = llmbda * (R)S.dllambi;
            mi.i = (int)ai;
            ai -= (R)mi.i;
            xi = urand<R>(rng);
            if (xi < ai) mi.i += 1;
            if (mi.i > 63) mi.i = 63;    // the reference reads past its table here
            if (q1 < (R)QMIN_MS) {
                mi.j = 0;
            } else if (q1 < (R)QMAX_MS) {
                R aj = q1 * (R)S.dqmsi;
                mi.j = (int)aj;
                aj -= (R)mi.j;
                xi = urand<R>(rng);
                if (xi < aj) mi.j += 1;
            } else {
                mi.j = 7;
            }
            if (llmbda < (R)2.2299)
                omega2 = chia2 * (lambda + (R)4.0) *
                         ((R)1.347006 + llmbda * ((R)0.209364 - llmbda * ((R)0.45525 - llmbda * ((R)0.50142 - (R)0.081234 * llmbda))));
            else
                omega2 = chia2 * (lambda + (R)4.0) *
                         ((R)-2.77164 + llmbda * ((R)2.94874 - llmbda * ((R)0.1535754 - llmbda * (R)0.00552888)));
            find_index = false;
        }
        const MsAlias<R> *row = T.ms + mi.i * 256 + mi.j * 32;
        do {
            xi = urand<R>(rng);
            R ak = xi * (R)31;
            int k = (int)ak;
            ak -= (R)k;
            const MsAlias<R> me = row[k];
            // node k, or the node it aliases when ak is above its weight; the
            // reference's interpolation step has du = ums[k] - ums[k] = 0, so
            // u is the node value; its random number is still drawn
            R u = ak > me.wms ? me.ualias : me.ums;
            (void)urand<R>(rng);
            xi = omega2 * u / ((R)1.0 + (R)0.5 * omega2 - u);
            if (xi > (R)1.99999) xi = (R)1.99999;
            cost = (R)1.0 - xi;
            rejf = spin_rejection<R>(T, rng, imed, qel, elke, beta2, q1, cost, spin_index, false, si);
            rnno = urand<R>(rng);
        } while (rnno > rejf);
        sint = m_sqrt(xi * ((R)2.0 - xi));
    }
}

// msdist (ompmc.c:3787-3976): PRESTA-II substeps; moves p to the step end
// position and direction, returns the straight-line distance.
template <typename R>
__device__ __forceinline__ R msdist(const DevTables<R> &T, Particle<R> &p, int imed, R rhof, R de, R tustep, R eke)
{
    const MediumScalars &M = T.small->med[imed];
    const int qel = (1 + p.iq) / 2;
    const R rm = (R)RM;
    R blccc = (R)M.blcc, xcccc = (R)M.xcc;
    R e = eke - (R)0.5 * de;
    R tau = e / rm;
    R tau2 = sq(tau);
    R epsilon = de / eke;
    R epsilonp = de / e;
    e *= ((R)1.0 - sq(epsilonp) * ((R)6.0 + (R)10.0 * tau + (R)5.0 * tau2) / ((R)24.0 * tau2 + (R)72.0 * tau + (R)48.0));
    R p2 = e * (e + (R)2.0 * rm);
    R beta2 = p2 / (p2 + sq(rm));
    R chia2 = xcccc / ((R)4.0 * p2 * blccc);
    R lambda = (R)0.5 * tustep * rhof * blccc / beta2;
    R temp2 = (R)0.166666 * ((R)4.0 + tau * ((R)6.0 + tau * ((R)7.0 + tau * ((R)4.0 + tau)))) *
              (epsilonp / ((tau + (R)1.0) * (tau + (R)2.0))) * (epsilonp / ((tau + (R)1.0) * (tau + (R)2.0)));
    lambda *= ((R)1.0 - temp2);
    R elke = m_log(e);
    int lelke = pwlf_idx(elke, M.eke1, M.eke0) - 1;
    if (lelke < 0) {
        lelke = 0;
        elke = ((R)1.0 - (R)M.eke0) / (R)M.eke1;
    }
    const ElecBin<R> &eb = T.el[(imed * 2 + qel) * MXEKE + lelke];
    R etap = elke * eb.eta1 + eb.eta0;
    R xi_corr = elke * eb.q1c1 + eb.q1c0;
    R gamma = elke * eb.q2c1 + eb.q2c0;
    R ms_corr = elke * eb.blc1 + eb.blc0;
    chia2 *= etap;
    lambda /= (etap * ((R)1.0 + chia2));
    lambda *= ms_corr;
    R chilog = m_log((R)1.0 + (R)1.0 / chia2);
    R q1 = (R)2.0 * chia2 * (chilog * ((R)1.0 + chia2) - (R)1.0);
    gamma = (R)6.0 * chia2 * ((R)1.0 + chia2) * (chilog * ((R)1.0 + (R)2.0 * chia2) - (R)2.0) / q1 * gamma;
    R xi = q1 * lambda;

    bool find_index = true, spin_index = true;
    MsIdx mi = {0, 0};
    SpinIdx si = {0, 0};
    R omega2 = 0;
    R w1, sint1, cphi1, sphi1, w2, sint2, cphi2, sphi2;
    mscat<R>(T, p.rng, imed, qel, spin_index, find_index, elke, beta2, xi, lambda, chia2, w1, sint1, mi, omega2, si);
    azimuth<R>(p.rng, cphi1, sphi1);
    mscat<R>(T, p.rng, imed, qel, spin_index, find_index, elke, beta2, xi, lambda, chia2, w2, sint2, mi, omega2, si);
    azimuth<R>(p.rng, cphi2, sphi2);

    R u2 = sint2 * cphi2;
    R v2 = sint2 * sphi2;
    R u2p = w1 * u2 + sint1 * w2;
    R us = u2p * cphi1 - v2 * sphi1;
    R vs = u2p * sphi1 + v2 * cphi1;
    R ws = w1 * w2 - sint1 * u2;

    xi *= (R)2 * xi_corr;
    R eta = urand<R>(p.rng);
    R eta1 = (R)0.5 * ((R)1.0 - eta);
    R delta = (R)0.9082483 - ((R)0.1020621 - (R)0.0263747 * gamma) * xi;
    R temp1 = (R)2.0 + tau;
    R temp = ((R)2.0 + tau * temp1) / ((tau + (R)1.0) * temp1);
    temp -= (tau + (R)1.0) / ((tau + (R)2.0) * (chilog * ((R)1.0 + chia2) - (R)1.0));
    temp *= epsilonp;
    temp1 = (R)1.0 - temp;
    delta += (R)0.40824829 * (epsilon * (tau + (R)1.0) /
                              ((tau + (R)2.0) * (chilog * ((R)1.0 + chia2) - (R)1.0) * (chilog * ((R)1.0 + (R)2.0 * chia2) - (R)2.0)) -
                              (R)0.25 * sq(temp));
    R b = eta * delta;
    R c = eta * ((R)1.0 - delta);
    R w1v2 = w1 * v2;
    R ut = b * sint1 * cphi1 + c * (cphi1 * u2 - sphi1 * w1v2) + eta1 * us * temp1;
    R vt = b * sint1 * sphi1 + c * (sphi1 * u2 + cphi1 * w1v2) + eta1 * vs * temp1;
    R wt = eta1 * ((R)1.0 + temp) + b * w1 + c * w2 + eta1 * ws * temp1;
    R ustep = tustep * m_sqrt(ut * ut + vt * vt + wt * wt);

    R u0 = p.u, v0 = p.v, w0 = p.w;
    R sint02 = u0 * u0 + v0 * v0;
    if (sint02 > (R)1.0E-20) {
        R sint0 = m_sqrt(sint02);
        R sint0i = (R)1.0 / sint0;
        R cphi0 = sint0i * u0, sphi0 = sint0i * v0;
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
    p.x = p.x + tustep * ut;
    p.y = p.y + tustep * vt;
    p.z = p.z + tustep * wt;
    p.u = us;
    p.v = vs;
    p.w = ws;
    return ustep;
}

// rannih (ompmc.c:4111-4167): positron at rest -> a and b are the two photons
template <typename R>
__device__ __forceinline__ void rannih(Particle<R> &a, Particle<R> &b)
{
    R rnno = urand<R>(a.rng);
    R costhe = (R)2.0 * rnno - (R)1;
    R sinthe = m_sqrt(m_fmax((R)0.0, ((R)1.0 - costhe) * ((R)1.0 + costhe)));
    (void)urand<R>(a.rng);     // drawn and unused in the reference
    R cphi, sphi;
    azimuth<R>(a.rng, cphi, sphi);
    a.e = (R)RM;
    a.iq = 0;
    a.u = sinthe * cphi;
    a.v = sinthe * sphi;
    a.w = costhe;
    b.e = (R)RM;
    b.iq = 0;
    b.x = a.x; b.y = a.y; b.z = a.z; b.ir = a.ir; b.wt = a.wt;
    b.u = (R)-1.0 * a.u;
    b.v = (R)-1.0 * a.v;
    b.w = (R)-1.0 * a.w;
    spawn<R>(a, b);
}

// brems (ompmc.c:4170-4356): a = electron (reduced energy), b = photon
template <typename R>
__device__ void brems(const DevTables<R> &T, Particle<R> &a, Particle<R> &b, int imed)
{
    const MediumScalars &M = T.small->med[imed];
    const R rm = (R)RM;
    R eie = a.e, phi1, phi2;
    int l = (eie < (R)50.0) ? 1 : 3;
    int l1 = l + 1;
    R ekin = eie - rm;
    R brmin = (R)M.ap / ekin;
    R waux = -m_log(brmin);
    R A = a.u, B = a.v, Cc = a.w;
    R sinpsi = A * A + B * B, sindel = 0, cosdel = 0;
    if (sinpsi > (R)1.0E-20) {
        sinpsi = m_sqrt(sinpsi);
        sindel = B / sinpsi;
        cosdel = A / sinpsi;
    }
    R ztarg = (R)M.zbrang;
    R tteie = eie / rm;
    R beta = m_sqrt((tteie - (R)1.0) * (tteie + (R)1.0)) / tteie;
    R y2max = (R)2.0 * beta * ((R)1.0 + beta) * tteie * tteie;
    R y2maxi = (R)1.0 / y2max;
    R z2max = y2max + (R)1.0;
    R z2maxi = m_sqrt(z2max);
    R aux, br, delta, r6, r7, rejf, ese, esg;
    do {
        r6 = urand<R>(a.rng);
        r7 = urand<R>(a.rng);
        br = brmin * m_exp(r6 * waux);
        esg = ekin * br;
        ese = eie - esg;
        delta = esg / eie / ese * (R)M.delcm;
        aux = ese / eie;
        if (delta < (R)1.0) {
            phi1 = (R)M.dl1[l - 1] + delta * ((R)M.dl2[l - 1] + delta * (R)M.dl3[l - 1]);
            phi2 = (R)M.dl1[l1 - 1] + delta * ((R)M.dl2[l1 - 1] + delta * (R)M.dl3[l1 - 1]);
        } else {
            phi1 = (R)M.dl4[l - 1] + (R)M.dl5[l - 1] * m_log(delta + (R)M.dl6[l - 1]);
            phi2 = phi1;
        }
        rejf = ((R)1.0 + sq(aux)) * phi1 - (R)2.0 * aux * phi2 / (R)3.0;
    } while (r7 >= rejf);

    b.e = esg;
    b.iq = 0;
    b.x = a.x; b.y = a.y; b.z = a.z; b.ir = a.ir; b.wt = a.wt;

    R y2tst = 0;
    R ttese = ese / rm;
    R esedei = ttese / tteie;
    R rjarg1 = (R)1.0 + esedei * esedei;
    R rjarg2 = rjarg1 + (R)2.0 * esedei;
    R rjarg3;
    R rtest = (R)1.0, rejtst = (R)0.0;
    aux = (R)2.0 * ese * tteie / esg;
    aux = aux * aux;
    R aux1 = aux * ztarg;
    if (aux1 > (R)10.0) rjarg3 = -m_log(ztarg) + ((R)1.0 - aux1) / sq(aux1);
    else rjarg3 = m_log(aux / ((R)1.0 + aux1));
    R rejmax = rjarg1 * rjarg3 - rjarg2;
    while (rtest >= rejtst) {
        y2tst = urand<R>(a.rng);
        rtest = urand<R>(a.rng);
        R aux3 = z2maxi / (y2tst + ((R)1.0 - y2tst) * z2maxi);
        rtest = rtest * aux3 * rejmax;
        y2tst = sq(aux3) - (R)1.0;
        R aux3_4 = m_pow4(aux3);
        R y2tst1 = esedei * y2tst / aux3_4;
        R aux4 = (R)16.0 * y2tst1 - rjarg2;
        R aux5 = rjarg1 - (R)4.0 * y2tst1;
        if (rtest < aux4 + aux5 * rjarg3) break;
        R aux2 = m_log(aux / ((R)1.0 + aux1 / aux3_4));
        rejtst = aux4 + aux5 * aux2;
    }
    R costhe = (R)1.0 - (R)2.0 * y2tst * y2maxi;
    R sinthe = m_sqrt(m_fmax((R)0.0, ((R)1.0 - sq(costhe))));
    R cphi, sphi;
    azimuth<R>(a.rng, cphi, sphi);
    if (sinpsi >= (R)1.0E-10) {
        R us = sinthe * cphi, vs = sinthe * sphi;
        b.u = Cc * cosdel * us - sindel * vs + A * costhe;
        b.v = Cc * sindel * us + cosdel * vs + B * costhe;
        b.w = Cc * costhe - sinpsi * us;
    } else {
        b.u = sinthe * cphi;
        b.v = sinthe * sphi;
        b.w = Cc * costhe;
    }
    a.e = ese;
    spawn<R>(a, b);
}

// moller (ompmc.c:4359-4435).  Returns false when below the threshold (no
// interaction, the electron simply re-enters transport).
template <typename R>
__device__ bool moller(const DevTables<R> &T, Particle<R> &a, Particle<R> &b, int imed)
{
    const MediumScalars &M = T.small->med[imed];
    const R rm = (R)RM;
    R eie = a.e, ekin = eie - rm;
    if (ekin <= (R)2.0 * (R)M.te) return false;
    R t0 = ekin / rm, e0 = t0 + (R)1.0;
    R extrae = eie - (R)M.thmoll;
    R g2 = sq(t0) / sq(e0);
    R g3 = ((R)2.0 * t0 + (R)1.0) / sq(e0);
    R gmax = ((R)1.0 + (R)1.25 * g2), br, rejf4, r, r27, r28;
    do {
        r27 = urand<R>(a.rng);
        br = (R)M.te / (ekin - extrae * r27);
        r = br / ((R)1.0 - br);
        r28 = urand<R>(a.rng);
        rejf4 = ((R)1.0 + g2 * sq(br) + r * (r - g3));
        r28 *= gmax;
    } while (r28 > rejf4);
    R ekse2 = br * ekin, ese1 = eie - ekse2, ese2 = ekse2 + rm;
    a.e = ese1;
    b.e = ese2;
    R h1 = (eie + rm) / ekin;
    R costh = h1 * (ese1 - rm) / (ese1 + rm);
    R sinthe = m_sqrt((R)1.0 - costh), costhe = m_sqrt(costh);
    Uphi<R> h;
    uphi21<R>(a, h, costhe, sinthe);
    b.iq = -1;
    costh = h1 * (ese2 - rm) / (ese2 + rm);
    sinthe = -m_sqrt((R)1.0 - costh);
    costhe = m_sqrt(costh);
    uphi32<R>(a, b, h, costhe, sinthe);
    spawn<R>(a, b);
    return true;
}

// bhabha (ompmc.c:4438-4525): slot a keeps the higher-energy particle
template <typename R>
__device__ void bhabha(const DevTables<R> &T, Particle<R> &a, Particle<R> &b, int imed)
{
    const MediumScalars &M = T.small->med[imed];
    const R rm = (R)RM;
    R eip = a.e, ekin = eip - rm;
    R t0 = ekin / rm, e0 = t0 + (R)1.0;
    R yy = (R)1.0 / (t0 + (R)2.0);
    R beta2 = (sq(e0) - (R)1.0) / sq(e0);
    R ep0 = (R)M.te / ekin, ep0c = (R)1.0 - ep0;
    R yp = (R)1.0 - (R)2.0 * yy;
    R b4 = m_pow3(yp);
    R b3 = b4 + sq(yp);
    R b2 = yp * ((R)3.0 + sq(yy));
    R b1 = (R)2.0 - sq(yy);
    R r3, r4, br, rejf2;
    do {
        r3 = urand<R>(a.rng);
        br = ep0 / ((R)1.0 - ep0c * r3);
        r4 = urand<R>(a.rng);
        rejf2 = ((R)1.0 - beta2 * br * (b1 - br * (b2 - br * (b3 - br * b4))));
    } while (r4 > rejf2);
    if (br < (R)0.5) {
        b.iq = -1;
    } else {
        a.iq = -1;
        b.iq = 1;
        br = (R)1.0 - br;
    }
    br = m_fmax(br, (R)0.0);
    R ekse2 = br * ekin, ese1 = eip - ekse2, ese2 = ekse2 + rm;
    a.e = ese1;
    b.e = ese2;
    R h1 = (eip + rm) / ekin;
    R costh = m_fmin((R)1.0, h1 * (ese1 - rm) / (ese1 + rm));
    R sinthe = m_sqrt((R)1.0 - costh), costhe = m_sqrt(costh);
    Uphi<R> h;
    uphi21<R>(a, h, costhe, sinthe);
    costh = h1 * (ese2 - rm) / (ese2 + rm);
    sinthe = -m_sqrt((R)1.0 - costh);
    costhe = m_sqrt(costh);
    uphi32<R>(a, b, h, costhe, sinthe);
    spawn<R>(a, b);
}

// annih (ompmc.c:4528-4645): positron in flight -> two photons a, b
template <typename R>
__device__ void annih(Particle<R> &a, Particle<R> &b)
{
    const R rm = (R)RM;
    R avip = a.e + rm;
    R A = avip / rm;
    R g = A - (R)1.0, t = g - (R)1.0, p = m_sqrt(A * t);
    R pot = p / t;
    R ep0 = (R)1.0 / (A + p);
    R wsamp = m_log(((R)1.0 - ep0) / ep0);
    R aa = a.u, bb = a.v, cc = a.w;
    R sinpsi = sq(aa) + sq(bb), sindel = 0, cosdel = 0;
    if (sinpsi > (R)1.0E-20) {
        sinpsi = m_sqrt(sinpsi);
        sindel = bb / sinpsi;
        cosdel = aa / sinpsi;
    }
    R ep, rejf, r1, r2;
    do {
        r1 = urand<R>(a.rng);
        ep = ep0 * m_exp(r1 * wsamp);
        r2 = urand<R>(a.rng);
        rejf = (R)1.0 - sq(ep * A - (R)1.0) / (ep * (sq(A) - (R)2.0));
    } while (r2 > rejf);
    R esg1 = avip * ep;
    a.e = esg1;
    a.iq = 0;
    R costhe = m_fmin((R)1.0, (esg1 - rm) * pot / esg1);
    R sinthe = m_sqrt((R)1.0 - sq(costhe));
    R sphi, cphi;
    azimuth<R>(a.rng, cphi, sphi);
    R us, vs;
    if (sinpsi >= (R)1.0E-10) {
        us = sinthe * cphi;
        vs = sinthe * sphi;
        a.u = cc * cosdel * us - sindel * vs + aa * costhe;
        a.v = cc * sindel * us + cosdel * vs + bb * costhe;
        a.w = cc * costhe - sinpsi * us;
    } else {
        a.u = sinthe * cphi;
        a.v = sinthe * sphi;
        a.w = cc * costhe;
    }
    R esg2 = avip - esg1;
    b.e = esg2;
    b.iq = 0;
    b.x = a.x; b.y = a.y; b.z = a.z; b.ir = a.ir; b.wt = a.wt;
    costhe = m_fmin((R)1.0, (esg2 - rm) * pot / esg2);
    sinthe = -m_sqrt((R)1.0 - sq(costhe));
    if (sinpsi >= (R)1.0E-10) {
        us = sinthe * cphi;
        vs = sinthe * sphi;
        b.u = cc * cosdel * us - sindel * vs + aa * costhe;
        b.v = cc * sindel * us + cosdel * vs + bb * costhe;
        b.w = cc * costhe - sinpsi * us;
    } else {
        b.u = sinthe * cphi;
        b.v = sinthe * sphi;
        b.w = cc * costhe;
    }
    spawn<R>(a, b);
}

}  // namespace omc
