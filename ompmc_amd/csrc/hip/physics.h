// Device physics of the shower() path for one particle (one lane).
//
// Each function is the device form of the ompMC routine named beside it, in
// the same floating-point operation order (the FP64 build disables FMA
// contraction), so that with the same random numbers a lane follows the
// same history as the reference algorithm.  Interactions write the
// particle that continues in the stack slot the reference keeps (`a`) and
// the secondary it pushes on top (`b`).
#pragma once

#include "common.h"

namespace omc {

// ------------------------------------------------------------ math ------
__device__ __forceinline__ double m_log(double x) { return log(x); }
__device__ __forceinline__ float  m_log(float x)  { return __logf(x); }
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ float  m_exp(float x)  { return __expf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float  m_sqrt(float x)  { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double m_sin(double x) { return sin(x); }
__device__ __forceinline__ float  m_sin(float x)  { return __sinf(x); }
__device__ __forceinline__ double m_cos(double x) { return cos(x); }
__device__ __forceinline__ float  m_cos(float x)  { return __cosf(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float  m_pow(float x, float y)  { return __powf(x, y); }
// x^3 and x^4 where the reference calls pow(x, 3.0) and pow(x, 4.0): the FP64
// engine keeps pow, whose rounding is the reference's; the FP32 engine
// multiplies, which is more accurate than __powf and a few instructions
// instead of an inlined pow inside a rejection loop.
__device__ __forceinline__ double m_pow3(double x) { return pow(x, 3.0); }
__device__ __forceinline__ float  m_pow3(float x)  { return x * x * x; }
__device__ __forceinline__ double m_pow4(double x) { return pow(x, 4.0); }
__device__ __forceinline__ float  m_pow4(float x)  { const float x2 = x * x; return x2 * x2; }
__device__ __forceinline__ double m_fmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float  m_fmax(float a, float b)  { return fmaxf(a, b); }
__device__ __forceinline__ double m_fmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float  m_fmin(float a, float b)  { return fminf(a, b); }
__device__ __forceinline__ double m_fabs(double a) { return fabs(a); }
__device__ __forceinline__ float  m_fabs(float a)  { return fabsf(a); }

template <typename R> __device__ __forceinline__ R sq(R x) { return x * x; }

// pwlfInterval / pwlfEval (ompmc.c:203-211)
template <typename R, typename T>
__device__ __forceinline__ int pwlf_idx(R l, T c1, T c0) { return (int)(l * (R)c1 + (R)c0); }

// Table bin, the reference's pwlfInterval - 1 (ompmc.c:1954, 4729), clamped to
// [0, nbins - 1].  l*c1 and c0 nearly cancel at a table's lower edge (c0 is
// about 1070 for the photon tables), so in FP32 an energy within ~5e-7
// (relative) above the edge -- a photon just above pcut, an electron just
// above ecut - rm -- rounds to bin -1 and would read the preceding medium's
// row, or memory before the table for medium 0.  The exact index there is
// bin 0, which the clamp returns.
template <typename R, typename T>
__device__ __forceinline__ int pwlf_bin(R l, T c1, T c0, int nbins)
{
    int i = (int)(l * (R)c1 + (R)c0) - 1;
    return min(max(i, 0), nbins - 1);
}

// --------------------------------------------------------- geometry -----
// n / d for 0 <= n < 2^24, d > 0 without an integer divide: the float
// quotient through the hardware reciprocal is off by at most a few units
// for the smallest d (one for d > 4), and the integer correction loops
// make it exact.
__device__ __forceinline__ int small_div(int n, int d)
{
    int q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    int r = n - q * d;
    while (r < 0) { q -= 1; r += d; }
    while (r >= d) { q += 1; r -= d; }
    return q;
}

// Voxel indices of region irl > 0 (region = 1 + ix + iy*imax + iz*ijmax),
// the decomposition omc_dosxyz.c:205-207 computes with % and /.
__device__ __forceinline__ void voxel_of(int irl, int imax, int ijmax, int &irx, int &iry, int &irz)
{
    int n = irl - 1;
    irz = small_div(n, ijmax);
    int rem = n - irz * ijmax;
    iry = small_div(rem, imax);
    irx = rem - iry * imax;
}
// howfar (omc_dosxyz.c:187-297) for a particle inside region irl > 0.  The
// boundary each axis heads for is loaded up front (the three loads are in
// flight together; an axis with zero direction cosine loads one it does not
// use, always inside the array), then the axes are tested in the
// reference's order z, x, y.
template <typename R, typename Geo>
__device__ __forceinline__ void howfar(const Geo &g, int irl, R x, R y, R z, R u, R v, R w,
                                       int &irnew, R &ustep)
{
    int ijmax = g.isize * g.jsize, imax = g.isize;
    int irx, iry, irz;
    voxel_of(irl, imax, ijmax, irx, iry, irz);
    const R bz = (R)glob(g.zb)[irz + (w > (R)0 ? 1 : 0)];
    const R bx = (R)glob(g.xb)[irx + (u > (R)0 ? 1 : 0)];
    const R by = (R)glob(g.yb)[iry + (v > (R)0 ? 1 : 0)];
    R dist;
    if (w > (R)0) {
        dist = (bz - z) / w;
        if (dist < ustep) { ustep = dist; irnew = (irz != g.ksize - 1) ? irl + ijmax : 0; }
    } else if (w < (R)0) {
        dist = -(z - bz) / w;
        if (dist < ustep) { ustep = dist; irnew = (irz != 0) ? irl - ijmax : 0; }
    }
    if (u > (R)0) {
        dist = (bx - x) / u;
        if (dist < ustep) { ustep = dist; irnew = (irx != g.isize - 1) ? irl + 1 : 0; }
    } else if (u < (R)0) {
        dist = -(x - bx) / u;
        if (dist < ustep) { ustep = dist; irnew = (irx != 0) ? irl - 1 : 0; }
    }
    if (v > (R)0) {
        dist = (by - y) / v;
        if (dist < ustep) { ustep = dist; irnew = (iry != g.jsize - 1) ? irl + imax : 0; }
    } else if (v < (R)0) {
        dist = -(y - by) / v;
        if (dist < ustep) { ustep = dist; irnew = (iry != 0) ? irl - imax : 0; }
    }
}

// hownear (omc_dosxyz.c:299-334)
template <typename R, typename Geo>
__device__ __forceinline__ R hownear(const Geo &g, int irl, R x, R y, R z)
{
    if (irl == 0) return (R)0;
    int ijmax = g.isize * g.jsize, imax = g.isize;
    int irx, iry, irz;
    voxel_of(irl, imax, ijmax, irx, iry, irz);
    const gmem<const double> *xb = glob(g.xb), *yb = glob(g.yb), *zb = glob(g.zb);
    R t = (R)1.0E10;
    t = m_fmin(t, (R)xb[irx + 1] - x);
    t = m_fmin(t, x - (R)xb[irx]);
    t = m_fmin(t, (R)yb[iry + 1] - y);
    t = m_fmin(t, y - (R)yb[iry]);
    t = m_fmin(t, (R)zb[irz + 1] - z);
    t = m_fmin(t, z - (R)zb[irz]);
    return t;
}

// -------------------------------------------------------- rotations -----
template <typename R> struct Uphi { R A, B, C, cosphi, sinphi; };

// selectAzimuthalAngle (ompmc.c:101-121)
template <typename R>
__device__ __forceinline__ void azimuth(Rng &rng, R &cphi, R &sphi)
{
    R xphi, xphi2, yphi, yphi2, rhophi2;
    do {
        xphi = urand<R>(rng);
        xphi = (R)2.0 * xphi - (R)1.0;
        xphi2 = xphi * xphi;
        yphi = urand<R>(rng);
        yphi2 = yphi * yphi;
        rhophi2 = xphi2 + yphi2;
    } while (rhophi2 > (R)1.0);
    rhophi2 = (R)1 / rhophi2;
    cphi = (xphi2 - yphi2) * rhophi2;
    sphi = (R)2.0 * xphi * yphi * rhophi2;
}

template <typename R>
__device__ __forceinline__ void rotate(R &u, R &v, R &w, const Uphi<R> &h, R costhe, R sinthe)
{
    R sinps2 = h.A * h.A + h.B * h.B;
    if (sinps2 < (R)1.0E-20) {
        u = sinthe * h.cosphi;
        v = sinthe * h.sinphi;
        w = h.C * costhe;
    } else {
        R sinpsi = m_sqrt(sinps2);
        R us = sinthe * h.cosphi, vs = sinthe * h.sinphi;
        R sindel = h.B / sinpsi, cosdel = h.A / sinpsi;
        u = h.C * cosdel * us - sindel * vs + h.A * costhe;
        v = h.C * sindel * us + cosdel * vs + h.B * costhe;
        w = -sinpsi * us + h.C * costhe;
    }
}

// uphi21 (ompmc.c:127-164)
template <typename R>
__device__ __forceinline__ void uphi21(Particle<R> &p, Uphi<R> &h, R costhe, R sinthe)
{
    azimuth<R>(p.rng, h.cosphi, h.sinphi);
    h.A = p.u; h.B = p.v; h.C = p.w;
    rotate<R>(p.u, p.v, p.w, h, costhe, sinthe);
}

// uphi32 (ompmc.c:166-201): b inherits position, region and weight of a
template <typename R>
__device__ __forceinline__ void uphi32(const Particle<R> &a, Particle<R> &b, const Uphi<R> &h, R costhe, R sinthe)
{
    b.x = a.x; b.y = a.y; b.z = a.z;
    b.ir = a.ir; b.wt = a.wt;
    b.u = a.u; b.v = a.v; b.w = a.w;
    rotate<R>(b.u, b.v, b.w, h, costhe, sinthe);
}

// a new particle on top of the stack takes the identity of its parent's job
template <typename R>
__device__ __forceinline__ void spawn(const Particle<R> &parent, Particle<R> &b)
{
    b.job = parent.job;
    b.hist = parent.hist;
    parent.rng.child(b.rng.k0, b.rng.k1);
    b.rng.ctr = 0u;
    b.rng.reset_cache();
}

// -------------------------------------------------- photon interactions --
constexpr double HC_INVERSE = 80.65506856998;
constexpr double TWICE_HC2 = 0.000307444456;

// rayleigh (ompmc.c:1102-1145); form factors of medium 0 (reference quirk)
template <typename R>
__device__ void rayleigh(const DevTables<R> &T, Particle<R> &p, int imed, R eig, R gle, int lgle)
{
    const SmallTables &S = *T.small;
    R2<R> pm = T.ph_pmax[imed * MXGE + lgle];
    R pmax = gle * pm.c1 + pm.c0;
    R xmax = (R)HC_INVERSE * eig;
    R rnno0, rnno1, xv, costhe, csqthe;
    do {
        rnno1 = urand<R>(p.rng);
        do {
            rnno0 = urand<R>(p.rng);
            rnno0 *= pmax;
            int ib = S.i_array[0] - 1;          // ibin = (int)rnno0 * dwi = 0
            if ((S.i_array[1] - 1) > ib)
                while (rnno0 >= (R)S.fcum[ib + 1]) ib++;
            rnno0 = (rnno0 - (R)S.fcum[ib]) * (R)S.c_array[ib];
            xv = (R)S.xgrid[ib] * m_exp(m_log((R)1.0 + rnno0) * (R)S.b_array[ib]);
        } while (xv >= xmax);
        xv /= eig;
        costhe = (R)1.0 - (R)TWICE_HC2 * sq(xv);
        csqthe = sq(costhe);
    } while ((R)2.0 * rnno1 >= ((R)1.0 + csqthe));
    R sinthe = m_sqrt((R)1.0 - csqthe);
    Uphi<R> h;
    uphi21<R>(p, h, costhe, sinthe);
}

// setPairRejectionFunction (ompmc.c:1418-1427)
template <typename R>
__device__ __forceinline__ R pair_rej(R zbrang, R xi, R esedei, R eseder, R tteig)
{
    return (R)2.0 + (R)3.0 * (esedei + eseder) -
           (R)4.0 * (esedei + eseder + (R)1.0 - (R)4.0 * sq(xi - (R)0.5)) *
               ((R)1.0 + (R)0.25 * m_log(sq(((R)1.0 + eseder) * ((R)1.0 + esedei) / ((R)2.0 * tteig)) +
                                         zbrang * sq(xi)));
}

// pair (ompmc.c:1429-1667): a becomes e-/e+ ese1, b gets ese2
template <typename R>
__device__ void pair(const DevTables<R> &T, Particle<R> &a, Particle<R> &b, int imed)
{
    const MediumScalars &M = T.small->med[imed];
    R eig = a.e;
    R ese1, ese2;
    int iq1, iq2, l, l1;
    const R rm = (R)RM;
    if (eig <= (R)2.1) {
        R r0 = urand<R>(a.rng), r1 = urand<R>(a.rng);
        ese2 = rm + (R)0.5 * r0 * (eig - (R)2.0 * rm);
        ese1 = eig - ese2;
        if (r1 < (R)0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    } else {
        R Amax, Bmax, delta, aux;
        if (eig < (R)50.0) {
            l = 4;
            l1 = l + 1;
            delta = (R)4.0 * (R)M.delcm / eig;
            if (delta < (R)1.0) {
                Amax = (R)M.dl1[l] + delta * ((R)M.dl2[l] + delta * (R)M.dl3[l]);
                Bmax = (R)M.dl1[l1] + delta * ((R)M.dl2[l1] + delta * (R)M.dl3[l1]);
            } else {
                aux = m_log(delta + (R)M.dl6[l]);
                Amax = (R)M.dl4[l] + (R)M.dl5[l] * aux;
                Bmax = (R)M.dl4[l1] + (R)M.dl5[l1] * aux;
            }
            aux = (R)1.0 - (R)2.0 * rm / eig;
            aux = aux * aux;
            aux *= Amax / (R)3.0;
            aux /= (Bmax + aux);
        } else {
            l = 6;
            Amax = (R)M.dl1[l];
            Bmax = (R)M.dl1[l + 1];
            aux = (R)M.bpar1 * ((R)1.0 - (R)M.bpar0 * rm / eig);
        }
        R eavail = eig - (R)2.0 * rm, rejf, rejmax, r4;
        do {
            R br, r0 = urand<R>(a.rng), r1 = urand<R>(a.rng);
            r4 = urand<R>(a.rng);
            if (r0 > aux) {
                br = (R)0.5 * r1;
                rejmax = Bmax;
                l1 = l + 1;
            } else {
                R r2 = urand<R>(a.rng), r3 = urand<R>(a.rng);
                br = (R)0.5 * ((R)1.0 - m_fmax(m_fmax(r1, r2), r3));
                rejmax = Amax;
                l1 = l;
            }
            ese2 = br * eavail + rm;
            ese1 = eig - ese2;
            delta = (eig * (R)M.delcm) / (ese2 * ese1);
            if (delta < (R)1.0) rejf = (R)M.dl1[l1] + delta * ((R)M.dl2[l1] + delta * (R)M.dl3[l1]);
            else rejf = (R)M.dl4[l1] + (R)M.dl5[l1] * m_log(delta + (R)M.dl6[l1]);
        } while (r4 * rejmax > rejf);
        ese1 = eig - ese2;
        r4 = urand<R>(a.rng);
        if (r4 < (R)0.5) { iq1 = -1; iq2 = 1; } else { iq1 = 1; iq2 = -1; }
    }
    a.e = ese1;
    b.e = ese2;
    const R zbr = (R)M.zbrang;
    Uphi<R> h;
    for (int j = 0; j < 2; ++j) {
        R ese = (j == 0) ? ese1 : ese2;
        R tteig = eig / rm, ttese = ese / rm;
        R esedei = ttese / (tteig - ttese);
        R eseder = (R)1.0 / esedei;
        R ximin = (R)1.0 / ((R)1.0 + sq((R)M_PI * ttese));
        R rejmin = pair_rej<R>(zbr, ximin, esedei, eseder, tteig);
        R ya = sq((R)2.0 / tteig);
        R xitry = m_fmax((R)0.01, m_fmax(ximin, m_fmin((R)0.5, m_sqrt(ya / zbr))));
        R galpha = (R)1.0 + (R)0.25 * m_log(ya + zbr * xitry * xitry);
        R gbeta = (R)0.5 * zbr * xitry / (ya + zbr * xitry * xitry);
        galpha -= gbeta * (xitry - (R)0.5);
        R ximid = galpha / ((R)3.0 * gbeta);
        if (galpha >= (R)0.0) ximid = (R)0.5 - ximid + m_sqrt(sq(ximid) + (R)0.25);
        else ximid = (R)0.5 - ximid - m_sqrt(sq(ximid) + (R)0.25);
        ximid = m_fmax((R)0.01, m_fmax(ximin, m_fmin((R)0.5, ximid)));
        R rejmid = pair_rej<R>(zbr, ximid, esedei, eseder, tteig);
        R rejtop = (R)1.0 * m_fmax(rejmin, rejmid);
        R theta, rejtst, rtest;
        do {
            R xitst = urand<R>(a.rng);
            rejtst = pair_rej<R>(zbr, xitst, esedei, eseder, tteig);
            rtest = urand<R>(a.rng);
            theta = m_sqrt((R)1.0 / xitst - (R)1.0) / ttese;
        } while ((rtest > rejtst / rejtop) && (theta >= (R)M_PI));   // reference's && kept
        R sinthe = m_sin(theta), costhe = m_cos(theta);
        if (j == 0) {
            uphi21<R>(a, h, costhe, sinthe);
        } else {
            sinthe = -sinthe;
            uphi32<R>(a, b, h, costhe, sinthe);
        }
    }
    a.iq = iq1;
    b.iq = iq2;
    spawn<R>(a, b);
}

// compton (ompmc.c:1670-1783): a = scattered photon, b = electron
template <typename R>
__device__ void compton(Particle<R> &a, Particle<R> &b)
{
    const R rm = (R)RM;
    R eig = a.e;
    R ko = a.e / rm;
    R broi = (R)1.0 + (R)2.0 * ko, bro = (R)1.0 / broi;
    bool first = true;
    R sinthe = 0, costhe, br, r1, r2, r3, aux, rejf3, temp = 0;
    R alph1 = 0, alph2 = 0, alpha = 0, rejmax = 0;
    do {
        if (ko > (R)2.0) {
            if (first) {
                alph1 = m_log(broi);
                alph2 = ko * (broi + (R)1.0) * sq(bro);
                alpha = alph1 + alph2;
            }
            do {
                r1 = urand<R>(a.rng);
                r2 = urand<R>(a.rng);
                if (r1 * alpha < alph1) br = m_exp(alph1 * r2) * bro;
                else br = m_sqrt(r2 * sq(broi) + ((R)1.0 - r2)) * bro;
                temp = ((R)1.0 - br) / (ko * br);
                sinthe = m_fmax((R)0.0, temp * ((R)2.0 - temp));
                aux = (R)1.0 + sq(br);
                rejf3 = aux - br * sinthe;
                r3 = urand<R>(a.rng);
            } while (r3 * aux > rejf3);
        } else {
            if (first) rejmax = broi + bro;
            do {
                r1 = urand<R>(a.rng);
                r2 = urand<R>(a.rng);
                br = bro + ((R)1.0 - bro) * r1;
                temp = ((R)1.0 - br) / (ko * br);
                sinthe = m_fmax((R)0.0, temp * ((R)2.0 - temp));
                rejf3 = (R)1.0 + br * br - br * sinthe;
            } while (r2 * br * rejmax > rejf3);
        }
        first = false;
    } while ((br < bro) || (br > (R)1));
    costhe = (R)1.0 - temp;
    sinthe = m_sqrt(sinthe);
    R esg = br * eig, ese = eig - esg + rm;
    a.e = esg;
    Uphi<R> h;
    uphi21<R>(a, h, costhe, sinthe);
    aux = (R)1.0 + br * br - (R)2.0 * br * costhe;
    if (aux > (R)1.0E-8) {
        costhe = ((R)1.0 - br * costhe) / m_sqrt(aux);
        sinthe = ((R)1.0 - costhe) * ((R)1.0 + costhe);
        sinthe = (sinthe > (R)0.0) ? -m_sqrt(sinthe) : (R)0.0;
    } else {
        costhe = (R)0.0;
        sinthe = (R)-1.0;
    }
    uphi32<R>(a, b, h, costhe, sinthe);
    b.e = ese;
    b.iq = -1;
    spawn<R>(a, b);
}

// photo (ompmc.c:1786-1846): the photon becomes the photo-electron
template <typename R>
__device__ void photo(Particle<R> &a, R ecut_here)
{
    const R rm = (R)RM;
    a.e += rm;
    a.iq = -1;
    R eelec = a.e;
    if (eelec > ecut_here) {
        R beta = m_sqrt((eelec - rm) * (eelec + rm)) / eelec;
        R costhe, sinth2, xi, rnpht2;
        R gamma = eelec / rm;
        R alpha = (R)0.5 * gamma - (R)0.5 + (R)1.0 / gamma;
        R ratio = beta / alpha;
        do {
            R rnpht = urand<R>(a.rng);
            rnpht = (R)2.0 * rnpht - (R)1.0;
            if (ratio <= (R)0.2) {
                R fkappa = rnpht + (R)0.5 * ratio * ((R)1.0 - rnpht) * ((R)1.0 + rnpht);
                if (gamma < (R)100.0) {
                    costhe = (beta + fkappa) / ((R)1.0 + beta * fkappa);
                } else if (fkappa > (R)0.0) {
                    costhe = (R)1.0 - ((R)1.0 - fkappa) * (gamma - (R)3.0) /
                                          ((R)2.0 * ((R)1.0 + fkappa) * m_pow3(gamma - (R)1.0));
                } else {
                    costhe = (beta + fkappa) / ((R)1.0 + beta * fkappa);
                }
                xi = ((R)1.0 + beta * fkappa) * sq(gamma);
            } else {
                xi = sq(gamma) * ((R)1.0 + alpha * (m_sqrt((R)1.0 + ratio * ((R)2.0 * rnpht + ratio)) - (R)1.0));
                costhe = ((R)1.0 - (R)1.0 / xi) / beta;
            }
            sinth2 = m_fmax(((R)1.0 - costhe) * ((R)1.0 + costhe), (R)0.0);
            rnpht2 = urand<R>(a.rng);
        } while (rnpht2 > (R)0.5 * ((R)1.0 + gamma) * sinth2 * xi / gamma);
        R sinthe = m_sqrt(sinth2);
        Uphi<R> h;
        uphi21<R>(a, h, costhe, sinthe);
    }
}

}  // namespace omc
