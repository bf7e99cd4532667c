/*
 * Test-only harness linked against the reference ompMC sources (oracle/_ref).
 *
 *   ref_dump tables  <inp-stem> <outdir>
 *       runs the reference initialisation (initPhantom, initMediaData,
 *       initSource, initRegions) and writes every table the transport reads
 *       as raw little-endian arrays <outdir>/<name>.bin
 *   ref_dump vectors <inp-stem> <outdir> <n>
 *       calls individual reference functions on n generated states each
 *       (a fixed xorshift sequence, so the vectors are reproducible) and
 *       writes inputs and outputs side by side (per-function parity
 *       vectors): howfar and hownear on the geometry's voxels (positions on
 *       and between boundaries, axis-parallel and oblique directions, the
 *       outside region), computeDrange and computeEloss at the arguments
 *       electron() passes (ompmc.c:4832-4926, 4983), and the photon table
 *       lookups pwlfInterval / pwlfEval (ompmc.c:1952-1969)
 *
 * The outputs are consumed by tests/golden/make_golden.py, which turns them
 * into hashes and small fixtures.
 */
#include "ompmc.h"
#include "omc_utilities.h"
#include "omc_random.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int     refacc_isize(void);
int     refacc_jsize(void);
int     refacc_ksize(void);
double *refacc_bounds(int d);
double *refacc_densities(void);
int     refacc_src_spectrum(void);
double *refacc_cdfinv(int w);
double  refacc_src_scalar(int w);
void    refacc_init_source(void);
void    refacc_init_phantom(void);
void    refacc_init_regions(void);
void    refacc_init_score(void);
void    refacc_init_history(void);
double *refacc_endep(void);

static const char *g_out;

static void dump(const char *name, const void *p, size_t bytes)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", g_out, name);
    FILE *fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fwrite(p, 1, bytes, fp);
    fclose(fp);
}

#define DD(name, ptr, n) dump(name, ptr, (size_t)(n) * sizeof(double))
#define DI(name, ptr, n) dump(name, ptr, (size_t)(n) * sizeof(int))

static void init_all(const char *stem)
{
    parseInputFile((char *)stem);
    refacc_init_phantom();
    initMediaData();
    refacc_init_source();
    refacc_init_regions();
    initVrt();
    refacc_init_score();
}

static int dump_tables(void)
{
    int nmed = media.nmed;
    size_t G = (size_t)nmed * MXGE, K = (size_t)nmed * MXEKE;
    DD("ge0", photon_data.ge0, nmed); DD("ge1", photon_data.ge1, nmed);
    DD("gmfp0", photon_data.gmfp0, G); DD("gmfp1", photon_data.gmfp1, G);
    DD("gbr10", photon_data.gbr10, G); DD("gbr11", photon_data.gbr11, G);
    DD("gbr20", photon_data.gbr20, G); DD("gbr21", photon_data.gbr21, G);
    DD("cohe0", photon_data.cohe0, G); DD("cohe1", photon_data.cohe1, G);
    DD("xgrid", rayleigh_data.xgrid, nmed * MXRAYFF); DD("fcum", rayleigh_data.fcum, nmed * MXRAYFF);
    /* b_array[99] is never written by the reference (uninitialised) */
    for (int m = 0; m < nmed; ++m) rayleigh_data.b_array[m * MXRAYFF + MXRAYFF - 1] = 0.0;
    for (int m = 0; m < nmed; ++m) rayleigh_data.c_array[m * MXRAYFF + MXRAYFF - 1] = 0.0;
    DD("b_array", rayleigh_data.b_array, nmed * MXRAYFF); DD("c_array", rayleigh_data.c_array, nmed * MXRAYFF);
    DI("i_array", rayleigh_data.i_array, nmed * RAYCDFSIZE);
    DD("pmax0", rayleigh_data.pmax0, G); DD("pmax1", rayleigh_data.pmax1, G);
    DD("dl1", pair_data.dl1, nmed * 8); DD("dl2", pair_data.dl2, nmed * 8); DD("dl3", pair_data.dl3, nmed * 8);
    DD("dl4", pair_data.dl4, nmed * 8); DD("dl5", pair_data.dl5, nmed * 8); DD("dl6", pair_data.dl6, nmed * 8);
    DD("bpar0", pair_data.bpar0, nmed); DD("bpar1", pair_data.bpar1, nmed);
    DD("delcm", pair_data.delcm, nmed); DD("zbrang", pair_data.zbrang, nmed);
    DD("eke0", electron_data.eke0, nmed); DD("eke1", electron_data.eke1, nmed);
    DD("blcc", electron_data.blcc, nmed); DD("xcc", electron_data.xcc, nmed);
    DD("esig_e", electron_data.esig_e, nmed); DD("psig_e", electron_data.psig_e, nmed);
    DD("expeke1", electron_data.expeke1, nmed);
    DI("sig_ismonotone", electron_data.sig_ismonotone, 2 * nmed);
#define EK(f) DD(#f, electron_data.f, K)
    EK(esig0); EK(esig1); EK(psig0); EK(psig1); EK(ededx0); EK(ededx1); EK(pdedx0); EK(pdedx1);
    EK(ebr10); EK(ebr11); EK(pbr10); EK(pbr11); EK(pbr20); EK(pbr21); EK(tmxs0); EK(tmxs1);
    EK(etae_ms0); EK(etae_ms1); EK(etap_ms0); EK(etap_ms1); EK(q1ce_ms0); EK(q1ce_ms1);
    EK(q1cp_ms0); EK(q1cp_ms1); EK(q2ce_ms0); EK(q2ce_ms1); EK(q2cp_ms0); EK(q2cp_ms1);
    EK(blcce0); EK(blcce1); EK(e_array);
#undef EK
    DD("range_ep", electron_data.range_ep, 2 * K);
    double sp[5] = { spin_data.b2spin_min, spin_data.dbeta2i, spin_data.espml, spin_data.dleneri, spin_data.dqq1i };
    DD("spin_scalars", sp, 5);
    DD("spin_rej", spin_data.spin_rej, (size_t)nmed * 2 * 32 * 16 * 32);
    /* wms/ims entry 31 of each row is never written by the reference */
    for (int i = 0; i < 64 * 8; ++i) { mscat_data.wms_array[i * 32 + 31] = 0.0; mscat_data.ims_array[i * 32 + 31] = 0; }
    DD("ums", mscat_data.ums_array, 64 * 8 * 32); DD("fms", mscat_data.fms_array, 64 * 8 * 32);
    DD("wms", mscat_data.wms_array, 64 * 8 * 32); DI("ims", mscat_data.ims_array, 64 * 8 * 32);
    double ms[2] = { mscat_data.dllambi, mscat_data.dqmsi };
    DD("ms_scalars", ms, 2);
    double pg[8 * MXMED];
    for (int m = 0; m < nmed; ++m) {
        pg[8 * m + 0] = pegs_data.rho[m]; pg[8 * m + 1] = pegs_data.rlc[m];
        pg[8 * m + 2] = pegs_data.ae[m];  pg[8 * m + 3] = pegs_data.ap[m];
        pg[8 * m + 4] = pegs_data.ue[m];  pg[8 * m + 5] = pegs_data.up[m];
        pg[8 * m + 6] = pegs_data.te[m];  pg[8 * m + 7] = pegs_data.thmoll[m];
    }
    DD("pegs_scalars", pg, 8 * nmed);
    int nreg = refacc_isize() * refacc_jsize() * refacc_ksize() + 1;
    DI("region_med", region.med, nreg);
    DD("region_rhof", region.rhof, nreg);
    DD("region_pcut", region.pcut, nreg);
    DD("region_ecut", region.ecut, nreg);
    if (refacc_src_spectrum()) { DD("cdfinv1", refacc_cdfinv(0), 1000); DD("cdfinv2", refacc_cdfinv(1), 1000); }
    double ss[11];
    for (int w = 0; w < 11; ++w) ss[w] = refacc_src_scalar(w);
    DD("source_scalars", ss, 11);
    return 0;
}

/* ---- per-function vectors -------------------------------------------------- */
static uint64_t g_rs = 0x9E3779B97F4A7C15ull;

static double uni(void)
{
    g_rs ^= g_rs >> 12;
    g_rs ^= g_rs << 25;
    g_rs ^= g_rs >> 27;
    return (double)((g_rs * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
}

static int pick(int n) { int k = (int)(uni() * n); return k < n ? k : n - 1; }

static FILE *vopen(const char *name)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", g_out, name);
    FILE *fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    return fp;
}

/* one record = doubles (integers stored as doubles, exact below 2^53) */
static void rec(FILE *fp, const double *v, int n) { fwrite(v, sizeof(double), (size_t)n, fp); }

static int dump_vectors(int n)
{
    const int nx = refacc_isize(), ny = refacc_jsize(), nz = refacc_ksize();
    const double *xb = refacc_bounds(0), *yb = refacc_bounds(1), *zb = refacc_bounds(2);
    const int nmed = media.nmed;
    initStack();
    stack.np = 0;
    /* howfar / hownear: record irl x y z u v w ustep_in | idisc irnew ustep_out tperp */
    FILE *fg = vopen("vec_geometry");
    for (int k = 0; k < n; ++k) {
        int ix = pick(nx), iy = pick(ny), iz = pick(nz);
        int irl = (k % 17 == 0) ? 0 : 1 + ix + iy * nx + iz * nx * ny;
        double f[3];
        for (int d = 0; d < 3; ++d) {
            const int mode = pick(8);          /* 0: lower plane, 1: upper plane, else inside */
            f[d] = mode == 0 ? 0.0 : mode == 1 ? 1.0 : uni();
        }
        double x = xb[ix] + f[0] * (xb[ix + 1] - xb[ix]);
        double y = yb[iy] + f[1] * (yb[iy + 1] - yb[iy]);
        double z = zb[iz] + f[2] * (zb[iz + 1] - zb[iz]);
        double u, v, w;
        if (k % 5 == 0) {                      /* axis-parallel, either sense */
            const int a = pick(3);
            const double sgn = uni() < 0.5 ? -1.0 : 1.0;
            u = a == 0 ? sgn : 0.0; v = a == 1 ? sgn : 0.0; w = a == 2 ? sgn : 0.0;
        } else {
            const double ct = 2.0 * uni() - 1.0, ph = 6.283185307179586 * uni();
            const double st = sqrt(1.0 - ct * ct);
            u = st * cos(ph); v = st * sin(ph); w = ct;
        }
        const double dx = xb[ix + 1] - xb[ix];
        const double ustep_in = (k % 7 == 0) ? 1.0E8 : 3.0 * dx * uni();
        stack.ir[0] = irl;
        stack.x[0] = x; stack.y[0] = y; stack.z[0] = z;
        stack.u[0] = u; stack.v[0] = v; stack.w[0] = w;
        int idisc = 0, irnew = irl;
        double ustep = ustep_in;
        howfar(&idisc, &irnew, &ustep);
        const double tperp = hownear();
        const double r[12] = { irl, x, y, z, u, v, w, ustep_in, idisc, irnew, ustep, tperp };
        rec(fg, r, 12);
    }
    fclose(fg);
    /* computeDrange at electron()'s range call and computeEloss at its CH
     * step: imed iq lelke eke elke ekei elkei rhof tustep range | drange de */
    FILE *fe = vopen("vec_csda");
    for (int k = 0; k < n; ++k) {
        const int imed = pick(nmed), iq = uni() < 0.5 ? -1 : 1, qel = (1 + iq) / 2;
        const double lo = log(pegs_data.ae[imed] - RM), hi = log(pegs_data.ue[imed] - RM);
        const double eke = exp(lo + (hi - lo) * (0.001 + 0.998 * uni()));
        const double elke = log(eke);
        const int lelke = pwlfInterval(imed, elke, electron_data.eke1, electron_data.eke0) - 1;
        const double ekei = electron_data.e_array[imed * MXEKE + lelke];
        const double elkei = (lelke + 1 - electron_data.eke0[imed]) / electron_data.eke1[imed];
        const double dr = computeDrange(imed, iq, lelke, eke, ekei, elke, elkei);
        const double rhof = 0.2 + 1.8 * uni();
        double range = dr + electron_data.range_ep[qel * nmed * MXEKE + imed * MXEKE + lelke];
        range /= rhof;
        const double tustep = range * (k % 9 == 0 ? 1.0 : uni());
        const double de = computeEloss(imed, iq, 1, rhof, tustep, range, eke, elke, lelke);
        const double r[12] = { imed, iq, lelke, eke, elke, ekei, elkei, rhof, tustep, range, dr, de };
        rec(fe, r, 12);
    }
    fclose(fe);
    /* photon table lookups: imed e | lgle gmfp cohfac gbr1 gbr2 */
    FILE *fp = vopen("vec_pwlf");
    for (int k = 0; k < n; ++k) {
        const int imed = pick(nmed);
        const double lo = log(pegs_data.ap[imed]), hi = log(pegs_data.up[imed]);
        const double e = exp(lo + (hi - lo) * (k % 11 == 0 ? 1e-9 : uni()));
        const double gle = log(e);
        const int lgle = pwlfInterval(imed, gle, photon_data.ge1, photon_data.ge0) - 1;
        const int ix = imed * MXGE + lgle;
        const double r[7] = { imed, e, lgle, pwlfEval(ix, gle, photon_data.gmfp1, photon_data.gmfp0),
                              pwlfEval(ix, gle, photon_data.cohe1, photon_data.cohe0),
                              pwlfEval(ix, gle, photon_data.gbr11, photon_data.gbr10),
                              pwlfEval(ix, gle, photon_data.gbr21, photon_data.gbr20) };
        rec(fp, r, 7);
    }
    fclose(fp);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: ref_dump tables <inp-stem> <outdir>\n"
                        "       ref_dump vectors <inp-stem> <outdir> <n>\n");
        return 2;
    }
    g_out = argv[3];
    init_all(argv[2]);
    if (!strcmp(argv[1], "tables")) return dump_tables();
    if (!strcmp(argv[1], "vectors")) return dump_vectors(argc > 4 ? atoi(argv[4]) : 1000);
    fprintf(stderr, "unknown mode %s\n", argv[1]);
    return 2;
}
