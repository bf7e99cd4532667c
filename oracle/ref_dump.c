/*
 * Test-only harness linked against the reference ompMC sources (oracle/_ref).
 *
 *   ref_dump tables  <inp-stem> <outdir>
 *       runs the reference initialisation (initPhantom, initMediaData,
 *       initSource, initRegions) and writes every table the transport reads
 *       as raw little-endian arrays <outdir>/<name>.bin
 *   ref_dump vectors <inp-stem> <outdir> <tape-file> <n>
 *       calls individual reference transport functions on seeded states with
 *       the random-number stream replaced by a tape of 24-bit integers, and
 *       writes their outputs (per-function parity vectors)
 *
 * The outputs are consumed by tests/golden/make_golden.py, which turns them
 * into hashes and small fixtures.
 */
#include "ompmc.h"
#include "omc_utilities.h"
#include "omc_random.h"

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

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: ref_dump tables <inp-stem> <outdir>\n");
        return 2;
    }
    g_out = argv[3];
    init_all(argv[2]);
    if (!strcmp(argv[1], "tables")) return dump_tables();
    fprintf(stderr, "unknown mode %s\n", argv[1]);
    return 2;
}
