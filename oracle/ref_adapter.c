/*
 * Test-only checker linked against the reference ompMC sources (oracle/_ref):
 *
 *   ref_adapter <inp-stem> <outdir>
 *
 * runs the reference initialisation (initPhantom, initMediaData), fills an
 * omc_tables from the reference's globals through include/
 * omc_ompmc_adapter.h -- the adapter an ompMC program would use to feed the
 * device engine -- and writes every field of it as a raw little-endian
 * array <outdir>/<field>.bin.  tests/test_adapter.py compares them with
 * omc_tables_build for the same media.
 */
#include "ompmc.h"
#include "omc_utilities.h"
#include "omc_ompmc_adapter.h"

#include <stdio.h>
#include <stdlib.h>

void refacc_init_phantom(void);

static const char *g_out;

static void wr(const char *name, const void *p, size_t bytes)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%s.bin", g_out, name);
    FILE *fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fwrite(p, 1, bytes, fp);
    fclose(fp);
}

#define WD(f, n) wr(#f, t.f, (size_t)(n) * sizeof(double))
#define WI(f, n) wr(#f, t.f, (size_t)(n) * sizeof(int))

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: ref_adapter <inp-stem> <outdir>\n");
        return 2;
    }
    g_out = argv[2];
    parseInputFile(argv[1]);
    refacc_init_phantom();
    initMediaData();
    omc_tables t;
    const int rc = omc_tables_from_ompmc(&t);
    if (rc != OMC_OK) { fprintf(stderr, "adapter failed: %d\n", rc); return 1; }
    const int n = t.nmed;
    const size_t G = (size_t)n * OMC_MXGE, K = (size_t)n * OMC_MXEKE;
    wr("nmed", &t.nmed, sizeof(int));
    wr("names", t.names, sizeof t.names);
    WI(ne, n); WI(iunrst, n); WI(epstfl, n); WI(iaprim, n); WI(msge, n); WI(mge, n); WI(mseke, n);
    WI(meke, n); WI(mleke, n); WI(mcmfp, n); WI(mrange, n);
    WD(rho, n); WD(rlc, n); WD(ae, n); WD(ap, n); WD(ue, n); WD(up, n); WD(te, n); WD(thmoll, n);
    WD(pegs_delcm, n);
    WD(ge0, n); WD(ge1, n);
    WD(gmfp0, G); WD(gmfp1, G); WD(gbr10, G); WD(gbr11, G); WD(gbr20, G); WD(gbr21, G); WD(cohe0, G);
    WD(cohe1, G);
    WD(xgrid, n * OMC_MXRAYFF); WD(fcum, n * OMC_MXRAYFF); WD(b_array, n * OMC_MXRAYFF);
    WD(c_array, n * OMC_MXRAYFF); WI(i_array, n * OMC_RAYCDFSIZE); WD(pmax0, G); WD(pmax1, G);
    WD(dl1, n * 8); WD(dl2, n * 8); WD(dl3, n * 8); WD(dl4, n * 8); WD(dl5, n * 8); WD(dl6, n * 8);
    WD(bpar0, n); WD(bpar1, n); WD(delcm, n); WD(zbrang, n);
    WD(eke0, n); WD(eke1, n); WD(blcc, n); WD(xcc, n); WD(esig_e, n); WD(psig_e, n); WD(expeke1, n);
    int mono[2 * OMC_MXMED];
    for (int m = 0; m < n; ++m) { mono[m] = t.sig_ismonotone[0][m]; mono[n + m] = t.sig_ismonotone[1][m]; }
    wr("sig_ismonotone", mono, 2 * (size_t)n * sizeof(int));
    WD(esig0, K); WD(esig1, K); WD(psig0, K); WD(psig1, K); WD(ededx0, K); WD(ededx1, K); WD(pdedx0, K);
    WD(pdedx1, K); WD(ebr10, K); WD(ebr11, K); WD(pbr10, K); WD(pbr11, K); WD(pbr20, K); WD(pbr21, K);
    WD(tmxs0, K); WD(tmxs1, K); WD(etae_ms0, K); WD(etae_ms1, K); WD(etap_ms0, K); WD(etap_ms1, K);
    WD(q1ce_ms0, K); WD(q1ce_ms1, K); WD(q1cp_ms0, K); WD(q1cp_ms1, K); WD(q2ce_ms0, K); WD(q2ce_ms1, K);
    WD(q2cp_ms0, K); WD(q2cp_ms1, K); WD(blcce0, K); WD(blcce1, K); WD(range_ep, 2 * K); WD(e_array, K);
    double sp[5] = { t.b2spin_min, t.dbeta2i, t.espml, t.dleneri, t.dqq1i };
    wr("spin_scalars", sp, sizeof sp);
    WD(spin_rej, (size_t)n * OMC_SPIN_PER_MED);
    WD(ums, OMC_MS_SIZE); WD(fms, OMC_MS_SIZE); WD(wms, OMC_MS_SIZE); WI(ims, OMC_MS_SIZE);
    double ms[2] = { t.dllambi, t.dqmsi };
    wr("ms_scalars", ms, sizeof ms);
    return 0;
}
