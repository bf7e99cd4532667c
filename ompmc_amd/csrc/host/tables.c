/*
 * Builds all per-medium tables in the reference's order
 * (initMediaData, src/ompmc.c:5450-5487): pegs4 data, photon, Rayleigh,
 * pair, then multiple scattering / spin / electron step data.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <stdlib.h>
#include <string.h>

int omc_read_pegs(omc_tables *t, const char *path);
int omc_build_photon(omc_tables *t, const char *data_folder);
int omc_build_rayleigh(omc_tables *t, const char *pgs4form_file);
int omc_build_pair(omc_tables *t);
int omc_build_electron(omc_tables *t, const char *data_folder);

#define ALLOC(field, n, type) do { \
        t->field = (type *)calloc((size_t)(n), sizeof(type)); \
        if (!t->field) { omc_tables_free(t); return omc_set_error(OMC_ENOMEM, "out of memory"); } \
    } while (0)

void omc_tables_free(omc_tables *t)
{
    double **dp[] = {
        &t->gmfp0, &t->gmfp1, &t->gbr10, &t->gbr11, &t->gbr20, &t->gbr21, &t->cohe0, &t->cohe1,
        &t->xgrid, &t->fcum, &t->b_array, &t->c_array, &t->pmax0, &t->pmax1,
        &t->esig0, &t->esig1, &t->psig0, &t->psig1, &t->ededx0, &t->ededx1, &t->pdedx0, &t->pdedx1,
        &t->ebr10, &t->ebr11, &t->pbr10, &t->pbr11, &t->pbr20, &t->pbr21, &t->tmxs0, &t->tmxs1,
        &t->etae_ms0, &t->etae_ms1, &t->etap_ms0, &t->etap_ms1, &t->q1ce_ms0, &t->q1ce_ms1,
        &t->q1cp_ms0, &t->q1cp_ms1, &t->q2ce_ms0, &t->q2ce_ms1, &t->q2cp_ms0, &t->q2cp_ms1,
        &t->blcce0, &t->blcce1, &t->range_ep, &t->e_array, &t->spin_rej, &t->ums, &t->fms, &t->wms };
    for (size_t i = 0; i < sizeof dp / sizeof dp[0]; ++i) { free(*dp[i]); *dp[i] = NULL; }
    free(t->i_array); t->i_array = NULL;
    free(t->ims); t->ims = NULL;
}

int omc_tables_build(omc_tables *t, int nmed, const char names[][64],
                     const omc_data_paths *paths)
{
    if (nmed < 1 || nmed > OMC_MXMED)
        return omc_set_error(OMC_ERANGE, "number of media %d outside 1..%d", nmed, OMC_MXMED);
    if (!paths || !paths->pegs_file || !paths->pgs4form_file || !paths->data_folder)
        return omc_set_error(OMC_EARG, "data paths not given");
    memset(t, 0, sizeof *t);
    t->nmed = nmed;
    for (int i = 0; i < nmed; ++i) {
        strncpy(t->names[i], names[i], sizeof t->names[i] - 1);
        t->names[i][sizeof t->names[i] - 1] = '\0';
    }
    const size_t G = (size_t)nmed * OMC_MXGE, K = (size_t)nmed * OMC_MXEKE;
    ALLOC(gmfp0, G, double); ALLOC(gmfp1, G, double); ALLOC(gbr10, G, double); ALLOC(gbr11, G, double);
    ALLOC(gbr20, G, double); ALLOC(gbr21, G, double); ALLOC(cohe0, G, double); ALLOC(cohe1, G, double);
    ALLOC(xgrid, nmed * OMC_MXRAYFF, double); ALLOC(fcum, nmed * OMC_MXRAYFF, double);
    ALLOC(b_array, nmed * OMC_MXRAYFF, double); ALLOC(c_array, nmed * OMC_MXRAYFF, double);
    ALLOC(i_array, nmed * OMC_RAYCDFSIZE, int);
    ALLOC(pmax0, G, double); ALLOC(pmax1, G, double);
    ALLOC(esig0, K, double); ALLOC(esig1, K, double); ALLOC(psig0, K, double); ALLOC(psig1, K, double);
    ALLOC(ededx0, K, double); ALLOC(ededx1, K, double); ALLOC(pdedx0, K, double); ALLOC(pdedx1, K, double);
    ALLOC(ebr10, K, double); ALLOC(ebr11, K, double); ALLOC(pbr10, K, double); ALLOC(pbr11, K, double);
    ALLOC(pbr20, K, double); ALLOC(pbr21, K, double); ALLOC(tmxs0, K, double); ALLOC(tmxs1, K, double);
    ALLOC(etae_ms0, K, double); ALLOC(etae_ms1, K, double); ALLOC(etap_ms0, K, double); ALLOC(etap_ms1, K, double);
    ALLOC(q1ce_ms0, K, double); ALLOC(q1ce_ms1, K, double); ALLOC(q1cp_ms0, K, double); ALLOC(q1cp_ms1, K, double);
    ALLOC(q2ce_ms0, K, double); ALLOC(q2ce_ms1, K, double); ALLOC(q2cp_ms0, K, double); ALLOC(q2cp_ms1, K, double);
    ALLOC(blcce0, K, double); ALLOC(blcce1, K, double);
    ALLOC(range_ep, 2 * K, double); ALLOC(e_array, K, double);
    ALLOC(spin_rej, (size_t)nmed * OMC_SPIN_PER_MED, double);
    ALLOC(ums, OMC_MS_SIZE, double); ALLOC(fms, OMC_MS_SIZE, double); ALLOC(wms, OMC_MS_SIZE, double);
    ALLOC(ims, OMC_MS_SIZE, int);

    int rc = omc_read_pegs(t, paths->pegs_file);
    if (!rc) rc = omc_build_photon(t, paths->data_folder);
    if (!rc) rc = omc_build_rayleigh(t, paths->pgs4form_file);
    if (!rc) rc = omc_build_pair(t);
    if (!rc) rc = omc_build_electron(t, paths->data_folder);
    if (rc) omc_tables_free(t);
    return rc;
}
