/*
 * omc_ompmc_adapter.h - fills an omc_tables from the tables a running ompMC
 * program has built, so that ompMC's own table code can feed the device
 * engine unchanged.
 *
 * ompMC keeps every table the transport reads in globals that
 * initMediaData() fills (src/ompmc.c:5450-5487): media, pegs_data,
 * photon_data, rayleigh_data, pair_data, electron_data, spin_data and
 * mscat_data (struct definitions in src/ompmc.h:104-401).  This header reads
 * them.  Include it in one C file of the ompMC program, after ompmc.h and
 * the two headers of this library:
 *
 *     #include "ompmc.h"
 *     #include "omc_dev.h"
 *     #include "omc_ompmc_adapter.h"
 *     ...
 *     initMediaData();
 *     omc_tables t;
 *     if (omc_tables_from_ompmc(&t) == OMC_OK) omc_dev_upload_media(ctx, &t);
 *
 * The per-bin arrays are borrowed: t points into ompMC's own arrays, which
 * must stay allocated while t is used (omc_dev_upload_media copies them,
 * so they may be freed after it returns).  Scalars and per-medium values
 * are copied.  No omc_tables_free() for a table filled here.
 *
 * The field names below are ompMC's; their meaning and index conventions
 * are documented in omc_host.h, which follows the same layout.
 */
#ifndef OMC_OMPMC_ADAPTER_H
#define OMC_OMPMC_ADAPTER_H

#include <string.h>

#include "omc_host.h"

static inline int omc_tables_from_ompmc(omc_tables *t)
{
    const int nmed = media.nmed;
    if (!t) return OMC_EARG;
    if (nmed < 1 || nmed > OMC_MXMED) return OMC_ERANGE;
    if (!photon_data.gmfp0 || !electron_data.esig0 || !spin_data.spin_rej || !mscat_data.ums_array)
        return OMC_EARG;          /* initMediaData() has not run */
    memset(t, 0, sizeof *t);
    t->nmed = nmed;
    for (int m = 0; m < nmed; ++m) {
        strncpy(t->names[m], media.med_names[m], sizeof t->names[m] - 1);
        /* pegs4 header of the medium (src/ompmc.h:374-401) */
        t->ne[m] = pegs_data.ne[m];
        t->iunrst[m] = pegs_data.iunrst[m];
        t->epstfl[m] = pegs_data.epstfl[m];
        t->iaprim[m] = pegs_data.iaprim[m];
        t->msge[m] = pegs_data.msge[m];
        t->mge[m] = pegs_data.mge[m];
        t->mseke[m] = pegs_data.mseke[m];
        t->meke[m] = pegs_data.meke[m];
        t->mleke[m] = pegs_data.mleke[m];
        t->mcmfp[m] = pegs_data.mcmfp[m];
        t->mrange[m] = pegs_data.mrange[m];
        t->rho[m] = pegs_data.rho[m];
        t->rlc[m] = pegs_data.rlc[m];
        t->ae[m] = pegs_data.ae[m];
        t->ap[m] = pegs_data.ap[m];
        t->ue[m] = pegs_data.ue[m];
        t->up[m] = pegs_data.up[m];
        t->te[m] = pegs_data.te[m];
        t->thmoll[m] = pegs_data.thmoll[m];
        t->pegs_delcm[m] = pegs_data.delcm[m];
        for (int e = 0; e < OMC_MXELEMENT && e < MXELEMENT; ++e) {
            memcpy(t->elements[m][e].symbol, pegs_data.elements[m][e].symbol, 3);
            t->elements[m][e].z = pegs_data.elements[m][e].z;
            t->elements[m][e].wa = pegs_data.elements[m][e].wa;
            t->elements[m][e].pz = pegs_data.elements[m][e].pz;
            t->elements[m][e].rhoz = pegs_data.elements[m][e].rhoz;
        }
        /* photon and pair (src/ompmc.h:104-169) */
        t->ge0[m] = photon_data.ge0[m];
        t->ge1[m] = photon_data.ge1[m];
        for (int l = 0; l < 8; ++l) {
            t->dl1[m * 8 + l] = pair_data.dl1[m * 8 + l];
            t->dl2[m * 8 + l] = pair_data.dl2[m * 8 + l];
            t->dl3[m * 8 + l] = pair_data.dl3[m * 8 + l];
            t->dl4[m * 8 + l] = pair_data.dl4[m * 8 + l];
            t->dl5[m * 8 + l] = pair_data.dl5[m * 8 + l];
            t->dl6[m * 8 + l] = pair_data.dl6[m * 8 + l];
        }
        t->bpar0[m] = pair_data.bpar0[m];
        t->bpar1[m] = pair_data.bpar1[m];
        t->delcm[m] = pair_data.delcm[m];
        t->zbrang[m] = pair_data.zbrang[m];
        /* electron scalars (src/ompmc.h:202-252); sig_ismonotone is
         * [qel*nmed + imed] in ompMC, [qel][imed] here */
        t->eke0[m] = electron_data.eke0[m];
        t->eke1[m] = electron_data.eke1[m];
        t->blcc[m] = electron_data.blcc[m];
        t->xcc[m] = electron_data.xcc[m];
        t->esig_e[m] = electron_data.esig_e[m];
        t->psig_e[m] = electron_data.psig_e[m];
        t->expeke1[m] = electron_data.expeke1[m];
        t->sig_ismonotone[0][m] = electron_data.sig_ismonotone[0 * nmed + m];
        t->sig_ismonotone[1][m] = electron_data.sig_ismonotone[1 * nmed + m];
    }
    /* per-bin arrays, borrowed; ompMC's index conventions are kept */
    t->gmfp0 = photon_data.gmfp0;   t->gmfp1 = photon_data.gmfp1;
    t->gbr10 = photon_data.gbr10;   t->gbr11 = photon_data.gbr11;
    t->gbr20 = photon_data.gbr20;   t->gbr21 = photon_data.gbr21;
    t->cohe0 = photon_data.cohe0;   t->cohe1 = photon_data.cohe1;
    t->xgrid = rayleigh_data.xgrid; t->fcum = rayleigh_data.fcum;
    t->b_array = rayleigh_data.b_array; t->c_array = rayleigh_data.c_array;
    t->i_array = rayleigh_data.i_array;
    t->pmax0 = rayleigh_data.pmax0; t->pmax1 = rayleigh_data.pmax1;
    t->esig0 = electron_data.esig0;   t->esig1 = electron_data.esig1;
    t->psig0 = electron_data.psig0;   t->psig1 = electron_data.psig1;
    t->ededx0 = electron_data.ededx0; t->ededx1 = electron_data.ededx1;
    t->pdedx0 = electron_data.pdedx0; t->pdedx1 = electron_data.pdedx1;
    t->ebr10 = electron_data.ebr10;   t->ebr11 = electron_data.ebr11;
    t->pbr10 = electron_data.pbr10;   t->pbr11 = electron_data.pbr11;
    t->pbr20 = electron_data.pbr20;   t->pbr21 = electron_data.pbr21;
    t->tmxs0 = electron_data.tmxs0;   t->tmxs1 = electron_data.tmxs1;
    t->etae_ms0 = electron_data.etae_ms0; t->etae_ms1 = electron_data.etae_ms1;
    t->etap_ms0 = electron_data.etap_ms0; t->etap_ms1 = electron_data.etap_ms1;
    t->q1ce_ms0 = electron_data.q1ce_ms0; t->q1ce_ms1 = electron_data.q1ce_ms1;
    t->q1cp_ms0 = electron_data.q1cp_ms0; t->q1cp_ms1 = electron_data.q1cp_ms1;
    t->q2ce_ms0 = electron_data.q2ce_ms0; t->q2ce_ms1 = electron_data.q2ce_ms1;
    t->q2cp_ms0 = electron_data.q2cp_ms0; t->q2cp_ms1 = electron_data.q2cp_ms1;
    t->blcce0 = electron_data.blcce0; t->blcce1 = electron_data.blcce1;
    t->range_ep = electron_data.range_ep;
    t->e_array = electron_data.e_array;
    t->b2spin_min = spin_data.b2spin_min;
    t->dbeta2i = spin_data.dbeta2i;
    t->espml = spin_data.espml;
    t->dleneri = spin_data.dleneri;
    t->dqq1i = spin_data.dqq1i;
    t->spin_rej = spin_data.spin_rej;
    t->ums = mscat_data.ums_array;
    t->fms = mscat_data.fms_array;
    t->wms = mscat_data.wms_array;
    t->ims = mscat_data.ims_array;
    t->dllambi = mscat_data.dllambi;
    t->dqmsi = mscat_data.dqmsi;
    return OMC_OK;
}

#endif /* OMC_OMPMC_ADAPTER_H */
