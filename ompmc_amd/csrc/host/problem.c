/*
 * omc_dosxyz problem set-up from an .inp file, in the reference's order
 * (ucodes/omc_dosxyz/omc_dosxyz.c:1155-1231): phantom, media tables,
 * source, regions, splitting, history counts and seeds.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void omc_dosxyz_problem_free(omc_dosxyz_problem *p)
{
    omc_regions_free(&p->reg);
    omc_tables_free(&p->tab);
    omc_phantom_free(&p->ph);
}

static int need(const omc_inp *inp, const char *key, char *buf, size_t n)
{
    if (omc_inp_get(inp, key, buf, n) != 1)
        return omc_set_error(OMC_EKEY, "Can not find '%s' key on input file.", key);
    return OMC_OK;
}

int omc_problem_media(omc_dosxyz_problem *p, const omc_inp *inp)
{
    char pegs[OMC_INP_BUF], pgs4[OMC_INP_BUF], data[OMC_INP_BUF], buf[OMC_INP_BUF];
    OMC_TRY(need(inp, "pegs file", buf, sizeof buf));
    omc_inp_get_path(inp, "pegs file", pegs, sizeof pegs);
    OMC_TRY(need(inp, "pgs4form file", buf, sizeof buf));
    omc_inp_get_path(inp, "pgs4form file", pgs4, sizeof pgs4);
    OMC_TRY(need(inp, "data folder", buf, sizeof buf));
    omc_inp_get_path(inp, "data folder", data, sizeof data);
    omc_data_paths paths = { pegs, pgs4, data };
    return omc_tables_build(&p->tab, p->ph.nmed, (const char (*)[64])p->ph.names, &paths);
}

int omc_problem_spectrum(omc_spectrum *s, const omc_inp *inp)
{
    char buf[OMC_INP_BUF];
    if (omc_inp_get(inp, "spectrum file", buf, sizeof buf) == 1) {
        char path[OMC_INP_BUF];
        omc_inp_get_path(inp, "spectrum file", path, sizeof path);
        return omc_spectrum_read(s, path);
    }
    OMC_TRY(need(inp, "mono energy", buf, sizeof buf));
    omc_spectrum_mono(s, atof(buf));
    return OMC_OK;
}

int omc_dosxyz_problem_load(omc_dosxyz_problem *p, const char *inp_stem)
{
    memset(p, 0, sizeof *p);
    OMC_TRY(omc_inp_read(&p->inp, inp_stem));
    return omc_dosxyz_problem_from_inp(p);
}

int omc_dosxyz_problem_from_inp(omc_dosxyz_problem *p)
{
    const omc_inp *inp = &p->inp;
    char buf[OMC_INP_BUF], path[OMC_INP_BUF];
    int rc;

    OMC_TRY(need(inp, "phantom file", buf, sizeof buf));
    omc_inp_get_path(inp, "phantom file", path, sizeof path);
    OMC_TRY(omc_phantom_read(&p->ph, path));
    if ((rc = omc_problem_media(p, inp))) goto fail;

    /* source */
    if ((rc = omc_problem_spectrum(&p->src.spec, inp))) goto fail;
    double b[4] = { 0, 0, 0, 0 };
    if ((rc = need(inp, "collimator bounds", buf, sizeof buf))) goto fail;
    sscanf(buf, "%lf %lf %lf %lf", &b[0], &b[1], &b[2], &b[3]);
    if ((rc = need(inp, "charge", buf, sizeof buf))) goto fail;
    int charge = atoi(buf);
    if ((rc = need(inp, "ssd", buf, sizeof buf))) goto fail;
    double ssd = atof(buf);
    if ((rc = omc_source_dosxyz_setup(&p->src, &p->ph, b[0], b[1], b[2], b[3], ssd, charge))) goto fail;

    /* regions */
    if ((rc = need(inp, "global ecut", buf, sizeof buf))) goto fail;
    double ecut = atof(buf);
    if ((rc = need(inp, "global pcut", buf, sizeof buf))) goto fail;
    double pcut = atof(buf);
    if ((rc = omc_regions_build(&p->reg, &p->ph, &p->tab, ecut, pcut))) goto fail;

    if ((rc = need(inp, "nsplit", buf, sizeof buf))) goto fail;
    p->nsplit = atoi(buf);

    if ((rc = need(inp, "rng seeds", buf, sizeof buf))) goto fail;
    p->seeds[0] = p->seeds[1] = 0;
    sscanf(buf, "%d %d", &p->seeds[0], &p->seeds[1]);

    /* histories: ncase read with atoi, rounded down to whole batches */
    if ((rc = need(inp, "ncase", buf, sizeof buf))) goto fail;
    int nhist = atoi(buf);
    if ((rc = need(inp, "nbatch", buf, sizeof buf))) goto fail;
    int nbatch = atoi(buf);
    if (nbatch < 1) { rc = omc_set_error(OMC_EARG, "nbatch must be >= 1"); goto fail; }
    if (nhist / nbatch == 0) nhist = nbatch;
    p->nbatch = nbatch;
    p->nperbatch = nhist / nbatch;
    p->nhist = p->nperbatch * nbatch;

    if (omc_inp_get(inp, "output folder", buf, sizeof buf) == 1)
        omc_inp_get_path(inp, "output folder", p->output_folder, sizeof p->output_folder);
    return OMC_OK;
fail:
    omc_dosxyz_problem_free(p);
    return rc;
}
