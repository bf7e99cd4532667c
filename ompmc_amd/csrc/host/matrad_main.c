/*
 * omc_matrad - the matRad dose-influence computation of ompMC's MEX
 * function (omc_matrad.c:1258-1543) as a standalone program, transport on
 * the GPU.
 *
 *     omc_matrad -i <input stem> -o <output name> [--fp32] [--device N] [--verbose]
 *
 * MATLAB hands the MEX a CT cube and the mcSrc struct; here the phantom is an
 * egsphant file and mcSrc is a text file, with the same fields:
 *
 *   .inp keys: phantom file, pegs file, pgs4form file, data folder,
 *              global ecut, global pcut, spectrum file | mono energy,
 *              ncase (histories per beamlet), nbatch, rng seeds, nsplit,
 *              relative dose threshold, beamlet file, output folder
 *   beamlet file: "nBixels nBeams", then nBeams lines "xSource ySource
 *              zSource", then nBixels lines "iBeam xCorner yCorner zCorner
 *              xSide1 ySide1 zSide1 xSide2 ySide2 zSide2 [nHist]" with iBeam
 *              1-based as in matRad (omc_matrad.c:711-714); the optional
 *              nHist replaces ncase for that beamlet.  '#' starts a comment.
 *
 * Output <output folder><output name>.dij, the sparse matrix matRad gets from
 * the MEX (nvox rows, one column per beamlet, omc_matrad.c:1369-1479):
 *   char magic[8] = "OMCDIJ1", int64 nrows, ncols, nnz,
 *   int64 jcs[ncols+1], int64 irs[nnz] (0-based voxel), double vals[nnz].
 * All beamlets run in one device pass; each column is thresholded and
 * compacted on the device.
 */
#include "driver_common.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nb, nbeams;
    int *ibeam;
    double *src, *corner, *side1, *side2;
    int64_t *nhist;     /* -1 where the file gives none */
} beamlet_file;

/* next non-comment token */
static int next_token(FILE *fp, char *tok, int n)
{
    int c, k = 0;
    for (;;) {
        c = fgetc(fp);
        if (c == EOF) return 0;
        if (c == '#') { while (c != '\n' && c != EOF) c = fgetc(fp); continue; }
        if (c > ' ') break;
    }
    do {
        if (k < n - 1) tok[k++] = (char)c;
        c = fgetc(fp);
    } while (c != EOF && c > ' ' && c != '#');
    if (c == '#') ungetc(c, fp);
    tok[k] = 0;
    return 1;
}

static int read_line_values(FILE *fp, double *v, int maxv)
{
    /* reads the numbers of one logical line (up to newline) */
    char line[4096];
    for (;;) {
        if (!fgets(line, sizeof line, fp)) return -1;
        char *h = strchr(line, '#');
        if (h) *h = 0;
        int n = 0;
        char *s = line, *end;
        while (n < maxv) {
            double x = strtod(s, &end);
            if (end == s) break;
            v[n++] = x;
            s = end;
        }
        if (n > 0) return n;
    }
}

static int read_beamlets(const char *path, beamlet_file *b)
{
    memset(b, 0, sizeof *b);
    FILE *fp = fopen(path, "r");
    if (!fp) { printf("Unable to open file: %s\n", path); return OMC_EIO; }
    char tok[128];
    if (!next_token(fp, tok, sizeof tok)) goto bad;
    b->nb = atoi(tok);
    if (!next_token(fp, tok, sizeof tok)) goto bad;
    b->nbeams = atoi(tok);
    if (b->nb < 1 || b->nbeams < 1) goto bad;
    b->ibeam = malloc(sizeof(int) * (size_t)b->nb);
    b->src = malloc(sizeof(double) * 3 * (size_t)b->nbeams);
    b->corner = malloc(sizeof(double) * 3 * (size_t)b->nb);
    b->side1 = malloc(sizeof(double) * 3 * (size_t)b->nb);
    b->side2 = malloc(sizeof(double) * 3 * (size_t)b->nb);
    b->nhist = malloc(sizeof(int64_t) * (size_t)b->nb);
    if (!b->ibeam || !b->src || !b->corner || !b->side1 || !b->side2 || !b->nhist) { fclose(fp); return OMC_ENOMEM; }
    double v[16];
    for (int i = 0; i < b->nbeams; ++i) {
        if (read_line_values(fp, v, 16) != 3) goto bad;
        memcpy(b->src + 3 * i, v, 3 * sizeof(double));
    }
    for (int i = 0; i < b->nb; ++i) {
        int n = read_line_values(fp, v, 16);
        if (n != 10 && n != 11) goto bad;
        b->ibeam[i] = (int)v[0] - 1;            /* matRad's 1-based beam index */
        if (b->ibeam[i] < 0 || b->ibeam[i] >= b->nbeams) goto bad;
        memcpy(b->corner + 3 * i, v + 1, 3 * sizeof(double));
        memcpy(b->side1 + 3 * i, v + 4, 3 * sizeof(double));
        memcpy(b->side2 + 3 * i, v + 7, 3 * sizeof(double));
        b->nhist[i] = n == 11 ? (int64_t)v[10] : -1;
    }
    fclose(fp);
    return OMC_OK;
bad:
    fclose(fp);
    printf("Beamlet file %s is malformed.\n", path);
    return OMC_EFORMAT;
}

int main(int argc, char **argv)
{
    const double tbegin = drv_time();
    drv_args a;
    drv_parse_args(&a, argc, argv);

    omc_dosxyz_problem *p = calloc(1, sizeof *p);   /* holds inp, phantom, tables, regions */
    if (!p) return EXIT_FAILURE;
    int rc = omc_inp_read(&p->inp, a.input);
    if (rc) drv_die("reading the input", rc, 0);
    const omc_inp *inp = &p->inp;
    char buf[OMC_INP_BUF], path[OMC_INP_BUF];

    if (omc_inp_get(inp, "phantom file", buf, sizeof buf) != 1) { printf("Can not find 'phantom file' key on input file.\n"); return EXIT_FAILURE; }
    omc_inp_get_path(inp, "phantom file", path, sizeof path);
    if ((rc = omc_phantom_read(&p->ph, path))) drv_die("reading the phantom", rc, 0);
    if ((rc = omc_problem_media(p, inp))) drv_die("media data", rc, 0);
    omc_source_beamlets src;
    memset(&src, 0, sizeof src);
    if ((rc = omc_problem_spectrum(&src.spec, inp))) drv_die("source spectrum", rc, 0);
    double ecut, pcut;
    if (omc_inp_get(inp, "global ecut", buf, sizeof buf) != 1) { printf("Can not find 'global ecut' key on input file.\n"); return EXIT_FAILURE; }
    ecut = atof(buf);
    if (omc_inp_get(inp, "global pcut", buf, sizeof buf) != 1) { printf("Can not find 'global pcut' key on input file.\n"); return EXIT_FAILURE; }
    pcut = atof(buf);
    if ((rc = omc_regions_build(&p->reg, &p->ph, &p->tab, ecut, pcut))) drv_die("regions", rc, 0);
    int nsplit = 1;
    if (omc_inp_get(inp, "nsplit", buf, sizeof buf) == 1 && atoi(buf) > 1) nsplit = atoi(buf);
    int seeds[2] = { 97, 33 };
    if (omc_inp_get(inp, "rng seeds", buf, sizeof buf) == 1) sscanf(buf, "%d %d", &seeds[0], &seeds[1]);
    int nhist, nbatch, nperbatch;
    if ((rc = drv_histories(inp, &nhist, &nbatch, &nperbatch))) return EXIT_FAILURE;
    if (omc_inp_get(inp, "relative dose threshold", buf, sizeof buf) != 1) {
        printf("Can not find 'relative dose threshold' key on input file.\n");
        return EXIT_FAILURE;
    }
    const double rel_thr = atof(buf);
    if (omc_inp_get(inp, "beamlet file", buf, sizeof buf) != 1) { printf("Can not find 'beamlet file' key on input file.\n"); return EXIT_FAILURE; }
    omc_inp_get_path(inp, "beamlet file", path, sizeof path);
    beamlet_file bf;
    if ((rc = read_beamlets(path, &bf))) return EXIT_FAILURE;
    src.charge = 0;
    src.nbeamlets = bf.nb;
    src.nbeams = bf.nbeams;
    src.ibeam = bf.ibeam;
    src.src = bf.src;
    src.corner = bf.corner;
    src.side1 = bf.side1;
    src.side2 = bf.side2;
    int64_t *nh = malloc(sizeof(int64_t) * (size_t)bf.nb);
    int64_t total = 0;
    for (int i = 0; i < bf.nb; ++i) { nh[i] = bf.nhist[i] >= 0 ? bf.nhist[i] : nhist; total += nh[i]; }

    printf("Total Number of Beamlets:%d\n", bf.nb);
    printf("Histories per beamlet: %d (%lld in total)\n", nhist, (long long)total);
    printf("Number of statistical batches: %d\n", nbatch);
    printf("Using a relative dose cut-off of %f\n", rel_thr);

    omc_dev_ctx *ctx = drv_open_device(&a, &p->tab);
    int pid = -1;
    if ((rc = omc_dev_upload_phantom(ctx, &p->ph, &p->reg, &pid))) drv_die("phantom upload", rc, 1);
    printf("Execution time up to this point : %8.2f seconds\n", drv_time() - tbegin);

    omc_dev_stats st;
    omc_dev_run_params rp;
    memset(&rp, 0, sizeof rp);
    rp.phantom_id = pid;
    rp.nsplit = nsplit;
    rp.seed[0] = (uint32_t)seeds[0];
    rp.seed[1] = (uint32_t)seeds[1];
    rp.nbatch = nbatch;
    rp.stats = &st;
    const int64_t nvox = (int64_t)p->ph.isize * p->ph.jsize * p->ph.ksize;
    int64_t cap = nvox * bf.nb / 100 + 1, nnz = 0;
    int64_t *jcs = malloc(sizeof(int64_t) * (size_t)(bf.nb + 1));
    int64_t *irs = malloc(sizeof(int64_t) * (size_t)cap);
    double *vals = malloc(sizeof(double) * (size_t)cap);
    rc = omc_dev_beamlets_sparse(ctx, &src, nh, &rp, rel_thr, jcs, irs, vals, cap, &nnz);
    if (rc == OMC_ERANGE && nnz > cap) {
        cap = nnz;
        irs = realloc(irs, sizeof(int64_t) * (size_t)cap);
        vals = realloc(vals, sizeof(double) * (size_t)cap);
        rc = omc_dev_beamlets_sparse(ctx, NULL, NULL, NULL, rel_thr, jcs, irs, vals, cap, &nnz);
    }
    if (rc) drv_die("beamlet transport", rc, 1);
    printf("Simulation finished\n");
    printf("Transport: %.1f ms, %.4g histories/s, %lld nonzeros\n", st.transport_ms,
           st.transport_ms > 0 ? 1e3 * (double)st.histories / st.transport_ms : 0.0, (long long)nnz);

    char out[2 * OMC_INP_BUF];
    if (drv_output_path(inp, a.output, ".dij", out, sizeof out)) return EXIT_FAILURE;
    FILE *fp = fopen(out, "wb");
    if (!fp) { printf("Unable to open file: %s\n", out); return EXIT_FAILURE; }
    const char magic[8] = "OMCDIJ1";
    int64_t hdr[3] = { nvox, bf.nb, nnz };
    int ok = fwrite(magic, 1, 8, fp) == 8 && fwrite(hdr, sizeof(int64_t), 3, fp) == 3 &&
             fwrite(jcs, sizeof(int64_t), (size_t)bf.nb + 1, fp) == (size_t)bf.nb + 1 &&
             fwrite(irs, sizeof(int64_t), (size_t)nnz, fp) == (size_t)nnz &&
             fwrite(vals, sizeof(double), (size_t)nnz, fp) == (size_t)nnz;
    if (fclose(fp) != 0) ok = 0;
    if (!ok) { printf("Unable to write file: %s\n", out); return EXIT_FAILURE; }

    omc_dev_destroy(ctx);
    omc_dosxyz_problem_free(p);
    free(p); free(nh); free(jcs); free(irs); free(vals);
    free(bf.ibeam); free(bf.src); free(bf.corner); free(bf.side1); free(bf.side2); free(bf.nhist);
    printf("Total execution time : %8.5f seconds\n", drv_time() - tbegin);
    return EXIT_SUCCESS;
}
