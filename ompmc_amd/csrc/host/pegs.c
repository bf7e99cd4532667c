/*
 * pegs4dat reader: the subset of a PEGS4 medium block that the transport
 * uses (src/ompmc.c:5489-5943 readPegsFile).
 *
 * The file is held in memory and walked with two cursors that behave like
 * the stdio calls the reference mixes: `next_line` acts as fgets (a line of
 * at most 255 characters) and `next_number` as fscanf("%lf ") (skip blanks,
 * convert, skip trailing blanks including newlines).  Mixing them the same
 * way keeps every parsed double identical, and keeps the reference's
 * medium-search behaviour after the numeric block.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    const char *buf;
    size_t len, pos;
} cursor;

static int next_line(cursor *c, char *out, size_t outlen)
{
    if (c->pos >= c->len) return 0;
    size_t o = 0;
    while (c->pos < c->len && o + 1 < outlen) {
        char ch = c->buf[c->pos++];
        out[o++] = ch;
        if (ch == '\n') break;
    }
    out[o] = '\0';
    return 1;
}

static int next_number(cursor *c, double *v)
{
    while (c->pos < c->len && isspace((unsigned char)c->buf[c->pos])) c->pos++;
    if (c->pos >= c->len) return 0;
    char *end;
    *v = strtod(c->buf + c->pos, &end);
    if (end == c->buf + c->pos) return 0;
    c->pos = (size_t)(end - c->buf);
    while (c->pos < c->len && isspace((unsigned char)c->buf[c->pos])) c->pos++;
    return 1;
}

static void trim_copy(char *dst, size_t cap, const char *s, size_t n)
{
    size_t o = 0;
    for (size_t i = 0; i < n && s[i]; ++i)
        if (s[i] != ' ' && s[i] != '\t' && o + 1 < cap) dst[o++] = s[i];
    dst[o] = '\0';
}

/* Visits the comma-separated NAME=VALUE items of one header line. */
typedef int (*item_fn)(void *ctx, const char *name, const char *value);

static int for_each_item(const char *line, item_fn fn, void *ctx)
{
    char work[OMC_INP_BUF];
    snprintf(work, sizeof work, "%s", line);
    char *save = NULL;
    for (char *tok = strtok_r(work, ",", &save); tok; tok = strtok_r(NULL, ",", &save)) {
        char *eq = strchr(tok, '=');
        char name[64];
        if (!eq) {
            trim_copy(name, sizeof name, tok, strlen(tok));
            continue;
        }
        trim_copy(name, sizeof name, tok, (size_t)(eq - tok));
        /* the value ends at a further '=' (strtok on "=") */
        char value[OMC_INP_BUF];
        size_t vlen = strcspn(eq + 1, "=");
        snprintf(value, sizeof value, "%.*s", (int)vlen, eq + 1);
        int rc = fn(ctx, name, value);
        if (rc) return rc;
    }
    return 0;
}

typedef struct { omc_tables *t; int imed; } medium_ctx;

static int medium_item(void *vctx, const char *name, const char *value)
{
    medium_ctx *m = vctx;
    if (!strcmp(name, "RHO")) {
        if (sscanf(value, "%lf", &m->t->rho[m->imed]) != 1) return 1;
    } else if (!strcmp(name, "NE")) {
        unsigned u;
        if (sscanf(value, "%u", &u) != 1) return 1;
        m->t->ne[m->imed] = (int)u;
    } else if (!strcmp(name, "IUNRST")) {
        if (sscanf(value, "%d", &m->t->iunrst[m->imed]) != 1) return 1;
    } else if (!strcmp(name, "EPSTFL")) {
        if (sscanf(value, "%d", &m->t->epstfl[m->imed]) != 1) return 1;
    } else if (!strcmp(name, "IAPRIM")) {
        if (sscanf(value, "%d", &m->t->iaprim[m->imed]) != 1) return 1;
    }
    return 0;
}

static int element_item(void *vctx, const char *name, const char *value)
{
    omc_element *el = vctx;
    if (!strcmp(name, "ASYM")) {
        if (strlen(value) < 2) return 1;
        el->symbol[0] = value[0];
        el->symbol[1] = value[1];
        el->symbol[2] = '\0';
    } else if (!strcmp(name, "Z")) {
        if (sscanf(value, "%lf", &el->z) != 1) return 1;
    } else if (!strcmp(name, "A")) {
        if (sscanf(value, "%lf", &el->wa) != 1) return 1;
    } else if (!strcmp(name, "PZ")) {
        if (sscanf(value, "%lf", &el->pz) != 1) return 1;
    } else if (!strcmp(name, "RHOZ")) {
        if (sscanf(value, "%lf", &el->rhoz) != 1) return 1;
    }
    return 0;
}

/* Reads the numeric part of one medium; returns 1 when the medium is
 * complete, 0 when the block is malformed (the medium stays not found). */
static int read_medium_body(cursor *c, omc_tables *t, int imed)
{
    char line[OMC_INP_BUF];
    medium_ctx mctx = { t, imed };
    t->ne[imed] = 0;
    if (!next_line(c, line, sizeof line)) return 0;
    if (for_each_item(line, medium_item, &mctx)) return 0;
    if (t->ne[imed] > OMC_MXELEMENT) return 0;

    for (int m = 0; m < t->ne[imed]; ++m) {
        omc_element el;
        memset(&el, 0, sizeof el);
        if (!next_line(c, line, sizeof line)) return 0;
        if (for_each_item(line, element_item, &el)) return 0;
        t->elements[imed][m] = el;
    }

    if (!next_line(c, line, sizeof line)) return 0;
    if (sscanf(line, "%lf %lf %lf %lf %lf", &t->rlc[imed], &t->ae[imed],
               &t->ap[imed], &t->ue[imed], &t->up[imed]) != 5)
        return 0;
    t->te[imed] = t->ae[imed] - OMC_RM;
    t->thmoll[imed] = (t->te[imed]) * 2 + OMC_RM;

    if (!next_line(c, line, sizeof line)) return 0;
    if (sscanf(line, "%d %d %d %d %d %d %d", &t->msge[imed], &t->mge[imed],
               &t->mseke[imed], &t->meke[imed], &t->mleke[imed],
               &t->mcmfp[imed], &t->mrange[imed]) != 7)
        return 0;
    if (t->meke[imed] > OMC_MXEKE) return 0;

    /* seven lines of brems/pair fit parameters; the last one must hold five
     * numbers (the reference only validates it) */
    for (int i = 0; i < 7; ++i)
        if (!next_line(c, line, sizeof line)) return 0;
    double d[5];
    if (sscanf(line, "%lf %lf %lf %lf %lf", &d[0], &d[1], &d[2], &d[3], &d[4]) != 5)
        return 0;

    /* free-format numbers: dl6, delcm, alphi1, alphi2, bpar1, bpar2,
     * delpos1, delpos2, xr0, teff0, blcc, xcc, eke0, eke1 */
    double hdr[14];
    for (int i = 0; i < 14; ++i)
        if (!next_number(c, &hdr[i])) return 0;
    t->pegs_delcm[imed] = hdr[1];
    t->blcc[imed] = hdr[10];
    t->xcc[imed] = hdr[11];
    t->eke0[imed] = hdr[12];
    t->eke1[imed] = hdr[13];

    int neke = t->meke[imed];
    for (int k = 0; k < neke; ++k) {
        size_t ix = (size_t)imed * OMC_MXEKE + k;
        double *dst[16] = {
            &t->esig0[ix], &t->esig1[ix], &t->psig0[ix], &t->psig1[ix],
            &t->ededx0[ix], &t->ededx1[ix], &t->pdedx0[ix], &t->pdedx1[ix],
            &t->ebr10[ix], &t->ebr11[ix], &t->pbr10[ix], &t->pbr11[ix],
            &t->pbr20[ix], &t->pbr21[ix], &t->tmxs0[ix], &t->tmxs1[ix] };
        for (int q = 0; q < 16; ++q)
            if (!next_number(c, dst[q])) return 0;
    }

    /* lengths in cm: PEGS tabulates per radiation length */
    double dfacti = 1.0 / (t->rlc[imed]);
    t->blcc[imed] *= dfacti;
    for (int k = 0; k < neke; ++k) {
        size_t ix = (size_t)imed * OMC_MXEKE + k;
        t->esig0[ix] *= dfacti;
        t->psig0[ix] *= dfacti;
        t->ededx0[ix] *= dfacti;
        t->pdedx0[ix] *= dfacti;
        t->pdedx1[ix] *= dfacti;
        t->esig1[ix] *= dfacti;
        t->psig1[ix] *= dfacti;
        t->ededx1[ix] *= dfacti;
    }
    t->xcc[imed] *= sqrt(dfacti);
    return 1;
}

int omc_read_pegs(omc_tables *t, const char *path)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    rewind(fp);
    char *buf = malloc((size_t)n + 1);
    if (!buf) { fclose(fp); return omc_set_error(OMC_ENOMEM, "out of memory reading %s", path); }
    size_t got = fread(buf, 1, (size_t)n, fp);
    fclose(fp);
    buf[got] = '\0';

    cursor c = { buf, got, 0 };
    int found[OMC_MXMED] = { 0 };
    int nfound = 0;
    char line[OMC_INP_BUF];

    while (nfound < t->nmed && next_line(&c, line, sizeof line)) {
        if (strncmp(line, " MEDIUM=", 8) != 0) continue;
        char name[25];
        int j = 0;
        for (int i = 0; i < 24 && line[8 + i] && line[8 + i] != ' '; ++i)
            name[j++] = line[8 + i];
        name[j] = '\0';
        int imed = -1;
        for (int i = 0; i < t->nmed; ++i) {
            if (strncmp(name, t->names[i], 20) == 0 && strlen(t->names[i]) == strlen(name)) {
                imed = i;
                break;
            }
        }
        if (imed < 0) continue;
        if (read_medium_body(&c, t, imed)) {
            if (!found[imed]) nfound++;
            found[imed] = 1;
        }
    }
    free(buf);

    if (nfound < t->nmed) {
        char missing[256] = "";
        for (int i = 0; i < t->nmed; ++i)
            if (!found[i]) {
                strncat(missing, " ", sizeof missing - strlen(missing) - 1);
                strncat(missing, t->names[i], sizeof missing - strlen(missing) - 1);
            }
        return omc_set_error(OMC_EFORMAT, "media not found in pegs file %s:%s", path, missing);
    }
    return OMC_OK;
}
