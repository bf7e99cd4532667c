/*
 * .inp key/value parser with ompMC's semantics
 * (src/omc_utilities.c:63-121, 141-156).
 *
 * Reference behaviour kept on purpose so existing input files read the same:
 *   - any line containing '#' anywhere is ignored, as are blank lines;
 *   - key = text before the first '=', value = the rest of the line without
 *     the line terminator (leading blanks kept, trailing blanks kept);
 *   - lookups match the first key that contains the query as a substring;
 *   - the reference decrements its pair counter after reading and returns
 *     "not found" when it is zero, so a file with a single pair is unusable.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];

const char *omc_last_error(void) { return g_err; }

int omc_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

static int line_is_blank(const char *s)
{
    for (; *s; ++s)
        if (!isspace((unsigned char)*s)) return 0;
    return 1;
}

/* Splits one raw line into the next pair slot; returns 1 if a pair was
 * stored. */
static int take_line(omc_inp *inp, char *line)
{
    if (strchr(line, '#') || line_is_blank(line)) return 0;
    if (inp->n >= OMC_INP_PAIRS) return 0;
    /* the key ends at the first '=' or line terminator */
    size_t klen = strcspn(line, "=\r\n");
    char *rest = line + klen;
    if (*rest == '=') ++rest;
    else if (*rest) ++rest;
    size_t vlen = strcspn(rest, "\r\n");
    if (klen >= OMC_INP_BUF) klen = OMC_INP_BUF - 1;
    if (vlen >= OMC_INP_BUF) vlen = OMC_INP_BUF - 1;
    memcpy(inp->key[inp->n], line, klen);
    inp->key[inp->n][klen] = '\0';
    memcpy(inp->value[inp->n], rest, vlen);
    inp->value[inp->n][vlen] = '\0';
    inp->n++;
    return 1;
}

int omc_inp_parse_text(omc_inp *inp, const char *text)
{
    memset(inp, 0, sizeof *inp);
    const char *p = text;
    char line[OMC_INP_BUF];
    while (*p) {
        size_t len = strcspn(p, "\n");
        size_t take = len + (p[len] == '\n' ? 1 : 0);
        /* fgets semantics: at most BUF-1 characters per read */
        while (take > 0) {
            size_t chunk = take < OMC_INP_BUF - 1 ? take : OMC_INP_BUF - 1;
            memcpy(line, p, chunk);
            line[chunk] = '\0';
            take_line(inp, line);
            p += chunk;
            take -= chunk;
        }
    }
    return OMC_OK;
}

int omc_inp_read(omc_inp *inp, const char *stem)
{
    char path[1024];
    snprintf(path, sizeof path, "%s.inp", stem);
    FILE *fp = fopen(path, "r");
    if (!fp) return omc_set_error(OMC_EIO, "Unable to open file: %s", path);
    memset(inp, 0, sizeof *inp);
    char line[OMC_INP_BUF];
    while (fgets(line, sizeof line, fp)) take_line(inp, line);
    fclose(fp);
    return OMC_OK;
}

int omc_inp_get(const omc_inp *inp, const char *key, char *dest, size_t destlen)
{
    /* reference: usable pairs are [0, n-1] after its post-decrement, and a
     * zero post-decrement count means nothing is usable */
    if (inp->n - 1 <= 0) return 0;
    for (int i = 0; i < inp->n; ++i) {
        if (strstr(inp->key[i], key)) {
            snprintf(dest, destlen, "%s", inp->value[i]);
            return 1;
        }
    }
    return 0;
}

int omc_inp_get_path(const omc_inp *inp, const char *key, char *dest, size_t destlen)
{
    char raw[OMC_INP_BUF];
    if (!omc_inp_get(inp, key, raw, sizeof raw)) return 0;
    size_t o = 0;
    for (const char *s = raw; *s && o + 1 < destlen; ++s)
        if (!isspace((unsigned char)*s)) dest[o++] = *s;
    dest[o] = '\0';
    return 1;
}
