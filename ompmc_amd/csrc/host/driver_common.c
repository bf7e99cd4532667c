/*
 * driver_common.c - shared pieces of the omc_dosxyz and omc_matrad drivers.
 */
#include "driver_common.h"

#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

double drv_time(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void drv_parse_args(drv_args *a, int argc, char **argv)
{
    /* the engine drives 6 pipeline streams by default (up to 12 after
     * tuning) plus a helper stream: give the process 16 hardware queues
     * (HIP's default is 4) before HIP starts, so that no two pipelines
     * share one */
    const char *q = getenv("GPU_MAX_HW_QUEUES");
    if (!q || atoi(q) < 16) setenv("GPU_MAX_HW_QUEUES", "16", 1);
    memset(a, 0, sizeof *a);
    const char *env = getenv("OMC_PRECISION");
    a->precision = (env && !strcmp(env, "fp32")) ? OMC_PREC_FP32 : OMC_PREC_FP64;
    const char *wc = getenv("OMC_WOODCOCK");
    if (wc && (!strcmp(wc, "0") || !strcmp(wc, "off"))) a->woodcock = -1;
    else if (wc && (!strcmp(wc, "1") || !strcmp(wc, "on"))) a->woodcock = 1;
    static int verbose_flag = 0;
    enum { OPT_FP32 = 1000, OPT_FP64, OPT_DEVICE };
    static struct option opts[] = {
        { "verbose", no_argument, &verbose_flag, 1 },
        { "brief", no_argument, &verbose_flag, 0 },
        { "input", required_argument, 0, 'i' },
        { "output", required_argument, 0, 'o' },
        { "fp32", no_argument, 0, OPT_FP32 },
        { "fp64", no_argument, 0, OPT_FP64 },
        { "device", required_argument, 0, OPT_DEVICE },
        { 0, 0, 0, 0 }
    };
    for (;;) {
        int idx = 0;
        int c = getopt_long(argc, argv, "i:o:", opts, &idx);
        if (c == -1) break;
        switch (c) {
        case 0: break;
        case 'i': a->input = optarg; printf("option -i with value `%s'\n", optarg); break;
        case 'o': a->output = optarg; printf("option -o with value `%s'\n", optarg); break;
        case OPT_FP32: a->precision = OMC_PREC_FP32; break;
        case OPT_FP64: a->precision = OMC_PREC_FP64; break;
        case OPT_DEVICE: a->device = atoi(optarg); break;
        case '?': break;
        default: exit(EXIT_FAILURE);
        }
    }
    a->verbose = verbose_flag;
    if (a->verbose) puts("verbose flag is set");
    if (optind < argc) {
        printf("non-option ARGV-elements: ");
        while (optind < argc) printf("%s ", argv[optind++]);
        putchar('\n');
    }
    if (!a->input || !a->output) {
        fprintf(stderr, "usage: %s -i <input stem> -o <output name> [--fp32|--fp64] [--device N] [--verbose]\n",
                argv[0]);
        exit(EXIT_FAILURE);
    }
}

void drv_die(const char *what, int rc, int device_error)
{
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, device_error ? omc_dev_last_error() : omc_last_error());
    exit(EXIT_FAILURE);
}

omc_dev_ctx *drv_open_device(const drv_args *a, const omc_tables *tab)
{
    omc_dev_opts o;
    memset(&o, 0, sizeof o);
    o.device = a->device;
    o.precision = a->precision;
    o.woodcock = a->woodcock;
    omc_dev_ctx *ctx = NULL;
    int rc = omc_dev_create(&ctx, &o);
    if (rc) drv_die("device set-up", rc, 1);
    rc = omc_dev_upload_media(ctx, tab);
    if (rc) drv_die("media upload", rc, 1);
    const int delta = a->woodcock > 0 || (a->woodcock == 0 && a->precision == OMC_PREC_FP32);
    printf("Transport on HIP device %d, %s, photon flights by %s\n", a->device,
           a->precision == OMC_PREC_FP32 ? "FP32 transport with FP64 tallies" : "FP64 transport",
           delta ? "delta tracking" : "voxel stepping");
    return ctx;
}

int drv_histories(const omc_inp *inp, int *nhist, int *nbatch, int *nperbatch)
{
    char buf[OMC_INP_BUF];
    if (omc_inp_get(inp, "ncase", buf, sizeof buf) != 1) {
        printf("Can not find 'ncase' key on input file.\n");
        return OMC_EFORMAT;
    }
    int n = atoi(buf);
    if (omc_inp_get(inp, "nbatch", buf, sizeof buf) != 1) {
        printf("Can not find 'nbatch' key on input file.\n");
        return OMC_EFORMAT;
    }
    int nb = atoi(buf);
    if (nb < 1) {
        printf("'nbatch' must be at least 1.\n");
        return OMC_EFORMAT;
    }
    if (n / nb == 0) n = nb;
    *nbatch = nb;
    *nperbatch = n / nb;
    *nhist = *nperbatch * nb;
    return OMC_OK;
}

int drv_output_path(const omc_inp *inp, const char *output, const char *ext, char *path, int n)
{
    char folder[OMC_INP_BUF];
    if (omc_inp_get_path(inp, "output folder", folder, sizeof folder) != 1) {
        printf("Can not find 'output folder' key on input file.\n");
        return OMC_EFORMAT;
    }
    if (snprintf(path, (size_t)n, "%s%s%s", folder, output, ext) >= n) return OMC_ERANGE;
    return OMC_OK;
}
