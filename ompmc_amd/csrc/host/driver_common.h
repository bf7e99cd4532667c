/*
 * driver_common.h - helpers shared by the omc_dosxyz and omc_matrad drivers:
 * command line, wall clock, device set-up and fatal-error reporting.
 */
#ifndef OMC_DRIVER_COMMON_H
#define OMC_DRIVER_COMMON_H

#include "omc_dev.h"
#include "omc_host.h"

typedef struct {
    const char *input;      /* -i / --input: .inp stem */
    const char *output;     /* -o / --output: output file stem */
    int verbose;            /* --verbose / --brief */
    int precision;          /* --fp32 / --fp64 (default from OMC_PRECISION, else FP64) */
    int device;             /* --device N */
    int woodcock;           /* OMC_WOODCOCK: 0/off voxel-stepped photon flights, 1/on delta
                               tracking; unset = the engine's default (on for FP32 only) */
} drv_args;

/* getopt_long over the reference's options (omc_dosxyz.c:1078-1150) plus
 * --fp32/--fp64/--device; echoes them the way the reference does. */
void   drv_parse_args(drv_args *a, int argc, char **argv);
double drv_time(void);
/* Prints "<what>: <message>" with the library's last error and exits. */
void   drv_die(const char *what, int rc, int device_error);
/* Creates a device context and uploads the media tables. */
omc_dev_ctx *drv_open_device(const drv_args *a, const omc_tables *tab);
/* ncase / nbatch as the reference reads them: atoi, at least one history
 * per batch, rounded down to whole batches (omc_dosxyz.c:1206-1225). */
int    drv_histories(const omc_inp *inp, int *nhist, int *nbatch, int *nperbatch);
/* Input stem -> output path: "<output folder><output><ext>". */
int    drv_output_path(const omc_inp *inp, const char *output, const char *ext, char *path, int n);

#endif
