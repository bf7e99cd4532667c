/*
 * omc_compare3ddose a.3ddose b.3ddose [-p phantom.egsphant] [-t rel_threshold]
 *
 * Compares two dose distributions written by omc_dosxyz (this engine's or
 * the reference's): the fraction of the voxels above the threshold (default
 * 10% of the maximum) whose doses agree within 2 combined sigma, z
 * statistics and the integral ratio (energy, with the phantom's masses).
 * Prints key = value lines; exit status 0 when the 95% / 0.5% gates pass,
 * 1 when they fail, 2 on an error.
 */
#include "omc_host.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int main(int argc, char **argv)
{
    const char *a = NULL, *b = NULL, *phantom = NULL;
    double thr = 0.1;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-p") && i + 1 < argc) phantom = argv[++i];
        else if (!strcmp(argv[i], "-t") && i + 1 < argc) thr = atof(argv[++i]);
        else if (!a) a = argv[i];
        else if (!b) b = argv[i];
    }
    if (!a || !b) {
        fprintf(stderr, "usage: omc_compare3ddose a.3ddose b.3ddose [-p phantom.egsphant] [-t threshold]\n");
        return 2;
    }
    omc_phantom ph;
    memset(&ph, 0, sizeof ph);
    if (phantom && omc_phantom_read(&ph, phantom)) {
        fprintf(stderr, "%s\n", omc_last_error());
        return 2;
    }
    omc_dose_comparison r;
    const int rc = omc_compare_3ddose(a, b, phantom ? &ph : NULL, thr, &r);
    if (phantom) omc_phantom_free(&ph);
    if (rc) {
        fprintf(stderr, "%s\n", omc_last_error());
        return 2;
    }
    printf("voxels = %lld\nfraction_within_2sigma = %.6f\nz_mean = %.4f\nz_rms = %.4f\nz_max_abs = %.3f\n"
           "integral_ratio = %.6f\npass = %d\n", (long long)r.voxels, r.frac_within_2sigma, r.z_mean, r.z_rms,
           r.z_max_abs, r.integral_ratio, r.pass);
    return r.pass ? 0 : 1;
}
