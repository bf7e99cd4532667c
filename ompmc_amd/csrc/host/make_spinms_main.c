/* omc_make_spinms <path>: writes the synthetic spinms.data described in
 * elec.c (omc_write_synthetic_spinms). */
#include "omc_host.h"
#include <stdio.h>

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: omc_make_spinms <path/spinms.data>\n");
        return 2;
    }
    int rc = omc_write_synthetic_spinms(argv[1]);
    if (rc) {
        fprintf(stderr, "omc_make_spinms: %s\n", omc_last_error());
        return 1;
    }
    return 0;
}
