/*
 * Test-only accessor unit for the reference build in oracle/_ref.
 *
 * This file is compiled only by oracle/Makefile when the ompMC sources are
 * present at REF (the reference checkout).  It includes the reference's
 * omc_dosxyz user code in place (nothing is copied) with its main() renamed,
 * and adds small accessors so the dump harness can read the file-scope
 * globals that omc_dosxyz.c keeps (geometry, source, score).
 */
#define main omc_dosxyz_reference_main
#include "omc_dosxyz.c"
#undef main

int     refacc_isize(void) { return geometry.isize; }
int     refacc_jsize(void) { return geometry.jsize; }
int     refacc_ksize(void) { return geometry.ksize; }
double *refacc_bounds(int d) { return d == 0 ? geometry.xbounds : d == 1 ? geometry.ybounds : geometry.zbounds; }
double *refacc_densities(void) { return geometry.med_densities; }
int    *refacc_med_indices(void) { return geometry.med_indices; }
int     refacc_src_spectrum(void) { return source.spectrum; }
double *refacc_cdfinv(int w) { return w == 0 ? source.cdfinv1 : source.cdfinv2; }
double  refacc_src_scalar(int w)
{
    switch (w) {
    case 0: return source.energy;
    case 1: return source.ssd;
    case 2: return source.xinl;
    case 3: return source.xinu;
    case 4: return source.yinl;
    case 5: return source.yinu;
    case 6: return source.xsize;
    case 7: return source.ysize;
    case 8: return source.ixinl;
    case 9: return source.iyinl;
    default: return source.charge;
    }
}
double *refacc_endep(void) { return score.endep; }
double *refacc_accum(int w) { return w == 0 ? score.accum_endep : score.accum_endep2; }
void    refacc_init_history(void) { initHistory(); }
void    refacc_accum_endep(void) { accumEndep(); }
void    refacc_init_source(void) { initSource(); }
void    refacc_init_phantom(void) { initPhantom(); }
void    refacc_init_regions(void) { initRegions(); }
void    refacc_init_score(void) { initScore(); }
