/*
 * omc_host.h - host-side (CPU, C99) half of the ompmc_amd drop-in.
 *
 * Everything the transport kernels read is built here on the host, once per
 * problem, exactly as ompMC builds it: the .inp parser, the pegs4 / xcom /
 * pgs4form / msnew / spinms loaders and the per-medium table builders, the
 * egsphant reader and region setup, the spectrum and field sources, and the
 * batch statistics plus .3ddose / sparse-Dij writers.  The device engine
 * (omc_dev.h) consumes the structures declared here.
 *
 * Table values are bit-identical to the globals ompMC fills in
 * src/ompmc.c (photon_data, rayleigh_data, pair_data, electron_data,
 * spin_data, mscat_data); the index conventions are the reference's so that
 * a lookup `tab[imed*OMC_MXGE + lgle]` means the same thing on both sides.
 *
 * Error convention: every function returns 0 on success or a negative
 * OMC_E* code and never calls exit(); omc_last_error() gives the message.
 */
#ifndef OMC_HOST_H
#define OMC_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants (values fixed by src/ompmc.h:46-437) ---------------------- */
#define OMC_RM          0.5109989461   /* electron rest energy, MeV */
#define OMC_MXGE        2000           /* photon log-energy bins */
#define OMC_MXEKE       500            /* electron log-energy bins (allocated) */
#define OMC_MXMED       9              /* media per problem */
#define OMC_MXELEMENT   50             /* elements per medium; xcom covers Z<=50 */
#define OMC_MXRAYFF     100            /* Rayleigh form-factor grid */
#define OMC_RAYCDFSIZE  100            /* Rayleigh uniform-CDF bins */
#define OMC_MXE_SPIN    15
#define OMC_NE_SPIN     (2*OMC_MXE_SPIN + 2)   /* 32 energy nodes */
#define OMC_NQ_SPIN     16
#define OMC_NU_SPIN     32
#define OMC_SPIN_PER_MED (2*OMC_NE_SPIN*OMC_NQ_SPIN*OMC_NU_SPIN)  /* 32768 */
#define OMC_NL_MS       64
#define OMC_NQ_MS       8
#define OMC_NU_MS       32
#define OMC_MS_SIZE     (OMC_NL_MS*OMC_NQ_MS*OMC_NU_MS)            /* 16384 */
#define OMC_MXEBIN      200            /* spectrum bins */
#define OMC_INVDIM      1000           /* inverse-CDF bins */

/* ---- error codes ----------------------------------------------------------- */
#define OMC_OK           0
#define OMC_EIO         -1   /* file missing / unreadable */
#define OMC_EFORMAT     -2   /* file content not as expected */
#define OMC_EKEY        -3   /* required .inp key absent */
#define OMC_EARG        -4   /* invalid argument */
#define OMC_ENOMEM      -5
#define OMC_ERANGE      -6   /* size beyond a fixed maximum */
#define OMC_EDEVICE     -7   /* HIP runtime error */
#define OMC_EOVERFLOW   -8   /* particle queue overflow on the device */
#define OMC_EBUDGET     -9   /* request does not fit the HBM budget */
#define OMC_ENOTSUP     -10  /* feature not supported by this path */

const char *omc_last_error(void);

/* ---- .inp parser (src/omc_utilities.c:63-121) ------------------------------ */
#define OMC_INP_PAIRS   80
#define OMC_INP_BUF     256
typedef struct {
    int  n;                                 /* number of key/value pairs read */
    char key[OMC_INP_PAIRS][OMC_INP_BUF];
    char value[OMC_INP_PAIRS][OMC_INP_BUF];
} omc_inp;

/* Reads "<stem>.inp".  Reference quirks kept: a line containing '#' anywhere
 * is skipped, blank lines are skipped, a file with fewer than two pairs
 * yields no usable keys. */
int omc_inp_read(omc_inp *inp, const char *stem);
/* Parses an in-memory text with the same rules (used by drivers/tests). */
int omc_inp_parse_text(omc_inp *inp, const char *text);
/* First key that CONTAINS `key` as a substring wins (reference semantics).
 * Returns 1 and copies the raw value (leading blanks kept) or returns 0. */
int omc_inp_get(const omc_inp *inp, const char *key, char *dest, size_t destlen);
/* Same, with all white space removed (file names, folders). */
int omc_inp_get_path(const omc_inp *inp, const char *key, char *dest, size_t destlen);

/* ---- media tables ------------------------------------------------------------ */
typedef struct {
    char   symbol[3];
    double z, wa, pz, rhoz;
} omc_element;

typedef struct {
    int    nmed;
    char   names[OMC_MXMED][64];

    /* pegs4 header data (src/ompmc.h:374-401) */
    int    ne[OMC_MXMED], iunrst[OMC_MXMED], epstfl[OMC_MXMED], iaprim[OMC_MXMED];
    int    msge[OMC_MXMED], mge[OMC_MXMED], mseke[OMC_MXMED], meke[OMC_MXMED];
    int    mleke[OMC_MXMED], mcmfp[OMC_MXMED], mrange[OMC_MXMED];
    double rho[OMC_MXMED], rlc[OMC_MXMED], ae[OMC_MXMED], ap[OMC_MXMED];
    double ue[OMC_MXMED], up[OMC_MXMED], te[OMC_MXMED], thmoll[OMC_MXMED];
    double pegs_delcm[OMC_MXMED];
    omc_element elements[OMC_MXMED][OMC_MXELEMENT];

    /* photon (src/ompmc.h:104-113): [imed*OMC_MXGE + j] */
    double ge0[OMC_MXMED], ge1[OMC_MXMED];
    double *gmfp0, *gmfp1, *gbr10, *gbr11, *gbr20, *gbr21, *cohe0, *cohe1;

    /* Rayleigh (src/ompmc.h:136-144) */
    double *xgrid, *fcum, *b_array, *c_array;   /* [imed*OMC_MXRAYFF + j] */
    int    *i_array;                            /* [imed*OMC_RAYCDFSIZE + j] */
    double *pmax0, *pmax1;                      /* [imed*OMC_MXGE + j] */

    /* pair / brems rejection (src/ompmc.h:156-169): [imed*8 + l] */
    double dl1[OMC_MXMED*8], dl2[OMC_MXMED*8], dl3[OMC_MXMED*8];
    double dl4[OMC_MXMED*8], dl5[OMC_MXMED*8], dl6[OMC_MXMED*8];
    double bpar0[OMC_MXMED], bpar1[OMC_MXMED], delcm[OMC_MXMED], zbrang[OMC_MXMED];

    /* electron (src/ompmc.h:202-252): [imed*OMC_MXEKE + k] */
    double eke0[OMC_MXMED], eke1[OMC_MXMED];
    double blcc[OMC_MXMED], xcc[OMC_MXMED];
    double esig_e[OMC_MXMED], psig_e[OMC_MXMED], expeke1[OMC_MXMED];
    int    sig_ismonotone[2][OMC_MXMED];        /* [qel][imed] */
    double *esig0, *esig1, *psig0, *psig1, *ededx0, *ededx1, *pdedx0, *pdedx1;
    double *ebr10, *ebr11, *pbr10, *pbr11, *pbr20, *pbr21, *tmxs0, *tmxs1;
    double *etae_ms0, *etae_ms1, *etap_ms0, *etap_ms1;
    double *q1ce_ms0, *q1ce_ms1, *q1cp_ms0, *q1cp_ms1;
    double *q2ce_ms0, *q2ce_ms1, *q2cp_ms0, *q2cp_ms1;
    double *blcce0, *blcce1;
    double *range_ep;   /* [qel*nmed*OMC_MXEKE + imed*OMC_MXEKE + k] */
    double *e_array;    /* [imed*OMC_MXEKE + k] */

    /* spin rejection (src/ompmc.h:262-269):
     * spin_rej[imed*32768 + qel*16384 + i*512 + j*32 + k] */
    double b2spin_min, dbeta2i, espml, dleneri, dqq1i;
    double *spin_rej;

    /* screened-Rutherford MS (src/ompmc.h:302-310): [i*256 + j*32 + k] */
    double *ums, *fms, *wms;
    int    *ims;
    double dllambi, dqmsi;
} omc_tables;

/* File locations for omc_tables_build. */
typedef struct {
    const char *pegs_file;      /* e.g. .../700icru.pegs4dat */
    const char *pgs4form_file;  /* .../pgs4form.dat */
    const char *data_folder;    /* folder with xcom_*.data, msnew.data, spinms.data; trailing '/' */
} omc_data_paths;

/* Builds every table for the `nmed` media named in `names` (pegs names, as
 * in an egsphant header).  Mirrors initMediaData (src/ompmc.c:5450-5487). */
int  omc_tables_build(omc_tables *t, int nmed, const char names[][64],
                      const omc_data_paths *paths);
void omc_tables_free(omc_tables *t);

/* Writes the project's synthetic spinms.data (the real EGSnrc file is not
 * distributed with ompMC).  Binary layout of src/ompmc.c:2401-2607. */
int  omc_write_synthetic_spinms(const char *path);

/* The four XCOM element data sets the photon tables are built from (photo
 * absorption, Rayleigh, pair, triplet; read as getXsecData reads them,
 * src/ompmc.c:287-345), flattened for a device builder: set s, element
 * Z = z + 1 has n[s][z] (log E, log sigma) points starting at off[s][z] in
 * le and ls. */
typedef struct {
    int     n[4][OMC_MXELEMENT];
    int64_t off[4][OMC_MXELEMENT];
    int64_t total;
    double *le, *ls;
} omc_xcom_data;
int  omc_xcom_load(omc_xcom_data *x, const char *data_folder);
void omc_xcom_free(omc_xcom_data *x);

/* Individual pieces, exposed for tests. */
double omc_kn_sigma0(double e);
void   omc_spline_setup(const double *x, const double *f, double *a, double *b,
                        double *c, double *d, int n);
double omc_spline_eval(double s, const double *x, const double *a,
                       const double *b, const double *c, const double *d, int n);

/* ---- phantom (egsphant) and regions -------------------------------------- */
typedef struct {
    int     nmed;
    char    names[OMC_MXMED][64];
    int     isize, jsize, ksize;
    double *xbounds, *ybounds, *zbounds;  /* [n+1] */
    int    *med;        /* [nvox] medium number as stored (1-based, 0 = vacuum) */
    double *rho;        /* [nvox] mass density */
} omc_phantom;

int  omc_phantom_read(omc_phantom *p, const char *path);
int  omc_phantom_write(const omc_phantom *p, const char *path);
int  omc_phantom_alloc(omc_phantom *p, int nmed, int isize, int jsize, int ksize);
void omc_phantom_free(omc_phantom *p);

/* Per-region transport data, region 0 = outside (src/ompmc.h:405-413). */
typedef struct {
    int     nreg;       /* nvox + 1 */
    int    *med;        /* -1 = vacuum */
    double *rhof, *pcut, *ecut;
} omc_regions;

/* initRegions (ucodes/omc_dosxyz/omc_dosxyz.c:890-962).  Two deliberate
 * fixes: a region whose PEGS AE exceeds the global ECUT gets AE (the
 * reference leaves it uninitialised), and vacuum voxels get zero cuts. */
int  omc_regions_build(omc_regions *r, const omc_phantom *p,
                       const omc_tables *t, double ecut, double pcut);
void omc_regions_free(omc_regions *r);

/* ---- sources ------------------------------------------------------------------ */
typedef struct {
    int     spectrum;   /* 1 = sample from cdfinv, 0 = mono */
    double  energy;     /* mono kinetic energy, MeV */
    double  deltak;     /* number of inverse-CDF bins (1000) */
    double  cdfinv1[OMC_INVDIM], cdfinv2[OMC_INVDIM];
} omc_spectrum;

/* Spectrum file reader + inverse CDF (omc_dosxyz.c:383-510). */
int omc_spectrum_read(omc_spectrum *s, const char *path);
/* The same inverse CDF from in-memory bins (upper edges, weights). */
int omc_spectrum_from_bins(omc_spectrum *s, int nbin, double enmin,
                           const double *e_upper, const double *weight, int imode);
void omc_spectrum_mono(omc_spectrum *s, double energy);

/* omc_dosxyz point source at SSD collimated to a rectangle on z=zbounds[0]
 * (omc_dosxyz.c:521-632). */
typedef struct {
    omc_spectrum spec;
    int    charge;                  /* 0 photon, -1 e-, +1 e+ */
    double ssd;
    double xinl, xinu, yinl, yinu;
    double xsize, ysize;
    int    ixinl, ixinu, iyinl, iyinu;
} omc_source_dosxyz;

int omc_source_dosxyz_setup(omc_source_dosxyz *s, const omc_phantom *p,
                            double xinl, double xinu, double yinl, double yinu,
                            double ssd, int charge);

/* matRad beamlets: bixel parallelogram at the isocentre, back-projected onto
 * the phantom surface (omc_matrad.c:696-751, 1084-1254).  Arrays are owned by
 * the caller; beam quantities indexed by ibeam[b] (0-based). */
typedef struct {
    omc_spectrum spec;
    int     charge;             /* the MEX code never sets it: always photons */
    int     nbeamlets, nbeams;
    const int    *ibeam;        /* [nbeamlets] */
    const double *src;          /* [nbeams*3]   source position */
    const double *corner;       /* [nbeamlets*3] */
    const double *side1;        /* [nbeamlets*3] */
    const double *side2;        /* [nbeamlets*3] */
} omc_source_beamlets;

/* ---- scoring and output --------------------------------------------------------- */
/* accumulateResults (omc_dosxyz.c:719-799): per-voxel mean dose (Gy per
 * history) and relative batch uncertainty from batch sums.  `sum`/`sum2`
 * are [nvox+1] with region 0 = outside; outputs dose/unc are [nvox]. */
int omc_batch_results(const omc_phantom *p, const double *sum, const double *sum2,
                      int nbatch, double inc_fluence, int iout,
                      double *dose, double *unc);
/* outputResults' file body (omc_dosxyz.c:841-879). */
int omc_write_3ddose(const char *path, const omc_phantom *p,
                     const double *dose, const double *unc);
int omc_read_3ddose(const char *path, int *isize, int *jsize, int *ksize,
                    double **dose, double **unc);

/* Equal-accuracy comparison of two dose distributions with their relative
 * uncertainties (the .3ddose contents, omc_dosxyz.c:801-886): per voxel
 * z = (d_a - d_b) / sqrt((u_a d_a)^2 + (u_b d_b)^2) over the voxels whose mean
 * of the two doses exceeds rel_thr times its maximum; pass = at least 95% of
 * them within |z| <= 2 and the integral (dose times mass if masses are given,
 * else dose) within 0.5%.  No counterpart in the reference; it is the check
 * the GPU results are held to against the reference's. */
typedef struct {
    int64_t voxels;               /* voxels compared */
    double  frac_within_2sigma;
    double  z_mean, z_rms, z_max_abs;
    double  integral_ratio;       /* a / b */
    int     pass;
} omc_dose_comparison;
int omc_compare_doses(int64_t nvox, const double *dose_a, const double *unc_a, const double *dose_b,
                      const double *unc_b, const double *mass, double rel_thr, omc_dose_comparison *out);
/* The same on two .3ddose files; `phantom` (may be NULL) gives the masses. */
int omc_compare_3ddose(const char *path_a, const char *path_b, const omc_phantom *phantom, double rel_thr,
                       omc_dose_comparison *out);

/* matRad sparse column: voxels with dose > thr*max (omc_matrad.c:1420-1479).
 * Writes row indices (0-based voxel) and values; returns the count. */
int64_t omc_sparse_column(const double *dose, int64_t nvox, double rel_thr,
                          int64_t *rows, double *vals, int64_t cap);

/* ---- omc_dosxyz problem from an .inp file ---------------------------------------- */
typedef struct {
    omc_inp           inp;
    omc_phantom       ph;
    omc_tables        tab;
    omc_regions       reg;
    omc_source_dosxyz src;
    int  nsplit;
    int  seeds[2];                 /* "rng seeds" */
    int  nhist, nbatch, nperbatch; /* nhist = nperbatch * nbatch */
    char output_folder[OMC_INP_BUF];
} omc_dosxyz_problem;

/* Reads <stem>.inp and performs initPhantom, initMediaData, initSource,
 * initRegions and initVrt (omc_dosxyz.c:1155-1231). */
int  omc_dosxyz_problem_load(omc_dosxyz_problem *p, const char *inp_stem);
/* Same, with p->inp already filled (e.g. by omc_inp_parse_text). */
int  omc_dosxyz_problem_from_inp(omc_dosxyz_problem *p);
void omc_dosxyz_problem_free(omc_dosxyz_problem *p);
/* Pieces shared with the matRad driver. */
int  omc_problem_media(omc_dosxyz_problem *p, const omc_inp *inp);
int  omc_problem_spectrum(omc_spectrum *s, const omc_inp *inp);

#ifdef __cplusplus
}
#endif
#endif /* OMC_HOST_H */
