/*
 * Comparison of two dose distributions with their uncertainties, as
 * omc_dosxyz writes them to .3ddose files (omc_dosxyz.c:801-886): the
 * equal-accuracy gates the device engine is held to against the FP64
 * reference, usable on any pair of results.
 *
 *   z = (d_a - d_b) / sqrt((u_a d_a)^2 + (u_b d_b)^2)     per voxel
 *
 * over the voxels whose mean of the two doses is above rel_thr times the
 * largest such mean (choosing on one of them would favour the voxels where
 * it came out high); pass = at least 95% of them with |z| <= 2 and the
 * integral energy (dose times voxel mass, or plain dose sums without the
 * masses) within 0.5%.
 */
#include "omc_host.h"
#include "omc_internal.h"

#include <math.h>
#include <stdlib.h>

int omc_compare_doses(int64_t nvox, const double *da, const double *ua, const double *db, const double *ub,
                      const double *mass, double rel_thr, omc_dose_comparison *out)
{
    if (!da || !ua || !db || !ub || !out || nvox < 1) return omc_set_error(OMC_EARG, "null argument");
    double top = 0.0;
    for (int64_t v = 0; v < nvox; ++v) {
        const double m = 0.5 * (da[v] + db[v]);
        if (m > top) top = m;
    }
    const double thr = rel_thr * top;
    int64_t n = 0, inside = 0;
    double zs = 0.0, zz = 0.0, zmax = 0.0, ea = 0.0, eb = 0.0;
    for (int64_t v = 0; v < nvox; ++v) {
        const double w = mass ? mass[v] : 1.0;
        ea += w * da[v];
        eb += w * db[v];
        if (!(0.5 * (da[v] + db[v]) > thr)) continue;
        const double sa = ua[v] * da[v], sb = ub[v] * db[v];
        const double s = sqrt(sa * sa + sb * sb);
        const double z = s > 0.0 ? (da[v] - db[v]) / s : (da[v] == db[v] ? 0.0 : INFINITY);
        ++n;
        if (fabs(z) <= 2.0) ++inside;
        zs += z;
        zz += z * z;
        if (fabs(z) > zmax) zmax = fabs(z);
    }
    out->voxels = n;
    out->frac_within_2sigma = n ? (double)inside / (double)n : 0.0;
    out->z_mean = n ? zs / (double)n : 0.0;
    out->z_rms = n ? sqrt(zz / (double)n) : 0.0;
    out->z_max_abs = zmax;
    out->integral_ratio = eb != 0.0 ? ea / eb : 0.0;
    out->pass = n > 0 && out->frac_within_2sigma >= 0.95 && fabs(out->integral_ratio - 1.0) <= 0.005;
    return OMC_OK;
}

int omc_compare_3ddose(const char *path_a, const char *path_b, const omc_phantom *phantom, double rel_thr,
                       omc_dose_comparison *out)
{
    int na[3], nb[3];
    double *da = NULL, *ua = NULL, *db = NULL, *ub = NULL, *mass = NULL;
    int rc = omc_read_3ddose(path_a, &na[0], &na[1], &na[2], &da, &ua);
    if (rc) return rc;
    rc = omc_read_3ddose(path_b, &nb[0], &nb[1], &nb[2], &db, &ub);
    if (rc) { free(da); free(ua); return rc; }
    if (na[0] != nb[0] || na[1] != nb[1] || na[2] != nb[2]) {
        rc = omc_set_error(OMC_EARG, "grids differ: %dx%dx%d and %dx%dx%d", na[0], na[1], na[2], nb[0], nb[1],
                           nb[2]);
        goto done;
    }
    const int64_t nvox = (int64_t)na[0] * na[1] * na[2];
    if (phantom) {
        if (phantom->isize != na[0] || phantom->jsize != na[1] || phantom->ksize != na[2]) {
            rc = omc_set_error(OMC_EARG, "phantom %dx%dx%d does not match the dose grids", phantom->isize,
                               phantom->jsize, phantom->ksize);
            goto done;
        }
        mass = malloc((size_t)nvox * sizeof(double));
        if (!mass) { rc = omc_set_error(OMC_ENOMEM, "oom"); goto done; }
        for (int iz = 0; iz < na[2]; ++iz)
            for (int iy = 0; iy < na[1]; ++iy)
                for (int ix = 0; ix < na[0]; ++ix) {
                    const int64_t v = ix + (int64_t)na[0] * (iy + (int64_t)na[1] * iz);
                    mass[v] = (phantom->xbounds[ix + 1] - phantom->xbounds[ix]) *
                              (phantom->ybounds[iy + 1] - phantom->ybounds[iy]) *
                              (phantom->zbounds[iz + 1] - phantom->zbounds[iz]) * phantom->rho[v];
                }
    }
    rc = omc_compare_doses(nvox, da, ua, db, ub, mass, rel_thr, out);
done:
    free(da); free(ua); free(db); free(ub); free(mass);
    return rc;
}
