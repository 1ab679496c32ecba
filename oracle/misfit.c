/*
 * misfit.c -- oracle for the Gaussian misfit of the body-wave likelihood.
 * TEST INFRASTRUCTURE ONLY (see mct_oracle.h).
 */
#include "mct_oracle.h"
#include <math.h>
#include <float.h>

/* PI2 = 6.283185 is a default-real literal stored in a double parameter
 * (src/likelihood_body.F90:32); EPS = tiny(0d0)*10 (:33). */
#define ORC_PI2 ((double)6.283185f)
#define ORC_EPS (DBL_MIN * 10.0)

/* body_likelihood misfit block, src/likelihood_body.F90:286-333 */
int orc_body_misfit(int nsrc, int nrev, int np, const double *phase_time,
                    const double *ttime, const int32_t *raystat,
                    const double *locations, int sigdep,
                    const double *bnoise0, const double *bnoise1, int nrays,
                    double *srdist, double *sigma, double *out)
{
    const long npair = (long)nsrc * nrev;
    /* srdist(nrr,:) = phaseTime(k,j,:) with nrr = k + (j-1)*nrev (:293-298) */
    for (int i = 0; i < np; i++)
        for (int j = 0; j < nsrc; j++)
            for (int k = 0; k < nrev; k++)
                srdist[(long)i * npair + (long)j * nrev + k] =
                    phase_time[((long)i * nsrc + j) * nrev + k];
    /* sigma (:302-308) */
    for (int i = 0; i < np; i++)
        for (long r = 0; r < npair; r++)
            sigma[i * npair + r] = sigdep ? bnoise0[i] * srdist[i * npair + r] + bnoise1[i]
                                          : ttime[(long)i * 3 * npair + npair + r];
    double like = 0, misfit = 0, unw = 0;
    int fatal = 0;
    for (int i = 0; i < np; i++) {
        long nrr = 0;
        for (int j = 0; j < nsrc; j++)
            for (int k = 0; k < nrev; k++, nrr++) {
                double *sg = &sigma[i * npair + nrr];
                if (raystat[(long)i * 2 * npair + nrr] == 1) {
                    if (*sg < ORC_EPS) fatal = 1;
                    double r = phase_time[((long)i * nsrc + j) * nrev + k] + locations[4 * j + 3]
                               - ttime[(long)i * 3 * npair + nrr];
                    like = like + (r * r) / (2 * ((*sg) * (*sg)));
                    misfit = misfit + (r * r) / ((*sg) * (*sg));
                    unw = unw + r * r;
                } else {
                    *sg = 1.0;
                }
            }
    }
    double slog = 0;
    for (long r = 0; r < (long)np * npair; r++) slog += log(sigma[r]);
    like = like + slog + (double)((float)nrays / 2.0f) * log(ORC_PI2);
    out[0] = like; out[1] = misfit; out[2] = unw;
    return fatal ? -1 : 0;
}

/* Interpolate3d with point2idx (src/likelihood_body.F90:976-1007,1038-1054):
 * field Fortran (nz,ny,nx) (C: field[(ix*ny + iy)*nz + iz]); grid as T_GRID
 * with zmin/dz scaled; the point in unscaled units.  The cell is
 * floor((p - min)/d) + 1 clamped to [1, n-1]. */
double orc_interpolate3d(int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx,
                         double dy, double dz, double scaling, const double *field, const double *point)
{
    int ix = (int)floor((point[0] - xmin) / dx) + 1;
    int iy = (int)floor((point[1] - ymin) / dy) + 1;
    int iz = (int)floor((point[2] - zmin / scaling) / (dz / scaling)) + 1;
    if (ix < 1) ix = 1;
    if (iy < 1) iy = 1;
    if (iz < 1) iz = 1;
    if (ix >= nx) ix = nx - 1;
    if (iy >= ny) iy = ny - 1;
    if (iz >= nz) iz = nz - 1;
    double dsx = point[0] - (xmin + (ix - 1) * dx);
    double dsy = point[1] - (ymin + (iy - 1) * dy);
    double dsz = point[2] - (zmin + (iz - 1) * dz) / scaling;
    double qv = 0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++)
            for (int k = 1; k <= 2; k++) {
                double weight = (1.0 - fabs((i - 1) * dx - dsx) / dx) * (1.0 - fabs((j - 1) * dy - dsy) / dy) *
                                (1.0 - fabs((k - 1) * dz / scaling - dsz) / (dz / scaling));
                qv = qv + weight * field[((long)(ix + i - 2) * ny + (iy + j - 2)) * nz + (iz + k - 2)];
            }
    return qv;
}
