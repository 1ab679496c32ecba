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
