/*
 * surf.c -- oracle for the surface-wave likelihood (datatype 2, and the
 * surface half of datatype 3).  TEST INFRASTRUCTURE ONLY (see mct_oracle.h).
 *
 * Composes the restated stages in the order of likelihood (src/likelihood.f90
 * :72-88) and surf_likelihood (src/likelihood_surf.F90:110-405):
 *   kdtree_to_grid (nearest.c) -> vs2vp_3d / vp2rho_3d (src/utils.f90:102-134;
 *   datatype 2 only, datatype 3 keeps the sites' vp and rho)
 *   -> check_model (:631-646) -> convert_to_layer (:523-629) -> surfmodes per
 *   column (surfdisp.c) -> ierr reject (:212-224) -> like%vel padding
 *   (:259-264) -> modrays per period (fm2d.c) -> srdist = phaseTime (:329-330)
 *   -> sigma (:356-372) -> misfit (:374-404).
 * With group velocities (phaseGroup = 1): modrays traces rays (uar = 0,
 * :245), srdist = the ray lengths (:326-327), a crazy ray rejects the model
 * (:341-345), and the group delays along the rays replace the phase times
 * (CalGroupTime, :348-351).
 * OpenMP runs over columns and over periods, as the reference does.
 */
#include "mct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SURF_EPS ((double)1.0e-10f)      /* src/likelihood_surf.F90:37, default-real literal */
#define SURF_PI2 ((double)6.283185f)     /* :36 */
#define SURF_NMAX 100                    /* :30 */

/* vp and rho at node k of a column: the model's own (gvp/grho non-NULL, the
 * joint likelihood, datatype 3: src/likelihood.f90:83-88) or vs2vp_3d /
 * vp2rho_3d of vs (datatype 2, src/likelihood.f90:75-76, src/utils.f90:102-134) */
static void node_vprho(const double *vs, const double *gvp, const double *grho, int k, double *vp, double *rho)
{
    if (gvp) {
        *vp = gvp[k];
        *rho = grho[k];
        return;
    }
    *vp = vs[k] * (double)1.730f;   /* POISSON */
    *rho = (double)1.74f * pow(*vp, 0.25);
}

/* convert_to_layer for one column (no vbnd): with water > EPS a water layer
 * first (vp = waterVel 1.5, vs = 0, rho = waterDensity 1, thickness water;
 * src/likelihood_surf.F90:32-33,546-551); then a layer starts where vs
 * changes by more than EPS and takes the values of its first node; the
 * half-space (thickness 0) takes the bottom node's.  Thicknesses are scaled,
 * except that with water > 0 the first layer's is water itself (:616-617:
 * the water layer's, or for 0 < water <= EPS the first solid layer's).
 * Returns the layer count. */
static int column_layers(const double *vs, const double *gvp, const double *grho, int nz, double dz,
                         double scaling, double water, double *thick, double *vp, double *vsl, double *rho)
{
    int nl = 0, last_k = 1;
    if (water > SURF_EPS) {
        vp[0] = 1.5; vsl[0] = 0.0; rho[0] = 1.0; thick[0] = water;
        nl = 1;
    }
    double last_vs = vs[0];
    for (int k = 2; k <= nz; k++) {
        if (fabs(vs[k - 1] - last_vs) > SURF_EPS) {
            vsl[nl] = last_vs;
            node_vprho(vs, gvp, grho, last_k - 1, &vp[nl], &rho[nl]);
            thick[nl] = ((double)(k - last_k) * dz) / scaling;
            nl++;
            last_vs = vs[k - 1];
            last_k = k;
        }
    }
    vsl[nl] = vs[nz - 1];
    node_vprho(vs, gvp, grho, nz - 1, &vp[nl], &rho[nl]);
    thick[nl] = 0.0;
    if (water > 0.0) thick[0] = water;
    return nl + 1;
}

int orc_column_layers_all(long ncol, int nz, const double *vs, double dz, double scaling, double water,
                          int maxl, double *lay, int32_t *nl)
{
    int bad = 0;
    #pragma omp parallel for schedule(static) reduction(|:bad)
    for (long c = 0; c < ncol; c++) {
        double th[SURF_NMAX + 2], a[SURF_NMAX + 2], b[SURF_NMAX + 2], r[SURF_NMAX + 2];
        int n = column_layers(vs + c * nz, NULL, NULL, nz, dz, scaling, water, th, a, b, r);
        if (n > maxl) { bad = 1; continue; }
        nl[c] = n;
        for (int k = 0; k < n; k++) {
            double *o = lay + ((size_t)c * maxl + k) * 4;
            o[0] = th[k]; o[1] = a[k]; o[2] = b[k]; o[3] = r[k];
        }
    }
    return bad ? -1 : 0;
}

int orc_surf_likelihood(int nx, int ny, int nz, double xmin, double ymin, double zmin,
                        double dx, double dy, double dz, double scaling,
                        const double *points, const double *params, int ncells,
                        int np, const double *freqs, int raylov, double dphase,
                        int nsrc, const double *src, int nrev, const double *rev,
                        const double *ttime, const int32_t *raystat, int sigdep,
                        const double *snoise0, const double *snoise1, int nrays,
                        int gridx, int gridy, int sgref, int sgdic, int sgext, int order,
                        double band, int nthreads,
                        double *out, double *phase_time, double *srdist, double *sigma,
                        double *phase_vel, int32_t *sites_id, int phase_group, double *group_vel,
                        int straight, int datatype, double water)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const long G = (long)nx * ny * nz, C = (long)nx * ny, npair = (long)nsrc * nrev;
    double *vp = malloc(sizeof(double) * G), *vs = malloc(sizeof(double) * G), *rho = malloc(sizeof(double) * G);
    int32_t *sid = sites_id ? sites_id : malloc(sizeof(int32_t) * G);
    for (long i = 0; i < G; i++) { vp[i] = 0; vs[i] = 0; rho[i] = 0; }
    orc_kdtree_to_grid(nx, ny, nz, xmin, ymin, zmin, dx, dy, dz, points, params, ncells,
                       NULL, NULL, sid, vp, vs, rho);
    int rc = 0;
    /* check_model: any column whose deeper vs is below the top vs */
    for (long c = 0; c < C && !rc; c++)
        for (int k = 1; k < nz; k++)
            if (vs[c * nz + k] < vs[c * nz]) { rc = 1; break; }
    double *pvel = phase_vel ? phase_vel : malloc(sizeof(double) * C * np);
    double *gvel = group_vel ? group_vel : malloc(sizeof(double) * C * np);
    for (long i = 0; i < C * np; i++) { pvel[i] = 100.0; gvel[i] = 100.0; }
    if (!rc) {
        int bad = 0;
        /* columns in Fortran order: c = iy + ny*ix; pvel (np, ny, nx) */
        #pragma omp parallel for schedule(dynamic, 16) reduction(|:bad)
        for (long c = 0; c < C; c++) {
            double th[SURF_NMAX + 2], a[SURF_NMAX + 2], b[SURF_NMAX + 2], r[SURF_NMAX + 2];
            int nl = datatype == 3
                         ? column_layers(vs + c * nz, vp + c * nz, rho + c * nz, nz, dz, scaling, water, th, a, b, r)
                         : column_layers(vs + c * nz, NULL, NULL, nz, dz, scaling, water, th, a, b, r);
            int ierr = orc_surfmodes(th, a, b, r, nl, freqs, np, raylov, phase_group, dphase,
                                     pvel + c * np, gvel + c * np, NULL);
            if (ierr) bad = 1;
        }
        if (bad) rc = 1;
    }
    if (!rc && straight) {
        /* straight rays (:233-243): the phase or group velocities integrated
         * along the source-receiver segments; srdist = the distances */
        double *gvp = malloc(sizeof(double) * C);
        double *scx = malloc(sizeof(double) * nsrc), *scz = malloc(sizeof(double) * nsrc);
        double *rcx = malloc(sizeof(double) * nrev), *rcz = malloc(sizeof(double) * nrev);
        for (int j = 0; j < nsrc; j++) { scx[j] = src[2 * j]; scz[j] = src[2 * j + 1]; }
        for (int k = 0; k < nrev; k++) { rcx[k] = rev[2 * k]; rcz[k] = rev[2 * k + 1]; }
        const double *vsel = phase_group ? gvel : pvel;
        for (int p = 0; p < np && !rc; p++) {
            for (long c = 0; c < C; c++) gvp[c] = vsel[c * np + p];
            if (orc_straight_rays(nx, ny, xmin, ymin, dx, dy, gvp, nsrc, scx, scz, nrev, rcx, rcz,
                                  raystat + (long)p * 2 * npair, phase_time + (long)p * npair,
                                  srdist + (long)p * npair))
                rc = -6;
        }
        free(gvp); free(scx); free(scz); free(rcx); free(rcz);
    }
    if (!rc && !straight) {
        /* like%vel(np, ny+2, nx+2) with replicated edges; per period a Fortran
         * velv(0:ny+1, 0:nx+1) grid */
        const int ldv = ny + 2;
        double *velv = malloc(sizeof(double) * (size_t)np * ldv * (nx + 2));
        for (int p = 0; p < np; p++)
            for (int ix = 0; ix < nx + 2; ix++)
                for (int iy = 0; iy < ny + 2; iy++) {
                    int cx = ix < 1 ? 1 : (ix > nx ? nx : ix), cy = iy < 1 ? 1 : (iy > ny ? ny : iy);
                    velv[((size_t)p * (nx + 2) + ix) * ldv + iy] = pvel[((long)(cx - 1) * ny + (cy - 1)) * np + p];
                }
        double *scx = malloc(sizeof(double) * nsrc), *scz = malloc(sizeof(double) * nsrc);
        double *rcx = malloc(sizeof(double) * nrev), *rcz = malloc(sizeof(double) * nrev);
        for (int j = 0; j < nsrc; j++) { scx[j] = src[2 * j]; scz[j] = src[2 * j + 1]; }
        for (int k = 0; k < nrev; k++) { rcx[k] = rev[2 * k]; rcz[k] = rev[2 * k + 1]; }
        /* group velocities per period as Fortran gvel(ny,nx) grids (like%gvel(i,:,:)) */
        double *gvp = phase_group ? malloc(sizeof(double) * (size_t)np * C) : NULL;
        if (gvp)
            for (int p = 0; p < np; p++)
                for (long c = 0; c < C; c++) gvp[(size_t)p * C + c] = gvel[c * np + p];
        double *rlen = phase_group ? malloc(sizeof(double) * (size_t)np * npair) : NULL;
        int fatal = 0, ncrazy = 0;
        #pragma omp parallel for schedule(dynamic, 1) reduction(|:fatal) reduction(+:ncrazy)
        for (int p = 0; p < np; p++) {
            int e;
            if (phase_group) {
                int32_t cr = 0;
                double *tp = malloc(sizeof(double) * npair);
                e = orc_fm2d_rays(nx, ny, xmin, ymin, dx, dy, velv + (size_t)p * (nx + 2) * ldv, gridx, gridy,
                                  sgref, sgdic, sgext, order, band, nsrc, scx, scz, nrev, rcx, rcz,
                                  raystat + (long)p * 2 * npair, tp, NULL, 1, gvp + (size_t)p * C,
                                  phase_time + (long)p * npair, rlen + (long)p * npair, &cr, 0, NULL, NULL);
                free(tp);
                ncrazy += cr;
            } else {
                e = orc_fm2d(nx, ny, xmin, ymin, dx, dy, velv + (size_t)p * (nx + 2) * ldv, gridx, gridy,
                             sgref, sgdic, sgext, order, band, nsrc, scx, scz, nrev, rcx, rcz,
                             raystat + (long)p * 2 * npair, phase_time + (long)p * npair, NULL);
            }
            if (e) fatal = 1;
        }
        free(velv); free(scx); free(scz); free(rcx); free(rcz); free(gvp);
        if (fatal) rc = -6;
        else if (ncrazy > 0) rc = 2;
        if (!rc) {
            /* srdist = phaseTime (phase velocities) or the ray length (group) */
            for (long i = 0; i < (long)np * npair; i++) srdist[i] = phase_group ? rlen[i] : phase_time[i];
        }
        free(rlen);
    }
    if (!rc) {
        /* sigma; misfit in the reference's loop order */
        for (int i = 0; i < np; i++)
            for (long r = 0; r < npair; r++) {
                long o = (long)i * npair + r;
                if (sigdep)
                    sigma[o] = raystat[(long)i * 2 * npair + r] == 1 ? snoise0[i] * srdist[o] + snoise1[i] : 1.0;
                else
                    sigma[o] = ttime[(long)i * 3 * npair + npair + r];
            }
        double like = 0, misfit = 0, unw = 0;
        for (int i = 0; i < np && !rc; i++)
            for (long r = 0; r < npair; r++) {
                long o = (long)i * npair + r;
                if (raystat[(long)i * 2 * npair + r] == 1) {
                    if (sigma[o] < SURF_EPS) { rc = -4; break; }
                    double d = phase_time[o] - ttime[(long)i * 3 * npair + r];
                    like = like + (d * d) / (2 * (sigma[o] * sigma[o]));
                    misfit = misfit + (d * d) / (sigma[o] * sigma[o]);
                    unw = unw + d * d;
                } else {
                    sigma[o] = 1.0;
                }
            }
        double slog = 0;
        for (long i = 0; i < (long)np * npair; i++) slog += log(sigma[i]);
        out[0] = like + slog + (double)((float)nrays / 2.0f) * log(SURF_PI2);
        out[1] = misfit;
        out[2] = unw;
    }
    free(vp); free(vs); free(rho);
    if (!sites_id) free(sid);
    if (!phase_vel) free(pvel);
    if (!group_vel) free(gvel);
    return rc;
}
