/*
 * mct_oracle.h -- CPU oracle for the MCTomo likelihood forward model.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference algorithms (each function cites the reference file:line it
 * follows).  It is the checker: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (libmctomo_hip.so)
 * never links, calls or falls back to it.
 *
 * Array conventions follow the reference's Fortran storage:
 *   grid arrays (nz,ny,nx) column-major  -> flat index k + nz*(j + ny*i)
 *   points / params (3,ncells)           -> p[3*c + d]
 *   indices written to sites_id are 1-based site numbers, as in Fortran.
 */
#ifndef MCT_ORACLE_H
#define MCT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- (a) Voronoi -> grid nearest site ---------------------------------- */
/* kdtree_to_grid (src/mcmc_loc2.f90:2002-2082) over the whole grid or a
 * box.  box = {x0,y0,z0,x1,y1,z1} (scaled z, like the reference's d3 box);
 * NULL means the whole grid.  pm = {vp,vs,rho} or NULL.  Returns 0. */
int orc_kdtree_to_grid(int nx, int ny, int nz,
                       double xmin, double ymin, double zmin,
                       double dx, double dy, double dz,
                       const double *points, const double *params, int ncells,
                       const double *box, const double *pm,
                       int32_t *sites_id, double *vp, double *vs, double *rho);

/* One kdtree2 1-NN query (src/kdtree2.f90:1028-1069) on a tree built from
 * `points`; returns the 1-based index.  Used by tests on single queries. */
int orc_kdtree_nearest(const double *points, int ncells, const double *q);

/* Brute-force nearest site with the kdtree2 distance summation order; returns
 * the smallest squared distance and the number of sites that attain it. */
int orc_brute_nearest(const double *points, int ncells, const double *q,
                      double *best_sd, int *n_at_best);

/* ---- (b) dispersion --------------------------------------------------------- */
/* surfdisp96 (surfmodes/surfdisp96.f:52-382), flat earth, fundamental mode.
 * iwave 1 Love, 2 Rayleigh; igr 0 phase, 1 group; t(kmax) periods in s.
 * cp/cg (kmax) as the Fortran writes them; returns ierr.  n_dltar (may be
 * NULL) receives the number of period-equation evaluations. */
int orc_surfdisp96(const float *thkm, const float *vpm, const float *vsm, const float *rhom,
                   int nlayer, int iwave, int igr, int kmax, const double *t, double dphase,
                   double *cp, double *cg, long *n_dltar);
/* surfmodes' surfdisp96 branch (surfmodes/surfmodes.f90:79-95): FP64 layers
 * cast to REAL*4, periods 1/freqs.  raylov 1 Rayleigh, 0 Love. */
int orc_surfmodes(const double *thick, const double *vp, const double *vs, const double *rho,
                  int nlayer, const double *freqs, int np, int raylov, int phase_group,
                  double dc, double *phase, double *group, long *n_dltar);

/* ---- (c) 2-D fast marching (fm2d, phase times) ----------------------------- */
/* modrays' phase-time path for one period (fm2d/fm2dray_cartesian.f90:67-480):
 * velv(nvz+2,nvx+2) padded phase-velocity grid (Fortran order), gdx/gdz
 * B-spline dicing, asgr/sgdl/sgs source refinement, fom order, snb band.
 * srs (nsrc*nrc) 1 = pair used (receiver fastest); ttime(nrc,nsrc) is written
 * for used pairs only.  tfield (nsrc, nnx, nnz) may be NULL.
 * Returns 0, -1 source outside the grid, -2 receiver outside (the reference
 * STOPs). */
int orc_fm2d(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
             const double *velv, int gdx, int gdz, int asgr, int sgdl, int sgs, int fom,
             double snb, int nsrc, const double *scx, const double *scz, int nrc,
             const double *rcx, const double *rcz, const int32_t *srs, double *ttime,
             double *tfield);
/* orc_fm2d plus, when rays != 0, the group-velocity leg (uar = 0): rpaths from
 * every receiver of a used pair, the ray length rlen (nsrc,nrc) and, when gvel
 * (the group velocity grid, Fortran gvel(nvz,nvx)) is given, the group delay
 * gtime (nsrc,nrc) (CalGroupTime); unused pairs get 0.  crazy counts rays that
 * hit the point limit.  npts/pts (nsrc*nrc, maxpts, 2) receive the ray points
 * when not NULL (a ray longer than maxpts keeps its first maxpts). */
int orc_fm2d_rays(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
                  const double *velv, int gdx, int gdz, int asgr, int sgdl, int sgs, int fom,
                  double snb, int nsrc, const double *scx, const double *scz, int nrc,
                  const double *rcx, const double *rcz, const int32_t *srs, double *ttime,
                  double *tfield, int rays, const double *gvel, double *gtime, double *rlen,
                  int32_t *crazy, int maxpts, int32_t *npts, double *pts);

/* Straight rays (isStraight = 1, src/likelihood_surf.F90:412-489): dist and
 * group delays gtime (nsrc,nrc) along the source-receiver segments through
 * gvel (Fortran gvel(nvz,nvx)).  Returns -1 for a point outside the model. */
int orc_straight_rays(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd, const double *gvel,
                      int nsrc, const double *scx, const double *scz, int nrc, const double *rcx,
                      const double *rcz, const int32_t *srs, double *gtime, double *dist);

/* ---- (c) 3-D fast marching (SFMM) --------------------------------------- */
/* Same signature and semantics as the reference C ABI fastMarching3D
 * (fastmarching/fastMarching.cpp:26-129). vel/tfield are Fortran (nz,ny,nx). */
int orc_fastMarching3D(int nx, int ny, int nz, double dx, double dy, double dz,
                       double xmin, double ymin, double zmin,
                       const double *vel, const double *src, const double *rev,
                       int nrev, int order, double *ttime, double *tfield);

/* Body-wave forward model for all sources (src/likelihood_body.F90:204-224,
 * nsrc <= nrev branch), OpenMP over sources like the reference.
 * locations(4,nsrc), rev(3,nrev); phase_time(nrev,nsrc).  nthreads<=0: all. */
int orc_body_traveltimes(int nx, int ny, int nz, double dx, double dy, double dz,
                         double xmin, double ymin, double zmin, double scaling,
                         const double *vel, int nsrc, const double *locations,
                         int nrev, const double *rev, int order,
                         double *phase_time, int nthreads);

/* Test hook: replay push / key-change / pop on the SFMM queue (fmm3d.c) */
int orc_heap_replay(int ncell, double *keys, int nops, const int *ops, const double *vals,
                    int *popped);

/* ---- (d) misfit ----------------------------------------------------------- */
/* Body-wave Gaussian misfit (src/likelihood_body.F90:286-333) for np data
 * sets (np=1 P, np=2 P&S).  phase_time(nrev,nsrc,np); ttime(nsrc*nrev,3,np);
 * raystat(nsrc*nrev,2,np); locations(4,nsrc); bnoise0/1(np).
 * srdist and sigma (nsrc*nrev,np) are written like the reference's
 * like%srdist / like%sigma.  out = {like, misfit, unweighted_misfit}.
 * Returns -1 if a used sigma is below EPS (the reference raises a fatal error),
 * else 0. */
int orc_body_misfit(int nsrc, int nrev, int np, const double *phase_time,
                    const double *ttime, const int32_t *raystat,
                    const double *locations, int sigdep,
                    const double *bnoise0, const double *bnoise1, int nrays,
                    double *srdist, double *sigma, double *out);


/* convert_to_layer (src/likelihood_surf.F90:523-629; water: T_GRID%waterDepth,
 * 0 for none) for every column of vs (ncol, nz): lay (ncol, maxl, 4) = thick, vp, vs, rho with vp and
 * rho from vs2vp_3d / vp2rho_3d, nl (ncol) layer counts.  -1 if a column has
 * more than maxl layers. */
int orc_column_layers_all(long ncol, int nz, const double *vs, double dz, double scaling, double water,
                          int maxl, double *lay, int32_t *nl);

/* Interpolate3d (src/likelihood_body.F90:976-1007): the trilinear value of a
 * traveltime field (Fortran (nz,ny,nx)) at point[3] (z unscaled), misfit.c. */
double orc_interpolate3d(int nx, int ny, int nz, double xmin, double ymin, double zmin, double dx,
                         double dy, double dz, double scaling, const double *field, const double *point);

/* Surface-wave likelihood (datatype 2), surf.c.  Grid as
 * T_GRID (dz scaled, scaling = grid%scaling); points/params (ncells,3);
 * freqs (np) Hz; raylov 1 Rayleigh, 0 Love; src (nsrc,2), rev (nrev,2) as
 * (x, y); ttime (np,3,nsrc*nrev) and raystat (np,2,nsrc*nrev) as T_DATA;
 * snoise0/1 (np).  Outputs: out {like, misfit, unweighted}, phase_time
 * (np,nsrc,nrev), srdist and sigma (np,nsrc*nrev), phase_vel (nx,ny,np) (may
 * be NULL), sites_id (nx,ny,nz) (may be NULL).  phase_group 1 computes group
 * velocities and times: group delays along the rays (CalGroupTime), srdist =
 * ray lengths; group_vel (nx,ny,np) may be NULL.  datatype 2: vp and rho of
 * the layers from vs2vp_3d / vp2rho_3d; 3 (the surface half of the joint
 * likelihood): the sites' own vp and rho.  Returns 0, 1 (model rejected:
 * check_model or a dispersion failure), 2 (a crazy ray), or < 0 (fatal). */
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
                        double *phase_vel, int32_t *sites_id, int phase_group,
                        double *group_vel, int straight, int datatype, double water);

#ifdef __cplusplus
}
#endif
#endif
