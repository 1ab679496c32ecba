/*
 * mctomo_hip.h -- C ABI of libmctomo_hip.so, the MI355X (gfx950) likelihood
 * forward model that sits under MCTomo's Fortran m_likelihood surface.
 *
 * Conventions (the reference's Fortran storage, passed with c_loc as in
 * src/fastMarching_wrapper.f90:51-74):
 *   grid arrays are Fortran (nz,ny,nx): flat index k + nz*(j + ny*i);
 *   points/params are (3,ncells); locations (4,nsrc); rev (3,nrev);
 *   sites_id holds 1-based site numbers.
 * Host pointers unless the name ends in _dev (device pointers, asynchronous on
 * the context's stream; call mct_sync before reading results).
 *
 * Status codes: 0 ok; >0 soft reject (the caller sets like = huge and bumps
 * RTI%num_bad_model / num_bad_ray as src/likelihood_surf.F90:161-164,212-224,
 * 341-345 do); <0 fatal (the caller raises exception_raiseError,
 * src/m_exception.f90:223-228).  No C++ exception crosses this boundary.
 */
#ifndef MCTOMO_HIP_H
#define MCTOMO_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MCT_OK = 0,
    MCT_REJECT_MODEL = 1,     /* check_model / dispersion failure -> num_bad_model */
    MCT_REJECT_RAY = 2,       /* crazy ray -> num_bad_ray */
    MCT_ERR_ARG = -1,         /* invalid argument or size */
    MCT_ERR_HIP = -2,         /* HIP runtime error (no device, launch failure) */
    MCT_ERR_NOMEM = -3,       /* device allocation failed */
    MCT_ERR_SIGMA = -4,       /* noise level below EPS ('The noise level is 0!') */
    MCT_ERR_NOCONV = -5,      /* eikonal solver hit its iteration cap */
    MCT_ERR_OUTSIDE = -6,     /* source or receiver outside the grid */
    MCT_ERR_IO = -7,          /* a data file could not be opened, read or written */
    MCT_ERR_SCRATCH = -8      /* the allocation-free entry's scratch cannot hold the request */
};

/* T_GRID, src/settings.f90:20-28.  zmin/dz are the SCALED values held by the
 * reference after grid_setup (src/settings.f90:100-123). */
typedef struct {
    int nx, ny, nz;
    double xmin, ymin, zmin;
    double dx, dy, dz;
    double scaling;
} mct_grid;

/* ---- (a) Voronoi -> grid ------------------------------------------------- */
/* Replaces kdtree_to_grid (src/mcmc_loc2.f90:2002-2082): for every grid node
 * inside box (NULL = whole grid; {x0,y0,z0,x1,y1,z1}) writes the 1-based index
 * of the nearest site with kdtree2's tie-breaking (src/kdtree2.f90:1369-1599)
 * and copies params(:,id) into vp, vs, rho.  With pm != NULL ({vp,vs,rho})
 * only nodes whose old (vp,vs) equal pm within 1e-8 are updated (:2055-2062). */
int mct_voronoi_to_grid(const mct_grid *grid, const double *points, const double *params,
                        int ncells, const double *box, const double *pm,
                        int32_t *sites_id, double *vp, double *vs, double *rho);

/* ---- (b) dispersion ---------------------------------------------------------- */
/* surfdisp96 (surfmodes/surfdisp96.f:52-382) as surfmodes calls it
 * (surfmodes/surfmodes.f90:79-95: flat earth, fundamental mode), for nmod
 * independent REAL*4 layer models at once.  layers (nmod,4,nl_max) holds
 * thick, vp, vs, rho per model (last layer = half-space, thick 0); nlayer
 * (nmod); periods (np) in s; iwave 1 Love, 2 Rayleigh; igr 0 phase, 1 group;
 * dphase the search step (LIKE_SET%dPhaseVel).  phase/group (nmod,np) and
 * ierr (nmod) receive cp, cg and ierr exactly as surfdisp96 returns them. */
int mct_surfdisp96_batch(int nmod, int nl_max, const int32_t *nlayer, const float *layers,
                         int np, const double *periods, int iwave, int igr, double dphase,
                         double *phase, double *group, int32_t *ierr);

/* The surface likelihood's dispersion stage on explicit columns: check_model,
 * convert_to_layer (thickness (k-last_k)*dz/scaling, vp = POISSON*vs, rho =
 * 1.74*vp^0.25; src/likelihood_surf.F90:523-646, src/utils.f90:102-134) and
 * surfdisp96 per column (surfmodes/surfmodes.f90:79-95).  vs (ncol, nz) holds
 * Fortran vs(nz) per column; phase/group (ncol, np); ierr (ncol); bad_model
 * (1): check_model failed for some column.  mode picks the kernel the
 * likelihood uses: 1 one lane per column, 2 one wave per column, 3 the column
 * queue in layer-count order with identical columns solved once (the default
 * for large batches).  water: T_GRID%waterDepth (km, not scaled; 0 for no
 * water layer), the water layer convert_to_layer puts on every column
 * (src/likelihood_surf.F90:546-551,616-617).  nz <= 100, np <= 60. */
int mct_dispersion_columns(int ncol, int nz, double dz, double scaling, double water, const double *vs,
                           int np, const double *periods, int iwave, int igr, double dphase,
                           int mode, double *phase, double *group, int32_t *ierr,
                           int32_t *bad_model);

/* ---- (c) 2-D phase traveltimes ---------------------------------------------- */
/* modrays' phase-time path (fm2d/fm2dray_cartesian.f90:67-480, with
 * crazyray = 0, uar = 1, earth = 6371) for np periods at once.  velv holds np
 * padded control grids, each Fortran velv(0:nvz+1, 0:nvx+1) as
 * fm2d_globalp's velv; goxd/gozd the origin and dvxd/dvzd the control
 * spacing (km); gdx, gdz, asgr, sgdl, sgs, fom as modrays' arguments
 * (LIKE_SET%gridx/gridy, sgref, sgdic, sgext, order).  scx/scz (nsrc),
 * rcx/rcz (nrc); srs (np, nsrc, nrc) marks the pairs whose time is computed
 * (receiver fastest); ttime (np, nsrc, nrc) receives their times, other
 * entries are left as passed in.  method 1: one lane solves one (period,
 * source) with the reference's sequential heap loop.  method 0: the heap loop
 * only where its history shows (the source seed, the refined grid's stop and
 * the coarse restart around the source), the rest of the grid by parallel
 * sweeps to the same fixed point; problems that do not fit fall back to
 * method 1, and so do problems whose converged field fails the check that the
 * heap loop would freeze it in time order (on for fom 1; the environment's
 * MCT_FM2D_VERIFY=0/1 overrides).  Returns MCT_ERR_OUTSIDE where modrays
 * would stop on a source or receiver outside the grid. */
int mct_modrays(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
                int np, const double *velv, int gdx, int gdz, int asgr, int sgdl, int sgs,
                int fom, int nsrc, const double *scx, const double *scz, int nrc,
                const double *rcx, const double *rcz, const int32_t *srs, double *ttime,
                int method);

/* Device bytes of the allocation-free entry's scratch array (one per GPU). */
#define MCT_NOALLOC_SCRATCH_BYTES (128L << 20)

/* No counterpart in the reference: mct_modrays for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The same computation as mct_modrays with
 * every array a DEVICE pointer of the caller's: velv_dev, scx/scz_dev,
 * rcx/rcz_dev, srs_dev as there, ttime_dev receives the marked pairs' times
 * and status_dev (one int32) the status (0, or MCT_ERR_OUTSIDE where modrays
 * stops).  Queued on `stream` (a hipStream_t, NULL the default stream) and
 * returns without waiting.  The solver's workspace is one device array of
 * MCT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; problems
 * run in chunks of what it holds, and MCT_ERR_SCRATCH is returned (nothing
 * launched) when the coarse velocities and one problem's workspace do not
 * fit.  Nothing is allocated.  Calls that overlap on one GPU share the array
 * and race; each call leaves it holding its own state. */
int mct_modrays_dev_noalloc(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
                            int np, const double *velv_dev, int gdx, int gdz, int asgr, int sgdl, int sgs,
                            int fom, int nsrc, const double *scx_dev, const double *scz_dev, int nrc,
                            const double *rcx_dev, const double *rcz_dev, const int32_t *srs_dev,
                            double *ttime_dev, int32_t *status_dev, int method, void *stream);
/* MCT_NOALLOC_SCRATCH_BYTES as the library was built with it. */
int mct_noalloc_scratch_bytes(int64_t *bytes);

/* Problems (period, source) the calling thread's last mct_modrays call handed
 * from the parallel path to the sequential solver.  No counterpart in the
 * reference: a measurement. */
int mct_fm2d_last_handover(long long *n);

/* ---- (c) 3-D eikonal: link-level drop-in ---------------------------------- */
/* Same symbol, signature and semantics as the reference C ABI
 * (fastmarching/fastMarching.cpp:26-30; Fortran interface
 * src/fastMarching_wrapper.f90:18-29).  vel/tfield are Fortran (nz,ny,nx);
 * tfield may be NULL.  Returns 0, or a negative MCT_ERR_* code. */
int fastMarching3D(int nx, int ny, int nz, double dx, double dy, double dz,
                   double xmin, double ymin, double zmin, double *vel,
                   const double *src, const double *rev, int nrev, int order,
                   double *ttime, double *tfield);

/* ---- body-wave likelihood context ---------------------------------------- */
typedef struct mct_ctx mct_ctx;

/* The immutable part of T_DATA (src/likelihood_settings.f90:19-29) for one
 * body-wave data set and the T_LIKE_SET knobs that the body path reads
 * (src/likelihood_settings.f90:47-79). */
typedef struct {
    int datatype;             /* 0: P only (np=1), 1: P & S (np=2), 3: P & S as the body half of
                                 the joint likelihood (dat(1)) */
    int order;                /* 0 first order, 1 mixed order (LIKE_SET%ORDER) */
    int sigdep;               /* MCMC_SET%SIGDEP: sigma = b0*t + b1 when != 0 */
    int nsrc, nrev;
    const double *rev;        /* (3,nrev) receiver coordinates */
    const double *ttime;      /* (nsrc*nrev,3,np) observed time, sigma, unused */
    const int32_t *raystat;   /* (nsrc*nrev,2,np) 1 = datum present */
    int nrays;                /* dat%nrays */
    int max_cells;            /* largest Voronoi site count a call may pass */
    int max_chains;           /* largest batch for the _batch_dev entry */
    int dynamic;              /* LIKE_SET%dynamic: != 0 traces from the receivers and keeps
                                 their traveltime fields for mct_body_location_likelihood
                                 (src/likelihood_body.F90:250-279) */
} mct_body_setup;

/* like_setup for body data (src/likelihood.f90:31-55): uploads the data to
 * HBM and returns an opaque context.  The context keeps no per-chain state
 * between calls, so the sampler's like_copy / restore on reject
 * (src/mcmc_loc2.f90:201,554-567) stays correct. */
int mct_body_create(const mct_grid *grid, const mct_body_setup *setup, mct_ctx **out);
int mct_destroy(mct_ctx *ctx);

/* likelihood() for body waves (src/likelihood.f90:57-93 ->
 * src/likelihood_body.F90:151-337) on one chain's Voronoi model:
 * nearest-site rasterisation, one 3-D eikonal per source (per wave type),
 * receiver picks, Gaussian misfit.  locations (4,nsrc) (x,y,z,t0);
 * bnoise0/bnoise1 (np).  out = {like, misfit, unweighted_misfit}.
 * phase_time (nrev,nsrc,np), srdist and sigma (nsrc*nrev,np) may be NULL;
 * when given they receive like%phaseTime / like%srdist / like%sigma. */
int mct_body_likelihood(mct_ctx *ctx, const double *points, const double *params, int ncells,
                        const double *locations, const double *bnoise0, const double *bnoise1,
                        double *out, double *phase_time, double *srdist, double *sigma);

/* Batched form for many chains whose models are already in HBM (the serving
 * and throughput path).  points_dev/params_dev: (nchain, max_cells, 3);
 * ncells_dev (nchain) int32; locations_dev (nchain, nsrc, 4);
 * noise_dev (nchain, 2, np) = (bnoise0, bnoise1); out_dev (nchain, 3).
 * Asynchronous on the context stream; status_dev (nchain) int32 receives the
 * per-chain status. */
int mct_body_likelihood_batch_dev(mct_ctx *ctx, int nchain, const double *points_dev,
                                  const double *params_dev, const int32_t *ncells_dev,
                                  const double *locations_dev, const double *noise_dev,
                                  double *out_dev, int32_t *status_dev);
/* The context's HIP stream (hipStream_t), for callers that order their own
 * work with the context's. */
int mct_body_stream(mct_ctx *ctx, void **stream);

/* noise_likelihood for datatype 0, 1 (src/likelihood.f90:131-146 ->
 * body_noise_likelihood, src/likelihood_body.F90:38-78): the likelihood of the
 * traveltimes kept from the context's last likelihood call under new noise
 * parameters, sigma = bnoise0*t + bnoise1 (np each), with the origin times
 * locations(4,nsrc) (NULL: zero).  No eikonal solve.  Returns the status. */
int mct_body_noise_likelihood(mct_ctx *ctx, const double *locations, const double *bnoise0,
                              const double *bnoise1, double *out);
/* The same for the first nchain chains of the last batch: locations_dev
 * (nchain, nsrc, 4) or NULL, noise_dev (nchain, 2, np); asynchronous. */
/* The same over traveltimes the caller passes: phase_time (nrev,nsrc,np)
 * Fortran = like%likelihoods(1)%phaseTime, as the likelihood call wrote it.
 * The reference's body_noise_likelihood reads that array (copied with
 * like_copy, so restored on reject); this form keeps the drop-in stateless.
 * sigma (nsrc*nrev,np) may be NULL; else it receives like%sigma. */
int mct_body_noise_likelihood_from(mct_ctx *ctx, const double *phase_time, const double *locations,
                                   const double *bnoise0, const double *bnoise1, double *out, double *sigma);
int mct_body_noise_likelihood_batch_dev(mct_ctx *ctx, int nchain, const double *locations_dev,
                                        const double *noise_dev, double *out_dev, int32_t *status_dev);

/* ---- data files (host only) ------------------------------------------------ */
/* read_sources (src/likelihood_settings.f90:198-228): "nsrc ndim", then one
 * record of ndim values per source, Fortran list-directed.  With src NULL only
 * n and ndim are returned; otherwise src (n, ndim) (= Fortran src(ndim,n)),
 * which holds cap doubles: MCT_ERR_ARG before anything is written when
 * n*ndim > cap.  bnd (nbnd doubles: min, max per coordinate) may be NULL; when
 * given, nbnd must be 2*ndim (out_bnd compares coordinate i with pair i,
 * src/settings.f90:143-154), and a point outside it is MCT_ERR_OUTSIDE. */
int mct_read_sources(const char *path, const double *bnd, int nbnd, long cap, int *n, int *ndim, double *src);
/* read_times_3 (src/likelihood_settings.f90:332-420) as read_data calls it (no
 * ds file): the number of periods (or wave types) np and their values, then per
 * (source, receiver) pair a validity flag and, for a valid pair, np records
 * "t sigma".  freqs (np), ttime (np,3,nsrc*nrev) = Fortran ttime(npair,3,np),
 * raystat (np,2,nsrc*nrev) = Fortran raystat(npair,2,np); nrays.  A used pair
 * (t /= -1) gets raystat = (1, pair index); an invalid pair ttime = (-1, 1, 0).
 * With freqs, ttime or raystat NULL only np is returned (cap_np is the
 * arrays' period capacity). */
int mct_read_times(const char *path, int nsrc, int nrev, int cap_np, int *np, double *freqs, double *ttime,
                   int32_t *raystat, int *nrays);
/* write_sources (src/likelihood_settings.f90:230-257). */
int mct_write_sources(const char *path, int n, int ndim, const double *src);
/* write_atimes (src/read_write.f90:484-530), the file write_data writes: np,
 * the freqs, then per pair 1 and np records "time sigma" when any period uses
 * the pair (time + origin time locations(4,src) on used periods when
 * locations (nsrc,4) is given), else 0.  phase_time (np,nsrc,nrev), sigma
 * (np,nsrc*nrev), raystat as in mct_read_times.  Readable by read_times_3. */
int mct_write_times(const char *path, int np, const double *freqs, int nsrc, int nrev, const int32_t *raystat,
                    const double *phase_time, const double *sigma, const double *locations);

/* location_likelihood with dynamic != 0 (src/likelihood.f90:162-191 ->
 * body_location_likelihood, src/likelihood_body.F90:82-149): the times from the
 * receivers' fields of the context's last likelihood call, interpolated at the
 * new source locations (Interpolate3d, :976-1007: trilinear over the clamped
 * cell), sigma = bnoise0*t + bnoise1, residuals with the new origin times.  No
 * eikonal solve.  locations (nsrc,4) = Fortran locations(4,nsrc).  Needs a
 * context created with dynamic != 0 (else MCT_ERR_ARG); with dynamic = 0 the
 * reference runs the full likelihood, which is mct_body_likelihood. */
int mct_body_location_likelihood(mct_ctx *ctx, const double *locations, const double *bnoise0,
                                 const double *bnoise1, double *out);
/* The same for the first nchain chains of the last batch: locations_dev
 * (nchain, nsrc, 4), noise_dev (nchain, 2, np); asynchronous. */
int mct_body_location_likelihood_batch_dev(mct_ctx *ctx, int nchain, const double *locations_dev,
                                           const double *noise_dev, double *out_dev, int32_t *status_dev);

/* Device memory for callers that hold no device allocator of their own (a
 * Fortran host, or Python without a GPU framework): plain hipMalloc / hipFree
 * and synchronous copies.  No counterpart in the reference, which has no
 * device memory; they exist so the _dev entry points are callable from the
 * same C ABI. */
int mct_dev_alloc(void **ptr, int64_t bytes);
int mct_dev_free(void *ptr);
int mct_copy_h2d(void *dst_dev, const void *src_host, int64_t bytes);
int mct_copy_d2h(void *dst_host, const void *src_dev, int64_t bytes);
/* Device-to-device copy (also the known-size copy that calibrates HBM
 * traffic counters). */
int mct_copy_d2d(void *dst_dev, const void *src_dev, int64_t bytes);

/* Eikonal schedule and precision of a context: band_tiles is how many 4^3
 * tiles the relaxation band advances per round (smaller = fewer re-visits,
 * more rounds); t64 = 1 stores arrival times in FP64 (the parity setting),
 * 0 in FP32.  A context starts with FP64 times, velocities and local solves
 * (the reference's precision; the environment variable MCT_EIK_F64=0 at
 * create keeps FP64 times with FP32 velocities and solves); t64 = 0 switches
 * all three to FP32 for the rest of the context's life.  Values <= 0 / < 0
 * leave the setting unchanged. */
int mct_set_tuning(mct_ctx *ctx, float band_tiles, int t64);

/* How order 1 (LIKE_SET%order = 1, highAccuracySolveEikonal) is solved.  The
 * reference's mixed-order times depend on the order in which its heap pops
 * cells (fmm/fmm.hpp:92-158 over fmm/sfmm.hpp:44-52), not on the field alone.
 * MCT_ORDER1_HEAP (the default) replays that heap loop exactly, one wave per
 * (chain, wave type, source): the times equal the reference's SFMM.
 * MCT_ORDER1_PARALLEL relaxes the same local update to its fixed point with
 * the tiled solver used for order 0: much faster, within about 1e-2 relative
 * of the reference.  The environment variable MCT_EIK_ORDER1=parallel sets the
 * default.  The heap replay needs FP64 times (mct_set_tuning with t64 = 0
 * selects the parallel method). */
enum { MCT_ORDER1_HEAP = 0, MCT_ORDER1_PARALLEL = 1 };
int mct_body_set_order1(mct_ctx *ctx, int method);

/* Wait for the context's stream. */
int mct_sync(mct_ctx *ctx);

/* Stage timing of the last synchronised call, measured with HIP events on
 * the context's stream: [0] nearest-site ms, [1] eikonal ms (all rounds with
 * the convergence checks between them), [2] picks+misfit ms, [3] eikonal
 * rounds; with kernel timing on, [4] summed duration of the eikonal round
 * kernels alone (ms) and [5] the number of round launches timed. */
int mct_last_timings(mct_ctx *ctx, double *ms6);

/* Turn per-launch events around every eikonal round kernel on or off (off by
 * default: they cost a little host time per round). */
int mct_set_kernel_timing(mct_ctx *ctx, int on);

/* ---- surface-wave likelihood context (datatype 2) ------------------------- */
/* Replaces likelihood() for datatype 2 (src/likelihood.f90:72-80 ->
 * surf_likelihood, src/likelihood_surf.F90:110-405) with phase or group
 * velocities (LIKE_SET%phaseGroup), with or without a water layer.  The setup mirrors T_DATA /
 * T_LIKE_SET (src/likelihood_settings.f90:19-79); arrays are copied at create
 * time.  With group velocities modrays traces a ray per used pair (rpaths,
 * fm2d/fm2dray_cartesian.f90:801-1420), the group delays along the rays
 * (CalGroupTime, likelihood_surf.F90:454-521) are the times, srdist holds the
 * ray lengths, and a ray that hits rpaths' point limit rejects the model
 * (MCT_REJECT_RAY, :341-345). */
typedef struct {
    int raylov;               /* 1 Rayleigh, 0 Love (LIKE_SET%raylov) */
    int phase_group;          /* 0 phase, 1 group (LIKE_SET%phaseGroup); 1 needs sgref = 1 and
                                 raystat(n,2) = n on used pairs, as the data readers set it */
    double dphase;            /* LIKE_SET%dPhaseVel */
    int np;                   /* periods, <= 60 */
    const double *freqs;      /* (np) Hz, dat%freqs */
    int nsrc, nrev;
    const double *src;        /* (nsrc,2) x, y: dat%src(1:2,:) */
    const double *rev;        /* (nrev,2) x, y: dat%rev(1:2,:) */
    const double *ttime;      /* (np,3,nsrc*nrev) = Fortran dat%ttime(nsrc*nrev,3,np) */
    const int32_t *raystat;   /* (np,2,nsrc*nrev) = Fortran dat%raystat(nsrc*nrev,2,np) */
    int sigdep;               /* LIKE_SET%sigdep */
    int nrays;                /* dat%nrays */
    int gridx, gridy;         /* LIKE_SET%gridx / gridy: B-spline dicing */
    int sgref, sgdic, sgext;  /* refined source grid switch, dicing, extent */
    int order;                /* LIKE_SET%order: 0 first order, 1 mixed order */
    double band;              /* LIKE_SET%band (narrow-band size fraction) */
    int max_cells;            /* largest Voronoi site count a call may pass */
    int max_chains;           /* largest batch for the _batch_dev entry */
    int straight;             /* LIKE_SET%isStraight: 1 = times along straight source-receiver
                                 rays (src/likelihood_surf.F90:233-243,412-452), no 2-D solve;
                                 srdist = the distances */
    int datatype;             /* LIKE_SET%datatype: 2 (or 0) surface waves alone, vp = 1.73 vs and
                                 rho = 1.74 vp^0.25 (vs2vp_3d / vp2rho_3d, src/likelihood.f90:
                                 75-76); 3 the surface half of the joint likelihood: vp and rho
                                 are the sites' own parameters (RTI%parameters(1,:), (3,:);
                                 src/likelihood.f90:83-88) */
    double water_depth;       /* T_GRID%waterDepth (src/settings.f90:26), km, not scaled: above
                                 EPS every column gets a water layer on top (vp 1.5, vs 0,
                                 rho 1; src/likelihood_surf.F90:32-33,546-551,616-617); 0: none */
} mct_surf_setup;

typedef struct mct_surf_ctx mct_surf_ctx;

/* like_setup for a surface-wave data set.  grid as in mct_voronoi_to_grid
 * (dz scaled); grid->nz <= 100 (NMAX, src/likelihood_surf.F90:30). */
int mct_surf_create(const mct_grid *grid, const mct_surf_setup *setup, mct_surf_ctx **ctx);
int mct_surf_destroy(mct_surf_ctx *ctx);

/* One likelihood() call for one chain (host arrays).  points/params
 * (ncells,3) as RTI%points / RTI%parameters (vp, vs, rho: only vs is used;
 * vp and rho follow vs2vp_3d / vp2rho_3d).  snoise0/snoise1 (np) (may be NULL
 * when sigdep = 0).  out = {like, misfit, unweighted_misfit}.  phase_time
 * (np,nsrc,nrev), srdist and sigma (np,nsrc*nrev), phase_vel (nx,ny,np)
 * (Fortran pvel(np,ny,nx)) are filled when not NULL.  Returns MCT_OK,
 * MCT_REJECT_MODEL (check_model or a dispersion failure: the shim sets
 * like = huge and counts RTI%num_bad_model) or a negative error. */
int mct_surf_likelihood(mct_surf_ctx *ctx, const double *points, const double *params, int ncells,
                        const double *snoise0, const double *snoise1, double *out,
                        double *phase_time, double *srdist, double *sigma, double *phase_vel);

/* nchain likelihood() calls at once on device arrays: points/params
 * (nchain,max_cells,3), ncells (nchain), noise (nchain,2,np) = snoise0,
 * snoise1; out (nchain,3), status (nchain).  Asynchronous on the context's
 * stream; mct_surf_sync waits.  A batch can run in parts (environment
 * MCT_SURF_SPLIT at create, default 1) on streams of the context, each part
 * started when the previous part's dispersion is done so that one part's
 * 2-D solves overlap the next part's dispersion; results do not depend on it. */
int mct_surf_likelihood_batch_dev(mct_surf_ctx *ctx, int nchain, const double *points_dev,
                                  const double *params_dev, const int32_t *ncells_dev,
                                  const double *noise_dev, double *out_dev, int32_t *status_dev);
int mct_surf_sync(mct_surf_ctx *ctx);
/* mct_surf_likelihood_batch_dev replayed from a HIP graph (no counterpart in
 * the reference): the first call with a given nchain and set of device
 * pointers captures the context's launch sequence for one batch (its flag
 * clears are kernels, not memsets) and later calls with the same arguments
 * replay it with one launch (a context keeps up to 64 such graphs; the
 * served form's lanes replay one per batch size).  Stage timings are not recorded in graph mode.
 * A batch that runs in parts (MCT_SURF_SPLIT) is MCT_ERR_ARG.  Results equal
 * the eager call's. */
int mct_surf_likelihood_batch_graph(mct_surf_ctx *ctx, int nchain, const double *points_dev,
                                    const double *params_dev, const int32_t *ncells_dev,
                                    const double *noise_dev, double *out_dev, int32_t *status_dev);

/* likelihood() for a proposal that changed the model only inside
 * perturbed_box (src/likelihood_surf.F90:110-270): convert_to_layer and
 * surfdisp96 run only for the box's columns widened by one (:155-169); every
 * other column keeps the phase velocities of the chain's accepted state, as
 * like%vel keeps them there; the 2-D solves and the misfit cover the whole
 * grid as in the reference.  box = {x0, y0, z0, x1, y1, z1}
 * (perturbed_box(1:2)).  The chain needs an accepted state (mct_surf_accept
 * after a call whose proposal the sampler accepted), else MCT_ERR_ARG.  The
 * result equals the whole-grid call's when the model outside the box is the
 * accepted one. */
int mct_surf_likelihood_box(mct_surf_ctx *ctx, const double *points, const double *params, int ncells,
                            const double *box, const double *snoise0, const double *snoise1, double *out,
                            double *phase_time, double *srdist, double *sigma, double *phase_vel);
/* The column range of box as the reference computes it: cols = {ix0, ix1,
 * iy0, iy1}, 1-based inclusive, widened by one and clamped to the grid. */
int mct_surf_box_columns(mct_surf_ctx *ctx, const double *box, int32_t *cols);
/* Batched form: cols (nchain, 4) host int32 column ranges (as
 * mct_surf_box_columns gives), the rest as mct_surf_likelihood_batch_dev. */
int mct_surf_likelihood_box_batch_dev(mct_surf_ctx *ctx, int nchain, const double *points_dev,
                                      const double *params_dev, const int32_t *ncells_dev, const int32_t *cols,
                                      const double *noise_dev, double *out_dev, int32_t *status_dev);
/* The sampler accepted the proposal of the context's last likelihood call for
 * chain (src/mcmc_loc2.f90:549-567 keeps like then; on reject it restores
 * like_copy): its phase velocities become the chain's accepted state, from
 * which later box calls start.  Queued on the context's stream. */
int mct_surf_accept(mct_surf_ctx *ctx, int chain);
/* The same for chains 0..nchain-1 with accepted[c] != 0. */
int mct_surf_accept_batch(mct_surf_ctx *ctx, int nchain, const int32_t *accepted);

/* Group velocities of chain's last call, (nx,ny,np) = Fortran gvel(np,ny,nx)
 * (surfmodes' gvel output, like%gvel; meaningful with phase_group = 1).
 * Waits for the context's stream. */
int mct_surf_group_velocities(mct_surf_ctx *ctx, int chain, double *group_vel);

/* noise_likelihood for datatype 2 (src/likelihood.f90:131-150 ->
 * surf_noise_likelihood, src/likelihood_surf.F90:56-99): the likelihood of the
 * phase times kept from the context's last likelihood call under new noise
 * parameters, sigma = snoise0*t + snoise1 (np each).  No dispersion or 2-D
 * solve.  MCT_ERR_SIGMA where the reference stops on a zero noise level. */
int mct_surf_noise_likelihood(mct_surf_ctx *ctx, const double *snoise0, const double *snoise1, double *out);
/* surf_noise_likelihood over traveltimes the caller passes (phase_time
 * (nrev,nsrc,np) Fortran = like%phaseTime; srdist (nsrc*nrev,np) = like%srdist,
 * needed with group velocities or straight rays, else NULL), so a sampler
 * that restored like_copy after a reject gets the accepted state's misfit.
 * sigma may be NULL. */
int mct_surf_noise_likelihood_from(mct_surf_ctx *ctx, const double *phase_time, const double *srdist,
                                   const double *snoise0, const double *snoise1, double *out, double *sigma);
/* The same for the first nchain chains of the last batch: noise_dev
 * (nchain, 2, np); asynchronous on the context's stream. */
int mct_surf_noise_likelihood_batch_dev(mct_surf_ctx *ctx, int nchain, const double *noise_dev,
                                        double *out_dev, int32_t *status_dev);

/* Stage times of the last synchronised call (ms): nearest site, dispersion,
 * 2-D traveltimes, misfit.  For a batch run in parts the first three are the
 * first part's stages and the last runs to the end of the whole batch. */
int mct_surf_last_timings(mct_surf_ctx *ctx, double *ms4);

/* The 2-D stage of the last synchronised call split at the sweep launch (ms):
 * velocity grids and the near-source kernel, then the sweep with the receiver
 * picks (first launch chunk; 0 when the stage did not run the parallel path). */
int mct_surf_last_timings2d(mct_surf_ctx *ctx, double *near_ms, double *sweep_ms);

/* Number of parts a batch of nchain chains runs in (return value, or < 0 on
 * error) and the chains of the first part. */
int mct_surf_split_parts(mct_surf_ctx *ctx, int nchain, int *first_part_chains);

/* 2-D traveltime method: 0 parallel sweeps with the heap loop near the
 * source (default), 1 the sequential heap loop per (period, source). */
int mct_surf_set_method(mct_surf_ctx *ctx, int fm2d_method);

/* Dispersion launch shape: 0 by size (default), 1 one lane per column
 * (throughput), 2 one wave per column (latency), 3 one lane per column with
 * columns taken from a queue in layer-count order.  Results are the same. */
int mct_surf_set_dispersion_mode(mct_surf_ctx *ctx, int mode);

/* Launch shape of a context, for tuning (no counterpart in the reference;
 * results do not depend on it): MCT_TUNE_DSP_WAVES waves of the queue
 * dispersion kernel (0: one per 64 columns), MCT_TUNE_SWEEP_KIND the 2-D
 * sweep (0 default, 1 strip-pipelined, 2 anti-diagonal), MCT_TUNE_SWEEP_THREADS
 * threads per problem of the anti-diagonal sweep (0 default, 64, 128, 256),
 * MCT_TUNE_FM2D_CHUNK 2-D problems per launch (0: all the workspace holds),
 * MCT_TUNE_DSP_DEDUP 1 (default) to solve each distinct Vs column profile of
 * a chain once and copy its curves to the identical columns, 0 to solve
 * every column.  MCT_ERR_ARG for an unknown key or value.  get returns the
 * value in use. */
enum { MCT_TUNE_DSP_WAVES = 1, MCT_TUNE_SWEEP_KIND = 2, MCT_TUNE_SWEEP_THREADS = 3, MCT_TUNE_FM2D_CHUNK = 4,
       MCT_TUNE_DSP_DEDUP = 5 };
int mct_surf_set_tuning(mct_surf_ctx *ctx, int key, int64_t value);
int mct_surf_get_tuning(mct_surf_ctx *ctx, int key, int64_t *value);

/* Work of the last call's dispersion stage over its first nchain chains:
 * evals = period-equation evaluations of surfdisp96's sequential root walk
 * (dltar calls), layer_evals = the same weighted by (layers - 1), the
 * per-evaluation layer recursion count. */
int mct_surf_dispersion_work(mct_surf_ctx *ctx, int nchain, int64_t *evals, int64_t *layer_evals);

/* The part of that work the last call actually solved: columns solved (the
 * others copied an identical column of the same chain) and their evaluations
 * weighted by (layers - 1).  No counterpart in the reference: a measurement. */
int mct_surf_dispersion_solved(mct_surf_ctx *ctx, int nchain, int64_t *columns, int64_t *layer_evals);

/* Work of the parallel 2-D path in the last launch chunk: {slots swept, mean
 * sweep rounds (four directions each), mean diagonal steps executed, mean
 * node updates} per (period, source) problem. */
int mct_surf_fm2d_stats(mct_surf_ctx *ctx, double *stats4);

/* Problems of the last launch chunk that the parallel 2-D path handed to the
 * sequential solver: [0] failed the heap-order check, [1] all hand-overs (the
 * check and a heap loop that outgrew its LDS neighbourhood), [2..8] problems
 * per reason of the check: no finite time, tied times, a key at or below the
 * next neighbour's time, a rising key, a last key that is not the time, an
 * invalid band heap at the hand-over, a sweep that did not settle.  No
 * counterpart in the reference: a measurement. */
int mct_surf_fm2d_handover(mct_surf_ctx *ctx, int64_t *counts9);

/* The context's HIP stream (hipStream_t) for callers that queue their own
 * copies in order with the context's work. */
int mct_surf_stream(mct_surf_ctx *ctx, void **stream);

/* ---- joint body + surface-wave likelihood (datatype 3) ---------------------- */
/* like_setup / likelihood for LIKE_SET%datatype = 3 (src/likelihood.f90:44-52,
 * 83-88): body_likelihood on dat(1) with P and S plus surf_likelihood on
 * dat(2), whose layers take the sites' own vp and rho, summed.  body must have
 * datatype 3, surf datatype 3, and both the same max_cells and max_chains; the
 * grid is shared.  A Fortran caller can equally keep the two contexts itself
 * and add the results, as the reference does; this context runs the two
 * pipelines of a batch at once on their own streams. */
typedef struct mct_joint_ctx mct_joint_ctx;
int mct_joint_create(const mct_grid *grid, const mct_body_setup *body, const mct_surf_setup *surf,
                     mct_joint_ctx **ctx);
int mct_joint_destroy(mct_joint_ctx *ctx);
/* The joint context's body and surface contexts (for their noise and
 * location likelihoods, arrays and timings; owned by the joint context). */
int mct_joint_parts(mct_joint_ctx *ctx, mct_ctx **body, mct_surf_ctx **surf);
/* One chain: points/params (ncells,3), locations (nsrc,4), bnoise0/bnoise1
 * (2: P, S), snoise0/snoise1 (np surface periods); any noise may be NULL
 * (zero).  out = {like, misfit, unweighted} summed; out_parts (6, may be NULL)
 * = the body then the surface triple.  Returns the status: 0, > 0 the surface
 * half rejected the model (like = huge), < 0 fatal. */
int mct_joint_likelihood(mct_joint_ctx *ctx, const double *points, const double *params, int ncells,
                         const double *locations, const double *bnoise0, const double *bnoise1,
                         const double *snoise0, const double *snoise1, double *out, double *out_parts);
/* nchain chains on device arrays: points/params (nchain,max_cells,3), ncells
 * (nchain), locations (nchain,nsrc,4), bnoise (nchain,2,2), snoise
 * (nchain,2,np); out (nchain,3), status (nchain).  Asynchronous; mct_joint_sync
 * waits. */
int mct_joint_likelihood_batch_dev(mct_joint_ctx *ctx, int nchain, const double *points_dev,
                                   const double *params_dev, const int32_t *ncells_dev,
                                   const double *locations_dev, const double *bnoise_dev,
                                   const double *snoise_dev, double *out_dev, int32_t *status_dev);
int mct_joint_sync(mct_joint_ctx *ctx);

/* ---- served form: continuous batching -------------------------------------- */
/* No counterpart in the reference (its sampler calls likelihood() one chain
 * at a time).  A server owns nlanes surface contexts, each with room for
 * max_batch chains; a dispatcher thread packs whatever requests are pending
 * into a batch for the first free lane, so batches on different lanes
 * overlap on the device and requests that arrive meanwhile are admitted as
 * soon as a lane frees.  Device memory is fixed at create time (admission is
 * bounded by nlanes x max_batch slots).  Each result equals the plain call's. */
typedef struct mct_server mct_server;
int mct_serve_create(const mct_grid *grid, const mct_surf_setup *setup, int max_batch, int nlanes,
                     mct_server **srv);
/* Copies the inputs (points/params (ncells,3), snoise0/1 (np) or NULL) and
 * returns at once with a ticket. */
int mct_serve_submit(mct_server *srv, const double *points, const double *params, int ncells,
                     const double *snoise0, const double *snoise1, int64_t *ticket);
/* Blocks until the ticket's evaluation is done: out {like, misfit,
 * unweighted}, status as mct_surf_likelihood's, latency submit -> done (ms). */
int mct_serve_wait(mct_server *srv, int64_t ticket, double *out, int32_t *status, double *latency_ms);
/* Batches launched and evaluations finished so far; returns the first error. */
int mct_serve_stats(mct_server *srv, int64_t *batches, int64_t *evals);
/* Batch fill (no counterpart in the reference): while some lane is running, a
 * batch smaller than its lane's max_batch waits until its oldest request has
 * waited fill_ms (default 25); with no lane running it is launched at once.
 * fill_ms < 0 launches whatever is pending as soon as a lane is free. */
int mct_serve_set_fill(mct_server *srv, double fill_ms);
/* A server over a mix of models (no counterpart in the reference): model m is
 * a data set on its own grid (grids[m], setups[m]) with lanes of max_batch[m]
 * chains, at most max_lanes lanes per model.  Lanes are admitted by device
 * footprint: each model starts with one lane (MCT_ERR_NOMEM if those exceed
 * hbm_budget bytes); more are created on demand while the bytes of all lanes
 * stay within hbm_budget (0: no limit), and when a model has requests waiting
 * and no room, idle lanes of models with nothing pending are released, least
 * recently used first.  The setups' arrays are copied. */
int mct_serve_create_mix(int nmodels, const mct_grid *grids, const mct_surf_setup *setups, const int *max_batch,
                         int max_lanes, int64_t hbm_budget, mct_server **srv);
/* mct_serve_submit for model `model` of a mix (mct_serve_submit is model 0). */
int mct_serve_submit_to(mct_server *srv, int model, const double *points, const double *params, int ncells,
                        const double *snoise0, const double *snoise1, int64_t *ticket);
/* Device bytes held by the lanes now and at most so far, the lanes now, and
 * how many lanes were released to make room. */
int mct_serve_memory(mct_server *srv, int64_t *bytes_in_use, int64_t *peak_bytes, int32_t *lanes,
                     int64_t *lane_releases);
/* Finishes the pending requests and frees everything. */
int mct_serve_destroy(mct_server *srv);

/* Device bytes a surface context holds (all its buffers). */
int mct_surf_device_bytes(mct_surf_ctx *ctx, int64_t *bytes);

/* Device pointers of the context's phase_time (nchain,np,nsrc,nrev) and
 * phase velocity (nchain,nx,ny,np) buffers (valid until destroy). */
int mct_surf_dev_arrays(mct_surf_ctx *ctx, double **phase_time_dev, double **phase_vel_dev);

/* Library version string and the device name the library runs on. */
const char *mct_version(void);

#ifdef __cplusplus
}
#endif
#endif
