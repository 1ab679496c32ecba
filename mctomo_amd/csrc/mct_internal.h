// mct_internal.h -- shared declarations of the HIP implementation of the
// MCTomo likelihood forward model (gfx950).  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "../../include/mctomo_hip.h"

#define MCT_TILE 4                 // eikonal tiles are 4x4x4 nodes = one wave64
#define MCT_TILE_N 64
#define MCT_KD_BUCKET 12           // kdtree2 bucket_size (src/kdtree2.f90:516)
#define MCT_MAX_FLAGGED 4096       // near-tie nodes replayed per chain before a full replay

#define MCT_CHECK(expr)                                                       \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return MCT_ERR_HIP;                             \
    } while (0)

// Geometry of one rasterised grid and its eikonal tiling.
struct mct_geom {
    int nx, ny, nz;
    double xmin, ymin, zmin;       // z scaled (Voronoi distances)
    double dx, dy, dz;             // dz scaled
    double scaling;
    int ntx, nty, ntz, ntiles;     // tiles of 4x4x4
};

static inline mct_geom mct_make_geom(const mct_grid *g)
{
    mct_geom G;
    G.nx = g->nx; G.ny = g->ny; G.nz = g->nz;
    G.xmin = g->xmin; G.ymin = g->ymin; G.zmin = g->zmin;
    G.dx = g->dx; G.dy = g->dy; G.dz = g->dz;
    G.scaling = g->scaling > 0 ? g->scaling : 1.0;
    G.ntx = (G.nx + MCT_TILE - 1) / MCT_TILE;
    G.nty = (G.ny + MCT_TILE - 1) / MCT_TILE;
    G.ntz = (G.nz + MCT_TILE - 1) / MCT_TILE;
    G.ntiles = G.ntx * G.nty * G.ntz;
    return G;
}

// Per-chain scratch of the exact kd-tree replay (rare near-tie nodes).
struct mct_kd_node {
    double cut_val, cut_val_left, cut_val_right;
    double lo[3], hi[3];
    int cut_dim, l, u, left, right, pad;
};

// ---- nearest-site stage ----------------------------------------------------
struct mct_nearest_args {
    mct_geom G;
    int nchain;
    int max_cells;
    const double *points;          // (nchain, max_cells, 3)
    const double *params;          // (nchain, max_cells, 3)
    const int32_t *ncells;         // (nchain) or NULL -> ncells_const
    int ncells_const;
    int ix0, ix1, iy0, iy1, iz0, iz1;   // 1-based inclusive box
    const double *pm;              // {vp,vs,rho} or NULL
    // outputs (any may be NULL)
    float *vel_tiles;              // (nchain, nwave, ntiles*64) eikonal velocity
    double *vel_tiles64;           // the same in FP64 (either may be NULL)
    int nwave;                     // 1: vp only, 2: vp and vs
    int32_t *sites_id;             // (nchain, nz*ny*nx) Fortran order
    double *vp, *vs, *rho;         // (nchain, nz*ny*nx) Fortran order
    // near-tie bookkeeping (per chain)
    int32_t *flag_count;           // (nchain)
    int32_t *flag_nodes;           // (nchain, MCT_MAX_FLAGGED) Fortran node index
    mct_kd_node *kd_nodes;         // (nchain, 2*max_cells+2)
    int32_t *kd_ind;               // (nchain, max_cells+1)
    double *kd_rdata;              // (nchain, 3*max_cells)
};

int mct_launch_nearest(const mct_nearest_args &a, hipStream_t s);

// ---- 3-D eikonal stage -------------------------------------------------------
struct mct_eik_problem_set {
    mct_geom G;
    int nprob;                     // problems = chains * waves * sources
    int nsrc_eff;                  // sources traced per (chain, wave)
    int nwave;
    int order;
    void *T;                       // (nprob, ntiles*64) arrival times, float or double
    int t64;                       // 1: arrival times stored as double
    const float *vel;              // (nchain*nwave, ntiles*64), or NULL with vel64
    const double *vel64;           // the same in FP64: local solves in FP64 (needs t64), or NULL
    uint8_t *flags[2];             // active-tile flags of the current / next round
    uint32_t *key[2];              // per tile: smallest activating time (float bits)
    uint32_t *any;                 // [3] rotating "some tile is active" flags
    const float *band_dt;          // (nchain*nwave) band time step per round
    int band_rounds;               // rounds with the band schedule, then plain FIM
    int32_t *status;               // (nprob) or NULL
};

// Initialise arrival times from source points (Solver::setInitialPoints,
// fastmarching/fmm/solver.hpp:72-92) and seed the first active list.
// Source s of chain c is at src_pts + c*src_chain_stride + s*src_stride
// (doubles); xmin/ymin/zmin are subtracted inside, leaf sizes lx, ly, lz.
int mct_launch_eik_init(const mct_eik_problem_set &P, const double *src_pts, int src_stride,
                        long src_chain_stride, double xmin, double ymin, double zmin,
                        double lx, double ly, double lz, hipStream_t s);
int mct_launch_eik_round(const mct_eik_problem_set &P, int round, double hx, double hy, double hz,
                         int grid_blocks, hipStream_t s);
// band_dt[cw] = tiles_per_round * 4 * min(h) / max velocity of (chain, wave) cw
// (from vel64 when given, else vel)
int mct_launch_band_dt(const float *vel, const double *vel64, long tsz, int ncw, float hmin, float tiles_per_round,
                       float *band_dt, hipStream_t s);

// Band schedule: the first band rounds advance tau by band_dt per round; then
// tau is lifted (plain fast-iterative rounds) until no tile is active.
static inline int mct_band_rounds(const mct_geom &G, float band_tiles)
{
    float bt = band_tiles > 0.05f ? band_tiles : 0.05f;
    return (int)(6.0f * (G.ntx + G.nty + G.ntz) / bt) + 8;
}
// Tiles per round of the band schedule (environment MCT_BAND_TILES, else 0.5).
float mct_default_band_tiles();
// Arrival-time storage: double unless the environment sets MCT_EIK_F32=1.
int mct_default_t64();
// FP64 velocities and local solves: on unless the environment sets MCT_EIK_F64=0.
int mct_default_f64();

// Exact replay of the reference's SFMM heap loop (fmm/fmm.hpp:92-158), one
// wave per problem, FP64 times and velocities.  T must hold +inf everywhere
// before the launch (mct_launch_eik_init clears it); the kernel writes the
// source cells itself.  heap: (nprob, hcap) scratch; err: one word of error
// bits (read back with mct_sfmm_status); pops: (nprob) pop counts or NULL.
struct mct_sfmm_args {
    mct_geom G;
    int nprob, nsrc_eff, nwave, order;
    double *T;                     // (nprob, ntiles*64)
    const double *V;               // (nchain*nwave, ntiles*64)
    const double *src;             // source s of chain c at src + c*src_chain_stride + s*src_stride
    int src_stride;
    long src_chain_stride;
    double xmin, ymin, zmin;       // subtracted from the source coordinates
    double lx, ly, lz;             // leaf sizes
    uint32_t *heap;
    long hcap;
    uint32_t *err;
    uint32_t *pops;
};
long mct_sfmm_heap_cap(const mct_geom &G);
int mct_launch_sfmm(const mct_sfmm_args &a, hipStream_t s);
int mct_sfmm_status(uint32_t err_bits);
// Order 1 replays the heap unless the environment sets MCT_EIK_ORDER1=parallel.
int mct_default_heap_order1();

// ---- picks + misfit -------------------------------------------------------------
struct mct_pick_args {
    mct_geom G;
    const void *T;                 // (nprob, ntiles*64)
    int t64;                       // T holds doubles
    int nchain, nwave, nsrc_eff, npick;
    const double *pick_pts;        // points to pick at, (npick,3) shared or per chain
    int pick_stride_chain;         // 0 if shared across chains, else doubles per chain
    int pick_point_stride;         // doubles between points (3 or 4)
    double xmin, ymin, zmin, lx, ly, lz;
    double *out;                   // (nchain, nwave, nsrc_eff, npick)
    int interp3d;                  // Interpolate3d (likelihood_body.F90:976-1007) instead of getTime
};
int mct_launch_picks(const mct_pick_args &a, hipStream_t s);

// n int32 set to v by a kernel (a stream's memset as a kernel node, so that
// launch sequences that clear flags can be captured in a HIP graph)
int mct_fill_i32(int32_t *p, long n, int32_t v, hipStream_t s);

struct mct_misfit_args {
    int nchain, nsrc, nrev, np;
    int reciprocal;                // picks are (rev-major) when traced from receivers
    const double *picks;           // (nchain, np, nsrc_eff, npick)
    const double *ttime;           // (np, 3, nsrc*nrev)  observed
    const int32_t *raystat;        // (np, 2, nsrc*nrev)
    const double *locations;       // (nchain, nsrc, 4): origin time locations[..][3]; NULL = none
    const double *noise;           // (nchain, 2, np) or NULL
    int sigdep, nrays;
    double eps;                    // smallest admissible sigma
    double *out;                   // (nchain, 3); like = DBL_MAX for a chain whose status is > 0
    double *phase_time;            // (nchain, np, nsrc, nrev) or NULL
    double *srdist, *sigma;        // (nchain, np, nsrc*nrev) or NULL
    int32_t *status;               // (nchain) or NULL
    const double *dist;            // (nchain, np, nsrc*nrev) srdist when it is not the pick (ray
                                   // lengths for group times), or NULL
};
int mct_launch_misfit(const mct_misfit_args &a, hipStream_t s);

// ---- 2-D phase traveltimes (fm2d) -------------------------------------------------
struct mct_fm2d_args {
    int nvx, nvz;               // control grid nodes (grid nx, ny)
    double gox, goz, dvx, dvz;  // origin and control spacing
    int gdx, gdz, asgr, sgdl, sgs, fom;
    int nnx, nnz;               // coarse (B-spline) grid
    int ld, nc;                 // working arrays: ld rows (z) x nc columns (x)
    int nchain, np, nsrc, nrc;
    const double *velv;         // (nchain*np, (nvx+2)*(nvz+2)) padded control grids, Fortran (nvz+2,nvx+2)
    double *veln;               // (nchain*np, nnx*nnz) coarse velocity, Fortran (nnz,nnx)
    double *sln;                // the same layout: 1.0 / veln (the sweeps' slowness), or NULL
    const double *scx, *scz, *rcx, *rcz;
    const int32_t *srs;         // per period srs_pstride entries; the first nsrc*nrc mark the pairs used (receiver fastest)
    long srs_pstride;
    double *ttn, *ttr;          // (nprob, ld*nc) coarse/refined times (ttr = refined copy)
    double *vr;                 // (nprob, ld*nc) refined velocity
    int32_t *nsts, *nstr;       // (nprob, ld*nc)
    int32_t *heap;              // (nprob, ld*nc+8) packed (px<<16)|pz
    double *ttime;              // (nchain, np, nsrc, nrc): phase times, receiver fastest
    int32_t *status;            // (nchain): MCT_ERR_OUTSIDE for a source or receiver outside the
                                // grid; chains whose status is > 0 (rejected) are skipped
    long prob0, nrun;           // problems [prob0, prob0+nrun) use workspace slots 0..nrun-1
    const int32_t *only_flag;   // sequential solver: when set, only slots whose flag is 1
    int32_t *rounds_out;        // parallel path: sweep rounds per slot, or NULL
    hipEvent_t ev_sweep;        // parallel path: recorded before the first chunk's sweep, or NULL
    // group velocities (modrays with uar = 0): after each chunk's solve, rays
    // from the receivers of every used pair; the group delays along them
    // replace ttime and their lengths go to rlen.  The solvers then also keep
    // each problem's refined source grid (times in ttr, states in nstr).
    int rays;
    const double *gvel;         // (nchain, nvx*nvz, np) group velocities, column iy + nvz*ix (0-based)
    double *rlen;               // (nchain, np, nsrc, nrc) ray lengths, receiver fastest
    // launch shape of the parallel path (0: the environment's or the default)
    int sweep_kind;             // 1 strip-pipelined, 2 anti-diagonal
    int sweep_threads;          // anti-diagonal sweep: 64, 128 or 256 threads per problem
    // parallel path: check each converged field against the heap loop's freeze
    // order and hand the problems that fail to the sequential solver (1 on,
    // -1 off, 0 the default: the environment's MCT_FM2D_VERIFY, else on for
    // the mixed-order update, fom 1)
    int verify;
};
// Coarse B-spline velocity per (chain, period), then one lane per
// (chain, period, source) running the sequential modrays phase solve.
// Workspace slots (ttn..heap) are sized for `chunk` problems; problems run in
// chunks of that many.
int mct_launch_fm2d(const mct_fm2d_args &a, long chunk, hipStream_t s);
int mct_launch_fm2d_gridder(const mct_fm2d_args &a, hipStream_t s);
int mct_launch_fm2d_solve(const mct_fm2d_args &b, hipStream_t s);
// The problems whose flag is 1 (handed over by the parallel path), one wave
// each with the heap in LDS; a solved problem's flag becomes 3, one whose band
// outgrew the LDS heap keeps 1 (for mct_launch_fm2d_solve).
int mct_launch_fm2d_wave(const mct_fm2d_args &b, int32_t *pflag, hipStream_t s);
// Parallel path: seed and source neighbourhood in LDS per problem, coarse
// sweeps per problem, the sequential solver for flagged problems.  pflag
// holds one int per workspace slot.  mct_fm2d_fast_ok tells whether the
// geometry fits the LDS budget of the parallel path.
int mct_fm2d_fast_ok(const mct_fm2d_args &a);
int mct_launch_fm2d_fast(const mct_fm2d_args &a, long chunk, int32_t *pflag, hipStream_t s);
// rpaths + ray length + CalGroupTime for the chunk [b.prob0, b.prob0 + b.nrun):
// one thread per (problem, receiver)
int mct_launch_fm2d_rays(const mct_fm2d_args &b, hipStream_t s);
// straight rays (isStraight = 1): group delays along the source-receiver
// segments for a.nchain chains into a.ttime, through a.gvel
int mct_launch_straight_rays(const mct_fm2d_args &a, hipStream_t s);
// Bytes of fm2d workspace per problem slot.
static inline long mct_fm2d_slot_bytes(long ld, long nc)
{
    return ld * nc * (8 + 8 + 8 + 4 + 4) + 4 * (ld * nc + 8);
}

// ---- dispersion (surfdisp96 per column) ------------------------------------------
struct mct_dsp_args {
    int ncol;                     // columns per chain (nx*ny, Fortran order j + ny*i)
    int nz;
    int nchain;
    const double *vs;             // (nchain, nx, ny, nz) Fortran (nz,ny,nx) per chain
    double dz, scaling;           // grid%dz (scaled) and grid%scaling
    double water;                 // T_GRID%waterDepth: a water layer on every column above it (0: none)
    const double *periods;        // (np) periods in s (1/freqs)
    int np;
    int iwave;                    // 1 Love, 2 Rayleigh
    int igr;                      // 0 phase, 1 group
    double dphase;
    double *phase, *group;        // (nchain, ncol, np)
    int32_t *ierr;                // (nchain, ncol)
    int32_t *bad_model;           // (nchain): check_model failed
    int32_t *ndltar;              // (nchain, ncol) or NULL: period-equation evaluations of the
                                  // sequential walk (surfdisp96's own count)
    int32_t *nlayer;              // (nchain, ncol) or NULL: layers of the column
    float *lay;                   // lane-per-column scratch (mct_dispersion_lay_floats) or NULL
                                  // for one wave per column
    unsigned long long *queue;    // column queue counter: with lay, persistent waves
    long lay_waves;               // waves the scratch holds (queue mode)
    int32_t *jobs, *keys;         // queue mode: (nchain*ncol) columns ordered by layer count, or NULL
    int32_t *hist;                // (nz + 1) bins for that order
    // sorted queue mode, optional: columns whose Vs profiles are bit-identical
    // to an earlier column of the same chain are not solved but copied from it
    unsigned long long *hkey;     // (nchain, 1 << hlog) profile-hash table keys, or NULL
    int32_t *hrep;                // (nchain, 1 << hlog) smallest queue position per key
    int32_t *dup;                 // per queue position: the position it copies, or -1
    int32_t *njobs;               // (1) queue positions solved
    int hlog;                     // table slots per chain = 1 << hlog >= 2 * ncol
    // queue mode: only the columns of a box per chain (a perturbed-box call), or NULL
    const int32_t *box;           // (nchain, 4): ix0, ix1, iy0, iy1, 1-based inclusive
    const int32_t *box_off;       // (nchain + 1): first compact index of each chain's box
    long box_total;               // box columns over all chains
    int ny;                       // columns are (ix-1)*ny + (iy-1)
    // joint likelihood (datatype 3): the sites' vp and rho per node, laid out
    // as vs; NULL derives them from vs (datatype 2)
    const double *vp, *rho;
};
// One wave per (column, chain): check_model, convert_to_layer, surfdisp96.
// Requires np <= 60 and nz <= 100.
int mct_launch_dispersion(const mct_dsp_args &a, hipStream_t s);
// Floats of layer scratch the lane-per-column kernel needs.
long mct_dispersion_lay_floats(int ncol, int nz, int nchain);
