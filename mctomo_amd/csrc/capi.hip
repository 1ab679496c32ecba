// capi.hip -- the C ABI of libmctomo_hip.so (declared in include/mctomo_hip.h).
//
// Host orchestration of the body-wave likelihood (src/likelihood.f90:57-93 ->
// src/likelihood_body.F90:151-337): nearest-site rasterisation, 3-D eikonal
// rounds, picks and misfit, all on one HIP stream per context.  Device memory
// is owned by the context; host arrays stay owned by the caller (Fortran) and
// are only read or written during a call, as with c_loc in
// src/fastMarching_wrapper.f90:51-74.
#include "mct_internal.h"
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <new>

#define MCT_VERSION "mctomo-hip 0.1 (gfx950)"

namespace {

struct dbuf {
    void *p = nullptr;
    size_t n = 0;
};

int dalloc(dbuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.n >= bytes) return MCT_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr; b.n = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) { b.p = nullptr; return MCT_ERR_NOMEM; }
    b.n = bytes;
    return MCT_OK;
}

void dfree(dbuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr; b.n = 0;
}

template <class T> T *P(dbuf &b) { return (T *)b.p; }

// Eikonal workspace for a set of problems on one geometry.
struct eik_ws {
    dbuf T, flags0, flags1, key0, key1, counts, band;
    dbuf heap, err;                  // heap replay (order 1): per-problem heaps, error bits
    uint32_t *h_count = nullptr;     // pinned
    int alloc(const mct_geom &G, long nprob)
    {
        long nt = nprob * G.ntiles;
        int rc;
        if ((rc = dalloc(T, sizeof(double) * nt * MCT_TILE_N))) return rc;
        if ((rc = dalloc(flags0, (size_t)((nt + 3) & ~3L)))) return rc;
        if ((rc = dalloc(flags1, (size_t)((nt + 3) & ~3L)))) return rc;
        if ((rc = dalloc(counts, 4 * sizeof(uint32_t)))) return rc;
        if ((rc = dalloc(key0, sizeof(uint32_t) * nt))) return rc;
        if ((rc = dalloc(key1, sizeof(uint32_t) * nt))) return rc;
        if ((rc = dalloc(band, sizeof(float) * (nprob + 1)))) return rc;
        if (!h_count && hipHostMalloc((void **)&h_count, 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
            return MCT_ERR_NOMEM;
        return MCT_OK;
    }
    void release()
    {
        dfree(T); dfree(flags0); dfree(flags1); dfree(key0); dfree(key1); dfree(counts); dfree(band);
        dfree(heap); dfree(err);
        if (h_count) (void)hipHostFree(h_count);
        h_count = nullptr;
    }
};

int round_grid_blocks()
{
    static int blocks = 0;
    if (!blocks) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
        }
        blocks = cus * 32;          // 32 single-wave tile workers per CU
    }
    return blocks;
}

#define MCT_MAX_TIMED_ROUNDS 4096

// Optional per-launch timing of the round kernel: events around every launch.
struct round_timer {
    bool on = false;
    hipEvent_t a[MCT_MAX_TIMED_ROUNDS], b[MCT_MAX_TIMED_ROUNDS];
    int created = 0, used = 0;
    double kernel_ms = 0.0;
    int ensure()
    {
        for (; created < MCT_MAX_TIMED_ROUNDS; created++) {
            if (hipEventCreate(&a[created]) != hipSuccess) return MCT_ERR_HIP;
            if (hipEventCreate(&b[created]) != hipSuccess) { (void)hipEventDestroy(a[created]); return MCT_ERR_HIP; }
        }
        return MCT_OK;
    }
    void release()
    {
        for (int i = 0; i < created; i++) { (void)hipEventDestroy(a[i]); (void)hipEventDestroy(b[i]); }
        created = 0;
    }
    void collect()
    {
        kernel_ms = 0.0;
        for (int i = 0; i < used; i++) {
            float f = 0.f;
            if (hipEventElapsedTime(&f, a[i], b[i]) == hipSuccess) kernel_ms += f;
        }
    }
};

// Run eikonal rounds until no tile is active.  Returns rounds used or <0.
int eik_solve(const mct_eik_problem_set &Pset, eik_ws &ws, double hx, double hy, double hz,
              hipStream_t s, int *rounds_out, round_timer *tm)
{
    const mct_geom &G = Pset.G;
    const int max_rounds = Pset.band_rounds + 8 * (G.ntx + G.nty + G.ntz) + 256;
    const int K = 8;
    int blocks = round_grid_blocks();
    int r = 0;
    for (;;) {
        for (int k = 0; k < K; k++, r++) {
            bool timed = tm && tm->on && r < MCT_MAX_TIMED_ROUNDS;
            if (timed) MCT_CHECK(hipEventRecord(tm->a[r], s));
            int rc = mct_launch_eik_round(Pset, r, hx, hy, hz, blocks, s);
            if (rc) return rc;
            if (timed) { MCT_CHECK(hipEventRecord(tm->b[r], s)); tm->used = r + 1; }
        }
        // any[r % 3] is set when some tile is active for round r
        MCT_CHECK(hipMemcpyAsync(ws.h_count, Pset.any + (r % 3), sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s));
        MCT_CHECK(hipStreamSynchronize(s));
        if (*ws.h_count == 0) break;
        if (r >= max_rounds) return MCT_ERR_NOCONV;
    }
    if (rounds_out) *rounds_out = r;
    return MCT_OK;
}

// The reference's heap loop replayed exactly (order 1: its result depends on
// the pop order).  T must be cleared and seeded by mct_launch_eik_init.
int sfmm_solve(const mct_eik_problem_set &E, eik_ws &ws, const double *src_pts, int sstride, long schain,
               double xmin, double ymin, double zmin, double lx, double ly, double lz, hipStream_t s)
{
    if (!E.t64 || !E.vel64) return MCT_ERR_ARG;
    mct_sfmm_args a;
    memset(&a, 0, sizeof a);
    a.G = E.G; a.nprob = E.nprob; a.nsrc_eff = E.nsrc_eff; a.nwave = E.nwave; a.order = E.order;
    a.T = (double *)E.T; a.V = E.vel64;
    a.src = src_pts; a.src_stride = sstride; a.src_chain_stride = schain;
    a.xmin = xmin; a.ymin = ymin; a.zmin = zmin; a.lx = lx; a.ly = ly; a.lz = lz;
    a.hcap = mct_sfmm_heap_cap(E.G);
    int rc;
    if ((rc = dalloc(ws.heap, sizeof(uint32_t) * (size_t)a.hcap * (size_t)E.nprob))) return rc;
    if ((rc = dalloc(ws.err, 4 * sizeof(uint32_t)))) return rc;
    a.heap = (uint32_t *)ws.heap.p; a.err = (uint32_t *)ws.err.p;
    if ((rc = mct_launch_sfmm(a, s))) return rc;
    MCT_CHECK(hipMemcpyAsync(ws.h_count, a.err, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    return mct_sfmm_status(*ws.h_count);
}

} // namespace

struct mct_ctx {
    mct_geom G;
    mct_body_setup S;
    int nwave, nsrc_eff, npick, reciprocal;
    float band_tiles;                // eikonal schedule: tiles the band advances per round
    int t64;                         // eikonal arrival times stored as double
    int f64;                         // eikonal velocities and local solves in FP64 (with t64)
    int heap_order1;                 // order 1: replay the reference's heap loop (needs f64)
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double ms[4] = {0, 0, 0, 0};
    // immutable data
    dbuf rev, ttime, raystat;
    // per-chain inputs (host entry point stages here)
    dbuf points, params, locations, noise, out, status;
    // stage buffers
    dbuf vel, flag_count, flag_nodes, kd_nodes, kd_ind, kd_rdata, picks;
    dbuf phase_time, srdist, sigma;
    eik_ws eik;
    round_timer *timer = nullptr;
    double *h_out = nullptr;
    int32_t *h_status = nullptr;
    double *h_in = nullptr;          // pinned staging of one chain's inputs
};

static int body_pipeline(mct_ctx *C, int nchain, const double *d_points, const double *d_params,
                         const int32_t *d_ncells, int ncells_const, const double *d_locations,
                         const double *d_noise, double *d_out, int32_t *d_status, bool want_arrays)
{
    const mct_geom &G = C->G;
    hipStream_t s = C->stream;
    int rc;
    MCT_CHECK(hipEventRecord(C->ev[0], s));
    // (a) nearest site -> FP32 velocity tiles (vp, and vs for P&S)
    mct_nearest_args na;
    memset(&na, 0, sizeof na);
    na.G = G; na.nchain = nchain; na.max_cells = C->S.max_cells;
    na.points = d_points; na.params = d_params; na.ncells = d_ncells; na.ncells_const = ncells_const;
    na.ix0 = 1; na.ix1 = G.nx; na.iy0 = 1; na.iy1 = G.ny; na.iz0 = 1; na.iz1 = G.nz;
    na.vel_tiles = C->f64 ? nullptr : P<float>(C->vel);
    na.vel_tiles64 = C->f64 ? P<double>(C->vel) : nullptr;
    na.nwave = C->nwave;
    na.flag_count = P<int32_t>(C->flag_count); na.flag_nodes = P<int32_t>(C->flag_nodes);
    na.kd_nodes = P<mct_kd_node>(C->kd_nodes); na.kd_ind = P<int32_t>(C->kd_ind);
    na.kd_rdata = P<double>(C->kd_rdata);
    if ((rc = mct_launch_nearest(na, s))) return rc;
    MCT_CHECK(hipEventRecord(C->ev[1], s));

    // (c) eikonal: one problem per (chain, wave, traced point)
    const double lz = G.dz / G.scaling, zmin = G.zmin / G.scaling;   // likelihood_body.F90:209-210
    mct_eik_problem_set E;
    memset(&E, 0, sizeof E);
    E.G = G; E.nsrc_eff = C->nsrc_eff; E.nwave = C->nwave; E.order = C->S.order;
    E.nprob = nchain * C->nwave * C->nsrc_eff;
    E.T = C->eik.T.p; E.t64 = C->t64;
    E.vel = C->f64 ? nullptr : P<float>(C->vel);
    E.vel64 = C->f64 ? P<double>(C->vel) : nullptr;
        E.flags[0] = P<uint8_t>(C->eik.flags0); E.flags[1] = P<uint8_t>(C->eik.flags1);
    E.any = P<uint32_t>(C->eik.counts);
    E.key[0] = P<uint32_t>(C->eik.key0); E.key[1] = P<uint32_t>(C->eik.key1);
    E.band_dt = P<float>(C->eik.band);
    E.band_rounds = mct_band_rounds(G, C->band_tiles);
    if ((rc = mct_launch_band_dt(E.vel, E.vel64, (long)G.ntiles * MCT_TILE_N, nchain * C->nwave,
                                 (float)fmin(fmin(G.dx, G.dy), lz), C->band_tiles,
                                 P<float>(C->eik.band), s)))
        return rc;
    const double *src_pts; int sstride; long schain;
    const double *pick_pts; int pstride; int pchain;
    if (!C->reciprocal) {
        src_pts = d_locations; sstride = 4; schain = (long)C->S.nsrc * 4;
        pick_pts = P<double>(C->rev); pstride = 3; pchain = 0;
    } else {   // nsrc > nrev: trace from receivers (likelihood_body.F90:226-249)
        src_pts = P<double>(C->rev); sstride = 3; schain = 0;
        pick_pts = d_locations; pstride = 4; pchain = C->S.nsrc * 4;
    }
    if ((rc = mct_launch_eik_init(E, src_pts, sstride, schain, G.xmin, G.ymin, zmin, G.dx, G.dy, lz, s)))
        return rc;
    int rounds = 0;
    if (C->timer) C->timer->used = 0;
    if (E.order == 1 && C->heap_order1 && C->f64) {
        if ((rc = sfmm_solve(E, C->eik, src_pts, sstride, schain, G.xmin, G.ymin, zmin, G.dx, G.dy, lz, s)))
            return rc;
    } else if ((rc = eik_solve(E, C->eik, G.dx, G.dy, lz, s, &rounds, C->timer))) {
        return rc;
    }
    MCT_CHECK(hipEventRecord(C->ev[2], s));

    // (d) picks + misfit
    mct_pick_args pa;
    memset(&pa, 0, sizeof pa);
    pa.G = G; pa.T = E.T; pa.t64 = E.t64; pa.nchain = nchain; pa.nwave = C->nwave; pa.nsrc_eff = C->nsrc_eff;
    pa.npick = C->npick; pa.pick_pts = pick_pts; pa.pick_stride_chain = pchain;
    pa.pick_point_stride = pstride; pa.xmin = G.xmin; pa.ymin = G.ymin; pa.zmin = zmin;
    pa.lx = G.dx; pa.ly = G.dy; pa.lz = lz; pa.out = P<double>(C->picks); pa.interp3d = 0;
    if ((rc = mct_launch_picks(pa, s))) return rc;
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = nchain; ma.nsrc = C->S.nsrc; ma.nrev = C->S.nrev; ma.np = C->nwave;
    ma.reciprocal = C->reciprocal; ma.picks = P<double>(C->picks);
    ma.ttime = P<double>(C->ttime); ma.raystat = P<int32_t>(C->raystat);
    ma.locations = d_locations; ma.noise = d_noise; ma.sigdep = C->S.sigdep; ma.nrays = C->S.nrays;
    ma.eps = DBL_MIN * 10.0;   // EPS = tiny(0d0)*10 (src/likelihood_body.F90:33)
    ma.out = d_out;
    ma.phase_time = want_arrays ? P<double>(C->phase_time) : nullptr;
    ma.srdist = want_arrays ? P<double>(C->srdist) : nullptr;
    ma.sigma = want_arrays ? P<double>(C->sigma) : nullptr;
    ma.status = d_status;
    if ((rc = mct_launch_misfit(ma, s))) return rc;
    MCT_CHECK(hipEventRecord(C->ev[3], s));
    C->ms[3] = rounds;
    return MCT_OK;
}

// body_noise_likelihood (src/likelihood_body.F90:38-78): the misfit of the
// picks kept from the last likelihood call with sigma = bnoise0*t + bnoise1
// (srdist is the traveltime, :296) and the origin times of d_locations; the
// reference checks no lower bound on sigma here.
static int body_noise(mct_ctx *C, int nchain, const double *d_locations, const double *d_noise,
                      double *d_out, int32_t *d_status)
{
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = nchain; ma.nsrc = C->S.nsrc; ma.nrev = C->S.nrev; ma.np = C->nwave;
    ma.reciprocal = C->reciprocal; ma.picks = P<double>(C->picks);
    ma.ttime = P<double>(C->ttime); ma.raystat = P<int32_t>(C->raystat);
    ma.locations = d_locations; ma.noise = d_noise; ma.sigdep = 1; ma.nrays = C->S.nrays;
    ma.eps = -DBL_MAX;
    ma.out = d_out; ma.status = d_status;
    return mct_launch_misfit(ma, C->stream);
}

// body_location_likelihood with dynamic != 0 (src/likelihood_body.F90:82-149):
// each receiver's field of the last call interpolated at the new source
// locations (Interpolate3d), then the misfit with srdist = that time, sigma =
// bnoise0*t + bnoise1 and the new origin times (no lower bound on sigma).
static int body_location(mct_ctx *C, int nchain, const double *d_locations, const double *d_noise,
                         double *d_out, int32_t *d_status)
{
    if (!C->S.dynamic) return MCT_ERR_ARG;
    mct_pick_args pa;
    memset(&pa, 0, sizeof pa);
    pa.G = C->G; pa.T = C->eik.T.p; pa.t64 = C->t64; pa.nchain = nchain; pa.nwave = C->nwave;
    pa.nsrc_eff = C->nsrc_eff; pa.npick = C->npick; pa.pick_pts = d_locations;
    pa.pick_stride_chain = C->S.nsrc * 4; pa.pick_point_stride = 4;
    pa.out = P<double>(C->picks); pa.interp3d = 1;
    int rc = mct_launch_picks(pa, C->stream);
    if (rc) return rc;
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = nchain; ma.nsrc = C->S.nsrc; ma.nrev = C->S.nrev; ma.np = C->nwave;
    ma.reciprocal = C->reciprocal; ma.picks = P<double>(C->picks);
    ma.ttime = P<double>(C->ttime); ma.raystat = P<int32_t>(C->raystat);
    ma.locations = d_locations; ma.noise = d_noise; ma.sigdep = 1; ma.nrays = C->S.nrays;
    ma.eps = -DBL_MAX;
    ma.out = d_out; ma.status = d_status;
    return mct_launch_misfit(ma, C->stream);
}

extern "C" {

const char *mct_version(void) { return MCT_VERSION; }

int mct_body_create(const mct_grid *grid, const mct_body_setup *setup, mct_ctx **out)
{
    if (!grid || !setup || !out) return MCT_ERR_ARG;
    if (grid->nx < 2 || grid->ny < 2 || grid->nz < 2 || setup->nsrc < 1 || setup->nrev < 1 ||
        setup->max_cells < 1 || setup->max_chains < 1 || !setup->rev || !setup->ttime || !setup->raystat)
        return MCT_ERR_ARG;
    // datatype 3 (the body half of the joint likelihood) traces P and S as 1
    // does (src/likelihood_body.F90:204-216, the branch for datatype != 0)
    if (setup->datatype != 0 && setup->datatype != 1 && setup->datatype != 3) return MCT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MCT_ERR_HIP;
    mct_ctx *C = new (std::nothrow) mct_ctx();
    if (!C) return MCT_ERR_NOMEM;
    C->G = mct_make_geom(grid);
    C->S = *setup;
    C->S.rev = nullptr; C->S.ttime = nullptr; C->S.raystat = nullptr;
    C->nwave = setup->datatype == 0 ? 1 : 2;
    // from the receivers when they are fewer (likelihood_body.F90:226-249), and
    // always with dynamic fields (:250-279)
    C->reciprocal = setup->dynamic != 0 || setup->nsrc > setup->nrev;
    C->nsrc_eff = C->reciprocal ? setup->nrev : setup->nsrc;
    C->npick = C->reciprocal ? setup->nsrc : setup->nrev;
    C->band_tiles = mct_default_band_tiles();
    C->f64 = mct_default_f64();
    C->t64 = C->f64 ? 1 : mct_default_t64();
    C->heap_order1 = mct_default_heap_order1();
    int rc = MCT_OK;
    const long npair = (long)setup->nsrc * setup->nrev;
    const int np = C->nwave, M = setup->max_chains, S = setup->max_cells;
    const long gsz = (long)C->G.ntiles * MCT_TILE_N;
    if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess) { delete C; return MCT_ERR_HIP; }
    for (int i = 0; i < 4; i++)
        if (hipEventCreate(&C->ev[i]) != hipSuccess) { mct_destroy(C); return MCT_ERR_HIP; }
#define ALLOC(buf, bytes) if (!rc) rc = dalloc(C->buf, (size_t)(bytes))
    ALLOC(rev, sizeof(double) * 3 * setup->nrev);
    ALLOC(ttime, sizeof(double) * 3 * npair * np);
    ALLOC(raystat, sizeof(int32_t) * 2 * npair * np);
    ALLOC(points, sizeof(double) * 3 * S * M);
    ALLOC(params, sizeof(double) * 3 * S * M);
    ALLOC(locations, sizeof(double) * 4 * setup->nsrc * M);
    ALLOC(noise, sizeof(double) * 2 * np * M);
    ALLOC(out, sizeof(double) * 3 * M);
    ALLOC(status, sizeof(int32_t) * M);
    ALLOC(vel, (C->f64 ? sizeof(double) : sizeof(float)) * gsz * np * M);
    ALLOC(flag_count, sizeof(int32_t) * M);
    ALLOC(flag_nodes, sizeof(int32_t) * MCT_MAX_FLAGGED * M);
    ALLOC(kd_nodes, sizeof(mct_kd_node) * (2 * S + 2) * M);
    ALLOC(kd_ind, sizeof(int32_t) * (S + 1) * M);
    ALLOC(kd_rdata, sizeof(double) * 3 * S * M);
    ALLOC(picks, sizeof(double) * npair * np * M);
    ALLOC(phase_time, sizeof(double) * npair * np);
    ALLOC(srdist, sizeof(double) * npair * np);
    ALLOC(sigma, sizeof(double) * npair * np);
#undef ALLOC
    if (!rc) rc = C->eik.alloc(C->G, (long)M * np * C->nsrc_eff);
    if (!rc && hipHostMalloc((void **)&C->h_out, sizeof(double) * 3, hipHostMallocDefault) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && hipHostMalloc((void **)&C->h_status, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && hipHostMalloc((void **)&C->h_in, sizeof(double) * (6 * (size_t)S + 4 * (size_t)setup->nsrc + 4),
                             hipHostMallocDefault) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && hipMemcpy(C->rev.p, setup->rev, sizeof(double) * 3 * setup->nrev, hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
    if (!rc && hipMemcpy(C->ttime.p, setup->ttime, sizeof(double) * 3 * npair * np, hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
    if (!rc && hipMemcpy(C->raystat.p, setup->raystat, sizeof(int32_t) * 2 * npair * np, hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
    if (rc) { mct_destroy(C); return rc; }
    *out = C;
    return MCT_OK;
}

int mct_destroy(mct_ctx *C)
{
    if (!C) return MCT_OK;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    dbuf *bufs[] = {&C->rev, &C->ttime, &C->raystat, &C->points, &C->params, &C->locations,
                    &C->noise, &C->out, &C->status, &C->vel, &C->flag_count, &C->flag_nodes,
                    &C->kd_nodes, &C->kd_ind, &C->kd_rdata, &C->picks, &C->phase_time,
                    &C->srdist, &C->sigma};
    for (dbuf *b : bufs) dfree(*b);
    C->eik.release();
    if (C->timer) { C->timer->release(); delete C->timer; }
    if (C->h_out) (void)hipHostFree(C->h_out);
    if (C->h_status) (void)hipHostFree(C->h_status);
    if (C->h_in) (void)hipHostFree(C->h_in);
    for (int i = 0; i < 4; i++) if (C->ev[i]) (void)hipEventDestroy(C->ev[i]);
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
    return MCT_OK;
}

int mct_body_likelihood(mct_ctx *C, const double *points, const double *params, int ncells,
                        const double *locations, const double *bnoise0, const double *bnoise1,
                        double *out, double *phase_time, double *srdist, double *sigma)
{
    if (!C || !points || !params || !locations || !out) return MCT_ERR_ARG;
    if (ncells < 1 || ncells > C->S.max_cells) return MCT_ERR_ARG;
    if (C->S.sigdep && (!bnoise0 || !bnoise1)) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->nwave;
    const long npair = (long)C->S.nsrc * C->S.nrev;
    // stage through pinned memory: the caller's arrays may be reused on return
    MCT_CHECK(hipStreamSynchronize(s));
    double *hp = C->h_in, *hq = hp + 3 * (size_t)ncells, *hl = hq + 3 * (size_t)ncells,
           *hn = hl + 4 * (size_t)C->S.nsrc;
    memcpy(hp, points, sizeof(double) * 3 * ncells);
    memcpy(hq, params, sizeof(double) * 3 * ncells);
    memcpy(hl, locations, sizeof(double) * 4 * C->S.nsrc);
    for (int i = 0; i < 2 * np; i++) hn[i] = 0.0;
    if (bnoise0) for (int i = 0; i < np; i++) { hn[i] = bnoise0[i]; hn[np + i] = bnoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->points.p, hp, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->params.p, hq, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->locations.p, hl, sizeof(double) * 4 * C->S.nsrc, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    bool arrays = phase_time || srdist || sigma;
    int rc = body_pipeline(C, 1, P<double>(C->points), P<double>(C->params), nullptr, ncells,
                           P<double>(C->locations), P<double>(C->noise), P<double>(C->out),
                           P<int32_t>(C->status), arrays);
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (phase_time) MCT_CHECK(hipMemcpyAsync(phase_time, C->phase_time.p, sizeof(double) * npair * np, hipMemcpyDeviceToHost, s));
    if (srdist) MCT_CHECK(hipMemcpyAsync(srdist, C->srdist.p, sizeof(double) * npair * np, hipMemcpyDeviceToHost, s));
    if (sigma) MCT_CHECK(hipMemcpyAsync(sigma, C->sigma.p, sizeof(double) * npair * np, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    int src = mct_sync(C);
    if (src) return src;
    return *C->h_status;
}

int mct_body_likelihood_batch_dev(mct_ctx *C, int nchain, const double *points_dev,
                                  const double *params_dev, const int32_t *ncells_dev,
                                  const double *locations_dev, const double *noise_dev,
                                  double *out_dev, int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !points_dev || !params_dev || !ncells_dev ||
        !locations_dev || !out_dev)
        return MCT_ERR_ARG;
    if (status_dev) MCT_CHECK(hipMemsetAsync(status_dev, 0, sizeof(int32_t) * nchain, C->stream));
    return body_pipeline(C, nchain, points_dev, params_dev, ncells_dev, 0, locations_dev,
                         noise_dev, out_dev, status_dev, false);
}

int mct_body_noise_likelihood(mct_ctx *C, const double *locations, const double *bnoise0,
                              const double *bnoise1, double *out)
{
    if (!C || !bnoise0 || !bnoise1 || !out) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->nwave;
    MCT_CHECK(hipStreamSynchronize(s));
    double *hl = C->h_in, *hn = hl + 4 * (size_t)C->S.nsrc;
    for (int j = 0; j < 4 * C->S.nsrc; j++) hl[j] = locations ? locations[j] : 0.0;
    for (int i = 0; i < np; i++) { hn[i] = bnoise0[i]; hn[np + i] = bnoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->locations.p, hl, sizeof(double) * 4 * C->S.nsrc, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    int rc = body_noise(C, 1, P<double>(C->locations), P<double>(C->noise), P<double>(C->out), P<int32_t>(C->status));
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    return *C->h_status;
}

int mct_body_noise_likelihood_from(mct_ctx *C, const double *phase_time, const double *locations,
                                   const double *bnoise0, const double *bnoise1, double *out, double *sigma)
{
    if (!C || !phase_time || !bnoise0 || !bnoise1 || !out) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->nwave;
    const long npt = (long)np * C->S.nsrc * C->S.nrev;
    MCT_CHECK(hipStreamSynchronize(s));
    double *hl = C->h_in, *hn = hl + 4 * (size_t)C->S.nsrc;
    for (int j = 0; j < 4 * C->S.nsrc; j++) hl[j] = locations ? locations[j] : 0.0;
    for (int i = 0; i < np; i++) { hn[i] = bnoise0[i]; hn[np + i] = bnoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->locations.p, hl, sizeof(double) * 4 * C->S.nsrc, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->picks.p, phase_time, sizeof(double) * npt, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = 1; ma.nsrc = C->S.nsrc; ma.nrev = C->S.nrev; ma.np = np;
    ma.reciprocal = 0;                 // phaseTime(nrev,nsrc,np) as the caller holds it
    ma.picks = P<double>(C->picks); ma.ttime = P<double>(C->ttime); ma.raystat = P<int32_t>(C->raystat);
    ma.locations = P<double>(C->locations); ma.noise = P<double>(C->noise); ma.sigdep = 1;
    ma.nrays = C->S.nrays; ma.eps = -DBL_MAX;
    ma.out = P<double>(C->out); ma.status = P<int32_t>(C->status);
    ma.sigma = sigma ? P<double>(C->sigma) : nullptr;
    int rc = mct_launch_misfit(ma, s);
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (sigma)
        MCT_CHECK(hipMemcpyAsync(sigma, C->sigma.p, sizeof(double) * npt, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    return *C->h_status;
}

int mct_body_noise_likelihood_batch_dev(mct_ctx *C, int nchain, const double *locations_dev,
                                        const double *noise_dev, double *out_dev, int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !noise_dev || !out_dev) return MCT_ERR_ARG;
    if (status_dev) MCT_CHECK(hipMemsetAsync(status_dev, 0, sizeof(int32_t) * nchain, C->stream));
    return body_noise(C, nchain, locations_dev, noise_dev, out_dev, status_dev);
}

int mct_body_location_likelihood(mct_ctx *C, const double *locations, const double *bnoise0,
                                 const double *bnoise1, double *out)
{
    if (!C || !locations || !bnoise0 || !bnoise1 || !out) return MCT_ERR_ARG;
    if (!C->S.dynamic) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->nwave;
    MCT_CHECK(hipStreamSynchronize(s));
    double *hl = C->h_in, *hn = hl + 4 * (size_t)C->S.nsrc;
    for (int j = 0; j < 4 * C->S.nsrc; j++) hl[j] = locations[j];
    for (int i = 0; i < np; i++) { hn[i] = bnoise0[i]; hn[np + i] = bnoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->locations.p, hl, sizeof(double) * 4 * C->S.nsrc, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    int rc = body_location(C, 1, P<double>(C->locations), P<double>(C->noise), P<double>(C->out),
                           P<int32_t>(C->status));
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    return *C->h_status;
}

int mct_body_location_likelihood_batch_dev(mct_ctx *C, int nchain, const double *locations_dev,
                                           const double *noise_dev, double *out_dev, int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !locations_dev || !noise_dev || !out_dev) return MCT_ERR_ARG;
    if (status_dev) MCT_CHECK(hipMemsetAsync(status_dev, 0, sizeof(int32_t) * nchain, C->stream));
    return body_location(C, nchain, locations_dev, noise_dev, out_dev, status_dev);
}

int mct_dev_alloc(void **ptr, int64_t bytes)
{
    if (!ptr || bytes < 0) return MCT_ERR_ARG;
    *ptr = nullptr;
    if (hipMalloc(ptr, bytes > 0 ? (size_t)bytes : 16) != hipSuccess) { *ptr = nullptr; return MCT_ERR_NOMEM; }
    return MCT_OK;
}

int mct_dev_free(void *ptr)
{
    if (ptr) MCT_CHECK(hipFree(ptr));
    return MCT_OK;
}

int mct_copy_h2d(void *dst_dev, const void *src_host, int64_t bytes)
{
    if (bytes < 0 || (bytes > 0 && (!dst_dev || !src_host))) return MCT_ERR_ARG;
    if (bytes) MCT_CHECK(hipMemcpy(dst_dev, src_host, (size_t)bytes, hipMemcpyHostToDevice));
    return MCT_OK;
}

int mct_copy_d2d(void *dst_dev, const void *src_dev, int64_t bytes)
{
    if (bytes < 0 || (bytes > 0 && (!dst_dev || !src_dev))) return MCT_ERR_ARG;
    if (bytes) MCT_CHECK(hipMemcpy(dst_dev, src_dev, (size_t)bytes, hipMemcpyDeviceToDevice));
    return MCT_OK;
}

int mct_copy_d2h(void *dst_host, const void *src_dev, int64_t bytes)
{
    if (bytes < 0 || (bytes > 0 && (!dst_host || !src_dev))) return MCT_ERR_ARG;
    if (bytes) MCT_CHECK(hipMemcpy(dst_host, src_dev, (size_t)bytes, hipMemcpyDeviceToHost));
    return MCT_OK;
}

int mct_body_stream(mct_ctx *C, void **stream)
{
    if (!C || !stream) return MCT_ERR_ARG;
    *stream = (void *)C->stream;
    return MCT_OK;
}

int mct_sync(mct_ctx *C)
{
    if (!C) return MCT_ERR_ARG;
    MCT_CHECK(hipStreamSynchronize(C->stream));
    float f;
    for (int i = 0; i < 3; i++)
        if (hipEventElapsedTime(&f, C->ev[i], C->ev[i + 1]) == hipSuccess) C->ms[i] = f;
    if (C->timer && C->timer->on) C->timer->collect();
    return MCT_OK;
}

int mct_set_tuning(mct_ctx *C, float band_tiles, int t64)
{
    if (!C) return MCT_ERR_ARG;
    if (band_tiles > 0.f) C->band_tiles = band_tiles;
    if (t64 >= 0) C->t64 = t64 ? 1 : 0;
    if (!C->t64) C->f64 = 0;   // FP32 arrival times: FP32 velocities and local solves too
    return MCT_OK;
}

int mct_body_set_order1(mct_ctx *C, int method)
{
    if (!C || (method != MCT_ORDER1_HEAP && method != MCT_ORDER1_PARALLEL)) return MCT_ERR_ARG;
    C->heap_order1 = method == MCT_ORDER1_HEAP;
    return MCT_OK;
}

int mct_set_kernel_timing(mct_ctx *C, int on)
{
    if (!C) return MCT_ERR_ARG;
    if (on && !C->timer) {
        C->timer = new (std::nothrow) round_timer();
        if (!C->timer) return MCT_ERR_NOMEM;
        if (C->timer->ensure()) { C->timer->release(); delete C->timer; C->timer = nullptr; return MCT_ERR_HIP; }
    }
    if (C->timer) C->timer->on = on != 0;
    return MCT_OK;
}

int mct_last_timings(mct_ctx *C, double *ms6)
{
    if (!C || !ms6) return MCT_ERR_ARG;
    for (int i = 0; i < 4; i++) ms6[i] = C->ms[i];
    ms6[4] = (C->timer && C->timer->on) ? C->timer->kernel_ms : 0.0;
    ms6[5] = (C->timer && C->timer->on) ? (double)C->timer->used : 0.0;
    return MCT_OK;
}

} // extern "C"
