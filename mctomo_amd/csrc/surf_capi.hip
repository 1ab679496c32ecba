// surf_capi.hip -- surface-wave likelihood context (datatype 2, phase or
// group velocities) behind the C ABI declared in include/mctomo_hip.h.
//
// One call is the reference's likelihood() for datatype 2 (src/likelihood.f90
// :72-80 -> src/likelihood_surf.F90:110-405) over the whole grid or a
// perturbed box, for one chain (host arrays) or many chains (device arrays),
// all on the context's stream:
//   nearest site -> vs grid (Fortran (nz,ny,nx) per chain)
//   dispersion: check_model, convert_to_layer, surfdisp96 per column (the
//     box's columns only in a perturbed-box call)
//   k_surf_status: reject chains with check_model / dispersion failures
//   k_surf_pad: like%vel(np, ny+2, nx+2) with replicated edges
//   fm2d: B-spline gridder, the near-source kernel, the coarse sweeps (or the
//     sequential solve) per (chain, period, source), then rays and group
//     delays with group velocities; or straight rays and no 2-D solve
//   misfit: srdist = phaseTime (ray lengths / distances with rays), sigma,
//     Gaussian misfit
#include "mct_internal.h"
#include <vector>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <new>

namespace {

struct sbuf {
    void *p = nullptr;
    size_t n = 0;
    int alloc(size_t bytes)
    {
        if (bytes == 0) bytes = 16;
        if (n >= bytes) return MCT_OK;
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; return MCT_ERR_NOMEM; }
        n = bytes;
        return MCT_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

// status[c] = MCT_REJECT_MODEL when check_model failed or any column's
// surfdisp96 returned ierr (src/likelihood_surf.F90:161-164,212-224)
__global__ void k_surf_status(const int32_t *bad_model, const int32_t *ierr, int ncol, int32_t *status)
{
    const int c = blockIdx.x;
    __shared__ int bad;
    if (threadIdx.x == 0) bad = bad_model[c];
    __syncthreads();
    int mine = 0;
    for (int i = threadIdx.x; i < ncol; i += blockDim.x) mine |= ierr[(long)c * ncol + i];
    if (mine) atomicOr(&bad, 1);
    __syncthreads();
    if (threadIdx.x == 0 && bad && status[c] == 0) status[c] = MCT_REJECT_MODEL;
}

// like%vel(:, iy0+1:iy1+1, ix0+1:ix1+1) = pvel with the edge copies of
// src/likelihood_surf.F90:259-264 for the whole grid: node (ix, iy) of the
// padded grid takes column (clamp(ix,1,nx), clamp(iy,1,ny)).
__global__ void k_surf_pad(const double *phase, int nx, int ny, int np, int nchain, double *velv)
{
    const long per = (long)(nx + 2) * (ny + 2);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per * np * nchain) return;
    const long cp = t / per;
    const int node = (int)(t - cp * per);
    const int c = (int)(cp / np), p = (int)(cp % np);
    const int ix = node / (ny + 2), iy = node % (ny + 2);
    const int cx = ix < 1 ? 1 : (ix > nx ? nx : ix), cy = iy < 1 ? 1 : (iy > ny ? ny : iy);
    const long col = (long)(cx - 1) * ny + (cy - 1);
    velv[t] = phase[((long)c * nx * ny + col) * np + p];
}

} // namespace

#define MCT_SURF_SPLIT_MAX 4

struct mct_surf_ctx {
    mct_geom G;
    mct_surf_setup S;
    mct_fm2d_args F;                 // geometry part of the 2-D solve
    long fm2d_chunk;
    int fm2d_method;                 // 0 parallel sweeps, 1 sequential heap loop
    long dsp_waves;                  // queue mode: at most this many waves (0: one per 64 columns)
    long fm2d_run_chunk = 0;         // 2-D problems per launch (0: the workspace's capacity)
    // the batched call captured in HIP graphs, one per set of arguments (a
    // server lane replays one per batch size), oldest evicted beyond 64
    bool capturing = false;
    struct graph_entry {
        const void *key[6];
        int nchain;
        hipGraphExec_t exec;
    };
    std::vector<graph_entry> graphs;
    int dsp_unsorted;                // queue mode: columns in grid order instead of by layer count
    int dsp_dedup;                   // sorted queue mode: solve each distinct column profile once
    int dsp_hlog;                    // profile-hash slots per chain = 1 << dsp_hlog
    int dedup_last;                  // the last whole-grid call copied duplicate columns
    int dsp_mode;                    // 0 by size, 1 lane per column, 2 wave per column,
                                     // 3 lane per column from a column queue
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // a batch in nsplit parts on the sub streams (MCT_SURF_SPLIT, default 1)
    int nsplit = 1;
    hipStream_t sub[MCT_SURF_SPLIT_MAX] = {};
    hipEvent_t sev[MCT_SURF_SPLIT_MAX][2] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t ev_sweep = nullptr;   // between the 2-D near-source and sweep launches
    double ms[4] = {0, 0, 0, 0};
    double ms2d[2] = {0, 0};
    bool sweep_rec = false;          // ev_sweep was recorded by the last call
    // immutable data
    sbuf periods, scx, scz, rcx, rcz, ttime, raystat;
    // per-chain inputs staged by the host entry point
    sbuf points, params, noise, out, status;
    // stage buffers
    sbuf vs, flag_count, flag_nodes, kd_nodes, kd_ind, kd_rdata;
    sbuf phase, group, ierr, bad_model, ndltar, nlayer, lay, dsp_queue, dsp_jobs, dsp_keys, dsp_hist, velv, veln, phase_time, srdist, sigma;
    sbuf dsp_hkey, dsp_hrep, dsp_dup, dsp_njobs;
    sbuf ttn, ttr, vr, nsts, nstr, heap, pflag, rounds;
    // perturbed-box calls: accepted phase velocities per chain, the boxes of a call
    sbuf phase_keep, box;
    // group velocities: ray lengths (srdist) per chain and the accepted group velocities
    sbuf rlen, group_keep;
    // joint likelihood (datatype 3): the sites' vp and rho on the grid
    sbuf vpg, rhog;
    int32_t *h_box = nullptr;        // pinned: (nchain, 4) boxes then (nchain + 1) offsets
    std::vector<char> keep_valid;
    double *h_in = nullptr, *h_out = nullptr;
    int32_t *h_status = nullptr;
};

// The pipeline for chains [c0, c0 + nchain) of the batch on stream s: the
// chain-indexed buffers are offset by c0, and the part's own 2-D problem slots
// (nparts parts share the workspace), queue counter and histogram are used.
// ev_dsp (may be null) is recorded once the dispersion stage is queued.  The
// stage events ev[1..3] follow part 0; ev[0] and ev[4] bracket a one-part run
// here and a split run in the caller.
static int surf_pipeline(mct_surf_ctx *C, int c0, int nchain, int part, int nparts, hipStream_t s,
                         const double *d_points, const double *d_params,
                         const int32_t *d_ncells, int ncells_const, const double *d_noise,
                         double *d_out, int32_t *d_status, bool want_arrays, hipEvent_t ev_dsp,
                         const int32_t *d_box = nullptr, const int32_t *d_box_off = nullptr, long box_total = 0)
{
    const mct_geom &G = C->G;
    const mct_surf_setup &S = C->S;
    const int ncol = G.nx * G.ny;
    const long gsz = (long)ncol * G.nz;
    const int NC = S.max_cells, np = S.np;
    const long npair = (long)S.nsrc * S.nrev;
    const long wpc = (ncol + 63) / 64;          // waves of lane-per-column scratch per chain
    const bool timing = part == 0;
    int rc;
    const bool events = !C->capturing;   // no timing events inside a captured graph
    if (nparts == 1 && events) MCT_CHECK(hipEventRecord(C->ev[0], s));
    // (a) nearest site -> vs
    double *vs = C->vs.as<double>() + c0 * gsz;
    mct_nearest_args na;
    memset(&na, 0, sizeof na);
    na.G = G; na.nchain = nchain; na.max_cells = NC;
    na.points = d_points; na.params = d_params; na.ncells = d_ncells; na.ncells_const = ncells_const;
    na.ix0 = 1; na.ix1 = G.nx; na.iy0 = 1; na.iy1 = G.ny; na.iz0 = 1; na.iz1 = G.nz;
    na.vs = vs; na.nwave = 1;
    const bool joint = S.datatype == 3;
    if (joint) {   // the sites' own vp and rho (src/likelihood.f90:83-88)
        na.vp = C->vpg.as<double>() + c0 * gsz;
        na.rho = C->rhog.as<double>() + c0 * gsz;
    }
    na.flag_count = C->flag_count.as<int32_t>() + c0;
    na.flag_nodes = C->flag_nodes.as<int32_t>() + (long)c0 * MCT_MAX_FLAGGED;
    na.kd_nodes = C->kd_nodes.as<mct_kd_node>() + (long)c0 * (2 * NC + 2);
    na.kd_ind = C->kd_ind.as<int32_t>() + (long)c0 * (NC + 1);
    na.kd_rdata = C->kd_rdata.as<double>() + (long)c0 * 3 * NC;
    if ((rc = mct_launch_nearest(na, s))) return rc;
    if (timing && events) MCT_CHECK(hipEventRecord(C->ev[1], s));
    // (b) dispersion per column
    mct_dsp_args da;
    memset(&da, 0, sizeof da);
    da.ncol = ncol; da.nz = G.nz; da.nchain = nchain; da.vs = vs;
    da.dz = G.dz; da.scaling = G.scaling; da.water = S.water_depth; da.periods = C->periods.as<double>(); da.np = np;
    da.iwave = S.raylov == 1 ? 2 : 1; da.igr = S.phase_group; da.dphase = S.dphase;
    da.phase = C->phase.as<double>() + (long)c0 * ncol * np;
    da.group = C->group.as<double>() + (long)c0 * ncol * np;
    da.ierr = C->ierr.as<int32_t>() + (long)c0 * ncol;
    da.bad_model = C->bad_model.as<int32_t>() + c0;
    da.ndltar = C->ndltar.as<int32_t>() + (long)c0 * ncol;
    da.nlayer = C->nlayer.as<int32_t>() + (long)c0 * ncol;
    if (joint) { da.vp = na.vp; da.rho = na.rho; }
    // many columns: one lane per column from a column queue (throughput); few:
    // one wave per column (latency)
    int dmode = C->dsp_mode;
    if (dmode == 0) dmode = (long)ncol * nchain >= 16384 ? 3 : 2;
    if (d_box) {
        // a perturbed-box call (src/likelihood_surf.F90:155-170): the box's
        // columns are recomputed, the others keep the accepted phase velocities
        dmode = 3;
        MCT_CHECK(hipMemcpyAsync(da.phase, C->phase_keep.as<double>() + (long)c0 * ncol * np,
                                 sizeof(double) * ncol * np * nchain, hipMemcpyDeviceToDevice, s));
        if (S.phase_group)   // like%gvel keeps the accepted group velocities too (:227-229)
            MCT_CHECK(hipMemcpyAsync(da.group, C->group_keep.as<double>() + (long)c0 * ncol * np,
                                     sizeof(double) * ncol * np * nchain, hipMemcpyDeviceToDevice, s));
        if ((rc = mct_fill_i32(da.ierr, (long)ncol * nchain, 0, s))) return rc;
        da.box = d_box; da.box_off = d_box_off; da.box_total = box_total; da.ny = G.ny;
    }
    da.lay = (dmode == 1 || dmode == 3) ? C->lay.as<float>() + (long)c0 * wpc * 4 * (G.nz + 1) * 64 : nullptr;
    if (da.lay && dmode == 3) {
        da.queue = C->dsp_queue.as<unsigned long long>() + part;
        da.lay_waves = (long)nchain * wpc;
        if (C->dsp_waves > 0 && da.lay_waves > C->dsp_waves) da.lay_waves = C->dsp_waves;
        if (!C->dsp_unsorted || d_box) {
            da.jobs = C->dsp_jobs.as<int32_t>() + (long)c0 * ncol;
            da.keys = C->dsp_keys.as<int32_t>() + (long)c0 * ncol;
            da.hist = C->dsp_hist.as<int32_t>() + 128 * part;
            // identical-column detection hashes the Vs profile, which decides the
            // column's layers only when vp and rho follow from Vs
            if (C->dsp_dedup && !joint) {
                da.hlog = C->dsp_hlog;
                da.hkey = C->dsp_hkey.as<unsigned long long>() + ((long)c0 << C->dsp_hlog);
                da.hrep = C->dsp_hrep.as<int32_t>() + ((long)c0 << C->dsp_hlog);
                da.dup = C->dsp_dup.as<int32_t>() + (long)c0 * ncol;
                da.njobs = C->dsp_njobs.as<int32_t>() + part;
            }
        }
    }
    if ((rc = mct_launch_dispersion(da, s))) return rc;
    if (timing) C->dedup_last = !d_box && da.dup != nullptr;
    hipLaunchKernelGGL(k_surf_status, dim3(nchain), dim3(256), 0, s, (const int32_t *)da.bad_model,
                       (const int32_t *)da.ierr, ncol, d_status);
    MCT_CHECK(hipGetLastError());
    if (timing && events) MCT_CHECK(hipEventRecord(C->ev[2], s));
    if (ev_dsp) MCT_CHECK(hipEventRecord(ev_dsp, s));
    double *phase_time = C->phase_time.as<double>() + (long)c0 * npair * np;
    double *rlen = (S.phase_group || S.straight) ? C->rlen.as<double>() + (long)c0 * npair * np : nullptr;
    if (S.straight) {
        // (c') straight rays (likelihood_surf.F90:233-243): the phase (or group)
        // velocities integrated along the source-receiver segments; no 2-D solve
        mct_fm2d_args fa = C->F;
        fa.nchain = nchain;
        fa.scx = C->scx.as<double>(); fa.scz = C->scz.as<double>();
        fa.rcx = C->rcx.as<double>(); fa.rcz = C->rcz.as<double>();
        fa.srs = C->raystat.as<int32_t>(); fa.srs_pstride = 2L * S.nsrc * S.nrev;
        fa.ttime = phase_time; fa.status = d_status;
        fa.gvel = S.phase_group ? da.group : da.phase;
        if (timing) C->sweep_rec = false;
        if ((rc = mct_launch_straight_rays(fa, s))) return rc;
    } else {
    // (c) padded phase-velocity grids, then modrays per (chain, period, source)
    double *velv = C->velv.as<double>() + (long)c0 * (G.nx + 2) * (G.ny + 2) * np;
    {
        const long n = (long)(G.nx + 2) * (G.ny + 2) * np * nchain;
        hipLaunchKernelGGL(k_surf_pad, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           (const double *)da.phase, G.nx, G.ny, np, nchain, velv);
        MCT_CHECK(hipGetLastError());
    }
    mct_fm2d_args fa = C->F;
    const long wsz = (long)fa.ld * fa.nc;
    const long run = C->fm2d_run_chunk > 0 && C->fm2d_run_chunk < C->fm2d_chunk ? C->fm2d_run_chunk : C->fm2d_chunk;
    const long chunk = run / nparts, soff = chunk * part;
    fa.nchain = nchain;
    fa.velv = velv;
    fa.veln = C->veln.as<double>() + (long)c0 * fa.nnx * fa.nnz * np;
    fa.sln = fa.veln + (long)C->F.nnx * C->F.nnz * np * S.max_chains;
    fa.scx = C->scx.as<double>(); fa.scz = C->scz.as<double>();
    fa.rcx = C->rcx.as<double>(); fa.rcz = C->rcz.as<double>();
    fa.srs = C->raystat.as<int32_t>(); fa.srs_pstride = 2L * S.nsrc * S.nrev;
    fa.ttn = C->ttn.as<double>() + soff * wsz; fa.ttr = C->ttr.as<double>() + soff * wsz;
    fa.vr = C->vr.as<double>() + soff * wsz;
    fa.nsts = C->nsts.as<int32_t>() + soff * wsz; fa.nstr = C->nstr.as<int32_t>() + soff * wsz;
    fa.heap = C->heap.as<int32_t>() + soff * (wsz + 8);
    fa.ttime = phase_time; fa.status = d_status;
    fa.rounds_out = C->rounds.as<int32_t>() + 4 * soff;
    fa.ev_sweep = timing && events ? C->ev_sweep : nullptr;
    // group velocities: rays, group delays (CalGroupTime) and ray lengths
    // (likelihood_surf.F90:245,311-351)
    fa.rays = S.phase_group; fa.gvel = da.group; fa.rlen = rlen;
    const bool fast = C->fm2d_method == 0 && mct_fm2d_fast_ok(fa);
    if (timing && events) C->sweep_rec = fast;
    if (fast)
        rc = mct_launch_fm2d_fast(fa, chunk, C->pflag.as<int32_t>() + soff, s);
    else
        rc = mct_launch_fm2d(fa, chunk, s);
    if (rc) return rc;
    }
    if (timing && events) MCT_CHECK(hipEventRecord(C->ev[3], s));
    // (d) misfit: srdist = phaseTime, or the ray length for group times
    // (src/likelihood_surf.F90:324-331)
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = nchain; ma.nsrc = S.nsrc; ma.nrev = S.nrev; ma.np = np; ma.reciprocal = 0;
    ma.picks = phase_time; ma.ttime = C->ttime.as<double>();
    ma.raystat = C->raystat.as<int32_t>(); ma.locations = nullptr; ma.noise = d_noise;
    ma.sigdep = S.sigdep; ma.nrays = S.nrays;
    ma.eps = (double)1.0e-10f;   // EPS, src/likelihood_surf.F90:37
    ma.out = d_out; ma.status = d_status;
    ma.srdist = want_arrays ? C->srdist.as<double>() : nullptr;   // chain 0 only
    ma.sigma = want_arrays ? C->sigma.as<double>() : nullptr;
    ma.dist = rlen;
    if ((rc = mct_launch_misfit(ma, s))) return rc;
    if (nparts == 1 && events) MCT_CHECK(hipEventRecord(C->ev[4], s));
    return MCT_OK;
}

// surf_noise_likelihood (src/likelihood_surf.F90:56-99): the misfit of the
// phase times kept from the last likelihood call with sigma = snoise0*t +
// snoise1 (srdist is the phase time, :329-330, or the ray length for group
// times, :326-327); sigma below EPS on a used pair
// is the reference's stop.
static int surf_noise(mct_surf_ctx *C, int nchain, const double *d_noise, double *d_out, int32_t *d_status)
{
    const mct_surf_setup &S = C->S;
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = nchain; ma.nsrc = S.nsrc; ma.nrev = S.nrev; ma.np = S.np; ma.reciprocal = 0;
    ma.picks = C->phase_time.as<double>(); ma.ttime = C->ttime.as<double>();
    ma.raystat = C->raystat.as<int32_t>(); ma.locations = nullptr; ma.noise = d_noise;
    ma.sigdep = 1; ma.nrays = S.nrays;
    ma.eps = (double)1.0e-10f;   // EPS, src/likelihood_surf.F90:37
    ma.out = d_out; ma.status = d_status;
    ma.dist = (S.phase_group || S.straight) ? C->rlen.as<double>() : nullptr;
    return mct_launch_misfit(ma, C->stream);
}

// parts a batch of nchain chains runs in
static int surf_parts(const mct_surf_ctx *C, int nchain)
{
    int P = C->nsplit < nchain ? C->nsplit : nchain;
    if (P < 1 || C->fm2d_chunk / P < 1) P = 1;
    return P;
}

extern "C" {

int mct_surf_create(const mct_grid *grid, const mct_surf_setup *setup, mct_surf_ctx **out)
{
    if (!grid || !setup || !out) return MCT_ERR_ARG;
    *out = nullptr;
    const mct_surf_setup &S = *setup;
    if (grid->nx < 2 || grid->ny < 2 || grid->nz < 2 || S.np < 1 || S.np > 60 || !S.freqs ||
        S.nsrc < 1 || S.nrev < 1 || !S.src || !S.rev || !S.ttime || !S.raystat || S.max_cells < 1 ||
        S.max_chains < 1 || S.gridx < 1 || S.gridy < 1 || S.sgdic < 1 || S.sgext < 1 || grid->nz > 100)
        return MCT_ERR_ARG;
    if (S.raylov != 0 && S.raylov != 1) return MCT_ERR_ARG;
    if (!(S.water_depth >= 0.0)) return MCT_ERR_ARG;
    // group times trace rays from the refined source grid (rpaths reads it)
    if (S.phase_group != 0 && S.phase_group != 1) return MCT_ERR_ARG;
    if (S.straight != 0 && S.straight != 1) return MCT_ERR_ARG;
    if (S.datatype != 0 && S.datatype != 2 && S.datatype != 3) return MCT_ERR_ARG;
    if (S.phase_group == 1 && S.sgref != 1 && !S.straight) return MCT_ERR_ARG;
    if (S.phase_group == 1 && !S.straight) {
        // rpaths stores pair n's ray at slot raystat(n,2) and CalGroupTime reads
        // slot n for pair n (fm2dray_cartesian.f90:1403, likelihood_surf.F90:477):
        // the data readers set raystat(n,2) = n (likelihood_settings.f90:401)
        const long npair0 = (long)S.nsrc * S.nrev;
        for (int i = 0; i < S.np; i++)
            for (long n = 0; n < npair0; n++)
                if (S.raystat[(long)i * 2 * npair0 + n] != 0 && S.raystat[((long)i * 2 + 1) * npair0 + n] != n + 1)
                    return MCT_ERR_ARG;
    }
    const long nnx = (long)(grid->nx - 1) * S.gridx + 1, nnz = (long)(grid->ny - 1) * S.gridy + 1;
    const long rg = 2L * S.sgext * S.sgdic + 1;
    if (nnx > 32767 || nnz > 32767 || rg > 32767) return MCT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MCT_ERR_HIP;
    mct_surf_ctx *C = new (std::nothrow) mct_surf_ctx();
    if (!C) return MCT_ERR_NOMEM;
    C->G = mct_make_geom(grid);
    C->S = S;
    C->S.freqs = nullptr; C->S.src = nullptr; C->S.rev = nullptr; C->S.ttime = nullptr; C->S.raystat = nullptr;
    mct_fm2d_args &F = C->F;
    memset(&F, 0, sizeof F);
    F.nvx = grid->nx; F.nvz = grid->ny; F.gox = grid->xmin; F.goz = grid->ymin;
    F.dvx = grid->dx; F.dvz = grid->dy;
    F.gdx = S.gridx; F.gdz = S.gridy; F.asgr = S.sgref; F.sgdl = S.sgdic; F.sgs = S.sgext; F.fom = S.order;
    F.nnx = (int)nnx; F.nnz = (int)nnz;
    F.ld = (int)(nnz > rg ? nnz : rg); F.nc = (int)(nnx > rg ? nnx : rg);
    F.np = S.np; F.nsrc = S.nsrc; F.nrc = S.nrev;
    const int M = S.max_chains, NC = S.max_cells, np = S.np;
    const long npair = (long)S.nsrc * S.nrev, gsz = (long)grid->nx * grid->ny * grid->nz;
    const long ncol = (long)grid->nx * grid->ny;
    const long nprob = (long)M * np * S.nsrc;
    // fm2d workspace: problem slots for at most MCT_FM2D_WS_GB GiB (default 64:
    // a 64-chain batch at 201x201 with 20 periods and 32 sources in one launch),
    // whole waves of 64; more problems run in chunks
    const long slot = mct_fm2d_slot_bytes(F.ld, F.nc);
    long ws_gb = 64;
    {
        const char *e = getenv("MCT_FM2D_WS_GB");
        if (e && atol(e) > 0) ws_gb = atol(e);
    }
    long chunk = (ws_gb << 30) / slot;
    {
        const char *e = getenv("MCT_FM2D_CHUNK");     // problems per 2-D launch (tuning)
        long v = e ? atol(e) : 0;
        if (v > 0 && v < chunk) chunk = v;
    }
    chunk = chunk < 64 ? 64 : (chunk / 64) * 64;
    C->fm2d_chunk = chunk < nprob ? chunk : nprob;
    {
        const char *e = getenv("MCT_FM2D_METHOD");
        C->fm2d_method = e && atoi(e) == 1 ? 1 : 0;
        const char *d = getenv("MCT_DSP_MODE");
        C->dsp_mode = d ? atoi(d) : 0;
        if (C->dsp_mode < 0 || C->dsp_mode > 3) C->dsp_mode = 0;
        const char *w = getenv("MCT_DSP_WAVES");
        C->dsp_waves = w ? atol(w) : 0;
        const char *u = getenv("MCT_DSP_UNSORTED");     // columns in grid order (tuning)
        C->dsp_unsorted = u && atoi(u) != 0;
        const char *dd = getenv("MCT_DSP_DEDUP");       // 0: solve every column (tuning)
        C->dsp_dedup = !(dd && atoi(dd) == 0);
    }
    int rc = MCT_OK;
    if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess) { delete C; return MCT_ERR_HIP; }
    {
        const char *e = getenv("MCT_SURF_SPLIT");
        int v = e ? atoi(e) : 1;
        C->nsplit = v < 1 ? 1 : (v > MCT_SURF_SPLIT_MAX ? MCT_SURF_SPLIT_MAX : v);
        if (C->nsplit > 1) {
            bool ok = hipEventCreateWithFlags(&C->fork, hipEventDisableTiming) == hipSuccess;
            for (int p = 0; p < C->nsplit && ok; p++) {
                ok = hipStreamCreateWithFlags(&C->sub[p], hipStreamNonBlocking) == hipSuccess &&
                     hipEventCreateWithFlags(&C->sev[p][0], hipEventDisableTiming) == hipSuccess &&
                     hipEventCreateWithFlags(&C->sev[p][1], hipEventDisableTiming) == hipSuccess;
            }
            if (!ok) { mct_surf_destroy(C); return MCT_ERR_HIP; }
        }
    }
    for (int i = 0; i < 5; i++)
        if (hipEventCreate(&C->ev[i]) != hipSuccess) { mct_surf_destroy(C); return MCT_ERR_HIP; }
    if (hipEventCreate(&C->ev_sweep) != hipSuccess) { mct_surf_destroy(C); return MCT_ERR_HIP; }
    const long wsz = (long)F.ld * F.nc, ch = C->fm2d_chunk;
#define ALLOC(buf, bytes) if (!rc) rc = C->buf.alloc((size_t)(bytes))
    ALLOC(periods, sizeof(double) * np);
    ALLOC(scx, sizeof(double) * S.nsrc); ALLOC(scz, sizeof(double) * S.nsrc);
    ALLOC(rcx, sizeof(double) * S.nrev); ALLOC(rcz, sizeof(double) * S.nrev);
    ALLOC(ttime, sizeof(double) * 3 * npair * np);
    ALLOC(raystat, sizeof(int32_t) * 2 * npair * np);
    ALLOC(points, sizeof(double) * 3 * NC * M);
    ALLOC(params, sizeof(double) * 3 * NC * M);
    ALLOC(noise, sizeof(double) * 2 * np * M);
    ALLOC(out, sizeof(double) * 3 * M);
    ALLOC(status, sizeof(int32_t) * M);
    ALLOC(vs, sizeof(double) * gsz * M);
    if (S.datatype == 3) {
        ALLOC(vpg, sizeof(double) * gsz * M);
        ALLOC(rhog, sizeof(double) * gsz * M);
    }
    ALLOC(flag_count, sizeof(int32_t) * M);
    ALLOC(flag_nodes, sizeof(int32_t) * MCT_MAX_FLAGGED * M);
    ALLOC(kd_nodes, sizeof(mct_kd_node) * (2 * NC + 2) * M);
    ALLOC(kd_ind, sizeof(int32_t) * (NC + 1) * M);
    ALLOC(kd_rdata, sizeof(double) * 3 * NC * M);
    ALLOC(phase, sizeof(double) * ncol * np * M);
    ALLOC(group, sizeof(double) * ncol * np * M);
    ALLOC(ierr, sizeof(int32_t) * ncol * M);
    ALLOC(bad_model, sizeof(int32_t) * M);
    ALLOC(ndltar, sizeof(int32_t) * ncol * M);
    ALLOC(nlayer, sizeof(int32_t) * ncol * M);
    ALLOC(lay, sizeof(float) * mct_dispersion_lay_floats((int)ncol, grid->nz, M));
    ALLOC(phase_keep, sizeof(double) * ncol * np * M);
    if (S.phase_group || S.straight) ALLOC(rlen, sizeof(double) * npair * np * M);
    if (S.phase_group) ALLOC(group_keep, sizeof(double) * ncol * np * M);
    ALLOC(box, sizeof(int32_t) * (5 * (size_t)M + 1));
    ALLOC(dsp_queue, 256);
    ALLOC(dsp_jobs, sizeof(int32_t) * ncol * M);
    ALLOC(dsp_keys, sizeof(int32_t) * ncol * M);
    ALLOC(dsp_hist, sizeof(int32_t) * 128 * MCT_SURF_SPLIT_MAX);
    C->dsp_hlog = 1;
    while ((1L << C->dsp_hlog) < 2L * ncol) C->dsp_hlog++;
    ALLOC(dsp_hkey, sizeof(unsigned long long) * ((size_t)M << C->dsp_hlog));
    ALLOC(dsp_hrep, sizeof(int32_t) * ((size_t)M << C->dsp_hlog));
    ALLOC(dsp_dup, sizeof(int32_t) * ncol * M);
    ALLOC(dsp_njobs, sizeof(int32_t) * MCT_SURF_SPLIT_MAX);
    ALLOC(velv, sizeof(double) * (grid->nx + 2) * (grid->ny + 2) * np * M);
    ALLOC(veln, 2 * sizeof(double) * nnx * nnz * np * M);   // velocities, then slowness
    ALLOC(phase_time, sizeof(double) * npair * np * M);
    ALLOC(srdist, sizeof(double) * npair * np);
    ALLOC(sigma, sizeof(double) * npair * np);
    ALLOC(ttn, sizeof(double) * wsz * ch);
    ALLOC(ttr, sizeof(double) * wsz * ch);
    ALLOC(vr, sizeof(double) * wsz * ch);
    ALLOC(nsts, sizeof(int32_t) * wsz * ch);
    ALLOC(nstr, sizeof(int32_t) * wsz * ch);
    ALLOC(heap, sizeof(int32_t) * (wsz + 8) * ch);
    ALLOC(pflag, sizeof(int32_t) * ch);
    ALLOC(rounds, sizeof(int32_t) * 4 * ch);
    if (!rc && hipMemset(C->rounds.p, 0, sizeof(int32_t) * 4 * ch) != hipSuccess) rc = MCT_ERR_HIP;
#undef ALLOC
    if (!rc && hipHostMalloc((void **)&C->h_in, sizeof(double) * (6 * (size_t)NC + 2 * np + 8), hipHostMallocDefault) != hipSuccess)
        rc = MCT_ERR_NOMEM;
    if (!rc && hipHostMalloc((void **)&C->h_out, sizeof(double) * 3, hipHostMallocDefault) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && hipHostMalloc((void **)&C->h_status, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && hipHostMalloc((void **)&C->h_box, sizeof(int32_t) * (5 * (size_t)M + 1), hipHostMallocDefault) != hipSuccess)
        rc = MCT_ERR_NOMEM;
    if (!rc) C->keep_valid.assign((size_t)M, 0);
    if (!rc) {
        // periods = dble(1/freqs) (surfmodes/surfmodes.f90:81-83); source/receiver (x, y)
        double *per = (double *)malloc(sizeof(double) * np);
        double *xy = (double *)malloc(sizeof(double) * 2 * (S.nsrc > S.nrev ? S.nsrc : S.nrev));
        if (!per || !xy) rc = MCT_ERR_NOMEM;
        if (!rc) {
            for (int i = 0; i < np; i++) per[i] = 1 / S.freqs[i];
            if (hipMemcpy(C->periods.p, per, sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
            for (int j = 0; j < S.nsrc; j++) { xy[j] = S.src[2 * j]; xy[S.nsrc + j] = S.src[2 * j + 1]; }
            if (!rc && (hipMemcpy(C->scx.p, xy, sizeof(double) * S.nsrc, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(C->scz.p, xy + S.nsrc, sizeof(double) * S.nsrc, hipMemcpyHostToDevice) != hipSuccess))
                rc = MCT_ERR_HIP;
            for (int k = 0; k < S.nrev; k++) { xy[k] = S.rev[2 * k]; xy[S.nrev + k] = S.rev[2 * k + 1]; }
            if (!rc && (hipMemcpy(C->rcx.p, xy, sizeof(double) * S.nrev, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(C->rcz.p, xy + S.nrev, sizeof(double) * S.nrev, hipMemcpyHostToDevice) != hipSuccess))
                rc = MCT_ERR_HIP;
        }
        free(per); free(xy);
    }
    if (!rc && S.straight) {
        // setup_straightRays' distances (likelihood_surf.F90:426-432) are srdist for
        // every pair, period and chain
        std::vector<double> d((size_t)npair * np * M);
        for (int j = 0; j < S.nsrc; j++)
            for (int k = 0; k < S.nrev; k++) {
                double dx = S.rev[2 * k] - S.src[2 * j], dy = S.rev[2 * k + 1] - S.src[2 * j + 1];
                double ds = sqrt(dx * dx + dy * dy);
                for (long m = 0; m < (long)np * M; m++) d[m * npair + (long)j * S.nrev + k] = ds;
            }
        if (hipMemcpy(C->rlen.p, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
    }
    // like%phaseTime starts at zero; pairs that are never traced keep it
    if (!rc && hipMemset(C->phase_time.p, 0, sizeof(double) * npair * np * M) != hipSuccess) rc = MCT_ERR_HIP;
    if (!rc && hipMemcpy(C->ttime.p, S.ttime, sizeof(double) * 3 * npair * np, hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
    if (!rc && hipMemcpy(C->raystat.p, S.raystat, sizeof(int32_t) * 2 * npair * np, hipMemcpyHostToDevice) != hipSuccess) rc = MCT_ERR_HIP;
    if (rc) { mct_surf_destroy(C); return rc; }
    *out = C;
    return MCT_OK;
}

}   // extern "C"

// every device buffer of a context
static std::vector<sbuf *> surf_bufs(mct_surf_ctx *C)
{
    return {&C->periods, &C->scx, &C->scz, &C->rcx, &C->rcz, &C->ttime, &C->raystat,
            &C->points, &C->params, &C->noise, &C->out, &C->status, &C->vs, &C->flag_count,
            &C->flag_nodes, &C->kd_nodes, &C->kd_ind, &C->kd_rdata, &C->phase, &C->group,
            &C->ierr, &C->bad_model, &C->ndltar, &C->nlayer, &C->lay, &C->dsp_queue, &C->dsp_jobs,
            &C->dsp_keys, &C->dsp_hist, &C->velv, &C->veln, &C->phase_time, &C->srdist, &C->sigma,
            &C->ttn, &C->ttr, &C->vr, &C->nsts, &C->nstr, &C->heap, &C->pflag, &C->rounds,
            &C->phase_keep, &C->box, &C->rlen, &C->group_keep, &C->dsp_hkey, &C->dsp_hrep,
            &C->dsp_dup, &C->dsp_njobs, &C->vpg, &C->rhog};
}

extern "C" {

int mct_surf_device_bytes(mct_surf_ctx *C, int64_t *bytes)
{
    if (!C || !bytes) return MCT_ERR_ARG;
    int64_t b = 0;
    for (sbuf *s : surf_bufs(C)) b += (int64_t)s->n;
    *bytes = b;
    return MCT_OK;
}

int mct_surf_destroy(mct_surf_ctx *C)
{
    if (!C) return MCT_OK;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    for (sbuf *b : surf_bufs(C)) b->release();
    if (C->h_in) (void)hipHostFree(C->h_in);
    if (C->h_out) (void)hipHostFree(C->h_out);
    if (C->h_status) (void)hipHostFree(C->h_status);
    if (C->h_box) (void)hipHostFree(C->h_box);
    for (int i = 0; i < 5; i++) if (C->ev[i]) (void)hipEventDestroy(C->ev[i]);
    if (C->ev_sweep) (void)hipEventDestroy(C->ev_sweep);
    for (auto &g : C->graphs) (void)hipGraphExecDestroy(g.exec);
    C->graphs.clear();
    for (int p = 0; p < MCT_SURF_SPLIT_MAX; p++) {
        if (C->sub[p]) { (void)hipStreamSynchronize(C->sub[p]); (void)hipStreamDestroy(C->sub[p]); }
        for (int k = 0; k < 2; k++) if (C->sev[p][k]) (void)hipEventDestroy(C->sev[p][k]);
    }
    if (C->fork) (void)hipEventDestroy(C->fork);
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
    return MCT_OK;
}

int mct_surf_sync(mct_surf_ctx *C)
{
    if (!C) return MCT_ERR_ARG;
    MCT_CHECK(hipStreamSynchronize(C->stream));
    float f;
    for (int i = 0; i < 4; i++)
        if (hipEventElapsedTime(&f, C->ev[i], C->ev[i + 1]) == hipSuccess) C->ms[i] = f;
    C->ms2d[0] = C->ms2d[1] = 0.0;
    if (C->sweep_rec && hipEventElapsedTime(&f, C->ev[2], C->ev_sweep) == hipSuccess) {
        C->ms2d[0] = f;
        if (hipEventElapsedTime(&f, C->ev_sweep, C->ev[3]) == hipSuccess) C->ms2d[1] = f;
    }
    return MCT_OK;
}

int mct_surf_last_timings2d(mct_surf_ctx *C, double *near_ms, double *sweep_ms)
{
    if (!C || !near_ms || !sweep_ms) return MCT_ERR_ARG;
    *near_ms = C->ms2d[0];
    *sweep_ms = C->ms2d[1];
    return MCT_OK;
}

int mct_surf_last_timings(mct_surf_ctx *C, double *ms4)
{
    if (!C || !ms4) return MCT_ERR_ARG;
    for (int i = 0; i < 4; i++) ms4[i] = C->ms[i];
    return MCT_OK;
}

// One chain, host arrays; box (6: x0, y0, z0, x1, y1, z1) or NULL for the whole grid.
static int surf_plain(mct_surf_ctx *C, const double *points, const double *params, int ncells, const double *box,
                      const double *snoise0, const double *snoise1, double *out,
                      double *phase_time, double *srdist, double *sigma, double *phase_vel)
{
    if (!C || !points || !params || !out) return MCT_ERR_ARG;
    if (ncells < 1 || ncells > C->S.max_cells) return MCT_ERR_ARG;
    if (C->S.sigdep && (!snoise0 || !snoise1)) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->S.np;
    const long npair = (long)C->S.nsrc * C->S.nrev, ncol = (long)C->G.nx * C->G.ny;
    int32_t cols[4];
    if (box) {
        if (!C->keep_valid[0]) return MCT_ERR_ARG;      // no accepted state to start from
        int rcb = mct_surf_box_columns(C, box, cols);
        if (rcb) return rcb;
    }
    MCT_CHECK(hipStreamSynchronize(s));
    double *hp = C->h_in, *hq = hp + 3 * (size_t)ncells, *hn = hq + 3 * (size_t)ncells;
    memcpy(hp, points, sizeof(double) * 3 * ncells);
    memcpy(hq, params, sizeof(double) * 3 * ncells);
    for (int i = 0; i < 2 * np; i++) hn[i] = 0.0;
    if (snoise0) for (int i = 0; i < np; i++) { hn[i] = snoise0[i]; hn[np + i] = snoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->points.p, hp, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->params.p, hq, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    const int32_t *d_box = nullptr, *d_off = nullptr;
    long total = 0;
    if (box) {
        int32_t *hb = C->h_box;
        for (int k = 0; k < 4; k++) hb[k] = cols[k];
        total = (long)(cols[1] - cols[0] + 1) * (cols[3] - cols[2] + 1);
        hb[4] = 0;
        hb[5] = (int32_t)total;
        MCT_CHECK(hipMemcpyAsync(C->box.p, hb, sizeof(int32_t) * 6, hipMemcpyHostToDevice, s));
        d_box = C->box.as<int32_t>();
        d_off = d_box + 4;
    }
    int rc = surf_pipeline(C, 0, 1, 0, 1, s, C->points.as<double>(), C->params.as<double>(), nullptr, ncells,
                           C->noise.as<double>(), C->out.as<double>(), C->status.as<int32_t>(),
                           srdist || sigma, nullptr, d_box, d_off, total);
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (phase_time) MCT_CHECK(hipMemcpyAsync(phase_time, C->phase_time.p, sizeof(double) * npair * np, hipMemcpyDeviceToHost, s));
    if (phase_vel) MCT_CHECK(hipMemcpyAsync(phase_vel, C->phase.p, sizeof(double) * ncol * np, hipMemcpyDeviceToHost, s));
    if (srdist) MCT_CHECK(hipMemcpyAsync(srdist, C->srdist.p, sizeof(double) * npair * np, hipMemcpyDeviceToHost, s));
    if (sigma) MCT_CHECK(hipMemcpyAsync(sigma, C->sigma.p, sizeof(double) * npair * np, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    const int st = *C->h_status;
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    int src = mct_surf_sync(C);
    if (src) return src;
    return st;
}

int mct_surf_likelihood(mct_surf_ctx *C, const double *points, const double *params, int ncells,
                        const double *snoise0, const double *snoise1, double *out,
                        double *phase_time, double *srdist, double *sigma, double *phase_vel)
{
    return surf_plain(C, points, params, ncells, nullptr, snoise0, snoise1, out, phase_time, srdist, sigma,
                      phase_vel);
}

int mct_surf_box_columns(mct_surf_ctx *C, const double *box, int32_t *cols)
{
    if (!C || !box || !cols) return MCT_ERR_ARG;
    const mct_geom &G = C->G;
    const int expand = 1;   // src/likelihood_surf.F90:39
    // src/likelihood_surf.F90:155-158, then the clamp of :166-169
    int ix0 = (int)floor((box[0] - G.xmin) / G.dx) + 1 - expand;
    int ix1 = (int)floor((box[3] - G.xmin) / G.dx) + 1 + expand;
    int iy0 = (int)floor((box[1] - G.ymin) / G.dy) + 1 - expand;
    int iy1 = (int)floor((box[4] - G.ymin) / G.dy) + 1 + expand;
    if (ix0 < 1) ix0 = 1;
    if (iy0 < 1) iy0 = 1;
    if (ix1 > G.nx) ix1 = G.nx;
    if (iy1 > G.ny) iy1 = G.ny;
    if (ix0 > ix1 || iy0 > iy1) return MCT_ERR_ARG;
    cols[0] = ix0; cols[1] = ix1; cols[2] = iy0; cols[3] = iy1;
    return MCT_OK;
}

int mct_surf_likelihood_box(mct_surf_ctx *C, const double *points, const double *params, int ncells,
                            const double *box, const double *snoise0, const double *snoise1, double *out,
                            double *phase_time, double *srdist, double *sigma, double *phase_vel)
{
    if (!box) return MCT_ERR_ARG;
    return surf_plain(C, points, params, ncells, box, snoise0, snoise1, out, phase_time, srdist, sigma, phase_vel);
}

int mct_surf_likelihood_box_batch_dev(mct_surf_ctx *C, int nchain, const double *points_dev,
                                      const double *params_dev, const int32_t *ncells_dev, const int32_t *cols,
                                      const double *noise_dev, double *out_dev, int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !points_dev || !params_dev || !ncells_dev || !cols ||
        !noise_dev || !out_dev || !status_dev)
        return MCT_ERR_ARG;
    const mct_geom &G = C->G;
    for (int c = 0; c < nchain; c++) {
        const int32_t *b = cols + 4 * c;
        if (!C->keep_valid[c] || b[0] < 1 || b[1] > G.nx || b[0] > b[1] || b[2] < 1 || b[3] > G.ny || b[2] > b[3])
            return MCT_ERR_ARG;
    }
    hipStream_t s = C->stream;
    MCT_CHECK(hipStreamSynchronize(s));   // the pinned box staging is free again
    int32_t *hb = C->h_box, *ho = hb + 4 * (size_t)nchain;
    long total = 0;
    for (int c = 0; c < nchain; c++) {
        for (int k = 0; k < 4; k++) hb[4 * c + k] = cols[4 * c + k];
        ho[c] = (int32_t)total;
        total += (long)(cols[4 * c + 1] - cols[4 * c] + 1) * (cols[4 * c + 3] - cols[4 * c + 2] + 1);
    }
    ho[nchain] = (int32_t)total;
    MCT_CHECK(hipMemcpyAsync(C->box.p, hb, sizeof(int32_t) * (5 * (size_t)nchain + 1), hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(status_dev, 0, sizeof(int32_t) * nchain, s));
    const int32_t *d_box = C->box.as<int32_t>();
    return surf_pipeline(C, 0, nchain, 0, 1, s, points_dev, params_dev, ncells_dev, 0, noise_dev, out_dev,
                         status_dev, false, nullptr, d_box, d_box + 4 * nchain, total);
}

int mct_surf_accept(mct_surf_ctx *C, int chain)
{
    if (!C || chain < 0 || chain >= C->S.max_chains) return MCT_ERR_ARG;
    const long n = (long)C->G.nx * C->G.ny * C->S.np;
    MCT_CHECK(hipMemcpyAsync(C->phase_keep.as<double>() + chain * n, C->phase.as<double>() + chain * n,
                             sizeof(double) * n, hipMemcpyDeviceToDevice, C->stream));
    if (C->S.phase_group)
        MCT_CHECK(hipMemcpyAsync(C->group_keep.as<double>() + chain * n, C->group.as<double>() + chain * n,
                                 sizeof(double) * n, hipMemcpyDeviceToDevice, C->stream));
    C->keep_valid[chain] = 1;
    return MCT_OK;
}

int mct_surf_group_velocities(mct_surf_ctx *C, int chain, double *group_vel)
{
    if (!C || chain < 0 || chain >= C->S.max_chains || !group_vel) return MCT_ERR_ARG;
    const long n = (long)C->G.nx * C->G.ny * C->S.np;
    MCT_CHECK(hipStreamSynchronize(C->stream));
    MCT_CHECK(hipMemcpy(group_vel, C->group.as<double>() + chain * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    return MCT_OK;
}

int mct_surf_accept_batch(mct_surf_ctx *C, int nchain, const int32_t *accepted)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !accepted) return MCT_ERR_ARG;
    for (int c = 0; c < nchain; c++) {
        if (!accepted[c]) continue;
        int rc = mct_surf_accept(C, c);
        if (rc) return rc;
    }
    return MCT_OK;
}

int mct_surf_likelihood_batch_graph(mct_surf_ctx *C, int nchain, const double *points_dev,
                                    const double *params_dev, const int32_t *ncells_dev,
                                    const double *noise_dev, double *out_dev, int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !points_dev || !params_dev || !ncells_dev ||
        !noise_dev || !out_dev || !status_dev)
        return MCT_ERR_ARG;
    if (surf_parts(C, nchain) > 1) return MCT_ERR_ARG;   // one stream only
    const void *key[6] = {points_dev, params_dev, ncells_dev, noise_dev, out_dev, status_dev};
    hipGraphExec_t exec = nullptr;
    for (auto &g : C->graphs) {
        bool same = g.nchain == nchain;
        for (int i = 0; i < 6 && same; i++) same = g.key[i] == key[i];
        if (same) { exec = g.exec; break; }
    }
    if (!exec) {
        if (C->graphs.size() >= 64) {
            (void)hipGraphExecDestroy(C->graphs.front().exec);
            C->graphs.erase(C->graphs.begin());
        }
        MCT_CHECK(hipStreamSynchronize(C->stream));
        MCT_CHECK(hipStreamBeginCapture(C->stream, hipStreamCaptureModeThreadLocal));
        C->capturing = true;
        int rc = mct_fill_i32(status_dev, nchain, 0, C->stream);
        if (!rc)
            rc = surf_pipeline(C, 0, nchain, 0, 1, C->stream, points_dev, params_dev, ncells_dev, 0, noise_dev,
                               out_dev, status_dev, false, nullptr);
        C->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(C->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess || !g) return MCT_ERR_HIP;
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return MCT_ERR_HIP;
        mct_surf_ctx::graph_entry ge;
        for (int i = 0; i < 6; i++) ge.key[i] = key[i];
        ge.nchain = nchain;
        ge.exec = exec;
        C->graphs.push_back(ge);
    }
    MCT_CHECK(hipGraphLaunch(exec, C->stream));
    return MCT_OK;
}

int mct_surf_likelihood_batch_dev(mct_surf_ctx *C, int nchain, const double *points_dev,
                                  const double *params_dev, const int32_t *ncells_dev,
                                  const double *noise_dev, double *out_dev, int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !points_dev || !params_dev || !ncells_dev ||
        !noise_dev || !out_dev || !status_dev)
        return MCT_ERR_ARG;
    MCT_CHECK(hipMemsetAsync(status_dev, 0, sizeof(int32_t) * nchain, C->stream));
    const int P = surf_parts(C, nchain);
    if (P <= 1)
        return surf_pipeline(C, 0, nchain, 0, 1, C->stream, points_dev, params_dev, ncells_dev, 0, noise_dev,
                             out_dev, status_dev, false, nullptr);
    // parts of the batch on their own streams, each started when the previous
    // part's dispersion is done, so one part's 2-D solves overlap the next
    // part's dispersion; the context's stream waits for all of them
    const long NC = C->S.max_cells, np = C->S.np;
    MCT_CHECK(hipEventRecord(C->ev[0], C->stream));
    MCT_CHECK(hipEventRecord(C->fork, C->stream));
    int rc = MCT_OK;
    for (int p = 0; p < P && !rc; p++) {
        const int c0 = (int)((long)nchain * p / P), c1 = (int)((long)nchain * (p + 1) / P);
        hipStream_t sp = C->sub[p];
        MCT_CHECK(hipStreamWaitEvent(sp, C->fork, 0));
        if (p > 0) MCT_CHECK(hipStreamWaitEvent(sp, C->sev[p - 1][0], 0));
        rc = surf_pipeline(C, c0, c1 - c0, p, P, sp, points_dev + c0 * 3 * NC, params_dev + c0 * 3 * NC,
                           ncells_dev + c0, 0, noise_dev + c0 * 2 * np, out_dev + 3L * c0, status_dev + c0,
                           false, C->sev[p][0]);
        MCT_CHECK(hipEventRecord(C->sev[p][1], sp));
        MCT_CHECK(hipStreamWaitEvent(C->stream, C->sev[p][1], 0));
    }
    if (!rc) MCT_CHECK(hipEventRecord(C->ev[4], C->stream));
    return rc;
}

int mct_surf_noise_likelihood(mct_surf_ctx *C, const double *snoise0, const double *snoise1, double *out)
{
    if (!C || !snoise0 || !snoise1 || !out) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->S.np;
    MCT_CHECK(hipStreamSynchronize(s));
    double *hn = C->h_in;
    for (int i = 0; i < np; i++) { hn[i] = snoise0[i]; hn[np + i] = snoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    int rc = surf_noise(C, 1, C->noise.as<double>(), C->out.as<double>(), C->status.as<int32_t>());
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    return *C->h_status;
}

int mct_surf_noise_likelihood_from(mct_surf_ctx *C, const double *phase_time, const double *srdist,
                                   const double *snoise0, const double *snoise1, double *out, double *sigma)
{
    if (!C || !phase_time || !snoise0 || !snoise1 || !out) return MCT_ERR_ARG;
    const bool dist = C->S.phase_group || C->S.straight;
    if (dist && !srdist) return MCT_ERR_ARG;
    hipStream_t s = C->stream;
    const int np = C->S.np;
    const long npt = (long)np * C->S.nsrc * C->S.nrev;
    MCT_CHECK(hipStreamSynchronize(s));
    double *hn = C->h_in;
    for (int i = 0; i < np; i++) { hn[i] = snoise0[i]; hn[np + i] = snoise1[i]; }
    MCT_CHECK(hipMemcpyAsync(C->noise.p, hn, sizeof(double) * 2 * np, hipMemcpyHostToDevice, s));
    // the given traveltimes (and ray lengths) become chain 0's kept state
    MCT_CHECK(hipMemcpyAsync(C->phase_time.p, phase_time, sizeof(double) * npt, hipMemcpyHostToDevice, s));
    if (dist) MCT_CHECK(hipMemcpyAsync(C->rlen.p, srdist, sizeof(double) * npt, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemsetAsync(C->status.p, 0, sizeof(int32_t), s));
    const mct_surf_setup &S = C->S;
    mct_misfit_args ma;
    memset(&ma, 0, sizeof ma);
    ma.nchain = 1; ma.nsrc = S.nsrc; ma.nrev = S.nrev; ma.np = np; ma.reciprocal = 0;
    ma.picks = C->phase_time.as<double>(); ma.ttime = C->ttime.as<double>();
    ma.raystat = C->raystat.as<int32_t>(); ma.noise = C->noise.as<double>();
    ma.sigdep = 1; ma.nrays = S.nrays;
    ma.eps = (double)1.0e-10f;   // EPS, src/likelihood_surf.F90:37
    ma.out = C->out.as<double>(); ma.status = C->status.as<int32_t>();
    ma.dist = dist ? C->rlen.as<double>() : nullptr;
    ma.sigma = sigma ? C->sigma.as<double>() : nullptr;
    int rc = mct_launch_misfit(ma, s);
    if (rc) return rc;
    MCT_CHECK(hipMemcpyAsync(C->h_out, C->out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(C->h_status, C->status.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (sigma) MCT_CHECK(hipMemcpyAsync(sigma, C->sigma.p, sizeof(double) * npt, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < 3; i++) out[i] = C->h_out[i];
    return *C->h_status;
}

int mct_surf_noise_likelihood_batch_dev(mct_surf_ctx *C, int nchain, const double *noise_dev, double *out_dev,
                                        int32_t *status_dev)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains || !noise_dev || !out_dev || !status_dev) return MCT_ERR_ARG;
    MCT_CHECK(hipMemsetAsync(status_dev, 0, sizeof(int32_t) * nchain, C->stream));
    return surf_noise(C, nchain, noise_dev, out_dev, status_dev);
}

int mct_surf_fm2d_handover(mct_surf_ctx *C, int64_t *handover2)
{
    if (!C || !handover2) return MCT_ERR_ARG;
    const long n = C->fm2d_chunk;
    int32_t *r = (int32_t *)malloc(sizeof(int32_t) * 5 * n);
    if (!r) return MCT_ERR_NOMEM;
    if (hipStreamSynchronize(C->stream) != hipSuccess ||
        hipMemcpy(r, C->rounds.p, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(r + 4 * n, C->pflag.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess) {
        free(r);
        return MCT_ERR_HIP;
    }
    int64_t chk = 0, all = 0, why[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long i = 0; i < n; i++) {
        const int b = r[4 * i + 3];
        chk += b != 0;
        all += r[4 * n + i] == 1 || r[4 * n + i] == 3;
        for (int k = 0; k < 7; k++) why[k] += (b >> k) & 1;
    }
    free(r);
    handover2[0] = chk;
    handover2[1] = all;
    for (int k = 0; k < 7; k++) handover2[2 + k] = why[k];
    return MCT_OK;
}

int mct_surf_split_parts(mct_surf_ctx *C, int nchain, int *first_part_chains)
{
    if (!C || nchain < 1 || !first_part_chains) return MCT_ERR_ARG;
    const int P = surf_parts(C, nchain);
    *first_part_chains = (int)((long)nchain / P);
    return P;
}

int mct_surf_fm2d_stats(mct_surf_ctx *C, double *stats4)
{
    if (!C || !stats4) return MCT_ERR_ARG;
    const long n = C->fm2d_chunk;
    int32_t *r = (int32_t *)malloc(sizeof(int32_t) * 4 * n);
    if (!r) return MCT_ERR_NOMEM;
    if (hipStreamSynchronize(C->stream) != hipSuccess ||
        hipMemcpy(r, C->rounds.p, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost) != hipSuccess) {
        free(r);
        return MCT_ERR_HIP;
    }
    long cnt = 0;
    double sr = 0, ss = 0, se = 0;
    int mx = 0;
    for (long i = 0; i < n; i++)
        if (r[4 * i] > 0) {
            cnt++; sr += r[4 * i]; ss += r[4 * i + 1]; se += r[4 * i + 2];
            if (r[4 * i] > mx) mx = r[4 * i];
        }
    free(r);
    stats4[0] = (double)cnt;
    stats4[1] = cnt ? sr / cnt : 0.0;
    stats4[2] = cnt ? ss / cnt : 0.0;
    stats4[3] = cnt ? se / cnt : 0.0;
    (void)mx;
    return MCT_OK;
}

}   // extern "C"

// Captured batch graphs hold the launch shape of their capture: a setter that
// changes the shape drops them (after the context's queued work has finished).
static int surf_drop_graphs(mct_surf_ctx *C)
{
    if (C->graphs.empty()) return MCT_OK;
    MCT_CHECK(hipStreamSynchronize(C->stream));
    for (auto &g : C->graphs) (void)hipGraphExecDestroy(g.exec);
    C->graphs.clear();
    return MCT_OK;
}

extern "C" {

int mct_surf_set_method(mct_surf_ctx *C, int fm2d_method)
{
    if (!C || (fm2d_method != 0 && fm2d_method != 1)) return MCT_ERR_ARG;
    if (surf_drop_graphs(C)) return MCT_ERR_HIP;
    C->fm2d_method = fm2d_method;
    return MCT_OK;
}

int mct_surf_set_dispersion_mode(mct_surf_ctx *C, int mode)
{
    if (!C || mode < 0 || mode > 3) return MCT_ERR_ARG;
    if (surf_drop_graphs(C)) return MCT_ERR_HIP;
    C->dsp_mode = mode;
    return MCT_OK;
}

int mct_surf_set_tuning(mct_surf_ctx *C, int key, int64_t value)
{
    if (!C) return MCT_ERR_ARG;
    if (surf_drop_graphs(C)) return MCT_ERR_HIP;
    switch (key) {
    case MCT_TUNE_DSP_WAVES:
        if (value < 0) return MCT_ERR_ARG;
        C->dsp_waves = (long)value;
        return MCT_OK;
    case MCT_TUNE_SWEEP_KIND:
        if (value < 0 || value > 2) return MCT_ERR_ARG;
        C->F.sweep_kind = (int)value;
        return MCT_OK;
    case MCT_TUNE_SWEEP_THREADS:
        if (value != 0 && value != 64 && value != 128 && value != 256) return MCT_ERR_ARG;
        C->F.sweep_threads = (int)value;
        return MCT_OK;
    case MCT_TUNE_FM2D_CHUNK:
        if (value < 0) return MCT_ERR_ARG;
        C->fm2d_run_chunk = (long)value;
        return MCT_OK;
    case MCT_TUNE_DSP_DEDUP:
        if (value != 0 && value != 1) return MCT_ERR_ARG;
        C->dsp_dedup = (int)value;
        return MCT_OK;
    default:
        return MCT_ERR_ARG;
    }
}

int mct_surf_get_tuning(mct_surf_ctx *C, int key, int64_t *value)
{
    if (!C || !value) return MCT_ERR_ARG;
    switch (key) {
    case MCT_TUNE_DSP_WAVES: *value = C->dsp_waves; return MCT_OK;
    case MCT_TUNE_SWEEP_KIND: *value = C->F.sweep_kind; return MCT_OK;
    case MCT_TUNE_SWEEP_THREADS: *value = C->F.sweep_threads; return MCT_OK;
    case MCT_TUNE_FM2D_CHUNK: *value = C->fm2d_run_chunk > 0 ? C->fm2d_run_chunk : C->fm2d_chunk; return MCT_OK;
    case MCT_TUNE_DSP_DEDUP: *value = C->dsp_dedup; return MCT_OK;
    default: return MCT_ERR_ARG;
    }
}

int mct_surf_dispersion_work(mct_surf_ctx *C, int nchain, int64_t *evals, int64_t *layer_evals)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains) return MCT_ERR_ARG;
    const long n = (long)nchain * C->G.nx * C->G.ny;
    int32_t *a = (int32_t *)malloc(sizeof(int32_t) * n), *b = (int32_t *)malloc(sizeof(int32_t) * n);
    if (!a || !b) { free(a); free(b); return MCT_ERR_NOMEM; }
    int rc = MCT_OK;
    if (hipStreamSynchronize(C->stream) != hipSuccess ||
        hipMemcpy(a, C->ndltar.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b, C->nlayer.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = MCT_ERR_HIP;
    int64_t e = 0, le = 0;
    for (long i = 0; !rc && i < n; i++) { e += a[i]; le += (int64_t)a[i] * (b[i] > 1 ? b[i] - 1 : 0); }
    free(a); free(b);
    if (rc) return rc;
    if (evals) *evals = e;
    if (layer_evals) *layer_evals = le;
    return MCT_OK;
}

int mct_surf_dispersion_solved(mct_surf_ctx *C, int nchain, int64_t *columns, int64_t *layer_evals)
{
    if (!C || nchain < 1 || nchain > C->S.max_chains) return MCT_ERR_ARG;
    const long n = (long)nchain * C->G.nx * C->G.ny;
    int32_t *a = (int32_t *)malloc(sizeof(int32_t) * n), *b = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *d = (int32_t *)malloc(sizeof(int32_t) * n);
    if (!a || !b || !d) { free(a); free(b); free(d); return MCT_ERR_NOMEM; }
    int rc = MCT_OK;
    if (hipStreamSynchronize(C->stream) != hipSuccess ||
        hipMemcpy(a, C->ndltar.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b, C->nlayer.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        (C->dedup_last && hipMemcpy(d, C->dsp_dup.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess))
        rc = MCT_ERR_HIP;
    int64_t cols = 0, le = 0;
    for (long i = 0; !rc && i < n; i++) {
        if (C->dedup_last && d[i] >= 0) continue;       // copied from an identical column
        cols++;
        le += (int64_t)a[i] * (b[i] > 1 ? b[i] - 1 : 0);
    }
    free(a); free(b); free(d);
    if (rc) return rc;
    if (columns) *columns = cols;
    if (layer_evals) *layer_evals = le;
    return MCT_OK;
}

int mct_surf_stream(mct_surf_ctx *C, void **stream)
{
    if (!C || !stream) return MCT_ERR_ARG;
    *stream = (void *)C->stream;
    return MCT_OK;
}

int mct_surf_dev_arrays(mct_surf_ctx *C, double **phase_time_dev, double **phase_vel_dev)
{
    if (!C) return MCT_ERR_ARG;
    if (phase_time_dev) *phase_time_dev = C->phase_time.as<double>();
    if (phase_vel_dev) *phase_vel_dev = C->phase.as<double>();
    return MCT_OK;
}

} // extern "C"
