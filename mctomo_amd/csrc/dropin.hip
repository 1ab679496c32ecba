// dropin.hip -- link-level drop-ins for the reference's own entry points:
//   fastMarching3D      (fastmarching/fastMarching.cpp:26-129, C ABI)
//   mct_voronoi_to_grid (kdtree_to_grid, src/mcmc_loc2.f90:2002-2082)
// Both take host arrays in the reference's Fortran storage.  They share one
// process-wide workspace guarded by a mutex, so the reference's OpenMP loop
// over sources (src/likelihood_body.F90:204-224) may call fastMarching3D from
// several threads: the calls are serialised onto one stream.
#include "mct_internal.h"
#include <string.h>
#include <mutex>

namespace {

struct ws_buf {
    void *p = nullptr;
    size_t n = 0;
    int need(size_t bytes)
    {
        if (bytes == 0) bytes = 16;
        if (n >= bytes) return MCT_OK;
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return MCT_ERR_NOMEM;
        n = bytes;
        return MCT_OK;
    }
};

struct dropin_ws {
    std::mutex mu;
    hipStream_t s = nullptr;
    ws_buf vel_d, vel_t, T, fl0, fl1, k0, k1, band, counts, src, rev, tt, field, pts, prm,
        sid, vp, vs, rho, pm, fcount, fnodes, kdn, kdi, kdr, heap, err;
    uint32_t *h_count = nullptr;
    int init()
    {
        if (s) return MCT_OK;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return MCT_ERR_HIP;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MCT_ERR_HIP;
        if (hipHostMalloc((void **)&h_count, 4 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
            return MCT_ERR_NOMEM;
        return MCT_OK;
    }
};

dropin_ws &WS()
{
    static dropin_ws w;
    return w;
}

// Fortran (nz,ny,nx) double velocity -> FP32 eikonal tiles (padding lanes 0)
template <class VT>
__global__ void k_fortran_to_tiles(const double *v, VT *vt, mct_geom G)
{
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)G.ntiles * MCT_TILE_N;
    if (e >= n) return;
    int tile = (int)(e >> 6), lane = (int)(e & 63);
    int tx = tile % G.ntx, ty = (tile / G.ntx) % G.nty, tz = tile / (G.ntx * G.nty);
    int x = tx * 4 + (lane & 3), y = ty * 4 + ((lane >> 2) & 3), z = tz * 4 + (lane >> 4);
    VT out = 0;
    if (x < G.nx && y < G.ny && z < G.nz)
        out = (VT)v[(long)z + (long)G.nz * ((long)y + (long)G.ny * x)];
    vt[e] = out;
}

// tiles -> Fortran (nz,ny,nx) double field
__global__ void k_tiles_to_fortran(const void *tv, int t64, double *f, mct_geom G)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)G.nx * G.ny * G.nz;
    if (g >= n) return;
    int z = (int)(g % G.nz), y = (int)((g / G.nz) % G.ny), x = (int)(g / ((long)G.nz * G.ny));
    int tile = ((z >> 2) * G.nty + (y >> 2)) * G.ntx + (x >> 2);
    int lane = ((z & 3) << 4) | ((y & 3) << 2) | (x & 3);
    long i = ((long)tile << 6) | lane;
    f[g] = t64 ? ((const double *)tv)[i] : (double)((const float *)tv)[i];
}

int eik_rounds(const mct_eik_problem_set &E, uint32_t *h_count, double hx, double hy, double hz,
               hipStream_t s)
{
    const mct_geom &G = E.G;
    const int max_rounds = E.band_rounds + 8 * (G.ntx + G.nty + G.ntz) + 256;
    int dev = 0, cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        cus = prop.multiProcessorCount;
    int r = 0;
    for (;;) {
        for (int k = 0; k < 8; k++, r++) {
            int rc = mct_launch_eik_round(E, r, hx, hy, hz, cus * 32, s);
            if (rc) return rc;
        }
        MCT_CHECK(hipMemcpyAsync(h_count, E.any + (r % 3), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        MCT_CHECK(hipStreamSynchronize(s));
        if (*h_count == 0) return MCT_OK;
        if (r >= max_rounds) return MCT_ERR_NOCONV;
    }
}

} // namespace

extern "C" int fastMarching3D(int nx, int ny, int nz, double dx, double dy, double dz,
                              double xmin, double ymin, double zmin, double *vel,
                              const double *src, const double *rev, int nrev, int order,
                              double *ttime, double *tfield)
{
    if (nx < 2 || ny < 2 || nz < 2 || !(dx > 0) || !(dy > 0) || !(dz > 0) || !vel || !src ||
        nrev < 0 || (nrev > 0 && (!rev || !ttime)))
        return MCT_ERR_ARG;
    dropin_ws &W = WS();
    std::lock_guard<std::mutex> lock(W.mu);
    int rc = W.init();
    if (rc) return rc;
    mct_grid g = {nx, ny, nz, xmin, ymin, zmin, dx, dy, dz, 1.0};
    mct_geom G = mct_make_geom(&g);
    const long gn = (long)nx * ny * nz, tn = (long)G.ntiles * MCT_TILE_N;
    const int f64 = mct_default_f64();
    if ((rc = W.vel_d.need(sizeof(double) * gn)) || (rc = W.vel_t.need((f64 ? sizeof(double) : sizeof(float)) * tn)) ||
        (rc = W.T.need(sizeof(double) * tn)) || (rc = W.fl0.need((G.ntiles + 3) & ~3)) ||
        (rc = W.fl1.need((G.ntiles + 3) & ~3)) || (rc = W.counts.need(4 * sizeof(uint32_t))) ||
        (rc = W.k0.need(sizeof(uint32_t) * G.ntiles)) || (rc = W.k1.need(sizeof(uint32_t) * G.ntiles)) ||
        (rc = W.band.need(sizeof(float) * 4)) ||
        (rc = W.src.need(4 * sizeof(double))) || (rc = W.rev.need(sizeof(double) * 3 * (nrev + 1))) ||
        (rc = W.tt.need(sizeof(double) * (nrev + 1))) ||
        (tfield && (rc = W.field.need(sizeof(double) * gn))))
        return rc;
    hipStream_t s = W.s;
    MCT_CHECK(hipMemcpyAsync(W.vel_d.p, vel, sizeof(double) * gn, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(W.src.p, src, 3 * sizeof(double), hipMemcpyHostToDevice, s));
    if (nrev > 0) MCT_CHECK(hipMemcpyAsync(W.rev.p, rev, sizeof(double) * 3 * nrev, hipMemcpyHostToDevice, s));
    if (f64)
        hipLaunchKernelGGL(k_fortran_to_tiles<double>, dim3((unsigned)((tn + 255) / 256)), dim3(256), 0, s,
                           (const double *)W.vel_d.p, (double *)W.vel_t.p, G);
    else
        hipLaunchKernelGGL(k_fortran_to_tiles<float>, dim3((unsigned)((tn + 255) / 256)), dim3(256), 0, s,
                           (const double *)W.vel_d.p, (float *)W.vel_t.p, G);
    MCT_CHECK(hipGetLastError());
    mct_eik_problem_set E;
    memset(&E, 0, sizeof E);
    E.G = G; E.nprob = 1; E.nsrc_eff = 1; E.nwave = 1; E.order = order == 1 ? 1 : 0;
    E.T = W.T.p; E.t64 = f64 ? 1 : mct_default_t64();
    E.vel = f64 ? nullptr : (const float *)W.vel_t.p;
    E.vel64 = f64 ? (const double *)W.vel_t.p : nullptr;
        E.flags[0] = (uint8_t *)W.fl0.p; E.flags[1] = (uint8_t *)W.fl1.p;
    E.any = (uint32_t *)W.counts.p;
    E.key[0] = (uint32_t *)W.k0.p; E.key[1] = (uint32_t *)W.k1.p;
    E.band_dt = (const float *)W.band.p;
    E.band_rounds = mct_band_rounds(G, mct_default_band_tiles());
    if ((rc = mct_launch_band_dt(E.vel, E.vel64, tn, 1, (float)fmin(fmin(dx, dy), dz), mct_default_band_tiles(),
                                 (float *)W.band.p, s)))
        return rc;
    if ((rc = mct_launch_eik_init(E, (const double *)W.src.p, 3, 0, xmin, ymin, zmin, dx, dy, dz, s))) return rc;
    if (E.order == 1 && f64 && mct_default_heap_order1()) {
        // the reference's heap loop replayed (its mixed-order times depend on the pop order)
        mct_sfmm_args a;
        memset(&a, 0, sizeof a);
        a.G = G; a.nprob = 1; a.nsrc_eff = 1; a.nwave = 1; a.order = 1;
        a.T = (double *)W.T.p; a.V = (const double *)W.vel_t.p;
        a.src = (const double *)W.src.p; a.src_stride = 3; a.src_chain_stride = 0;
        a.xmin = xmin; a.ymin = ymin; a.zmin = zmin; a.lx = dx; a.ly = dy; a.lz = dz;
        a.hcap = mct_sfmm_heap_cap(G);
        if ((rc = W.heap.need(sizeof(uint32_t) * (size_t)a.hcap)) || (rc = W.err.need(4 * sizeof(uint32_t))))
            return rc;
        a.heap = (uint32_t *)W.heap.p; a.err = (uint32_t *)W.err.p;
        if ((rc = mct_launch_sfmm(a, s))) return rc;
        MCT_CHECK(hipMemcpyAsync(W.h_count, a.err, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        MCT_CHECK(hipStreamSynchronize(s));
        if ((rc = mct_sfmm_status(*W.h_count))) return rc;
    } else if ((rc = eik_rounds(E, W.h_count, dx, dy, dz, s))) {
        return rc;
    }
    if (nrev > 0) {
        mct_pick_args pa;
        memset(&pa, 0, sizeof pa);
        pa.G = G; pa.T = E.T; pa.t64 = E.t64; pa.nchain = 1; pa.nwave = 1; pa.nsrc_eff = 1; pa.npick = nrev;
        pa.pick_pts = (const double *)W.rev.p; pa.pick_stride_chain = 0; pa.pick_point_stride = 3;
        pa.xmin = xmin; pa.ymin = ymin; pa.zmin = zmin; pa.lx = dx; pa.ly = dy; pa.lz = dz;
        pa.out = (double *)W.tt.p;
        if ((rc = mct_launch_picks(pa, s))) return rc;
        MCT_CHECK(hipMemcpyAsync(ttime, W.tt.p, sizeof(double) * nrev, hipMemcpyDeviceToHost, s));
    }
    if (tfield) {
        hipLaunchKernelGGL(k_tiles_to_fortran, dim3((unsigned)((gn + 255) / 256)), dim3(256), 0, s,
                           (const void *)W.T.p, E.t64, (double *)W.field.p, G);
        MCT_CHECK(hipGetLastError());
        MCT_CHECK(hipMemcpyAsync(tfield, W.field.p, sizeof(double) * gn, hipMemcpyDeviceToHost, s));
    }
    MCT_CHECK(hipStreamSynchronize(s));
    return MCT_OK;
}

extern "C" int mct_voronoi_to_grid(const mct_grid *grid, const double *points, const double *params,
                                   int ncells, const double *box, const double *pm,
                                   int32_t *sites_id, double *vp, double *vs, double *rho)
{
    if (!grid || !points || !params || ncells < 1 || !sites_id || !vp || !vs || !rho) return MCT_ERR_ARG;
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1 || !(grid->dx > 0) || !(grid->dy > 0) || !(grid->dz > 0))
        return MCT_ERR_ARG;
    dropin_ws &W = WS();
    std::lock_guard<std::mutex> lock(W.mu);
    int rc = W.init();
    if (rc) return rc;
    mct_geom G = mct_make_geom(grid);
    const long gn = (long)G.nx * G.ny * G.nz;
    if ((rc = W.pts.need(sizeof(double) * 3 * ncells)) || (rc = W.prm.need(sizeof(double) * 3 * ncells)) ||
        (rc = W.sid.need(sizeof(int32_t) * gn)) || (rc = W.vp.need(sizeof(double) * gn)) ||
        (rc = W.vs.need(sizeof(double) * gn)) || (rc = W.rho.need(sizeof(double) * gn)) ||
        (rc = W.pm.need(sizeof(double) * 3)) || (rc = W.fcount.need(sizeof(int32_t))) ||
        (rc = W.fnodes.need(sizeof(int32_t) * MCT_MAX_FLAGGED)) ||
        (rc = W.kdn.need(sizeof(mct_kd_node) * (2 * ncells + 2))) ||
        (rc = W.kdi.need(sizeof(int32_t) * (ncells + 1))) || (rc = W.kdr.need(sizeof(double) * 3 * ncells)))
        return rc;
    hipStream_t s = W.s;
    MCT_CHECK(hipMemcpyAsync(W.pts.p, points, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(W.prm.p, params, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, s));
    // vp/vs/rho are inout: nodes outside the box (or not selected by pm) keep their values
    MCT_CHECK(hipMemcpyAsync(W.sid.p, sites_id, sizeof(int32_t) * gn, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(W.vp.p, vp, sizeof(double) * gn, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(W.vs.p, vs, sizeof(double) * gn, hipMemcpyHostToDevice, s));
    MCT_CHECK(hipMemcpyAsync(W.rho.p, rho, sizeof(double) * gn, hipMemcpyHostToDevice, s));
    if (pm) MCT_CHECK(hipMemcpyAsync(W.pm.p, pm, sizeof(double) * 3, hipMemcpyHostToDevice, s));
    mct_nearest_args a;
    memset(&a, 0, sizeof a);
    a.G = G; a.nchain = 1; a.max_cells = ncells; a.points = (const double *)W.pts.p;
    a.params = (const double *)W.prm.p; a.ncells = nullptr; a.ncells_const = ncells;
    a.ix0 = 1; a.ix1 = G.nx; a.iy0 = 1; a.iy1 = G.ny; a.iz0 = 1; a.iz1 = G.nz;
    if (box) {   // index range of the perturbed box (src/mcmc_loc2.f90:2034-2045)
        a.ix0 = (int)floor((box[0] - G.xmin) / G.dx) + 1; a.ix1 = (int)floor((box[3] - G.xmin) / G.dx) + 1;
        a.iy0 = (int)floor((box[1] - G.ymin) / G.dy) + 1; a.iy1 = (int)floor((box[4] - G.ymin) / G.dy) + 1;
        a.iz0 = (int)floor((box[2] - G.zmin) / G.dz) + 1; a.iz1 = (int)floor((box[5] - G.zmin) / G.dz) + 1;
        if (a.ix0 < 1) a.ix0 = 1;
        if (a.iy0 < 1) a.iy0 = 1;
        if (a.iz0 < 1) a.iz0 = 1;
        if (a.ix1 > G.nx) a.ix1 = G.nx;
        if (a.iy1 > G.ny) a.iy1 = G.ny;
        if (a.iz1 > G.nz) a.iz1 = G.nz;
    }
    a.pm = pm ? (const double *)W.pm.p : nullptr;
    a.vel_tiles = nullptr; a.nwave = 1;
    a.sites_id = (int32_t *)W.sid.p; a.vp = (double *)W.vp.p; a.vs = (double *)W.vs.p; a.rho = (double *)W.rho.p;
    a.flag_count = (int32_t *)W.fcount.p; a.flag_nodes = (int32_t *)W.fnodes.p;
    a.kd_nodes = (mct_kd_node *)W.kdn.p; a.kd_ind = (int32_t *)W.kdi.p; a.kd_rdata = (double *)W.kdr.p;
    if ((rc = mct_launch_nearest(a, s))) return rc;
    MCT_CHECK(hipMemcpyAsync(sites_id, W.sid.p, sizeof(int32_t) * gn, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(vp, W.vp.p, sizeof(double) * gn, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(vs, W.vs.p, sizeof(double) * gn, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipMemcpyAsync(rho, W.rho.p, sizeof(double) * gn, hipMemcpyDeviceToHost, s));
    MCT_CHECK(hipStreamSynchronize(s));
    return MCT_OK;
}
