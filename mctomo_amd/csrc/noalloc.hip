// noalloc.hip -- an allocation-free form of the 2-D phase-time entry.
//
// mct_modrays allocates its device buffers per call.  Callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc) get mct_modrays_dev_noalloc: the inputs and outputs are the
// caller's device arrays, and the solver's workspace -- the coarse B-spline
// velocities, the per-problem slots of the parallel and sequential solvers,
// the hand-over flags -- lives in one file-scope device array of
// MCT_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
// process ends.  Problems run in chunks of as many slots as the array holds;
// a call whose velocities and one slot do not fit is MCT_ERR_SCRATCH.  The
// array is shared by every call on a GPU: calls that overlap on one GPU race.
#include "mct_internal.h"

#include <mutex>
#include <string.h>

#define MCT_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[MCT_NOALLOC_SCRATCH_BYTES];

namespace {

// the device address of this GPU's scratch, looked up once per device (the
// host-side symbol is only a stand-in for the device array)
std::mutex g_mu;
void *g_ptr[MCT_NOALLOC_MAX_DEVICES] = {nullptr};

int scratch_ptr(unsigned char **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MCT_NOALLOC_MAX_DEVICES) return MCT_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MCT_ERR_HIP;
        g_ptr[dev] = p;
    }
    *out = (unsigned char *)g_ptr[dev];
    return MCT_OK;
}

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}   // namespace

extern "C" int mct_noalloc_scratch_bytes(int64_t *bytes)
{
    if (!bytes) return MCT_ERR_ARG;
    *bytes = (int64_t)MCT_NOALLOC_SCRATCH_BYTES;
    return MCT_OK;
}

extern "C" int mct_modrays_dev_noalloc(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
                                       int np, const double *velv_dev, int gdx, int gdz, int asgr, int sgdl,
                                       int sgs, int fom, int nsrc, const double *scx_dev, const double *scz_dev,
                                       int nrc, const double *rcx_dev, const double *rcz_dev,
                                       const int32_t *srs_dev, double *ttime_dev, int32_t *status_dev,
                                       int method, void *stream)
{
    if ((long)(nvx - 1) * gdx + 1 > 32767 || (long)(nvz - 1) * gdz + 1 > 32767 || 2L * sgs * sgdl + 1 > 32767)
        return MCT_ERR_ARG;
    if (nvx < 2 || nvz < 2 || np < 1 || gdx < 1 || gdz < 1 || sgdl < 1 || sgs < 1 || nsrc < 1 || nrc < 1 ||
        !velv_dev || !scx_dev || !scz_dev || !rcx_dev || !rcz_dev || !srs_dev || !ttime_dev || !status_dev ||
        !(dvxd > 0) || !(dvzd > 0) || (method != 0 && method != 1))
        return MCT_ERR_ARG;
    unsigned char *base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    mct_fm2d_args a;
    memset(&a, 0, sizeof a);
    a.nvx = nvx; a.nvz = nvz; a.gox = goxd; a.goz = gozd; a.dvx = dvxd; a.dvz = dvzd;
    a.gdx = gdx; a.gdz = gdz; a.asgr = asgr; a.sgdl = sgdl; a.sgs = sgs; a.fom = fom;
    a.nnx = (nvx - 1) * gdx + 1; a.nnz = (nvz - 1) * gdz + 1;
    const int rg = 2 * sgs * sgdl + 1;
    a.ld = a.nnz > rg ? a.nnz : rg;
    a.nc = a.nnx > rg ? a.nnx : rg;
    a.nchain = 1; a.np = np; a.nsrc = nsrc; a.nrc = nrc;
    const long nprob = (long)np * nsrc, wsz = (long)a.ld * a.nc;
    // scratch layout: coarse velocities, then `chunk` problem slots of each array
    const size_t veln_b = up256(2 * 8 * (size_t)a.nnx * a.nnz * np);   // velocities, then slowness
    const size_t per = (size_t)mct_fm2d_slot_bytes(a.ld, a.nc) + 4;          // + the hand-over flag
    if (veln_b + 6 * 256 + per > (size_t)MCT_NOALLOC_SCRATCH_BYTES) return MCT_ERR_SCRATCH;
    long chunk = (long)(((size_t)MCT_NOALLOC_SCRATCH_BYTES - veln_b - 6 * 256) / per);
    if (chunk > nprob) chunk = nprob;
    unsigned char *p = base;
    a.veln = (double *)p; a.sln = a.veln + (long)a.nnx * a.nnz * np; p += veln_b;
    a.ttn = (double *)p; p += up256(8 * (size_t)chunk * wsz);
    a.ttr = (double *)p; p += up256(8 * (size_t)chunk * wsz);
    a.vr = (double *)p; p += up256(8 * (size_t)chunk * wsz);
    a.nsts = (int32_t *)p; p += up256(4 * (size_t)chunk * wsz);
    a.nstr = (int32_t *)p; p += up256(4 * (size_t)chunk * wsz);
    a.heap = (int32_t *)p; p += up256(4 * (size_t)chunk * (wsz + 8));
    int32_t *pflag = (int32_t *)p; p += up256(4 * (size_t)chunk);
    if ((size_t)(p - base) > (size_t)MCT_NOALLOC_SCRATCH_BYTES) return MCT_ERR_SCRATCH;
    a.velv = velv_dev;
    a.scx = scx_dev; a.scz = scz_dev; a.rcx = rcx_dev; a.rcz = rcz_dev;
    a.srs = srs_dev; a.srs_pstride = (long)nsrc * nrc;
    a.ttime = ttime_dev; a.status = status_dev;
    // the status starts at zero (a kernel, not a memset, so the call can be captured)
    if ((rc = mct_fill_i32(status_dev, 1, 0, s))) return rc;
    const bool fast = method == 0 && mct_fm2d_fast_ok(a);
    return fast ? mct_launch_fm2d_fast(a, chunk, pflag, s) : mct_launch_fm2d(a, chunk, s);
}
