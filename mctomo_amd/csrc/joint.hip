// joint.hip -- the joint body + surface-wave likelihood (datatype 3).
//
// likelihood() with LIKE_SET%datatype = 3 (src/likelihood.f90:83-88) is
// body_likelihood on dat(1), P and S (src/likelihood_body.F90:204-216: the
// else branch of datatype 0), plus surf_likelihood on dat(2) with the model's
// own vp and rho (no vs2vp_3d / vp2rho_3d), the three sums added; like_setup
// builds both (:44-52), noise_likelihood (:146-151) and location_likelihood
// (:163-184) add them the same way.  A joint context holds a body context and
// a surface context in datatype 3 on the same grid.  A batch runs the two
// pipelines on their own streams at once -- the 3-D eikonal rounds and the
// dispersion / 2-D kernels share the device -- and one kernel adds them.
#include "mct_internal.h"

#include <float.h>
#include <new>
#include <string.h>

struct mct_joint_ctx {
    mct_ctx *body = nullptr;
    mct_surf_ctx *surf = nullptr;
    hipStream_t bs = nullptr, ss = nullptr;     // the contexts' own streams
    hipEvent_t ev_body = nullptr;
    int max_chains = 0, max_cells = 0, nsrc = 0, nwave = 2, np_s = 0;
    // per-part results of a batch and the plain call's device inputs
    double *d_out2 = nullptr;                    // (2, max_chains, 3)
    int32_t *d_st2 = nullptr;                    // (2, max_chains)
    double *d_in = nullptr;                      // one chain: points, params, locations, bnoise, snoise
    int32_t *d_nc = nullptr;
    double *d_out = nullptr;
    int32_t *d_st = nullptr;
    double *h_in = nullptr;                      // pinned staging
};

namespace {

// like = body + surface (a rejected surface model has like = huge, and huge
// plus a finite body term rounds to huge, as in the reference's sum); status:
// the first fatal (< 0) of the two, else the surface's soft reject.
__global__ void k_joint_sum(int nchain, const double *ob, const double *os, const int32_t *sb,
                            const int32_t *ss, double *out, int32_t *status)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchain) return;
    for (int k = 0; k < 3; k++) out[3 * c + k] = ob[3 * c + k] + os[3 * c + k];
    if (status) {
        const int a = sb ? sb[c] : 0, b = ss ? ss[c] : 0;
        status[c] = a < 0 ? a : (b < 0 ? b : (a > 0 ? a : b));
    }
}

int joint_batch(mct_joint_ctx *J, int nchain, const double *points_dev, const double *params_dev,
                const int32_t *ncells_dev, const double *locations_dev, const double *bnoise_dev,
                const double *snoise_dev, double *out_dev, int32_t *status_dev)
{
    double *ob = J->d_out2, *os = J->d_out2 + 3L * J->max_chains;
    int32_t *sb = J->d_st2, *ss = J->d_st2 + J->max_chains;
    int rc = mct_body_likelihood_batch_dev(J->body, nchain, points_dev, params_dev, ncells_dev, locations_dev,
                                           bnoise_dev, ob, sb);
    if (rc) return rc;
    MCT_CHECK(hipEventRecord(J->ev_body, J->bs));
    rc = mct_surf_likelihood_batch_dev(J->surf, nchain, points_dev, params_dev, ncells_dev, snoise_dev, os, ss);
    if (rc) return rc;
    MCT_CHECK(hipStreamWaitEvent(J->ss, J->ev_body, 0));
    hipLaunchKernelGGL(k_joint_sum, dim3((unsigned)((nchain + 63) / 64)), dim3(64), 0, J->ss, nchain,
                       (const double *)ob, (const double *)os, (const int32_t *)sb, (const int32_t *)ss, out_dev,
                       status_dev);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

}   // namespace

extern "C" {

int mct_joint_create(const mct_grid *grid, const mct_body_setup *body, const mct_surf_setup *surf,
                     mct_joint_ctx **out)
{
    if (!grid || !body || !surf || !out) return MCT_ERR_ARG;
    *out = nullptr;
    // datatype 3: P and S body waves; the surface half keeps the sites' vp, rho
    if (body->datatype != 3 || (surf->datatype != 3) || body->max_chains != surf->max_chains ||
        body->max_cells != surf->max_cells || body->max_chains < 1)
        return MCT_ERR_ARG;
    mct_joint_ctx *J = new (std::nothrow) mct_joint_ctx();
    if (!J) return MCT_ERR_NOMEM;
    J->max_chains = body->max_chains;
    J->max_cells = body->max_cells;
    J->nsrc = body->nsrc;
    J->np_s = surf->np;
    int rc = mct_body_create(grid, body, &J->body);
    if (!rc) rc = mct_surf_create(grid, surf, &J->surf);
    void *bs = nullptr, *ss = nullptr;
    if (!rc) rc = mct_body_stream(J->body, &bs);
    if (!rc) rc = mct_surf_stream(J->surf, &ss);
    J->bs = (hipStream_t)bs;
    J->ss = (hipStream_t)ss;
    const long M = J->max_chains, NC = J->max_cells;
    const long nin = 6 * NC + 4L * J->nsrc + 4 + 2L * J->np_s;
    if (!rc && (hipEventCreateWithFlags(&J->ev_body, hipEventDisableTiming) != hipSuccess ||
                hipMalloc((void **)&J->d_out2, sizeof(double) * 6 * M) != hipSuccess ||
                hipMalloc((void **)&J->d_st2, sizeof(int32_t) * 2 * M) != hipSuccess ||
                hipMalloc((void **)&J->d_in, sizeof(double) * nin) != hipSuccess ||
                hipMalloc((void **)&J->d_nc, sizeof(int32_t)) != hipSuccess ||
                hipMalloc((void **)&J->d_out, sizeof(double) * 3) != hipSuccess ||
                hipMalloc((void **)&J->d_st, sizeof(int32_t)) != hipSuccess ||
                hipHostMalloc((void **)&J->h_in, sizeof(double) * (nin + 8)) != hipSuccess))
        rc = MCT_ERR_NOMEM;
    if (rc) {
        mct_joint_destroy(J);
        return rc;
    }
    *out = J;
    return MCT_OK;
}

int mct_joint_destroy(mct_joint_ctx *J)
{
    if (!J) return MCT_ERR_ARG;
    if (J->bs) (void)hipStreamSynchronize(J->bs);
    if (J->ss) (void)hipStreamSynchronize(J->ss);
    if (J->body) mct_destroy(J->body);
    if (J->surf) mct_surf_destroy(J->surf);
    if (J->ev_body) (void)hipEventDestroy(J->ev_body);
    void *d[] = {J->d_out2, J->d_st2, J->d_in, J->d_nc, J->d_out, J->d_st};
    for (void *p : d)
        if (p) (void)hipFree(p);
    if (J->h_in) (void)hipHostFree(J->h_in);
    delete J;
    return MCT_OK;
}

int mct_joint_parts(mct_joint_ctx *J, mct_ctx **body, mct_surf_ctx **surf)
{
    if (!J || !body || !surf) return MCT_ERR_ARG;
    *body = J->body;
    *surf = J->surf;
    return MCT_OK;
}

int mct_joint_likelihood_batch_dev(mct_joint_ctx *J, int nchain, const double *points_dev,
                                   const double *params_dev, const int32_t *ncells_dev,
                                   const double *locations_dev, const double *bnoise_dev,
                                   const double *snoise_dev, double *out_dev, int32_t *status_dev)
{
    if (!J || nchain < 1 || nchain > J->max_chains || !points_dev || !params_dev || !ncells_dev ||
        !locations_dev || !out_dev)
        return MCT_ERR_ARG;
    return joint_batch(J, nchain, points_dev, params_dev, ncells_dev, locations_dev, bnoise_dev, snoise_dev,
                       out_dev, status_dev);
}

int mct_joint_sync(mct_joint_ctx *J)
{
    if (!J) return MCT_ERR_ARG;
    int rc = mct_sync(J->body);
    int rc2 = mct_surf_sync(J->surf);
    return rc ? rc : rc2;
}

int mct_joint_likelihood(mct_joint_ctx *J, const double *points, const double *params, int ncells,
                         const double *locations, const double *bnoise0, const double *bnoise1,
                         const double *snoise0, const double *snoise1, double *out, double *out_parts)
{
    if (!J || !points || !params || !locations || !out || ncells < 1 || ncells > J->max_cells) return MCT_ERR_ARG;
    const long NC = J->max_cells, ns = J->nsrc, nps = J->np_s;
    // device layout of the one-chain batch: points, params (max_cells x 3 each),
    // locations (nsrc x 4), body noise (2 x nwave), surface noise (2 x np)
    double *h = J->h_in;
    const long o_par = 3 * NC, o_loc = 6 * NC, o_bn = o_loc + 4 * ns, o_sn = o_bn + 4;
    memset(h, 0, sizeof(double) * (o_sn + 2 * nps));
    memcpy(h, points, sizeof(double) * 3 * ncells);
    memcpy(h + o_par, params, sizeof(double) * 3 * ncells);
    memcpy(h + o_loc, locations, sizeof(double) * 4 * ns);
    for (int w = 0; w < 2; w++) {
        h[o_bn + w] = bnoise0 ? bnoise0[w] : 0.0;
        h[o_bn + 2 + w] = bnoise1 ? bnoise1[w] : 0.0;
    }
    for (long p = 0; p < nps; p++) {
        h[o_sn + p] = snoise0 ? snoise0[p] : 0.0;
        h[o_sn + nps + p] = snoise1 ? snoise1[p] : 0.0;
    }
    int32_t nc = ncells;
    MCT_CHECK(hipMemcpyAsync(J->d_in, h, sizeof(double) * (o_sn + 2 * nps), hipMemcpyHostToDevice, J->bs));
    MCT_CHECK(hipMemcpyAsync(J->d_nc, &nc, sizeof(int32_t), hipMemcpyHostToDevice, J->bs));
    MCT_CHECK(hipStreamSynchronize(J->bs));
    int rc = joint_batch(J, 1, J->d_in, J->d_in + o_par, J->d_nc, J->d_in + o_loc, J->d_in + o_bn, J->d_in + o_sn,
                         J->d_out, J->d_st);
    if (rc) return rc;
    double res[9];
    int32_t st = 0;
    MCT_CHECK(hipMemcpyAsync(res, J->d_out, sizeof(double) * 3, hipMemcpyDeviceToHost, J->ss));
    MCT_CHECK(hipMemcpyAsync(res + 3, J->d_out2, sizeof(double) * 3, hipMemcpyDeviceToHost, J->ss));
    MCT_CHECK(hipMemcpyAsync(res + 6, J->d_out2 + 3L * J->max_chains, sizeof(double) * 3, hipMemcpyDeviceToHost,
                             J->ss));
    MCT_CHECK(hipMemcpyAsync(&st, J->d_st, sizeof(int32_t), hipMemcpyDeviceToHost, J->ss));
    rc = mct_joint_sync(J);
    if (rc) return rc;
    memcpy(out, res, sizeof(double) * 3);
    if (out_parts) memcpy(out_parts, res + 3, sizeof(double) * 6);
    return st;
}

}   // extern "C"
