// picks_misfit.hip -- receiver traveltime picks and the Gaussian misfit.
//
// Picks follow nDGridMap::getTime (fastmarching/ndgridmap/ndgridmap.hpp
// :462-489): trilinear weights over the 8 clamped corner nodes of the cell
// holding the point (getPointCoord truncates and clamps, :194-206).  The
// weights and the sum are FP64, the node times are the FP32 solver output.
//
// The misfit is the body_likelihood block (src/likelihood_body.F90:286-333):
// srdist = traveltime, sigma = b0*srdist + b1 (sigdep) or the data sigma,
// like = sum r^2/(2 sigma^2) + sum log sigma + nrays/2 * log(PI2) with the
// residual r = t + origin time - observed, summed as an FP64 block reduction.
#include "mct_internal.h"
#include <float.h>

#pragma clang fp contract(off)

__device__ __forceinline__ long tl_of(const mct_geom &G, int x, int y, int z)
{
    int tile = ((z >> 2) * G.nty + (y >> 2)) * G.ntx + (x >> 2);
    int lane = ((z & 3) << 4) | ((y & 3) << 2) | (x & 3);
    return ((long)tile << 6) | lane;
}

__global__ void k_picks(mct_pick_args a)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)a.nchain * a.nwave * a.nsrc_eff * a.npick;
    if (i >= total) return;
    const mct_geom &G = a.G;
    int r = (int)(i % a.npick);
    long prob = i / a.npick;                       // (chain, wave, src)
    int c = (int)(prob / ((long)a.nwave * a.nsrc_eff));
    const double *pp = a.pick_pts + (long)c * a.pick_stride_chain + (long)r * a.pick_point_stride;
    double p[3] = {pp[0] - a.xmin, pp[1] - a.ymin, pp[2] - a.zmin};
    double leaf[3] = {a.lx, a.ly, a.lz};
    int dims[3] = {G.nx, G.ny, G.nz};
    int pc[3];
    for (int d = 0; d < 3; d++) {
        // int(p/h) truncates toward zero, then clamps (getPointCoord);
        // done in double so out-of-range inputs never reach the conversion
        double qd = p[d] / leaf[d];
        int ci = !(qd >= 1.0) ? 0 : (qd >= (double)dims[d] ? dims[d] - 1 : (int)qd);
        pc[d] = ci;
    }
    const long toff = prob * (long)G.ntiles * MCT_TILE_N;
    const float *Tf = (const float *)a.T + toff;
    const double *Td = (const double *)a.T + toff;
    if (a.interp3d) {
        // Interpolate3d / point2idx (src/likelihood_body.F90:976-1007,1038-1054):
        // the cell floor((p - min)/d) + 1 clamped to [1, n-1] (NaN to 1), z in
        // unscaled units, trilinear weights in x, y, z loop order
        const double dzs = G.dz / G.scaling;
        const double fq[3] = {floor((pp[0] - G.xmin) / G.dx) + 1, floor((pp[1] - G.ymin) / G.dy) + 1,
                              floor((pp[2] - G.zmin / G.scaling) / dzs) + 1};
        int ic[3];
        for (int d = 0; d < 3; d++) ic[d] = !(fq[d] >= 1.0) ? 1 : (fq[d] >= (double)dims[d] ? dims[d] - 1 : (int)fq[d]);
        const double dsx = pp[0] - (G.xmin + (ic[0] - 1) * G.dx);
        const double dsy = pp[1] - (G.ymin + (ic[1] - 1) * G.dy);
        const double dsz = pp[2] - (G.zmin + (ic[2] - 1) * G.dz) / G.scaling;
        double qv = 0;
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++)
                for (int k = 1; k <= 2; k++) {
                    double weight = (1.0 - fabs((i - 1) * G.dx - dsx) / G.dx) *
                                    (1.0 - fabs((j - 1) * G.dy - dsy) / G.dy) *
                                    (1.0 - fabs((k - 1) * G.dz / G.scaling - dsz) / dzs);
                    long ti = tl_of(G, ic[0] + i - 2, ic[1] + j - 2, ic[2] + k - 2);
                    qv = qv + weight * (a.t64 ? Td[ti] : (double)Tf[ti]);
                }
        a.out[i] = qv;
        return;
    }
    double acc = 0.0;
    for (int v = 0; v < 8; v++) {
        int vc[3];
        for (int d = 0; d < 3; d++) {
            vc[d] = pc[d] + ((v >> (2 - d)) & 1);
            if (vc[d] >= dims[d]) vc[d] = dims[d] - 1;
        }
        double w = 1.0;
        for (int d = 0; d < 3; d++) {
            double s = p[d] - leaf[d] * pc[d];
            double wj = fabs((double)(vc[d] - pc[d]) * leaf[d] - leaf[d] + s);
            w = w * wj / leaf[d];
        }
        long ti = tl_of(G, vc[0], vc[1], vc[2]);
        acc = acc + w * (a.t64 ? Td[ti] : (double)Tf[ti]);
    }
    a.out[i] = acc;
}

int mct_launch_picks(const mct_pick_args &a, hipStream_t s)
{
    long total = (long)a.nchain * a.nwave * a.nsrc_eff * a.npick;
    if (total == 0) return MCT_OK;
    hipLaunchKernelGGL(k_picks, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

#define MIS_BLOCK 256

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < MIS_BLOCK / 64; i++) t += sh[i];
    return t;
}

__global__ __launch_bounds__(MIS_BLOCK) void k_misfit(mct_misfit_args a)
{
    __shared__ double sh[MIS_BLOCK / 64];
    __shared__ int fatal;
    const int c = blockIdx.x;
    const long npair = (long)a.nsrc * a.nrev;
    const double PI2 = (double)6.283185f;   // default-real literal (src/likelihood_body.F90:32)
    const double EPS = a.eps;
    if (a.status && a.status[c] > 0) {      // rejected model: like = huge(1d0)
        if (threadIdx.x == 0) { a.out[c * 3 + 0] = DBL_MAX; a.out[c * 3 + 1] = 0.0; a.out[c * 3 + 2] = 0.0; }
        return;
    }
    if (threadIdx.x == 0) fatal = 0;
    __syncthreads();
    double like = 0.0, misfit = 0.0, unw = 0.0, slog = 0.0;
    for (int i = 0; i < a.np; i++) {
        double b0 = a.noise ? a.noise[((long)c * 2 + 0) * a.np + i] : 0.0;
        double b1 = a.noise ? a.noise[((long)c * 2 + 1) * a.np + i] : 0.0;
        for (long nrr = threadIdx.x; nrr < npair; nrr += MIS_BLOCK) {
            int j = (int)(nrr / a.nrev), k = (int)(nrr % a.nrev);
            // phaseTime(k,j,i): traced from sources -> picks[c][i][j][k];
            // from receivers (nsrc > nrev) -> picks[c][i][k][j]
            long pidx = a.reciprocal ? (((long)c * a.np + i) * a.nrev + k) * a.nsrc + j
                                     : (((long)c * a.np + i) * a.nsrc + j) * a.nrev + k;
            double t = a.picks[pidx];
            // srdist: the phase time, or the ray length for group times (likelihood_surf.F90:324-331)
            double sd = a.dist ? a.dist[((long)c * a.np + i) * npair + nrr] : t;
            double sg = a.sigdep ? b0 * sd + b1 : a.ttime[((long)i * 3 + 1) * npair + nrr];
            if (a.raystat[((long)i * 2 + 0) * npair + nrr] == 1) {
                if (sg < EPS) fatal = 1;
                double r = t + (a.locations ? a.locations[((long)c * a.nsrc + j) * 4 + 3] : 0.0) - a.ttime[((long)i * 3 + 0) * npair + nrr];
                like = like + (r * r) / (2 * (sg * sg));
                misfit = misfit + (r * r) / (sg * sg);
                unw = unw + r * r;
            } else {
                sg = 1.0;
            }
            slog = slog + log(sg);
            if (a.phase_time) a.phase_time[(((long)c * a.np + i) * a.nsrc + j) * a.nrev + k] = t;
            if (a.srdist) a.srdist[((long)c * a.np + i) * npair + nrr] = sd;
            if (a.sigma) a.sigma[((long)c * a.np + i) * npair + nrr] = sg;
        }
    }
    like = block_sum(like, sh);
    misfit = block_sum(misfit, sh);
    unw = block_sum(unw, sh);
    slog = block_sum(slog, sh);
    __syncthreads();
    if (threadIdx.x == 0) {
        double half = (double)((float)a.nrays / 2.0f);   // dat%nrays/2.0 is default real
        a.out[c * 3 + 0] = like + slog + half * log(PI2);
        a.out[c * 3 + 1] = misfit;
        a.out[c * 3 + 2] = unw;
        if (a.status) a.status[c] = fatal ? MCT_ERR_SIGMA : a.status[c];
    }
}

__global__ void k_fill_i32(int32_t *p, long n, int32_t v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int mct_fill_i32(int32_t *p, long n, int32_t v, hipStream_t s)
{
    if (n <= 0) return MCT_OK;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_misfit(const mct_misfit_args &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_misfit, dim3(a.nchain), dim3(MIS_BLOCK), 0, s, a);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}
