// Group-velocity traveltimes of the surface-wave path: the rays of modrays'
// rpaths (fm2d/fm2dray_cartesian.f90:801-1420, uar = 0, no Frechet
// derivatives), their lengths (T_RAY%length, :51-65) and the group delays
// along them (CalGroupTime / GetVelocity, src/likelihood_surf.F90:454-521).
//
// One thread per (problem, receiver): a ray is a sequential walk of half-cell
// steps down the traveltime gradient from the receiver to the source, so the
// parallelism is across rays.  A thread streams the ray: each point is final
// when the walk moves past it, and the running length and group delay are
// summed in the reference's point order, so no ray points are stored.  The
// coarse field and the refined source grid of each problem are the solvers'
// workspace slots, read while they are still resident (after the chunk's
// solve, before the next chunk reuses the slots).
#include "mct_internal.h"

namespace {

// floor(v) + 1 as an int for a cell index (Fortran floor(...) + 1); a
// non-finite or huge v (a ray step from a zero gradient) gives an index far
// outside the grid, which every caller clamps or rejects, instead of an
// undefined conversion.
__device__ __forceinline__ int cell1(double v)
{
    const double f = floor(v);
    return (f >= -1.0e9 && f <= 1.0e9) ? (int)f + 1 : (f > 0 ? 1000000001 : -1000000000);
}

// GetVelocity (likelihood_surf.F90:491-521): bilinear value of one period's
// group velocities, Fortran gvel(ny,nx) strided by np in the dispersion output.
__device__ __forceinline__ double ray_velocity(const double *gv, int np, int nx, int ny, double xmin,
                                               double ymin, double dx, double dy, double px, double py)
{
    int ix = cell1((px - xmin) / dx);
    int iy = cell1((py - ymin) / dy);
    if (ix < 1) ix = 1;
    if (iy < 1) iy = 1;
    if (ix >= nx) ix = nx - 1;
    if (iy >= ny) iy = ny - 1;
    const double dsx = px - (xmin + (ix - 1) * dx);
    const double dsy = py - (ymin + (iy - 1) * dy);
    double qv = 0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            double weight = (1.0 - fabs((i - 1) * dx - dsx) / dx) * (1.0 - fabs((j - 1) * dy - dsy) / dy);
            qv = qv + weight * gv[((long)(ix + i - 2) * ny + (iy + j - 2)) * np];
        }
    return qv;
}

// Running sums of one ray: ray_length's sum of segment lengths and
// CalGroupTime's sum of dist*2/(vhead+vtail), both in point order.
struct ray_sum {
    int n;
    double px, py, len, gt, vhead;
};

} // namespace

__global__ __launch_bounds__(64) void k_fm2d_rays(mct_fm2d_args A)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nrun * A.nrc) return;
    const long slot = t / A.nrc;
    const int r = (int)(t % A.nrc) + 1;
    const long prob = A.prob0 + slot;
    const int s = (int)(prob % A.nsrc);
    const long cp = prob / A.nsrc;          // chain*np + period
    const int p = (int)(cp % A.np);
    const int c = (int)(cp / A.np);
    const long q = prob * A.nrc + (r - 1);
    if (A.status[c] > 0) return;            // model already rejected
    const int32_t *srs = A.srs + (long)p * A.srs_pstride + (long)s * A.nrc;
    if (srs[r - 1] == 0) {                  // no ray: CalGroupTime's zero, length 0
        A.ttime[q] = 0.0;
        A.rlen[q] = 0.0;
        return;
    }
    const long wsz = (long)A.ld * A.nc;
    const double *T = A.ttn + slot * wsz;   // converged coarse times, (ix-1)*ld + iz-1
    const double *TR = A.ttr + slot * wsz;  // refined source grid copy (ttnr)
    const int32_t *NR = A.nstr + slot * wsz; // its states (nstsr): 0 = frozen
    const int nnx = A.nnx, nnz = A.nnz;
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    const double gox = A.gox, goz = A.goz;
    const double scx = A.scx[s], scz = A.scz[s];
    const double rx = A.rcx[r - 1], rz = A.rcz[r - 1];
    const double *gv = A.gvel + (long)c * A.nvx * A.nvz * A.np + p;
#define TT_(iz, ix) T[(long)((ix) - 1) * A.ld + ((iz) - 1)]
#define TR_(iz, ix) TR[(long)((ix) - 1) * A.ld + ((iz) - 1)]
#define NR_(iz, ix) NR[(long)((ix) - 1) * A.ld + ((iz) - 1)]
    // the refined source grid of this problem (modrays :292-301,395-400)
    int isx0 = (int)((scx - gox) / dnx) + 1, isz0 = (int)((scz - goz) / dnz) + 1;
    if (isx0 == nnx) isx0 = isx0 - 1;
    if (isz0 == nnz) isz0 = isz0 - 1;
    const int vnl = isx0 - A.sgs > 1 ? isx0 - A.sgs : 1;
    const int vnr = isx0 + A.sgs < nnx ? isx0 + A.sgs : nnx;
    const int vnt = isz0 - A.sgs > 1 ? isz0 - A.sgs : 1;
    const int vnb = isz0 + A.sgs < nnz ? isz0 + A.sgs : nnz;
    const int nnxr = (vnr - vnl) * A.sgdl + 1, nnzr = (vnb - vnt) * A.sgdl + 1;
    const double dnxr = A.dvx / (double)(float)(A.gdx * A.sgdl);
    const double dnzr = A.dvz / (double)(float)(A.gdz * A.sgdl);
    const double goxr = gox + dnx * (vnl - 1), gozr = goz + dnz * (vnt - 1);
    const int asgr = A.asgr;

    ray_sum R;
    R.n = 0; R.px = 0; R.py = 0; R.len = 0; R.gt = 0; R.vhead = 0;
    auto point = [&](double x, double y) {
        const double vt = ray_velocity(gv, A.np, A.nvx, A.nvz, gox, goz, A.dvx, A.dvz, x, y);
        if (R.n > 0) {
            double dist = (x - R.px) * (x - R.px) + (y - R.py) * (y - R.py);
            dist = sqrt(dist);
            R.len = R.len + dist;
            R.gt = R.gt + dist * 2 / (R.vhead + vt);
        }
        R.vhead = vt;
        R.px = x; R.py = y;
        R.n++;
    };
    auto refined_cell = [&](double x, double z, int &ipxr, int &ipzr) -> int {
        ipxr = cell1((x - goxr) / dnxr);
        ipzr = cell1((z - gozr) / dnzr);
        int ig = 1;
        if (ipxr < 1 || ipxr >= nnxr) ig = 0;
        if (ipzr < 1 || ipzr >= nnzr) ig = 0;
        if (ig == 1) {
            if (NR_(ipzr, ipxr) != 0 || NR_(ipzr + 1, ipxr) != 0) ig = 0;
            if (NR_(ipzr, ipxr + 1) != 0 || NR_(ipzr + 1, ipxr + 1) != 0) ig = 0;
        }
        return ig;
    };

    double dpl = dnx, rd1 = dnz;
    if (rd1 < dpl) dpl = rd1;
    dpl = 0.5 * dpl;
    int isx, isz;
    if (asgr == 1) {
        isx = cell1((scx - goxr) / dnxr);
        isz = cell1((scz - gozr) / dnzr);
    } else {
        isx = cell1((scx - gox) / dnx);
        isz = cell1((scz - goz) / dnz);
    }
    int ipx = cell1((rx - gox) / dnx);
    int ipz = cell1((rz - goz) / dnz);
    if (ipx < 1 || ipx >= nnx || ipz < 1 || ipz >= nnz) {   // the reference stops (:953-960)
        atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE);
        return;
    }
    double cx = rx, cz = rz;
    point(cx, cz);
    int sw = 0;
    double sred = (scx - cx) * (scx - cx);
    sred = sred + (scz - cz) * (scz - cz);
    sred = sqrt(sred);
    if (sred < 2.0 * dpl) { point(scx, scz); sw = 1; }
    int ipxr = 0, ipzr = 0, igref = 0;
    if (asgr == 1) igref = refined_cell(rx, rz, ipxr, ipzr);
    if (sw == 0) {
        if (asgr == 1) {
            if (igref == 1 && ipxr == isx && ipzr == isz) { point(scx, scz); sw = 1; }
        } else if (ipx == isx && ipz == isz) {
            point(scx, scz);
            sw = 1;
        }
    }
    const long maxrp = (long)nnx * nnz;
    bool crazy = false;
    for (long j = 1; j <= maxrp && sw == 0; j++) {
        double dtx, dtz;
        if (igref == 1) {
            if (ipxr == 1) dtx = (TR_(ipzr, ipxr + 1) - TR_(ipzr, ipxr)) / dnxr;
            else if (ipxr == nnxr) dtx = (TR_(ipzr, ipxr) - TR_(ipzr, ipxr - 1)) / dnxr;
            else dtx = (TR_(ipzr, ipxr + 1) - TR_(ipzr, ipxr - 1)) / (2.0 * dnxr);
            if (ipzr == 1) dtz = (TR_(ipzr + 1, ipxr) - TR_(ipzr, ipxr)) / dnzr;
            else if (ipzr == nnzr) dtz = (TR_(ipzr, ipxr) - TR_(ipzr - 1, ipxr)) / dnzr;
            else dtz = (TR_(ipzr + 1, ipxr) - TR_(ipzr - 1, ipxr)) / (2.0 * dnzr);
        } else {
            if (ipx == 1) dtx = (TT_(ipz, ipx + 1) - TT_(ipz, ipx)) / dnx;
            else if (ipx == nnx) dtx = (TT_(ipz, ipx) - TT_(ipz, ipx - 1)) / dnx;
            else dtx = (TT_(ipz, ipx + 1) - TT_(ipz, ipx - 1)) / (2.0 * dnx);
            // the reference tests the refined-grid row here (:1090): a point in
            // the refined grid's last row takes the one-sided difference
            if (ipz == 1) dtz = (TT_(ipz + 1, ipx) - TT_(ipz, ipx)) / dnz;
            else if (asgr == 1 && ipzr == nnzr) dtz = (TT_(ipz, ipx) - TT_(ipz - 1, ipx)) / dnz;
            else dtz = (TT_(ipz + 1, ipx) - TT_(ipz - 1, ipx)) / (2.0 * dnz);
        }
        rd1 = sqrt(dtx * dtx + dtz * dtz);
        double nx_ = cx - dpl * dtx / rd1;
        double nz_ = cz - dpl * dtz / rd1;
        if (asgr == 1) igref = refined_cell(nx_, nz_, ipxr, ipzr);
        else igref = 0;
        ipx = cell1((nx_ - gox) / dnx);
        ipz = cell1((nz_ - goz) / dnz);
        sred = (scx - nx_) * (scx - nx_);
        sred = sred + (scz - nz_) * (scz - nz_);
        sred = sqrt(sred);
        bool done = sred < 2.0 * dpl;
        if (!done) done = asgr == 1 ? (igref == 1 && ipxr == isx && ipzr == isz) : (ipx == isx && ipz == isz);
        if (done) {
            point(nx_, nz_);
            point(scx, scz);
            break;
        }
        // the ray may leave the model: it is moved onto the edge (:1185-1204)
        if (ipx < 1) { nx_ = gox; ipx = 1; }
        if (ipx >= nnx) { nx_ = gox + (nnx - 1) * dnx; ipx = nnx - 1; }
        if (ipz < 1) { nz_ = goz; ipz = 1; }
        if (ipz >= nnz) { nz_ = goz + (nnz - 1) * dnz; ipz = nnz - 1; }
        if (j == maxrp - 1) { crazy = true; break; }   // point limit (:1209-1221)
        point(nx_, nz_);
        cx = nx_; cz = nz_;
    }
#undef TT_
#undef TR_
#undef NR_
    if (crazy) {
        // the ray keeps only the receiver; the model is rejected (num_bad_ray)
        atomicCAS(&A.status[c], 0, (int)MCT_REJECT_RAY);
        A.ttime[q] = 0.0;
        A.rlen[q] = 0.0;
        return;
    }
    A.ttime[q] = R.n >= 2 ? R.gt : 0.0;
    A.rlen[q] = R.len;
}

// Straight rays (isStraight = 1): setup_straightRays' points (likelihood_surf.F90
// :412-452) generated on the fly and CalGroupTime over them (:454-489), one
// thread per (chain, period, source, receiver) of a.nchain chains.  The
// distances (srdist) are fixed by the geometry and set at create time.
__global__ __launch_bounds__(256) void k_straight_rays(mct_fm2d_args A)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long npair = (long)A.nsrc * A.nrc;
    if (t >= (long)A.nchain * A.np * npair) return;
    const long q = t % npair;
    const long cp = t / npair;
    const int p = (int)(cp % A.np);
    const int c = (int)(cp / A.np);
    if (A.status[c] > 0) return;            // model already rejected
    const int j = (int)(q / A.nrc), k = (int)(q % A.nrc);
    if (A.srs[(long)p * A.srs_pstride + q] != 1) { A.ttime[t] = 0.0; return; }
    const double *gv = A.gvel + (long)c * A.nvx * A.nvz * A.np + p;
    const double xmax = A.gox + (A.nvx - 1) * A.dvx, ymax = A.goz + (A.nvz - 1) * A.dvz;
    double dl = A.dvx < A.dvz ? A.dvx : A.dvz;
    dl = dl / 2;
    const double sx = A.scx[j], sy = A.scz[j], rx = A.rcx[k], ry = A.rcz[k];
    const double dx = rx - sx, dy = ry - sy;
    const double ds = sqrt(dx * dx + dy * dy);
    const double fn = floor(ds / dl) + 1;
    const long npoints = fn < 1.0 ? 1 : (long)fn;
    ray_sum R;
    R.n = 0; R.px = 0; R.py = 0; R.len = 0; R.gt = 0; R.vhead = 0;
    auto point = [&](double x, double y) {
        const double vt = ray_velocity(gv, A.np, A.nvx, A.nvz, A.gox, A.goz, A.dvx, A.dvz, x, y);
        if (R.n > 0) {
            double dist = (x - R.px) * (x - R.px) + (y - R.py) * (y - R.py);
            dist = sqrt(dist);
            R.gt = R.gt + dist * 2 / (R.vhead + vt);
        }
        R.vhead = vt;
        R.px = x; R.py = y;
        R.n++;
    };
    for (long l = 1; l <= npoints - 1; l++) {
        const double px = sx + (l - 1) * dl * dx / ds;
        const double py = sy + (l - 1) * dl * dy / ds;
        if (px < A.gox || px > xmax || py < A.goz || py > ymax) {   // 'Straight ray out of area'
            atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE);
            return;
        }
        point(px, py);
    }
    point(rx, ry);
    A.ttime[t] = R.n >= 2 ? R.gt : 0.0;
}

int mct_launch_straight_rays(const mct_fm2d_args &a, hipStream_t s)
{
    if (!a.gvel) return MCT_ERR_ARG;
    const long n = (long)a.nchain * a.np * a.nsrc * a.nrc;
    if (n <= 0) return MCT_OK;
    hipLaunchKernelGGL(k_straight_rays, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_fm2d_rays(const mct_fm2d_args &b, hipStream_t s)
{
    if (!b.rays) return MCT_OK;
    if (!b.gvel || !b.rlen || !b.ttr || !b.nstr) return MCT_ERR_ARG;
    const long n = b.nrun * b.nrc;
    if (n <= 0) return MCT_OK;
    hipLaunchKernelGGL(k_fm2d_rays, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, b);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}
