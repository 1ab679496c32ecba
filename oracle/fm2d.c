/*
 * fm2d.c -- oracle for the 2-D phase-traveltime solve of the surface-wave
 * likelihood.  TEST INFRASTRUCTURE ONLY (see mct_oracle.h).
 *
 * Restates the phase-time path of modrays (fm2d/fm2dray_cartesian.f90:67-480):
 * gridder (:491-600, cubic B-spline resampling of the padded control grid),
 * bsplrefine (:608-674, the refined source grid), travel (fm2d/fm2d_ttime.f90
 * :41-238: corner initialisation, refined-grid early exit, remap restart),
 * fouds1 (:248-332) / fouds2 (:342-518), the narrow-band binary tree addtree /
 * downtree / updtree (:527-680), bilinear (:690-711) and srtimes
 * (fm2dray_cartesian.f90:682-797).  For group velocities (phaseGroup = 1,
 * uar = 0) it adds rpaths (:801-1420: the gradient ray from each receiver back
 * to the source, no Frechet derivatives), the ray length (T_RAY%length,
 * :51-65) and CalGroupTime / GetVelocity (src/likelihood_surf.F90:454-521:
 * the group delay along the ray, bilinear group velocity at its points).
 *
 * Arrays are Fortran column-major: a(iz,ix) with iz fastest, 1-based in the
 * comments; the C indices below are 0-based.
 */
#include "mct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int px, pz; } bp;   /* backpointer (1-based grid indices) */

typedef struct {
    /* current grid (coarse or refined) */
    int nnx, nnz;
    double gox, goz, dnx, dnz;
    int ld;                 /* leading dimension of ttn/nsts/veln (max nnz) */
    double *ttn, *veln;
    int *nsts;
    bp *btg;
    int ntr, maxbt;
    int fom;
    /* refined-grid bookkeeping */
    int vnl, vnr, vnt, vnb;
} fm2d_state;

#define TT(S, iz, ix) (S)->ttn[((ix) - 1) * (S)->ld + ((iz) - 1)]
#define NS(S, iz, ix) (S)->nsts[((ix) - 1) * (S)->ld + ((iz) - 1)]
#define VN(S, iz, ix) (S)->veln[((ix) - 1) * (S)->ld + ((iz) - 1)]

static void addtree(fm2d_state *S, int iz, int ix)
{
    S->ntr = S->ntr + 1;
    NS(S, iz, ix) = S->ntr;
    S->btg[S->ntr].px = ix;
    S->btg[S->ntr].pz = iz;
    int tpc = S->ntr, tpp = tpc / 2;
    while (tpp > 0) {
        if (TT(S, iz, ix) < TT(S, S->btg[tpp].pz, S->btg[tpp].px)) {
            NS(S, iz, ix) = tpp;
            NS(S, S->btg[tpp].pz, S->btg[tpp].px) = tpc;
            bp ex = S->btg[tpc]; S->btg[tpc] = S->btg[tpp]; S->btg[tpp] = ex;
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
}

static void downtree(fm2d_state *S)
{
    if (S->ntr == 1) { S->ntr = S->ntr - 1; return; }
    NS(S, S->btg[S->ntr].pz, S->btg[S->ntr].px) = 1;
    S->btg[1] = S->btg[S->ntr];
    S->ntr = S->ntr - 1;
    int tpp = 1, tpc = 2 * tpp;
    while (tpc < S->ntr) {
        double rd1 = TT(S, S->btg[tpc].pz, S->btg[tpc].px);
        double rd2 = TT(S, S->btg[tpc + 1].pz, S->btg[tpc + 1].px);
        if (rd1 > rd2) tpc = tpc + 1;
        rd1 = TT(S, S->btg[tpc].pz, S->btg[tpc].px);
        rd2 = TT(S, S->btg[tpp].pz, S->btg[tpp].px);
        if (rd1 < rd2) {
            NS(S, S->btg[tpp].pz, S->btg[tpp].px) = tpc;
            NS(S, S->btg[tpc].pz, S->btg[tpc].px) = tpp;
            bp ex = S->btg[tpc]; S->btg[tpc] = S->btg[tpp]; S->btg[tpp] = ex;
            tpp = tpc;
            tpc = 2 * tpp;
        } else {
            tpc = S->ntr + 1;
        }
    }
    if (tpc == S->ntr) {
        double rd1 = TT(S, S->btg[tpc].pz, S->btg[tpc].px);
        double rd2 = TT(S, S->btg[tpp].pz, S->btg[tpp].px);
        if (rd1 < rd2) {
            NS(S, S->btg[tpp].pz, S->btg[tpp].px) = tpc;
            NS(S, S->btg[tpc].pz, S->btg[tpc].px) = tpp;
            bp ex = S->btg[tpc]; S->btg[tpc] = S->btg[tpp]; S->btg[tpp] = ex;
        }
    }
}

static void updtree(fm2d_state *S, int iz, int ix)
{
    int tpc = NS(S, iz, ix), tpp = tpc / 2;
    while (tpp > 0) {
        if (TT(S, iz, ix) < TT(S, S->btg[tpp].pz, S->btg[tpp].px)) {
            NS(S, iz, ix) = tpp;
            NS(S, S->btg[tpp].pz, S->btg[tpp].px) = tpc;
            bp ex = S->btg[tpc]; S->btg[tpc] = S->btg[tpp]; S->btg[tpp] = ex;
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
}

/* first-order upwind update (FOUDS), min over the four quadrants */
static void fouds1(fm2d_state *S, int iz, int ix)
{
    int tsw1 = 0;
    double travm = 0.0;
    double slown = 1.0 / VN(S, iz, ix);
    for (int j = ix - 1; j <= ix + 1; j += 2)
        for (int k = iz - 1; k <= iz + 1; k += 2) {
            if (!(j >= 1 && j <= S->nnx && k >= 1 && k <= S->nnz)) continue;
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0;
            if (NS(S, iz, j) == 0) {
                swsol = 1;
                if (NS(S, k, ix) == 0) {
                    double u = S->dnx, v = S->dnz;
                    double em = TT(S, k, ix) - TT(S, iz, j);
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = TT(S, iz, j);
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (S->dnx * S->dnx));
                    tref = TT(S, iz, j);
                }
            } else if (NS(S, k, ix) == 0) {
                swsol = 1;
                a = 1.0; b = 0.0;
                c = -((slown * S->dnz) * (slown * S->dnz));
                tref = TT(S, k, ix);
            }
            if (swsol == 1) {
                double rd1 = b * b - 4.0 * a * c;
                if (rd1 < 0.0) rd1 = 0.0;
                double tdsh = (-b + sqrt(rd1)) / (2.0 * a);
                double trav = tref + tdsh;
                if (tsw1 == 1) travm = fmin(trav, travm);
                else { travm = trav; tsw1 = 1; }
            }
        }
    TT(S, iz, ix) = travm;
}

/* mixed-order upwind update: second order along an axis whose two upwind
 * nodes are alive and monotone */
static void fouds2(fm2d_state *S, int iz, int ix)
{
    int tsw1 = 0;
    double travm = 0.0;
    double slown = 1.0 / VN(S, iz, ix);
    for (int j = ix - 1; j <= ix + 1; j += 2) {
        if (!(j >= 1 && j <= S->nnx)) continue;
        int swj = -1, j2;
        if (j == ix - 1) {
            j2 = j - 1;
            if (j2 >= 1 && NS(S, iz, j2) == 0) swj = 0;
        } else {
            j2 = j + 1;
            if (j2 <= S->nnx && NS(S, iz, j2) == 0) swj = 0;
        }
        if (NS(S, iz, j) == 0 && swj == 0) {
            swj = -1;
            if (TT(S, iz, j) > TT(S, iz, j2)) swj = 0;
        } else {
            swj = -1;
        }
        for (int k = iz - 1; k <= iz + 1; k += 2) {
            if (!(k >= 1 && k <= S->nnz)) continue;
            int swk = -1, k2;
            if (k == iz - 1) {
                k2 = k - 1;
                if (k2 >= 1 && NS(S, k2, ix) == 0) swk = 0;
            } else {
                k2 = k + 1;
                if (k2 <= S->nnz && NS(S, k2, ix) == 0) swk = 0;
            }
            if (NS(S, k, ix) == 0 && swk == 0) {
                swk = -1;
                if (TT(S, k, ix) > TT(S, k2, ix)) swk = 0;
            } else {
                swk = -1;
            }
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0, tdiv = 1.0, u, v, em;
            if (swj == 0) {
                swsol = 1;
                if (swk == 0) {
                    u = 2.0 * S->dnx;
                    v = 2.0 * S->dnz;
                    em = 4.0 * TT(S, iz, j) - TT(S, iz, j2) - 4.0 * TT(S, k, ix);
                    em = em + TT(S, k2, ix);
                    a = v * v + u * u;
                    b = 2.0 * em * (u * u);
                    c = (u * u) * (em * em - (slown * slown) * (v * v));
                    tref = 4.0 * TT(S, iz, j) - TT(S, iz, j2);
                    tdiv = 3.0;
                } else if (NS(S, k, ix) == 0) {
                    u = S->dnz;
                    v = 2.0 * S->dnx;
                    em = 3.0 * TT(S, k, ix) - 4.0 * TT(S, iz, j) + TT(S, iz, j2);
                    a = v * v + 9.0 * (u * u);
                    b = 6.0 * em * (u * u);
                    c = (u * u) * (em * em - (slown * slown) * (v * v));
                    tref = TT(S, k, ix);
                    tdiv = 1.0;
                } else {
                    u = 2.0 * S->dnx;
                    a = 1.0; b = 0.0;
                    c = -((u * u) * (slown * slown));
                    tref = 4.0 * TT(S, iz, j) - TT(S, iz, j2);
                    tdiv = 3.0;
                }
            } else if (NS(S, iz, j) == 0) {
                swsol = 1;
                if (swk == 0) {
                    u = S->dnx;
                    v = 2.0 * S->dnz;
                    em = 3.0 * TT(S, iz, j) - 4.0 * TT(S, k, ix) + TT(S, k2, ix);
                    a = v * v + 9.0 * (u * u);
                    b = 6.0 * em * (u * u);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = TT(S, iz, j);
                    tdiv = 1.0;
                } else if (NS(S, k, ix) == 0) {
                    u = S->dnx;
                    v = S->dnz;
                    em = TT(S, k, ix) - TT(S, iz, j);
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = TT(S, iz, j);
                    tdiv = 1.0;
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (S->dnx * S->dnx));
                    tref = TT(S, iz, j);
                    tdiv = 1.0;
                }
            } else {
                if (swk == 0) {
                    swsol = 1;
                    u = 2.0 * S->dnz;
                    a = 1.0; b = 0.0;
                    c = -((u * u) * (slown * slown));
                    tref = 4.0 * TT(S, k, ix) - TT(S, k2, ix);
                    tdiv = 3.0;
                } else if (NS(S, k, ix) == 0) {
                    swsol = 1;
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (S->dnz * S->dnz));
                    tref = TT(S, k, ix);
                    tdiv = 1.0;
                }
            }
            if (swsol == 1) {
                double rd1 = b * b - 4.0 * a * c;
                if (rd1 < 0.0) rd1 = 0.0;
                double tdsh = (-b + sqrt(rd1)) / (2.0 * a);
                double trav = (tref + tdsh) / tdiv;
                if (tsw1 == 1) travm = fmin(trav, travm);
                else { travm = trav; tsw1 = 1; }
            }
        }
    }
    TT(S, iz, ix) = travm;
}

static void fouds(fm2d_state *S, int iz, int ix)
{
    if (S->fom == 0) fouds1(S, iz, ix);
    else fouds2(S, iz, ix);
}

static double bilinear(double nv[2][2], double dsx, double dsz, double dnx, double dnz)
{
    double biv = 0.0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            double produ = (1.0 - fabs(((i - 1) * dnx - dsx) / dnx)) * (1.0 - fabs(((j - 1) * dnz - dsz) / dnz));
            biv = biv + nv[i - 1][j - 1] * produ;
        }
    return biv;
}

/* travel: FMM from a source (urg 0/1) or from a remapped narrow band (urg 2).
 * Returns 0, or -1 when the source lies outside the grid (the reference STOPs). */
static int travel(fm2d_state *S, double scx, double scz, int urg)
{
    int isx = (int)((scx - S->gox) / S->dnx) + 1;
    int isz = (int)((scz - S->goz) / S->dnz) + 1;
    if (isx < 1 || isx > S->nnx || isz < 1 || isz > S->nnz) return -1;
    if (isx == S->nnx) isx = isx - 1;
    if (isz == S->nnz) isz = isz - 1;
    S->ntr = 0;
    if (urg == 2) {
        for (int i = 1; i <= S->nnx; i++)
            for (int j = 1; j <= S->nnz; j++)
                if (NS(S, j, i) > 0) addtree(S, j, i);
    } else {
        for (int i = 1; i <= S->nnx; i++)
            for (int j = 1; j <= S->nnz; j++) NS(S, j, i) = -1;
        double vss[2][2];
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) vss[i - 1][j - 1] = VN(S, isz - 1 + j, isx - 1 + i);
        double dsx = (scx - S->gox) - (isx - 1) * S->dnx;
        double dsz = (scz - S->goz) - (isz - 1) * S->dnz;
        double vsrc = bilinear(vss, dsx, dsz, S->dnx, S->dnz);
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) {
                double ex = dsx - (i - 1) * S->dnx, ez = dsz - (j - 1) * S->dnz;
                double ds = sqrt(ex * ex + ez * ez);
                TT(S, isz - 1 + j, isx - 1 + i) = 2.0 * ds / (vss[i - 1][j - 1] + vsrc);
                addtree(S, isz - 1 + j, isx - 1 + i);
            }
    }
    while (S->ntr > 0) {
        if (urg == 1) {   /* stop when the band reaches an inner edge of the refined grid */
            int ix = S->btg[1].px, iz = S->btg[1].pz, swrg = 0;
            if (ix == 1 && S->vnl != 1) swrg = 1;
            if (ix == S->nnx && S->vnr != S->nnx) swrg = 1;
            if (iz == 1 && S->vnt != 1) swrg = 1;
            if (iz == S->nnz && S->vnb != S->nnz) swrg = 1;
            if (swrg == 1) { NS(S, iz, ix) = 0; break; }
        }
        int ix = S->btg[1].px, iz = S->btg[1].pz;
        NS(S, iz, ix) = 0;
        downtree(S);
        for (int i = ix - 1; i <= ix + 1; i += 2) {
            if (i >= 1 && i <= S->nnx) {
                if (NS(S, iz, i) == -1) { fouds(S, iz, i); addtree(S, iz, i); }
                else if (NS(S, iz, i) > 0) { fouds(S, iz, i); updtree(S, iz, i); }
            }
        }
        for (int i = iz - 1; i <= iz + 1; i += 2) {
            if (i >= 1 && i <= S->nnz) {
                if (NS(S, i, ix) == -1) { fouds(S, i, ix); addtree(S, i, ix); }
                else if (NS(S, i, ix) > 0) { fouds(S, i, ix); updtree(S, i, ix); }
            }
        }
    }
    return 0;
}

/* cubic B-spline basis at u */
static void bspl(double u, double w[4])
{
    w[0] = (1.0 - u) * (1.0 - u) * (1.0 - u) / 6.0;
    w[1] = (4.0 - 6.0 * (u * u) + 3.0 * (u * u * u)) / 6.0;
    w[2] = (1.0 + 3.0 * u + 3.0 * (u * u) - 3.0 * (u * u * u)) / 6.0;
    w[3] = (u * u * u) / 6.0;
}

/* GetVelocity (src/likelihood_surf.F90:491-521): bilinear value of the grid
 * vel(ny,nx) (C: gv[(ix-1)*ny + iy-1]) at a point, the cell clamped inside. */
static double get_velocity(const double *gv, int nx, int ny, double xmin, double ymin, double dx,
                           double dy, double px, double py)
{
    int ix = (int)floor((px - xmin) / dx) + 1;
    int iy = (int)floor((py - ymin) / dy) + 1;
    if (ix < 1) ix = 1;
    if (iy < 1) iy = 1;
    if (ix >= nx) ix = nx - 1;
    if (iy >= ny) iy = ny - 1;
    double dsx = px - (xmin + (ix - 1) * dx);
    double dsy = py - (ymin + (iy - 1) * dy);
    double qv = 0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            double weight = (1.0 - fabs((i - 1) * dx - dsx) / dx) * (1.0 - fabs((j - 1) * dy - dsy) / dy);
            qv = qv + weight * gv[(long)(ix + i - 2) * ny + (iy + j - 2)];
        }
    return qv;
}

/* The refined source grid as rpaths sees it (module variables goxr, gozr,
 * dnxr, dnzr, nnxr, nnzr and the copies ttnr / nstsr made after the refined
 * solve, fm2dray_cartesian.f90:352-353,395-400). */
typedef struct {
    int nnxr, nnzr, ld;
    double goxr, gozr, dnxr, dnzr;
    const double *ttnr;
    const int *nstsr;
} ray_refined;

/* A ray's running sums: points (optional), length (ray_length, :51-65) and
 * group delay (CalGroupTime's loop body, likelihood_surf.F90:480-489). */
typedef struct {
    int n, maxpts;
    double *pts;
    double px, py, len, gt, vhead;
    const double *gv;
    int nx, ny;
    double xmin, ymin, dx, dy;
} ray_acc;

static void ray_point(ray_acc *A, double x, double y)
{
    if (A->pts && A->n < A->maxpts) { A->pts[2 * A->n] = x; A->pts[2 * A->n + 1] = y; }
    double vt = A->gv ? get_velocity(A->gv, A->nx, A->ny, A->xmin, A->ymin, A->dx, A->dy, x, y) : 0.0;
    if (A->n > 0) {
        double dist = (x - A->px) * (x - A->px) + (y - A->py) * (y - A->py);
        dist = sqrt(dist);
        A->len = A->len + dist;
        if (A->gv) A->gt = A->gt + dist * 2 / (A->vhead + vt);
    }
    A->vhead = vt;
    A->px = x; A->py = y;
    A->n++;
}

/* rpaths for one receiver (fm2dray_cartesian.f90:937-1397 with cfd = 0).
 * S is the converged coarse grid.  Returns 0, 1 for a crazy ray (the point
 * limit nnx*nnz was reached: the ray keeps only the receiver), or -2 for a
 * receiver outside the model (the reference STOPs). */
static int rpath_one(const fm2d_state *S, const ray_refined *R, int asgr, double scx, double scz,
                     double rx, double rz, ray_acc *A)
{
    const int nnx = S->nnx, nnz = S->nnz;
    const double gox = S->gox, goz = S->goz, dnx = S->dnx, dnz = S->dnz;
    const long maxrp = (long)nnx * nnz;
    double dpl = dnx, rd1 = dnz;
    if (rd1 < dpl) dpl = rd1;
    dpl = 0.5 * dpl;
    int isx, isz;
    if (asgr == 1) {
        isx = (int)floor((scx - R->goxr) / R->dnxr) + 1;
        isz = (int)floor((scz - R->gozr) / R->dnzr) + 1;
    } else {
        isx = (int)floor((scx - gox) / dnx) + 1;
        isz = (int)floor((scz - goz) / dnz) + 1;
    }
#define TR(iz, ix) R->ttnr[(long)((ix) - 1) * R->ld + ((iz) - 1)]
#define NR(iz, ix) R->nstsr[(long)((ix) - 1) * R->ld + ((iz) - 1)]
    int ipx = (int)floor((rx - gox) / dnx) + 1;
    int ipz = (int)floor((rz - goz) / dnz) + 1;
    if (ipx < 1 || ipx >= nnx || ipz < 1 || ipz >= nnz) return -2;
    double cx = rx, cz = rz;
    ray_point(A, cx, cz);
    int sw = 0;
    double sred = (scx - cx) * (scx - cx);
    sred = sred + (scz - cz) * (scz - cz);
    sred = sqrt(sred);
    if (sred < 2.0 * dpl) { ray_point(A, scx, scz); sw = 1; }
    int ipxr = 0, ipzr = 0, igref = 0;
    if (asgr == 1) {
        ipxr = (int)floor((rx - R->goxr) / R->dnxr) + 1;
        ipzr = (int)floor((rz - R->gozr) / R->dnzr) + 1;
        igref = 1;
        if (ipxr < 1 || ipxr >= R->nnxr) igref = 0;
        if (ipzr < 1 || ipzr >= R->nnzr) igref = 0;
        if (igref == 1) {
            if (NR(ipzr, ipxr) != 0 || NR(ipzr + 1, ipxr) != 0) igref = 0;
            if (NR(ipzr, ipxr + 1) != 0 || NR(ipzr + 1, ipxr + 1) != 0) igref = 0;
        }
    }
    if (sw == 0) {
        if (asgr == 1) {
            if (igref == 1 && ipxr == isx && ipzr == isz) { ray_point(A, scx, scz); sw = 1; }
        } else if (ipx == isx && ipz == isz) {
            ray_point(A, scx, scz);
            sw = 1;
        }
    }
    for (long j = 1; j <= maxrp; j++) {
        if (sw == 1) break;
        double dtx, dtz;
        if (igref == 1) {
            if (ipxr == 1) dtx = (TR(ipzr, ipxr + 1) - TR(ipzr, ipxr)) / R->dnxr;
            else if (ipxr == R->nnxr) dtx = (TR(ipzr, ipxr) - TR(ipzr, ipxr - 1)) / R->dnxr;
            else dtx = (TR(ipzr, ipxr + 1) - TR(ipzr, ipxr - 1)) / (2.0 * R->dnxr);
            if (ipzr == 1) dtz = (TR(ipzr + 1, ipxr) - TR(ipzr, ipxr)) / R->dnzr;
            else if (ipzr == R->nnzr) dtz = (TR(ipzr, ipxr) - TR(ipzr - 1, ipxr)) / R->dnzr;
            else dtz = (TR(ipzr + 1, ipxr) - TR(ipzr - 1, ipxr)) / (2.0 * R->dnzr);
        } else {
            if (ipx == 1) dtx = (TT(S, ipz, ipx + 1) - TT(S, ipz, ipx)) / dnx;
            else if (ipx == nnx) dtx = (TT(S, ipz, ipx) - TT(S, ipz, ipx - 1)) / dnx;
            else dtx = (TT(S, ipz, ipx + 1) - TT(S, ipz, ipx - 1)) / (2.0 * dnx);
            /* the reference tests the refined-grid index here (:1090): a point in
             * the refined grid's last row takes the one-sided difference */
            if (ipz == 1) dtz = (TT(S, ipz + 1, ipx) - TT(S, ipz, ipx)) / dnz;
            else if (asgr == 1 && ipzr == R->nnzr) dtz = (TT(S, ipz, ipx) - TT(S, ipz - 1, ipx)) / dnz;
            else dtz = (TT(S, ipz + 1, ipx) - TT(S, ipz - 1, ipx)) / (2.0 * dnz);
        }
        rd1 = sqrt(dtx * dtx + dtz * dtz);
        double nx_ = cx - dpl * dtx / rd1;
        double nz_ = cz - dpl * dtz / rd1;
        if (asgr == 1) {
            ipxr = (int)floor((nx_ - R->goxr) / R->dnxr) + 1;
            ipzr = (int)floor((nz_ - R->gozr) / R->dnzr) + 1;
            igref = 1;
            if (ipxr < 1 || ipxr >= R->nnxr) igref = 0;
            if (ipzr < 1 || ipzr >= R->nnzr) igref = 0;
            if (igref == 1) {
                if (NR(ipzr, ipxr) != 0 || NR(ipzr + 1, ipxr) != 0) igref = 0;
                if (NR(ipzr, ipxr + 1) != 0 || NR(ipzr + 1, ipxr + 1) != 0) igref = 0;
            }
        } else {
            igref = 0;
        }
        ipx = (int)floor((nx_ - gox) / dnx) + 1;
        ipz = (int)floor((nz_ - goz) / dnz) + 1;
        sred = (scx - nx_) * (scx - nx_);
        sred = sred + (scz - nz_) * (scz - nz_);
        sred = sqrt(sred);
        if (sred < 2.0 * dpl) { ray_point(A, nx_, nz_); ray_point(A, scx, scz); return 0; }
        if (asgr == 1) {
            if (igref == 1 && ipxr == isx && ipzr == isz) { ray_point(A, nx_, nz_); ray_point(A, scx, scz); return 0; }
        } else if (ipx == isx && ipz == isz) {
            ray_point(A, nx_, nz_);
            ray_point(A, scx, scz);
            return 0;
        }
        if (ipx < 1) { nx_ = gox; ipx = 1; }
        if (ipx >= nnx) { nx_ = gox + (nnx - 1) * dnx; ipx = nnx - 1; }
        if (ipz < 1) { nz_ = goz; ipz = 1; }
        if (ipz >= nnz) { nz_ = goz + (nnz - 1) * dnz; ipz = nnz - 1; }
        if (j == maxrp - 1) return 1;
        ray_point(A, nx_, nz_);
        cx = nx_; cz = nz_;
    }
#undef TR
#undef NR
    return 0;
}

int orc_fm2d(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
             const double *velv, int gdx, int gdz, int asgr, int sgdl, int sgs, int fom,
             double snb, int nsrc, const double *scx, const double *scz, int nrc,
             const double *rcx, const double *rcz, const int32_t *srs, double *ttime,
             double *tfield)
{
    return orc_fm2d_rays(nvx, nvz, goxd, gozd, dvxd, dvzd, velv, gdx, gdz, asgr, sgdl, sgs, fom, snb,
                         nsrc, scx, scz, nrc, rcx, rcz, srs, ttime, tfield, 0, NULL, NULL, NULL, NULL,
                         0, NULL, NULL);
}

int orc_fm2d_rays(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
                  const double *velv, int gdx, int gdz, int asgr, int sgdl, int sgs, int fom,
                  double snb, int nsrc, const double *scx, const double *scz, int nrc,
                  const double *rcx, const double *rcz, const int32_t *srs, double *ttime,
                  double *tfield, int rays, const double *gvel, double *gtime, double *rlen,
                  int32_t *crazy, int maxpts, int32_t *npts, double *pts)
{
    /* gridder: velv(0:nvz+1, 0:nvx+1) is the padded control grid */
    const int ldv = nvz + 2;
#define VV(i, j) velv[(j) * ldv + (i)]           /* velv(i,j), i in 0..nvz+1 */
    int nnx = (nvx - 1) * gdx + 1, nnz = (nvz - 1) * gdz + 1;
    double dnx = dvxd / gdx, dnz = dvzd / gdz;
    double gox = goxd, goz = gozd;
    double *velnc = (double *)malloc(sizeof(double) * nnx * nnz);
    {
        double ui[64][4], vi[64][4];
        for (int i = 1; i <= gdx + 1; i++) { double u = gdx; u = (i - 1) / u; bspl(u, ui[i - 1]); }
        for (int i = 1; i <= gdz + 1; i++) { double u = gdz; u = (i - 1) / u; bspl(u, vi[i - 1]); }
        for (int i = 1; i <= nvz - 1; i++) {
            int conz = (i == nvz - 1) ? gdz + 1 : gdz;
            for (int j = 1; j <= nvx - 1; j++) {
                int conx = (j == nvx - 1) ? gdx + 1 : gdx;
                for (int l = 1; l <= conz; l++) {
                    int stz = gdz * (i - 1) + l;
                    for (int m = 1; m <= conx; m++) {
                        int stx = gdx * (j - 1) + m;
                        double sumi = 0.0;
                        for (int i1 = 1; i1 <= 4; i1++) {
                            double sumj = 0.0;
                            for (int j1 = 1; j1 <= 4; j1++) sumj = sumj + ui[m - 1][j1 - 1] * VV(i - 2 + i1, j - 2 + j1);
                            sumi = sumi + vi[l - 1][i1 - 1] * sumj;
                        }
                        velnc[(stx - 1) * nnz + (stz - 1)] = sumi;
                    }
                }
            }
        }
    }
    /* working arrays sized for the larger of the coarse and refined grids */
    int sgsz = 2 * sgs * sgdl + 1;
    int ldm = nnz > sgsz ? nnz : sgsz, ncm = nnx > sgsz ? nnx : sgsz;
    fm2d_state S;
    memset(&S, 0, sizeof S);
    S.ld = ldm;
    S.ttn = (double *)calloc((size_t)ldm * ncm, sizeof(double));
    S.veln = (double *)calloc((size_t)ldm * ncm, sizeof(double));
    S.nsts = (int *)calloc((size_t)ldm * ncm, sizeof(int));
    double *ttnr = (double *)calloc((size_t)ldm * ncm, sizeof(double));
    int *nstsr = (int *)calloc((size_t)ldm * ncm, sizeof(int));
    S.maxbt = ncm * ldm + 8;   /* the reference sizes the tree NINT(snb*nnx*nnz); never exceeded */
    S.btg = (bp *)calloc((size_t)S.maxbt + 1, sizeof(bp));
    S.fom = fom;
    int rc = 0;
    ray_refined RR;
    memset(&RR, 0, sizeof RR);
    RR.ld = ldm; RR.ttnr = ttnr; RR.nstsr = nstsr;
    if (rays) {
        /* CalGroupTime starts from zero times; rays not traced keep no points */
        for (long q = 0; q < (long)nsrc * nrc; q++) {
            if (gtime) gtime[q] = 0.0;
            if (rlen) rlen[q] = 0.0;
            if (npts) npts[q] = 0;
        }
        if (crazy) *crazy = 0;
    }
    for (int i = 1; i <= nsrc; i++) {
        int any = 0;
        for (int j = 0; j < nrc; j++) any += srs[(i - 1) * nrc + j];
        if (any == 0 && i != 1) continue;
        double x = scx[i - 1], z = scz[i - 1];
        /* coarse grid */
        S.nnx = nnx; S.nnz = nnz; S.gox = gox; S.goz = goz; S.dnx = dnx; S.dnz = dnz;
        for (int a = 1; a <= nnx; a++)
            for (int b = 1; b <= nnz; b++) VN(&S, b, a) = velnc[(a - 1) * nnz + (b - 1)];
        int isx = (int)((x - gox) / dnx) + 1, isz = (int)((z - goz) / dnz) + 1;
        if (isx < 1 || isx > nnx || isz < 1 || isz > nnz) { rc = -1; break; }
        if (isx == nnx) isx = isx - 1;
        if (isz == nnz) isz = isz - 1;
        int vnl = isx - sgs, vnr = isx + sgs, vnt = isz - sgs, vnb = isz + sgs;
        if (vnl < 1) vnl = 1;
        if (vnr > nnx) vnr = nnx;
        if (vnt < 1) vnt = 1;
        if (vnb > nnz) vnb = nnz;
        if (asgr == 1) {
            /* refined source grid */
            int nrnx = (vnr - vnl) * sgdl + 1, nrnz = (vnb - vnt) * sgdl + 1;
            double drnx = dvxd / (double)(float)(gdx * sgdl);
            double drnz = dvzd / (double)(float)(gdz * sgdl);
            double gorx = gox + dnx * (vnl - 1), gorz = goz + dnz * (vnt - 1);
            S.nnx = nrnx; S.nnz = nrnz; S.dnx = drnx; S.dnz = drnz; S.gox = gorx; S.goz = gorz;
            S.vnl = vnl; S.vnr = vnr; S.vnt = vnt; S.vnb = vnb;
            /* bsplrefine */
            {
                int nrxr = gdx * sgdl, nrzr = gdz * sgdl;
                double uiw[64][4], viw[64][4];
                for (int jj = 1; jj <= nrxr + 1; jj++) { double u = nrxr; u = (jj - 1) / u; bspl(u, uiw[jj - 1]); }
                for (int ii = 1; ii <= nrzr + 1; ii++) { double v = nrzr; v = (ii - 1) / v; bspl(v, viw[ii - 1]); }
                int origx = (vnl - 1) * sgdl + 1, origz = (vnt - 1) * sgdl + 1;
                for (int ii = 1; ii <= nvz - 1; ii++) {
                    int conz = (ii == nvz - 1) ? nrzr + 1 : nrzr;
                    for (int jj = 1; jj <= nvx - 1; jj++) {
                        int conx = (jj == nvx - 1) ? nrxr + 1 : nrxr;
                        for (int k = 1; k <= conz; k++) {
                            int st1 = gdz * (ii - 1) + (k - 1) / sgdl + 1;
                            if (st1 < vnt || st1 > vnb) continue;
                            st1 = nrzr * (ii - 1) + k;
                            for (int l = 1; l <= conx; l++) {
                                int st2 = gdx * (jj - 1) + (l - 1) / sgdl + 1;
                                if (st2 < vnl || st2 > vnr) continue;
                                st2 = nrxr * (jj - 1) + l;
                                double sum[4];
                                for (int i1 = 1; i1 <= 4; i1++) {
                                    sum[i1 - 1] = 0.0;
                                    for (int j1 = 1; j1 <= 4; j1++)
                                        sum[i1 - 1] = sum[i1 - 1] + uiw[l - 1][j1 - 1] * VV(ii - 2 + i1, jj - 2 + j1);
                                    sum[i1 - 1] = viw[k - 1][i1 - 1] * sum[i1 - 1];
                                }
                                int idm1 = st1 - origz + 1, idm2 = st2 - origx + 1;
                                if (idm1 < 1 || idm1 > S.nnz || idm2 < 1 || idm2 > S.nnx) continue;
                                VN(&S, idm1, idm2) = sum[0] + sum[1] + sum[2] + sum[3];
                            }
                        }
                    }
                }
            }
            if (travel(&S, x, z, 1)) { rc = -1; break; }
            RR.nnxr = S.nnx; RR.nnzr = S.nnz; RR.goxr = S.gox; RR.gozr = S.goz; RR.dnxr = S.dnx; RR.dnzr = S.dnz;
            /* remap the refined grid onto the coarse grid */
            memcpy(ttnr, S.ttn, sizeof(double) * (size_t)ldm * ncm);
            memcpy(nstsr, S.nsts, sizeof(int) * (size_t)ldm * ncm);
            for (int a = 0; a < ldm * ncm; a++) S.nsts[a] = -1;
            for (int k = 1; k <= S.nnz; k += sgdl) {
                int idm1 = vnt + (k - 1) / sgdl;
                for (int l = 1; l <= S.nnx; l += sgdl) {
                    int idm2 = vnl + (l - 1) / sgdl;
                    int st = nstsr[(l - 1) * ldm + (k - 1)];
                    S.nsts[(idm2 - 1) * ldm + (idm1 - 1)] = st;
                    if (st >= 0) S.ttn[(idm2 - 1) * ldm + (idm1 - 1)] = ttnr[(l - 1) * ldm + (k - 1)];
                }
            }
            S.nnx = nnx; S.nnz = nnz; S.dnx = dnx; S.dnz = dnz; S.gox = gox; S.goz = goz;
            for (int a = 1; a <= nnx; a++)
                for (int b = 1; b <= nnz; b++) VN(&S, b, a) = velnc[(a - 1) * nnz + (b - 1)];
            /* complete the narrow band: alive nodes next to far nodes become close */
            for (int k = 1; k <= nnx; k++)
                for (int l = 1; l <= nnz; l++) {
                    if (NS(&S, l, k) != 0) continue;
                    if (l - 1 >= 1 && NS(&S, l - 1, k) == -1) NS(&S, l, k) = 1;
                    if (l + 1 <= nnz && NS(&S, l + 1, k) == -1) NS(&S, l, k) = 1;
                    if (k - 1 >= 1 && NS(&S, l, k - 1) == -1) NS(&S, l, k) = 1;
                    if (k + 1 <= nnx && NS(&S, l, k + 1) == -1) NS(&S, l, k) = 1;
                }
            if (travel(&S, x, z, 2)) { rc = -1; break; }
        } else {
            if (travel(&S, x, z, 0)) { rc = -1; break; }
        }
        if (tfield)
            for (int a = 1; a <= nnx; a++)
                for (int b = 1; b <= nnz; b++)
                    tfield[((size_t)(i - 1) * nnx + (a - 1)) * nnz + (b - 1)] = TT(&S, b, a);
        /* srtimes */
        for (int r = 1; r <= nrc; r++) {
            if (srs[(i - 1) * nrc + (r - 1)] == 0) continue;
            int irx = (int)floor((rcx[r - 1] - gox) / dnx) + 1;
            int irz = (int)floor((rcz[r - 1] - goz) / dnz) + 1;
            int sw = 0;
            if (irx < 1 || irx > nnx || irz < 1 || irz > nnz) { rc = -2; break; }
            if (irx == nnx) irx = irx - 1;
            if (irz == nnz) irz = irz - 1;
            int jsx = (int)floor((x - gox) / dnx) + 1, jsz = (int)floor((z - goz) / dnz) + 1;
            double dpl = dnx, rd1 = dnz;
            if (rd1 < dpl) dpl = rd1;
            double sred = (x - rcx[r - 1]) * (x - rcx[r - 1]);
            sred = sred + (z - rcz[r - 1]) * (z - rcz[r - 1]);
            sred = sqrt(sred);
            if (sred < dpl) sw = 1;
            if (jsx == irx && jsz == irz) sw = 1;
            double trr;
            if (sw == 1) {
                double vss[2][2];
                /* the reference reads veln(isz-1+l, isx-1+k) with floor() indices and
                 * no clamp; a source on the last grid line, or just below the origin
                 * (accepted by travel's truncation), would read outside the grid, so
                 * the cell index is clamped here (no effect anywhere else) */
                int csx = jsx < nnx ? jsx : nnx - 1, csz = jsz < nnz ? jsz : nnz - 1;
                csx = csx < 1 ? 1 : csx;
                csz = csz < 1 ? 1 : csz;
                for (int k = 1; k <= 2; k++)
                    for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = VN(&S, csz - 1 + l, csx - 1 + k);
                double drx = (x - gox) - (jsx - 1) * dnx, drz = (z - goz) - (jsz - 1) * dnz;
                double vels = bilinear(vss, drx, drz, dnx, dnz);
                for (int k = 1; k <= 2; k++)
                    for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = VN(&S, irz - 1 + l, irx - 1 + k);
                drx = (rcx[r - 1] - gox) - (irx - 1) * dnx;
                drz = (rcz[r - 1] - goz) - (irz - 1) * dnz;
                double velr = bilinear(vss, drx, drz, dnx, dnz);
                trr = 2.0 * sred / (vels + velr);
            } else {
                double drx = (rcx[r - 1] - gox) - (irx - 1) * dnx;
                double drz = (rcz[r - 1] - goz) - (irz - 1) * dnz;
                trr = 0.0;
                for (int k = 1; k <= 2; k++)
                    for (int l = 1; l <= 2; l++) {
                        double produ = (1.0 - fabs(((l - 1) * dnz - drz) / dnz)) * (1.0 - fabs(((k - 1) * dnx - drx) / dnx));
                        trr = trr + TT(&S, irz - 1 + l, irx - 1 + k) * produ;
                    }
            }
            ttime[(size_t)(i - 1) * nrc + (r - 1)] = trr;
        }
        if (rc) break;
        if (rays) {
            /* rpaths from every receiver of a used pair back to source i */
            for (int r = 1; r <= nrc && !rc; r++) {
                long q = (long)(i - 1) * nrc + (r - 1);
                if (srs[q] == 0) continue;
                ray_acc A;
                memset(&A, 0, sizeof A);
                A.maxpts = maxpts;
                A.pts = pts ? pts + 2 * (size_t)maxpts * q : NULL;
                A.gv = gvel; A.nx = nvx; A.ny = nvz; A.xmin = goxd; A.ymin = gozd; A.dx = dvxd; A.dy = dvzd;
                int e = rpath_one(&S, &RR, asgr, x, z, rcx[r - 1], rcz[r - 1], &A);
                if (e < 0) { rc = e; break; }
                if (e == 1) {
                    if (crazy) (*crazy)++;
                    A.n = 1; A.len = 0.0; A.gt = 0.0;
                }
                if (npts) npts[q] = A.n;
                if (rlen) rlen[q] = A.len;
                if (gtime) gtime[q] = A.n >= 2 ? A.gt : 0.0;
            }
            if (rc) break;
        }
    }
#undef VV
    free(velnc); free(S.ttn); free(S.veln); free(S.nsts); free(ttnr); free(nstsr); free(S.btg);
    return rc;
}

/* Straight rays (isStraight = 1): setup_straightRays (src/likelihood_surf.F90
 * :412-452) and CalGroupTime over them (:454-489).  dist (nsrc,nrc) = the
 * source-receiver distance for every pair; for a used pair the points are the
 * source + (l-1)*dl along the segment, l = 1..floor(dist/dl), then the
 * receiver, dl = min(dx,dy)/2; gtime (nsrc,nrc) the group delay along them
 * (0 for unused pairs).  Returns -1 when a point leaves [xmin,xmax]x[ymin,ymax]
 * (the reference stops: 'Straight ray out of area'). */
int orc_straight_rays(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd, const double *gvel,
                      int nsrc, const double *scx, const double *scz, int nrc, const double *rcx,
                      const double *rcz, const int32_t *srs, double *gtime, double *dist)
{
    const double xmax = goxd + (nvx - 1) * dvxd, ymax = gozd + (nvz - 1) * dvzd;
    double dl = dvxd < dvzd ? dvxd : dvzd;
    dl = dl / 2;
    for (int j = 1; j <= nsrc; j++)
        for (int k = 1; k <= nrc; k++) {
            const long q = (long)(j - 1) * nrc + (k - 1);
            double dx = rcx[k - 1] - scx[j - 1];
            double dy = rcz[k - 1] - scz[j - 1];
            double ds = sqrt(dx * dx + dy * dy);
            dist[q] = ds;
            gtime[q] = 0.0;
            if (srs[q] != 1) continue;
            long npoints = (long)floor(ds / dl) + 1;
            ray_acc A;
            memset(&A, 0, sizeof A);
            A.gv = gvel; A.nx = nvx; A.ny = nvz; A.xmin = goxd; A.ymin = gozd; A.dx = dvxd; A.dy = dvzd;
            for (long l = 1; l <= npoints - 1; l++) {
                double px = scx[j - 1] + (l - 1) * dl * dx / ds;
                double py = scz[j - 1] + (l - 1) * dl * dy / ds;
                if (px < goxd || px > xmax || py < gozd || py > ymax) return -1;
                ray_point(&A, px, py);
            }
            ray_point(&A, rcx[k - 1], rcz[k - 1]);
            gtime[q] = A.n >= 2 ? A.gt : 0.0;
        }
    return 0;
}
