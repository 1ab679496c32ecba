// fm2d.hip -- 2-D phase traveltimes of the surface-wave likelihood on gfx950.
//
// Replaces modrays' phase-time path (fm2d/fm2dray_cartesian.f90:67-480):
// gridder (:491-600), bsplrefine (:608-674), travel (fm2d/fm2d_ttime.f90:41-238)
// with fouds1 / fouds2 (:248-518) and its binary-tree narrow band (:527-680),
// and srtimes (fm2dray_cartesian.f90:682-797).
//
// The reference's answer depends on the order in which its heap freezes nodes:
// the refined source grid stops at the first popped node on an inner edge
// (fm2d_ttime.f90:143-163), the coarse solve restarts from the remapped band,
// and the mixed-order stencil only uses nodes already frozen.  A parallel
// relaxation cannot reproduce that order, so each (chain, period, source)
// problem runs the reference's own sequential algorithm in one lane: 64
// independent problems per wave, each with its grid, state and heap in a
// private slice of HBM.  Results are bit-identical to the sequential code.
// The coarse B-spline velocity of a period is computed once by a separate
// kernel (one thread per node) and shared by all sources of that period.
#include "mct_internal.h"
#include <stdlib.h>
#include <string.h>

#pragma clang fp contract(off)


// ---- coarse B-spline velocity ---------------------------------------------------
__device__ __forceinline__ void bspl_w(double u, double w[4])
{
    w[0] = (1.0 - u) * (1.0 - u) * (1.0 - u) / 6.0;
    w[1] = (4.0 - 6.0 * (u * u) + 3.0 * (u * u * u)) / 6.0;
    w[2] = (1.0 + 3.0 * u + 3.0 * (u * u) - 3.0 * (u * u * u)) / 6.0;
    w[3] = (u * u * u) / 6.0;
}

// gridder: veln(stz,stx) = sum_i1 vi(l,i1) * sum_j1 ui(m,j1) velv(i-2+i1, j-2+j1)
__global__ void k_fm2d_gridder(mct_fm2d_args A)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nn = (long)A.nnx * A.nnz;
    if (t >= nn * A.nchain * A.np) return;
    const int cp = (int)(t / nn);
    const int node = (int)(t - (long)cp * nn);
    const int stx = node / A.nnz + 1, stz = node % A.nnz + 1;
    // the control cell (i,j) and sub-node (l,m) that produce (stz,stx); the last
    // cell of each axis also owns its far edge node
    int i = (stz - 1) / A.gdz + 1, l = (stz - 1) % A.gdz + 1;
    if (i > A.nvz - 1) { i = A.nvz - 1; l = A.gdz + 1; }
    int j = (stx - 1) / A.gdx + 1, m = (stx - 1) % A.gdx + 1;
    if (j > A.nvx - 1) { j = A.nvx - 1; m = A.gdx + 1; }
    double u = A.gdx; u = (m - 1) / u;
    double v = A.gdz; v = (l - 1) / v;
    double ui[4], vi[4];
    bspl_w(u, ui);
    bspl_w(v, vi);
    const int ldv = A.nvz + 2;
    const double *vv = A.velv + (long)cp * ldv * (A.nvx + 2);
    double sumi = 0.0;
    for (int i1 = 1; i1 <= 4; i1++) {
        double sumj = 0.0;
        for (int j1 = 1; j1 <= 4; j1++) sumj = sumj + ui[j1 - 1] * vv[(j - 2 + j1) * ldv + (i - 2 + i1)];
        sumi = sumi + vi[i1 - 1] * sumj;
    }
    A.veln[(long)cp * nn + (long)(stx - 1) * A.nnz + (stz - 1)] = sumi;
    // the slowness fouds takes (slown = 1.0 / veln), once per node instead of per update
    if (A.sln) A.sln[(long)cp * nn + (long)(stx - 1) * A.nnz + (stz - 1)] = 1.0 / sumi;
}

// ---- per-lane sequential solver ----------------------------------------------------
// One lane runs one problem's heap loop in global memory, so a pop is a chain
// of dependent loads.  Two things keep that chain short: the heap holds each
// entry's key beside its node (hk[p] == the node's time for every position p,
// so a sift level is one load, not two), and a pop reads its four neighbours'
// stencils before it moves the heap, all at once.  The heap operations are the
// reference's (same comparisons, same ties, same final arrangement: a moving
// hole instead of pairwise swaps).
struct lane_grid {
    int nnx, nnz, ld;
    double gox, goz, dnx, dnz;
    double *t;              // times
    int32_t *s;             // state: -1 far, 0 alive, >0 heap position
    const double *v;        // velocity
    int vld;                // leading dimension of v
    int32_t *btg;           // heap, 1-based
    double *hk;             // heap keys, hk[p] = time of btg[p]
    int ntr, fom;
    int vnl, vnr, vnt, vnb;
};

#define LT(G, iz, ix) (G).t[((ix) - 1) * (G).ld + ((iz) - 1)]
#define LS(G, iz, ix) (G).s[((ix) - 1) * (G).ld + ((iz) - 1)]
#define LV(G, iz, ix) (G).v[((ix) - 1) * (G).vld + ((iz) - 1)]
#define BPX(e) ((e) >> 16)
#define BPZ(e) ((e) & 0xffff)

// addtree (fm2d_ttime.f90:527-564) for node (iz,ix) with time tv
__device__ static void l_addtree(lane_grid &G, int iz, int ix, double tv)
{
    G.ntr = G.ntr + 1;
    int tpc = G.ntr, tpp = tpc / 2;
    while (tpp > 0) {
        const int ep = G.btg[tpp];
        const double kp = G.hk[tpp];
        if (tv < kp) {           // the parent moves down into the hole
            LS(G, BPZ(ep), BPX(ep)) = tpc;
            G.btg[tpc] = ep;
            G.hk[tpc] = kp;
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
    LS(G, iz, ix) = tpc;
    G.btg[tpc] = (ix << 16) | iz;
    G.hk[tpc] = tv;
}

// downtree (fm2d_ttime.f90:575-644): remove the root
__device__ static void l_downtree(lane_grid &G)
{
    if (G.ntr == 1) { G.ntr = G.ntr - 1; return; }
    const int el = G.btg[G.ntr];
    const double kl = G.hk[G.ntr];
    G.ntr = G.ntr - 1;
    int tpp = 1, tpc = 2 * tpp;
    while (tpc < G.ntr) {
        const int e1 = G.btg[tpc], e2 = G.btg[tpc + 1];
        const double k1 = G.hk[tpc], k2 = G.hk[tpc + 1];
        int ec = e1;
        double kc = k1;
        if (k1 > k2) { tpc = tpc + 1; ec = e2; kc = k2; }
        if (kc < kl) {           // the smaller child moves up into the hole
            LS(G, BPZ(ec), BPX(ec)) = tpp;
            G.btg[tpp] = ec;
            G.hk[tpp] = kc;
            tpp = tpc;
            tpc = 2 * tpp;
        } else {
            tpc = G.ntr + 1;
        }
    }
    if (tpc == G.ntr) {
        const int ec = G.btg[tpc];
        const double kc = G.hk[tpc];
        if (kc < kl) {
            LS(G, BPZ(ec), BPX(ec)) = tpp;
            G.btg[tpp] = ec;
            G.hk[tpp] = kc;
            tpp = tpc;
        }
    }
    LS(G, BPZ(el), BPX(el)) = tpp;
    G.btg[tpp] = el;
    G.hk[tpp] = kl;
}

// updtree (fm2d_ttime.f90:653-680): node (iz,ix), already in the heap, has the
// new time tv; it moves up while smaller than its parent
__device__ static void l_updtree(lane_grid &G, int iz, int ix, double tv)
{
    int tpc = LS(G, iz, ix), tpp = tpc / 2;
    while (tpp > 0) {
        const int ep = G.btg[tpp];
        const double kp = G.hk[tpp];
        if (tv < kp) {
            LS(G, BPZ(ep), BPX(ep)) = tpc;
            G.btg[tpc] = ep;
            G.hk[tpc] = kp;
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
    LS(G, iz, ix) = tpc;
    G.btg[tpc] = (ix << 16) | iz;
    G.hk[tpc] = tv;
}

// A node's axis stencil read ahead: times one and two nodes away on each side
// ([0] the x-1 / z-1 side), whether each is inside the grid and alive, and
// the node's velocity.  Out-of-grid entries read the node itself.
struct l_sten {
    double tx1[2], tx2[2], tz1[2], tz2[2], v;
    bool ix1[2], iz1[2];
    bool ax1[2], ax2[2], az1[2], az2[2];
};

__device__ __forceinline__ void l_load_sten(const lane_grid &G, int iz, int ix, l_sten &S)
{
    int sx1[2], sx2[2], sz1[2], sz2[2];
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
        const int d = sd ? 1 : -1;
        const int x1 = ix + d, x2 = ix + 2 * d, z1 = iz + d, z2 = iz + 2 * d;
        S.ix1[sd] = x1 >= 1 && x1 <= G.nnx;
        S.iz1[sd] = z1 >= 1 && z1 <= G.nnz;
        const bool ix2 = G.fom != 0 && x2 >= 1 && x2 <= G.nnx, iz2 = G.fom != 0 && z2 >= 1 && z2 <= G.nnz;
        const int cx1 = S.ix1[sd] ? x1 : ix, cz1 = S.iz1[sd] ? z1 : iz;
        const int cx2 = ix2 ? x2 : ix, cz2 = iz2 ? z2 : iz;
        S.tx1[sd] = LT(G, iz, cx1);
        S.tz1[sd] = LT(G, cz1, ix);
        sx1[sd] = LS(G, iz, cx1);
        sz1[sd] = LS(G, cz1, ix);
        S.tx2[sd] = G.fom ? LT(G, iz, cx2) : 0.0;
        S.tz2[sd] = G.fom ? LT(G, cz2, ix) : 0.0;
        sx2[sd] = G.fom ? LS(G, iz, cx2) : 1;
        sz2[sd] = G.fom ? LS(G, cz2, ix) : 1;
        S.ax1[sd] = S.ix1[sd] && sx1[sd] == 0;
        S.az1[sd] = S.iz1[sd] && sz1[sd] == 0;
        S.ax2[sd] = ix2 && sx2[sd] == 0;
        S.az2[sd] = iz2 && sz2[sd] == 0;
    }
    S.v = LV(G, iz, ix);
}

// fouds1 (fm2d_ttime.f90:248-332) over a read-ahead stencil
__device__ static double l_fouds1s(const l_sten &S, double dnx, double dnz)
{
    int tsw1 = 0;
    double travm = 0.0;
    const double slown = 1.0 / S.v;
    for (int sj = 0; sj < 2; sj++)
        for (int sk = 0; sk < 2; sk++) {
            if (!(S.ix1[sj] && S.iz1[sk])) continue;
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0;
            if (S.ax1[sj]) {
                swsol = 1;
                if (S.az1[sk]) {
                    double u = dnx, v = dnz;
                    double em = S.tz1[sk] - S.tx1[sj];
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = S.tx1[sj];
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (dnx * dnx));
                    tref = S.tx1[sj];
                }
            } else if (S.az1[sk]) {
                swsol = 1;
                a = 1.0; b = 0.0;
                c = -((slown * dnz) * (slown * dnz));
                tref = S.tz1[sk];
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
    return travm;
}

// fouds2 (fm2d_ttime.f90:342-518) over a read-ahead stencil
__device__ static double l_fouds2s(const l_sten &S, double dnx, double dnz)
{
    int tsw1 = 0;
    double travm = 0.0;
    const double slown = 1.0 / S.v;
    for (int sj = 0; sj < 2; sj++) {
        if (!S.ix1[sj]) continue;
        const double tj = S.tx1[sj], tj2 = S.tx2[sj];
        int swj = S.ax2[sj] ? 0 : -1;
        if (S.ax1[sj] && swj == 0) {
            swj = -1;
            if (tj > tj2) swj = 0;
        } else {
            swj = -1;
        }
        for (int sk = 0; sk < 2; sk++) {
            if (!S.iz1[sk]) continue;
            const double tk = S.tz1[sk], tk2 = S.tz2[sk];
            int swk = S.az2[sk] ? 0 : -1;
            if (S.az1[sk] && swk == 0) {
                swk = -1;
                if (tk > tk2) swk = 0;
            } else {
                swk = -1;
            }
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0, tdiv = 1.0, u, v, em;
            if (swj == 0) {
                swsol = 1;
                if (swk == 0) {
                    u = 2.0 * dnx;
                    v = 2.0 * dnz;
                    em = 4.0 * tj - tj2 - 4.0 * tk;
                    em = em + tk2;
                    a = v * v + u * u;
                    b = 2.0 * em * (u * u);
                    c = (u * u) * (em * em - (slown * slown) * (v * v));
                    tref = 4.0 * tj - tj2;
                    tdiv = 3.0;
                } else if (S.az1[sk]) {
                    u = dnz;
                    v = 2.0 * dnx;
                    em = 3.0 * tk - 4.0 * tj + tj2;
                    a = v * v + 9.0 * (u * u);
                    b = 6.0 * em * (u * u);
                    c = (u * u) * (em * em - (slown * slown) * (v * v));
                    tref = tk;
                    tdiv = 1.0;
                } else {
                    u = 2.0 * dnx;
                    a = 1.0; b = 0.0;
                    c = -((u * u) * (slown * slown));
                    tref = 4.0 * tj - tj2;
                    tdiv = 3.0;
                }
            } else if (S.ax1[sj]) {
                swsol = 1;
                if (swk == 0) {
                    u = dnx;
                    v = 2.0 * dnz;
                    em = 3.0 * tj - 4.0 * tk + tk2;
                    a = v * v + 9.0 * (u * u);
                    b = 6.0 * em * (u * u);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = tj;
                    tdiv = 1.0;
                } else if (S.az1[sk]) {
                    u = dnx;
                    v = dnz;
                    em = tk - tj;
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = tj;
                    tdiv = 1.0;
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (dnx * dnx));
                    tref = tj;
                    tdiv = 1.0;
                }
            } else {
                if (swk == 0) {
                    swsol = 1;
                    u = 2.0 * dnz;
                    a = 1.0; b = 0.0;
                    c = -((u * u) * (slown * slown));
                    tref = 4.0 * tk - tk2;
                    tdiv = 3.0;
                } else if (S.az1[sk]) {
                    swsol = 1;
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (dnz * dnz));
                    tref = tk;
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
    return travm;
}

__device__ __forceinline__ double l_bilinear(const double nv[2][2], double dsx, double dsz, double dnx, double dnz)
{
    double biv = 0.0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            double produ = (1.0 - fabs(((i - 1) * dnx - dsx) / dnx)) * (1.0 - fabs(((j - 1) * dnz - dsz) / dnz));
            biv = biv + nv[i - 1][j - 1] * produ;
        }
    return biv;
}

// travel (fm2d_ttime.f90:41-238); returns -1 if the source is outside the grid
__device__ static int l_travel(lane_grid &G, double scx, double scz, int urg)
{
    int isx = (int)((scx - G.gox) / G.dnx) + 1;
    int isz = (int)((scz - G.goz) / G.dnz) + 1;
    if (isx < 1 || isx > G.nnx || isz < 1 || isz > G.nnz) return -1;
    if (isx == G.nnx) isx = isx - 1;
    if (isz == G.nnz) isz = isz - 1;
    G.ntr = 0;
    if (urg == 2) {
        for (int i = 1; i <= G.nnx; i++)
            for (int j = 1; j <= G.nnz; j++)
                if (LS(G, j, i) > 0) l_addtree(G, j, i, LT(G, j, i));
    } else {
        for (int i = 1; i <= G.nnx; i++)
            for (int j = 1; j <= G.nnz; j++) LS(G, j, i) = -1;
        double vss[2][2];
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) vss[i - 1][j - 1] = LV(G, isz - 1 + j, isx - 1 + i);
        double dsx = (scx - G.gox) - (isx - 1) * G.dnx;
        double dsz = (scz - G.goz) - (isz - 1) * G.dnz;
        double vsrc = l_bilinear(vss, dsx, dsz, G.dnx, G.dnz);
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) {
                double ex = dsx - (i - 1) * G.dnx, ez = dsz - (j - 1) * G.dnz;
                double ds = sqrt(ex * ex + ez * ez);
                const double tc = 2.0 * ds / (vss[i - 1][j - 1] + vsrc);
                LT(G, isz - 1 + j, isx - 1 + i) = tc;
                l_addtree(G, isz - 1 + j, isx - 1 + i, tc);
            }
    }
    while (G.ntr > 0) {
        int e = G.btg[1];
        int ix = BPX(e), iz = BPZ(e);
        if (urg == 1) {   // stop when the band reaches an inner edge of the refined grid
            int swrg = 0;
            if (ix == 1 && G.vnl != 1) swrg = 1;
            if (ix == G.nnx && G.vnr != G.nnx) swrg = 1;
            if (iz == 1 && G.vnt != 1) swrg = 1;
            if (iz == G.nnz && G.vnb != G.nnz) swrg = 1;
            if (swrg == 1) { LS(G, iz, ix) = 0; break; }
        }
        LS(G, iz, ix) = 0;
        // the four neighbours (x-1, x+1, z-1, z+1: the reference's order) and their
        // stencils, read before the heap moves.  Processing one neighbour changes
        // no alive flag and no alive node's time, which is all fouds reads, and
        // far/close of the others stays as read; their heap positions are read
        // again when updtree needs them.
        const int nbx[4] = {ix - 1, ix + 1, ix, ix}, nbz[4] = {iz, iz, iz - 1, iz + 1};
        int st[4];
        l_sten S[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool in = nbx[k] >= 1 && nbx[k] <= G.nnx && nbz[k] >= 1 && nbz[k] <= G.nnz;
            const int cx = in ? nbx[k] : ix, cz = in ? nbz[k] : iz;
            st[k] = in ? LS(G, cz, cx) : 0;          // 0: alive or outside, nothing to do
            l_load_sten(G, cz, cx, S[k]);
        }
        l_downtree(G);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (st[k] == 0) continue;
            const double tv = G.fom == 0 ? l_fouds1s(S[k], G.dnx, G.dnz) : l_fouds2s(S[k], G.dnx, G.dnz);
            LT(G, nbz[k], nbx[k]) = tv;
            if (st[k] == -1) l_addtree(G, nbz[k], nbx[k], tv);
            else l_updtree(G, nbz[k], nbx[k], tv);
        }
    }
    return 0;
}

// One lane per (chain, period, source): modrays' loop body for that source.
__global__ __launch_bounds__(64) void k_fm2d_solve(mct_fm2d_args A)
{
    const long slot = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= A.nrun) return;
    if (A.only_flag && A.only_flag[slot] != 1) return;   // only problems handed over by the parallel path
    const long prob = A.prob0 + slot;
    const int s = (int)(prob % A.nsrc);
    const long cp = prob / A.nsrc;          // chain*np + period
    const int p = (int)(cp % A.np);
    const int c = (int)(cp / A.np);
    if (A.status[c] > 0) return;            // model already rejected
    const int32_t *srs = A.srs + (long)p * A.srs_pstride + (long)s * A.nrc;
    int any = 0;
    for (int r = 0; r < A.nrc; r++) any += srs[r];
    if (any == 0 && s != 0) return;        // modrays skips sources without pairs (:211)
    const long wsz = (long)A.ld * A.nc;
    const double *velnc = A.veln + cp * (long)A.nnx * A.nnz;
    const double *velv = A.velv + cp * (long)(A.nvz + 2) * (A.nvx + 2);
    const int ldv = A.nvz + 2;
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    lane_grid G;
    G.ld = A.ld;
    G.t = A.ttn + slot * wsz;
    G.s = A.nsts + slot * wsz;
    G.btg = A.heap + slot * (wsz + 8);
    // heap keys, 1-based, in arrays the stage leaves free: ttr during the refined
    // stage (its copy of the refined grid is written after it), vr during the
    // coarse one (the refined velocities are done with)
    G.hk = A.vr + slot * wsz - 1;
    G.fom = A.fom;
    const double x = A.scx[s], z = A.scz[s];
    int isx = (int)((x - A.gox) / dnx) + 1, isz = (int)((z - A.goz) / dnz) + 1;
    if (isx < 1 || isx > A.nnx || isz < 1 || isz > A.nnz) { atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE); return; }
    if (isx == A.nnx) isx = isx - 1;
    if (isz == A.nnz) isz = isz - 1;
    int vnl = isx - A.sgs, vnr = isx + A.sgs, vnt = isz - A.sgs, vnb = isz + A.sgs;
    if (vnl < 1) vnl = 1;
    if (vnr > A.nnx) vnr = A.nnx;
    if (vnt < 1) vnt = 1;
    if (vnb > A.nnz) vnb = A.nnz;
    if (A.asgr == 1) {
        // refined source grid with B-spline velocities (bsplrefine)
        double *vr = A.vr + slot * wsz;
        G.nnx = (vnr - vnl) * A.sgdl + 1;
        G.nnz = (vnb - vnt) * A.sgdl + 1;
        G.dnx = A.dvx / (double)(float)(A.gdx * A.sgdl);
        G.dnz = A.dvz / (double)(float)(A.gdz * A.sgdl);
        G.gox = A.gox + dnx * (vnl - 1);
        G.goz = A.goz + dnz * (vnt - 1);
        G.vnl = vnl; G.vnr = vnr; G.vnt = vnt; G.vnb = vnb;
        G.v = vr; G.vld = A.ld;
        const int nrxr = A.gdx * A.sgdl, nrzr = A.gdz * A.sgdl;
        const int origx = (vnl - 1) * A.sgdl + 1, origz = (vnt - 1) * A.sgdl + 1;
        for (int ii = 1; ii <= A.nvz - 1; ii++) {
            int conz = (ii == A.nvz - 1) ? nrzr + 1 : nrzr;
            for (int jj = 1; jj <= A.nvx - 1; jj++) {
                int conx = (jj == A.nvx - 1) ? nrxr + 1 : nrxr;
                for (int k = 1; k <= conz; k++) {
                    int st1 = A.gdz * (ii - 1) + (k - 1) / A.sgdl + 1;
                    if (st1 < vnt || st1 > vnb) continue;
                    st1 = nrzr * (ii - 1) + k;
                    double v = nrzr; v = (k - 1) / v;
                    double vi[4];
                    bspl_w(v, vi);
                    for (int l = 1; l <= conx; l++) {
                        int st2 = A.gdx * (jj - 1) + (l - 1) / A.sgdl + 1;
                        if (st2 < vnl || st2 > vnr) continue;
                        st2 = nrxr * (jj - 1) + l;
                        int idm1 = st1 - origz + 1, idm2 = st2 - origx + 1;
                        if (idm1 < 1 || idm1 > G.nnz || idm2 < 1 || idm2 > G.nnx) continue;
                        double u = nrxr; u = (l - 1) / u;
                        double ui[4];
                        bspl_w(u, ui);
                        double sum[4];
                        for (int i1 = 1; i1 <= 4; i1++) {
                            sum[i1 - 1] = 0.0;
                            for (int j1 = 1; j1 <= 4; j1++)
                                sum[i1 - 1] = sum[i1 - 1] + ui[j1 - 1] * velv[(jj - 2 + j1) * ldv + (ii - 2 + i1)];
                            sum[i1 - 1] = vi[i1 - 1] * sum[i1 - 1];
                        }
                        vr[(idm2 - 1) * A.ld + (idm1 - 1)] = sum[0] + sum[1] + sum[2] + sum[3];
                    }
                }
            }
        }
        G.hk = A.ttr + slot * wsz - 1;
        if (l_travel(G, x, z, 1)) { atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE); return; }
        G.hk = A.vr + slot * wsz - 1;
        // remap the refined grid onto the coarse grid
        double *tr = A.ttr + slot * wsz;
        int32_t *nr = A.nstr + slot * wsz;
        const int rnx = G.nnx, rnz = G.nnz;
        for (int a = 1; a <= rnx; a++)
            for (int b = 1; b <= rnz; b++) {
                long o = (long)(a - 1) * A.ld + (b - 1);
                tr[o] = G.t[o];
                nr[o] = G.s[o];
            }
        for (int a = 1; a <= A.nnx; a++)
            for (int b = 1; b <= A.nnz; b++) G.s[(long)(a - 1) * A.ld + (b - 1)] = -1;
        for (int k = 1; k <= rnz; k += A.sgdl) {
            int idm1 = vnt + (k - 1) / A.sgdl;
            for (int l = 1; l <= rnx; l += A.sgdl) {
                int idm2 = vnl + (l - 1) / A.sgdl;
                int st = nr[(long)(l - 1) * A.ld + (k - 1)];
                G.s[(long)(idm2 - 1) * A.ld + (idm1 - 1)] = st;
                if (st >= 0) G.t[(long)(idm2 - 1) * A.ld + (idm1 - 1)] = tr[(long)(l - 1) * A.ld + (k - 1)];
            }
        }
        G.nnx = A.nnx; G.nnz = A.nnz; G.dnx = dnx; G.dnz = dnz; G.gox = A.gox; G.goz = A.goz;
        G.v = velnc; G.vld = A.nnz;
        // complete the narrow band: alive nodes next to far nodes become close
        for (int k = 1; k <= A.nnx; k++)
            for (int l = 1; l <= A.nnz; l++) {
                if (LS(G, l, k) != 0) continue;
                if ((l - 1 >= 1 && LS(G, l - 1, k) == -1) || (l + 1 <= A.nnz && LS(G, l + 1, k) == -1) ||
                    (k - 1 >= 1 && LS(G, l, k - 1) == -1) || (k + 1 <= A.nnx && LS(G, l, k + 1) == -1))
                    LS(G, l, k) = 1;
            }
        if (l_travel(G, x, z, 2)) { atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE); return; }
    } else {
        G.nnx = A.nnx; G.nnz = A.nnz; G.dnx = dnx; G.dnz = dnz; G.gox = A.gox; G.goz = A.goz;
        G.v = velnc; G.vld = A.nnz;
        if (l_travel(G, x, z, 0)) { atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE); return; }
    }
    // srtimes: receiver picks
    double *out = A.ttime + prob * A.nrc;
    for (int r = 1; r <= A.nrc; r++) {
        if (srs[r - 1] == 0) continue;
        const double rx = A.rcx[r - 1], rz = A.rcz[r - 1];
        int irx = (int)floor((rx - A.gox) / dnx) + 1;
        int irz = (int)floor((rz - A.goz) / dnz) + 1;
        if (irx < 1 || irx > A.nnx || irz < 1 || irz > A.nnz) { atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE); return; }
        if (irx == A.nnx) irx = irx - 1;
        if (irz == A.nnz) irz = irz - 1;
        int jsx = (int)floor((x - A.gox) / dnx) + 1, jsz = (int)floor((z - A.goz) / dnz) + 1;
        double dpl = dnx, rd1 = dnz;
        if (rd1 < dpl) dpl = rd1;
        double sred = (x - rx) * (x - rx);
        sred = sred + (z - rz) * (z - rz);
        sred = sqrt(sred);
        int sw = 0;
        if (sred < dpl) sw = 1;
        if (jsx == irx && jsz == irz) sw = 1;
        double trr;
        if (sw == 1) {
            double vss[2][2];
            // the source cell, kept inside the grid (the reference reads past the
            // grid for a source on its last line or just below its origin)
            int csx = jsx < A.nnx ? jsx : A.nnx - 1, csz = jsz < A.nnz ? jsz : A.nnz - 1;
            csx = csx < 1 ? 1 : csx;
            csz = csz < 1 ? 1 : csz;
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = velnc[(long)(csx - 1 + k - 1) * A.nnz + (csz - 1 + l - 1)];
            double drx = (x - A.gox) - (jsx - 1) * dnx, drz = (z - A.goz) - (jsz - 1) * dnz;
            double vels = l_bilinear(vss, drx, drz, dnx, dnz);
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = velnc[(long)(irx - 1 + k - 1) * A.nnz + (irz - 1 + l - 1)];
            drx = (rx - A.gox) - (irx - 1) * dnx;
            drz = (rz - A.goz) - (irz - 1) * dnz;
            double velr = l_bilinear(vss, drx, drz, dnx, dnz);
            trr = 2.0 * sred / (vels + velr);
        } else {
            double drx = (rx - A.gox) - (irx - 1) * dnx;
            double drz = (rz - A.goz) - (irz - 1) * dnz;
            trr = 0.0;
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) {
                    double produ = (1.0 - fabs(((l - 1) * dnz - drz) / dnz)) * (1.0 - fabs(((k - 1) * dnx - drx) / dnx));
                    trr = trr + LT(G, irz - 1 + l, irx - 1 + k) * produ;
                }
        }
        out[r - 1] = trr;
    }
}

// Lanes per problem in the grouped heap-loop kernel (64 / GRP_LANES problems per
// wave) and the waves per SIMD it is compiled for.  LDS bounds the problems in
// flight (a 512-entry heap is 6.4 KB with 32-bit entries, 5.3 KB with 16-bit
// ones: 24 or 30 per CU), so registers are not the limit.  Measured on a 64-chain order-1 batch at 201x201 (40,960 problems):
// 16 lanes 1.84 s, 8 lanes 1.90 s, 4 lanes 2.69 s, one problem per wave 2.31 s;
// heaps of 384, 640 and 768 entries 2.55, 2.13 and 2.53 s (a smaller heap
// sends more problems to the lane solver, a larger one runs fewer at once).
#ifndef WAVE_EU
#define WAVE_EU 2
#endif
#ifndef GRP_LANES
#define GRP_LANES 16
#endif
namespace {

// ---- the heap loop, several problems per wave ----------------------------------------
// The reference's heap loop (travel, fm2d_ttime.f90:41-238) for the mixed-order
// problems, with each problem's heap and keys in LDS.  A wave runs 64/GL
// problems, one group of GL lanes each: every lane of a group holds its
// problem's state and runs the heap operations (so one vector instruction
// advances every problem of the wave), the group's first lane alone writes
// global memory, and lanes 0-3 of a group load and solve the popped node's four
// neighbours.  The state array keeps each heap node's position, as the
// reference's does.  A neighbour already in the heap needs that position for
// updtree: it arrives with the stencil, the downtree's moves are logged in LDS
// while the loads are in flight and replayed on it, and the neighbours'
// addtree / updtree moves are applied as they happen, so it is loaded once.  A
// band that outgrows the LDS heap leaves the problem to the lane solver.
struct wave_grid {
    int nnx, nnz, ld;
    double gox, goz, dnx, dnz;
    double *t;              // times (global)
    int32_t *s;             // state: -1 far, 0 alive, >0 heap position (global)
    const double *v;
    int vld;
    void *hb;               // the group's LDS heap, 1-based (entries of type ET below)
    double *hk;             // its keys, hk[p] = time of hb[p]
    int2 *mv;               // its log of a downtree's moves (at most 16)
    int cap, ntr, full;
    int vnl, vnr, vnt, vnb;
    int base;               // the group's first lane
    bool lead;              // this lane is it
};

// heap positions of up to four neighbours, followed through the moves of a pop
struct w_track {
    int p[4];
    __device__ __forceinline__ void move(int from, int to)
    {
#pragma unroll
        for (int k = 0; k < 4; k++) p[k] = p[k] == from ? to : p[k];
    }
};

__device__ __forceinline__ double w_shfl_d(double v, int src)
{
    return __shfl(v, src, 64);
}

// Lanes of one wave read what a group's first lane wrote: keep the accesses in order.
__device__ __forceinline__ void w_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A heap entry names its node as (ix << 16) | iz (ET = int32_t), or, where the
// slot's arrays hold fewer than 65,536 words, as the node's offset in them
// (ET = uint16_t: half the LDS, so more problems fit on a CU).
template <class ET>
__device__ __forceinline__ int w_get(const wave_grid &G, int p) { return (int)((const ET *)G.hb)[p]; }
template <class ET>
__device__ __forceinline__ void w_put(const wave_grid &G, int p, int e) { ((ET *)G.hb)[p] = (ET)e; }
template <class ET>
__device__ __forceinline__ int w_enc(const wave_grid &G, int iz, int ix)
{
    return sizeof(ET) == 2 ? (ix - 1) * G.ld + (iz - 1) : ((ix << 16) | iz);
}
template <class ET>
__device__ __forceinline__ void w_dec(const wave_grid &G, int e, int &iz, int &ix)
{
    if (sizeof(ET) == 2) { ix = e / G.ld + 1; iz = e % G.ld + 1; }
    else { ix = BPX(e); iz = BPZ(e); }
}
template <class ET>
__device__ __forceinline__ void w_setpos(const wave_grid &G, int e, int pos)
{
    if (!G.lead) return;
    if (sizeof(ET) == 2) G.s[e] = pos;
    else LS(G, BPZ(e), BPX(e)) = pos;
}

// addtree (fm2d_ttime.f90:527-564)
template <class ET>
__device__ static void w_addtree(wave_grid &G, int iz, int ix, double tv, w_track &trk)
{
    if (G.ntr + 1 > G.cap) { G.full = 1; return; }
    G.ntr = G.ntr + 1;
    int tpc = G.ntr, tpp = tpc / 2;
    while (tpp > 0) {
        const int ep = w_get<ET>(G, tpp);
        const double kp = G.hk[tpp];
        if (tv < kp) {
            w_setpos<ET>(G, ep, tpc);
            w_put<ET>(G, tpc, ep);
            G.hk[tpc] = kp;
            trk.move(tpp, tpc);
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
    const int e = w_enc<ET>(G, iz, ix);
    w_setpos<ET>(G, e, tpc);
    w_put<ET>(G, tpc, e);
    G.hk[tpc] = tv;
}

// downtree (fm2d_ttime.f90:575-644); its moves (from, to) go to the group's
// log, their count to nmv
template <class ET>
__device__ static void w_downtree(wave_grid &G, int &nmv)
{
    nmv = 0;
    if (G.ntr == 1) { G.ntr = G.ntr - 1; return; }
    const int old = G.ntr;
    const int el = w_get<ET>(G, old);
    const double kl = G.hk[old];
    G.ntr = G.ntr - 1;
    int tpp = 1, tpc = 2 * tpp;
    while (tpc < G.ntr) {
        const int e1 = w_get<ET>(G, tpc), e2 = w_get<ET>(G, tpc + 1);
        const double k1 = G.hk[tpc], k2 = G.hk[tpc + 1];
        int ec = e1;
        double kc = k1;
        if (k1 > k2) { tpc = tpc + 1; ec = e2; kc = k2; }
        if (kc < kl) {
            w_setpos<ET>(G, ec, tpp);
            w_put<ET>(G, tpp, ec);
            G.hk[tpp] = kc;
            G.mv[nmv] = make_int2(tpc, tpp);
            nmv++;
            tpp = tpc;
            tpc = 2 * tpp;
        } else {
            tpc = G.ntr + 1;
        }
    }
    if (tpc == G.ntr) {
        const int ec = w_get<ET>(G, tpc);
        const double kc = G.hk[tpc];
        if (kc < kl) {
            w_setpos<ET>(G, ec, tpp);
            w_put<ET>(G, tpp, ec);
            G.hk[tpp] = kc;
            G.mv[nmv] = make_int2(tpc, tpp);
            nmv++;
            tpp = tpc;
        }
    }
    w_setpos<ET>(G, el, tpp);
    w_put<ET>(G, tpp, el);
    G.hk[tpp] = kl;
    G.mv[nmv] = make_int2(old, tpp);
    nmv++;
}

// updtree (fm2d_ttime.f90:653-680) from heap position tpc
template <class ET>
__device__ static void w_updtree(wave_grid &G, int tpc, int iz, int ix, double tv, w_track &trk)
{
    int tpp = tpc / 2;
    while (tpp > 0) {
        const int ep = w_get<ET>(G, tpp);
        const double kp = G.hk[tpp];
        if (tv < kp) {
            w_setpos<ET>(G, ep, tpc);
            w_put<ET>(G, tpc, ep);
            G.hk[tpc] = kp;
            trk.move(tpp, tpc);
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
    const int e = w_enc<ET>(G, iz, ix);
    w_setpos<ET>(G, e, tpc);
    w_put<ET>(G, tpc, e);
    G.hk[tpc] = tv;
}

// travel (fm2d_ttime.f90:41-238) on a group; -1 source outside, -2 heap full.
// Every lane of the group takes the same path through it.
template <int FOM, int GL, class ET>
__device__ static int w_travel(wave_grid &G, double scx, double scz, int urg)
{
    const int gl = threadIdx.x - G.base;
    int isx = (int)((scx - G.gox) / G.dnx) + 1;
    int isz = (int)((scz - G.goz) / G.dnz) + 1;
    if (isx < 1 || isx > G.nnx || isz < 1 || isz > G.nnz) return -1;
    if (isx == G.nnx) isx = isx - 1;
    if (isz == G.nnz) isz = isz - 1;
    G.ntr = 0;
    G.full = 0;
    w_track none;
    none.p[0] = none.p[1] = none.p[2] = none.p[3] = 0;
    if (urg == 2) {
        // the band in the reference's scan order (x outer, z inner), GL nodes at a time
        const long n = (long)G.nnx * G.nnz;
        const unsigned long long gmask = GL == 64 ? ~0ull : ((1ull << GL) - 1ull);
        for (long q0 = 0; q0 < n; q0 += GL) {
            const long q = q0 + gl;
            int i = 1, j = 1;
            bool band = false;
            double tq = 0.0;
            if (q < n) {
                i = (int)(q / G.nnz) + 1;
                j = (int)(q % G.nnz) + 1;
                band = LS(G, j, i) > 0;
                if (band) tq = LT(G, j, i);
            }
            unsigned long long m = (__ballot(band) >> G.base) & gmask;
            while (m) {
                const int b = __builtin_ctzll(m);
                m &= m - 1;
                const int bi = __shfl(i, G.base + b, 64), bj = __shfl(j, G.base + b, 64);
                w_addtree<ET>(G, bj, bi, w_shfl_d(tq, G.base + b), none);
                if (G.full) return -2;
            }
            w_sync();
        }
    } else {
        for (long q = gl; q < (long)G.nnx * G.nnz; q += GL) {
            const int i = (int)(q / G.nnz) + 1, j = (int)(q % G.nnz) + 1;
            LS(G, j, i) = -1;
        }
        w_sync();
        double vss[2][2];
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) vss[i - 1][j - 1] = LV(G, isz - 1 + j, isx - 1 + i);
        double dsx = (scx - G.gox) - (isx - 1) * G.dnx;
        double dsz = (scz - G.goz) - (isz - 1) * G.dnz;
        double vsrc = l_bilinear(vss, dsx, dsz, G.dnx, G.dnz);
        for (int i = 1; i <= 2; i++)
            for (int j = 1; j <= 2; j++) {
                double ex = dsx - (i - 1) * G.dnx, ez = dsz - (j - 1) * G.dnz;
                double ds = sqrt(ex * ex + ez * ez);
                const double tc = 2.0 * ds / (vss[i - 1][j - 1] + vsrc);
                if (G.lead) LT(G, isz - 1 + j, isx - 1 + i) = tc;
                w_addtree<ET>(G, isz - 1 + j, isx - 1 + i, tc, none);
            }
        w_sync();
    }
    // lane k < 4 of the group takes neighbour k (x-1, x+1, z-1, z+1: the reference's order)
    const int kn = gl & 3;
    const int dx = kn == 0 ? -1 : (kn == 1 ? 1 : 0), dz = kn == 2 ? -1 : (kn == 3 ? 1 : 0);
    while (G.ntr > 0) {
        int ix, iz;
        w_dec<ET>(G, w_get<ET>(G, 1), iz, ix);
        if (urg == 1) {   // stop when the band reaches an inner edge of the refined grid
            int swrg = 0;
            if (ix == 1 && G.vnl != 1) swrg = 1;
            if (ix == G.nnx && G.vnr != G.nnx) swrg = 1;
            if (iz == 1 && G.vnt != 1) swrg = 1;
            if (iz == G.nnz && G.vnb != G.nnz) swrg = 1;
            if (swrg == 1) {
                if (G.lead) LS(G, iz, ix) = 0;
                w_sync();
                break;
            }
        }
        if (G.lead) LS(G, iz, ix) = 0;
        w_sync();
        // the neighbours' states (heap positions) and stencils, read before the heap moves
        const int nx = ix + dx, nz = iz + dz;
        const bool in = gl < 4 && nx >= 1 && nx <= G.nnx && nz >= 1 && nz <= G.nnz;
        const int cx = in ? nx : ix, cz = in ? nz : iz;
        const int st = in ? LS(G, cz, cx) : 0;
        lane_grid L;
        L.nnx = G.nnx; L.nnz = G.nnz; L.ld = G.ld; L.t = G.t; L.s = G.s; L.v = G.v; L.vld = G.vld; L.fom = FOM;
        l_sten S;
        l_load_sten(L, cz, cx, S);
        // the heap moves while the loads are in flight; the neighbours'
        // positions then follow the logged moves, each lane its own
        int nmv = 0;
        w_downtree<ET>(G, nmv);
        w_sync();
        int mypos = st > 0 ? st : 0;
        for (int q = 0; q < nmv; q++) {
            const int2 m = G.mv[q];
            mypos = mypos == m.x ? m.y : mypos;
        }
        const int st0 = __shfl(st, G.base + 0, 64), st1 = __shfl(st, G.base + 1, 64);
        const int st2 = __shfl(st, G.base + 2, 64), st3 = __shfl(st, G.base + 3, 64);
        w_track trk;
        trk.p[0] = __shfl(mypos, G.base + 0, 64); trk.p[1] = __shfl(mypos, G.base + 1, 64);
        trk.p[2] = __shfl(mypos, G.base + 2, 64); trk.p[3] = __shfl(mypos, G.base + 3, 64);
        double tv = 0.0;
        if (st != 0) tv = FOM == 0 ? l_fouds1s(S, G.dnx, G.dnz) : l_fouds2s(S, G.dnx, G.dnz);
        const double tv0 = w_shfl_d(tv, G.base + 0), tv1 = w_shfl_d(tv, G.base + 1);
        const double tv2 = w_shfl_d(tv, G.base + 2), tv3 = w_shfl_d(tv, G.base + 3);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int stk = k == 0 ? st0 : (k == 1 ? st1 : (k == 2 ? st2 : st3));
            if (stk == 0) continue;
            const double tvk = k == 0 ? tv0 : (k == 1 ? tv1 : (k == 2 ? tv2 : tv3));
            const int xk = ix + (k == 0 ? -1 : (k == 1 ? 1 : 0)), zk = iz + (k == 2 ? -1 : (k == 3 ? 1 : 0));
            if (G.lead) LT(G, zk, xk) = tvk;
            if (stk == -1) {
                w_addtree<ET>(G, zk, xk, tvk, trk);
            } else {
                const int pos = trk.p[k];
                trk.p[k] = 0;
                w_updtree<ET>(G, pos, zk, xk, tvk, trk);
            }
        }
        if (G.full) return -2;
        w_sync();
    }
    return 0;
}

// One group of GL lanes per problem whose flag is 1: modrays' loop body for
// that source, as k_fm2d_solve runs it; the flag becomes 3 when the problem is
// solved here, 2 when there is nothing to solve, and stays 1 when its band
// outgrew the LDS heap (or its source lies outside the grid, which the lane
// solver reports).  Every branch below is taken by whole groups.
template <int FOM, int GL, class ET>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WAVE_EU))) void k_fm2d_wave(mct_fm2d_args A, int heap_cap, int32_t *pflag)
{
    extern __shared__ double w_lds[];
    constexpr int NG = 64 / GL;
    const int lane = threadIdx.x;
    const int g = lane / GL, gl = lane % GL;
    const long slot = (long)blockIdx.x * NG + g;
    // a group with nothing to do still runs with the others (its lanes mask out)
    bool live = slot < A.nrun;
    if (live) live = pflag[slot] == 1;
    const long prob = A.prob0 + (live ? slot : 0);
    const int s = (int)(prob % A.nsrc);
    const long cp = prob / A.nsrc;
    const int p = (int)(cp % A.np);
    const int c = (int)(cp / A.np);
    const int32_t *srs = A.srs + (long)p * A.srs_pstride + (long)s * A.nrc;
    {
        // a rejected model, or a source without pairs (modrays skips it, :211):
        // nothing to do
        int any = 0;
        if (live)
            for (int r = gl; r < A.nrc; r += GL) any |= srs[r];
        const unsigned long long gmask = GL == 64 ? ~0ull : ((1ull << GL) - 1ull);
        const bool anyg = ((__ballot(any != 0) >> (g * GL)) & gmask) != 0;
        if (live && (A.status[c] > 0 || (!anyg && s != 0))) {
            if (gl == 0) pflag[slot] = 2;
            live = false;
        }
    }
    if (!live) return;
    const long wsz = (long)A.ld * A.nc;
    const double *velnc = A.veln + cp * (long)A.nnx * A.nnz;
    const double *velv = A.velv + cp * (long)(A.nvz + 2) * (A.nvx + 2);
    const int ldv = A.nvz + 2;
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    wave_grid G;
    G.ld = A.ld;
    G.t = A.ttn + slot * wsz;
    G.s = A.nsts + slot * wsz;
    {
        const int hbw = (int)(((long)(heap_cap + 1) * sizeof(ET) + 7) / 8);   // doubles of entries
        const int per = heap_cap + 1 + hbw + 16;                              // doubles per group
        double *gb = w_lds + (long)g * per;
        G.hk = gb;
        G.hb = (void *)(gb + heap_cap + 1);
        G.mv = (int2 *)(gb + heap_cap + 1 + hbw);
    }
    G.cap = heap_cap;
    G.base = g * GL;
    G.lead = gl == 0;
    const double x = A.scx[s], z = A.scz[s];
    int isx = (int)((x - A.gox) / dnx) + 1, isz = (int)((z - A.goz) / dnz) + 1;
    if (isx < 1 || isx > A.nnx || isz < 1 || isz > A.nnz) return;   // the lane solver reports it
    if (isx == A.nnx) isx = isx - 1;
    if (isz == A.nnz) isz = isz - 1;
    int vnl = isx - A.sgs, vnr = isx + A.sgs, vnt = isz - A.sgs, vnb = isz + A.sgs;
    if (vnl < 1) vnl = 1;
    if (vnr > A.nnx) vnr = A.nnx;
    if (vnt < 1) vnt = 1;
    if (vnb > A.nnz) vnb = A.nnz;
    if (A.asgr == 1) {
        // refined source grid with B-spline velocities (bsplrefine), a node per lane
        double *vr = A.vr + slot * wsz;
        G.nnx = (vnr - vnl) * A.sgdl + 1;
        G.nnz = (vnb - vnt) * A.sgdl + 1;
        G.dnx = A.dvx / (double)(float)(A.gdx * A.sgdl);
        G.dnz = A.dvz / (double)(float)(A.gdz * A.sgdl);
        G.gox = A.gox + dnx * (vnl - 1);
        G.goz = A.goz + dnz * (vnt - 1);
        G.vnl = vnl; G.vnr = vnr; G.vnt = vnt; G.vnb = vnb;
        G.v = vr; G.vld = A.ld;
        const int nrxr = A.gdx * A.sgdl, nrzr = A.gdz * A.sgdl;
        const int origx = (vnl - 1) * A.sgdl + 1, origz = (vnt - 1) * A.sgdl + 1;
        const int rn = G.nnx * G.nnz;
        for (int q = gl; q < rn; q += GL) {
            const int idm2 = q / G.nnz + 1, idm1 = q % G.nnz + 1;
            const int st1 = idm1 + origz - 1, st2 = idm2 + origx - 1;
            int ii = (st1 - 1) / nrzr + 1, k = (st1 - 1) % nrzr + 1;
            if (ii > A.nvz - 1) { ii = A.nvz - 1; k = nrzr + 1; }
            int jj = (st2 - 1) / nrxr + 1, l = (st2 - 1) % nrxr + 1;
            if (jj > A.nvx - 1) { jj = A.nvx - 1; l = nrxr + 1; }
            double u = nrxr; u = (l - 1) / u;
            double v = nrzr; v = (k - 1) / v;
            double ui[4], vi[4];
            bspl_w(u, ui);
            bspl_w(v, vi);
            double sum[4];
            for (int i1 = 1; i1 <= 4; i1++) {
                sum[i1 - 1] = 0.0;
                for (int j1 = 1; j1 <= 4; j1++)
                    sum[i1 - 1] = sum[i1 - 1] + ui[j1 - 1] * velv[(jj - 2 + j1) * ldv + (ii - 2 + i1)];
                sum[i1 - 1] = vi[i1 - 1] * sum[i1 - 1];
            }
            vr[(long)(idm2 - 1) * A.ld + (idm1 - 1)] = sum[0] + sum[1] + sum[2] + sum[3];
        }
        w_sync();
        if (w_travel<FOM, GL, ET>(G, x, z, 1)) return;
        // remap the refined grid onto the coarse grid
        double *tr = A.ttr + slot * wsz;
        int32_t *nr = A.nstr + slot * wsz;
        const int rnx = G.nnx, rnz = G.nnz;
        for (int q = gl; q < rnx * rnz; q += GL) {
            const long o = (long)(q / rnz) * A.ld + (q % rnz);
            tr[o] = G.t[o];
            nr[o] = G.s[o];
        }
        w_sync();
        for (long q = gl; q < (long)A.nnx * A.nnz; q += GL)
            G.s[(q / A.nnz) * A.ld + (q % A.nnz)] = -1;
        w_sync();
        const int wkz = (rnz - 1) / A.sgdl + 1, wkx = (rnx - 1) / A.sgdl + 1;
        for (int q = gl; q < wkz * wkx; q += GL) {
            const int k = (q % wkz) * A.sgdl + 1, l = (q / wkz) * A.sgdl + 1;
            const int idm1 = vnt + (k - 1) / A.sgdl, idm2 = vnl + (l - 1) / A.sgdl;
            const int st = nr[(long)(l - 1) * A.ld + (k - 1)];
            G.s[(long)(idm2 - 1) * A.ld + (idm1 - 1)] = st;
            if (st >= 0) G.t[(long)(idm2 - 1) * A.ld + (idm1 - 1)] = tr[(long)(l - 1) * A.ld + (k - 1)];
        }
        w_sync();
        G.nnx = A.nnx; G.nnz = A.nnz; G.dnx = dnx; G.dnz = dnz; G.gox = A.gox; G.goz = A.goz;
        G.v = velnc; G.vld = A.nnz;
        // complete the narrow band: alive nodes next to far nodes become close
        for (long q = gl; q < (long)A.nnx * A.nnz; q += GL) {
            const int k = (int)(q / A.nnz) + 1, l = (int)(q % A.nnz) + 1;
            if (LS(G, l, k) != 0) continue;
            if ((l - 1 >= 1 && LS(G, l - 1, k) == -1) || (l + 1 <= A.nnz && LS(G, l + 1, k) == -1) ||
                (k - 1 >= 1 && LS(G, l, k - 1) == -1) || (k + 1 <= A.nnx && LS(G, l, k + 1) == -1))
                LS(G, l, k) = 1;
        }
        w_sync();
        if (w_travel<FOM, GL, ET>(G, x, z, 2)) return;
    } else {
        G.nnx = A.nnx; G.nnz = A.nnz; G.dnx = dnx; G.dnz = dnz; G.gox = A.gox; G.goz = A.goz;
        G.v = velnc; G.vld = A.nnz;
        if (w_travel<FOM, GL, ET>(G, x, z, 0)) return;
    }
    w_sync();
    // srtimes: receiver picks, a receiver per lane
    double *out = A.ttime + prob * A.nrc;
    for (int r = gl + 1; r <= A.nrc; r += GL) {
        if (srs[r - 1] == 0) continue;
        const double rx = A.rcx[r - 1], rz = A.rcz[r - 1];
        int irx = (int)floor((rx - A.gox) / dnx) + 1;
        int irz = (int)floor((rz - A.goz) / dnz) + 1;
        if (irx < 1 || irx > A.nnx || irz < 1 || irz > A.nnz) { atomicMin(&A.status[c], (int)MCT_ERR_OUTSIDE); continue; }
        if (irx == A.nnx) irx = irx - 1;
        if (irz == A.nnz) irz = irz - 1;
        int jsx = (int)floor((x - A.gox) / dnx) + 1, jsz = (int)floor((z - A.goz) / dnz) + 1;
        double dpl = dnx, rd1 = dnz;
        if (rd1 < dpl) dpl = rd1;
        double sred = (x - rx) * (x - rx);
        sred = sred + (z - rz) * (z - rz);
        sred = sqrt(sred);
        int sw = 0;
        if (sred < dpl) sw = 1;
        if (jsx == irx && jsz == irz) sw = 1;
        double trr;
        if (sw == 1) {
            double vss[2][2];
            int csx = jsx < A.nnx ? jsx : A.nnx - 1, csz = jsz < A.nnz ? jsz : A.nnz - 1;
            csx = csx < 1 ? 1 : csx;
            csz = csz < 1 ? 1 : csz;
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = velnc[(long)(csx - 1 + k - 1) * A.nnz + (csz - 1 + l - 1)];
            double drx = (x - A.gox) - (jsx - 1) * dnx, drz = (z - A.goz) - (jsz - 1) * dnz;
            double vels = l_bilinear(vss, drx, drz, dnx, dnz);
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = velnc[(long)(irx - 1 + k - 1) * A.nnz + (irz - 1 + l - 1)];
            drx = (rx - A.gox) - (irx - 1) * dnx;
            drz = (rz - A.goz) - (irz - 1) * dnz;
            double velr = l_bilinear(vss, drx, drz, dnx, dnz);
            trr = 2.0 * sred / (vels + velr);
        } else {
            double drx = (rx - A.gox) - (irx - 1) * dnx;
            double drz = (rz - A.goz) - (irz - 1) * dnz;
            trr = 0.0;
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) {
                    double produ = (1.0 - fabs(((l - 1) * dnz - drz) / dnz)) * (1.0 - fabs(((k - 1) * dnx - drx) / dnx));
                    trr = trr + LT(G, irz - 1 + l, irx - 1 + k) * produ;
                }
        }
        out[r - 1] = trr;
    }
    if (gl == 0) pflag[slot] = 3;
}

} // namespace

// LDS entries of each problem's heap in k_fm2d_wave (environment MCT_WAVE_HEAP,
// default 512)
static int w_heap_cap()
{
    const char *e = getenv("MCT_WAVE_HEAP");
    const int v = e ? atoi(e) : 512;
    return v >= 64 && v <= 4096 ? v : 512;
}

template <int GL, class ET>
static int w_launch(const mct_fm2d_args &b, int cap, int32_t *pflag, hipStream_t s)
{
    constexpr int NG = 64 / GL;
    static int lds_max = 0;
    if (!lds_max) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        lds_max = v;
    }
    // per problem: keys, heap entries, the downtree log (16 moves: at most 13 levels);
    // a heap too large for the workgroup's LDS is cut to what fits (bands that
    // outgrow it go to the lane solver like any other overflow)
    auto bytes = [&](int c) -> size_t {
        const long hbw = ((long)(c + 1) * (long)sizeof(ET) + 7) / 8;
        return (size_t)NG * 8 * (size_t)(c + 1 + hbw + 16);
    };
    while (cap > 64 && bytes(cap) > (size_t)lds_max) cap -= 64;
    const size_t lds = bytes(cap);
    if (lds > (size_t)lds_max) return MCT_ERR_ARG;
    const void *k0 = (const void *)k_fm2d_wave<0, GL, ET>, *k1 = (const void *)k_fm2d_wave<1, GL, ET>;
    if (lds > 64 * 1024)
        MCT_CHECK(hipFuncSetAttribute(b.fom == 0 ? k0 : k1, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const unsigned nb = (unsigned)((b.nrun + NG - 1) / NG);
    if (b.fom == 0) hipLaunchKernelGGL((k_fm2d_wave<0, GL, ET>), dim3(nb), dim3(64), lds, s, b, cap, pflag);
    else hipLaunchKernelGGL((k_fm2d_wave<1, GL, ET>), dim3(nb), dim3(64), lds, s, b, cap, pflag);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_fm2d_wave(const mct_fm2d_args &b, int32_t *pflag, hipStream_t s)
{
    if (!pflag || b.nrun < 1) return MCT_ERR_ARG;
    // bands that outgrow the heap go to the lane solver, which keeps all of them
    // in flight at once (a second pass with 2,048-entry heaps, four problems per
    // CU, was slower: 9.5 s instead of 2.6 s when a 384-entry first pass left a
    // large share of the 40,960 problems over)
    // 16-bit heap entries (a node's offset in the slot's arrays) where they fit.
    // A launch with few problems gives each a wave of its own, which shortens its
    // chain (a plain 201x201 order-1 call, 640 problems: 194.6 ms against 242.8
    // with four problems per wave); many problems share waves, which is the more
    // work per instruction (40,960 problems: 1.63 s against 2.31 s)
    const bool e16 = (long)b.ld * b.nc <= 65536;
    const int cap = w_heap_cap();
    const char *es = getenv("MCT_WAVE_SPLIT");      // the largest launch with a wave per problem
    const long split = es ? atol(es) : 4096;
    if (b.nrun <= split)
        return e16 ? w_launch<64, uint16_t>(b, cap, pflag, s) : w_launch<64, int32_t>(b, cap, pflag, s);
    return e16 ? w_launch<GRP_LANES, uint16_t>(b, cap, pflag, s) : w_launch<GRP_LANES, int32_t>(b, cap, pflag, s);
}

int mct_launch_fm2d_gridder(const mct_fm2d_args &a, hipStream_t s)
{
    const long nn = (long)a.nnx * a.nnz * a.nchain * a.np;
    hipLaunchKernelGGL(k_fm2d_gridder, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, a);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_fm2d_solve(const mct_fm2d_args &b, hipStream_t s)
{
    // problems per wave (one per lane): 32 measured fastest for the order-1
    // hand-over at 201x201 (5.2 s against 5.5 s with 64 per wave, 6.2 with 16);
    // environment MCT_SOLVE_LANES = 8, 16, 32 or 64
    const char *e = getenv("MCT_SOLVE_LANES");
    int lanes = e ? atoi(e) : 32;
    if (lanes != 8 && lanes != 16 && lanes != 64) lanes = 32;
    hipLaunchKernelGGL(k_fm2d_solve, dim3((unsigned)((b.nrun + lanes - 1) / lanes)), dim3(lanes), 0, s, b);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_fm2d(const mct_fm2d_args &a, long chunk, hipStream_t s)
{
    const long nn = (long)a.nnx * a.nnz * a.nchain * a.np;
    hipLaunchKernelGGL(k_fm2d_gridder, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, a);
    MCT_CHECK(hipGetLastError());
    const long nprob = (long)a.nchain * a.np * a.nsrc;
    if (chunk < 1) return MCT_ERR_ARG;
    for (long p0 = 0; p0 < nprob; p0 += chunk) {
        mct_fm2d_args b = a;
        b.prob0 = p0;
        b.nrun = nprob - p0 < chunk ? nprob - p0 : chunk;
        hipLaunchKernelGGL(k_fm2d_solve, dim3((unsigned)((b.nrun + 63) / 64)), dim3(64), 0, s, b);
        MCT_CHECK(hipGetLastError());
        int rc = mct_launch_fm2d_rays(b, s);
        if (rc) return rc;
    }
    return MCT_OK;
}

// problems the calling thread's last mct_modrays call handed from the parallel
// path to the sequential solver
static thread_local long long g_modrays_handover = 0;

extern "C" int mct_fm2d_last_handover(long long *n)
{
    if (!n) return MCT_ERR_ARG;
    *n = g_modrays_handover;
    return MCT_OK;
}

// ---- C ABI: modrays' phase times for one or more periods --------------------------
extern "C" int mct_modrays(int nvx, int nvz, double goxd, double gozd, double dvxd, double dvzd,
                           int np, const double *velv, int gdx, int gdz, int asgr, int sgdl, int sgs,
                           int fom, int nsrc, const double *scx, const double *scz, int nrc,
                           const double *rcx, const double *rcz, const int32_t *srs, double *ttime,
                           int method)
{
    if ((long)(nvx - 1) * gdx + 1 > 32767 || (long)(nvz - 1) * gdz + 1 > 32767 || 2L * sgs * sgdl + 1 > 32767)
        return MCT_ERR_ARG;   // heap entries pack (ix, iz) into 16 bits each
    if (nvx < 2 || nvz < 2 || np < 1 || gdx < 1 || gdz < 1 || sgdl < 1 || sgs < 1 || nsrc < 1 || nrc < 1 ||
        !velv || !scx || !scz || !rcx || !rcz || !srs || !ttime || !(dvxd > 0) || !(dvzd > 0))
        return MCT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MCT_ERR_HIP;
    mct_fm2d_args a;
    memset(&a, 0, sizeof a);
    a.nvx = nvx; a.nvz = nvz; a.gox = goxd; a.goz = gozd; a.dvx = dvxd; a.dvz = dvzd;
    a.gdx = gdx; a.gdz = gdz; a.asgr = asgr; a.sgdl = sgdl; a.sgs = sgs; a.fom = fom;
    a.nnx = (nvx - 1) * gdx + 1; a.nnz = (nvz - 1) * gdz + 1;
    int rg = 2 * sgs * sgdl + 1;
    a.ld = a.nnz > rg ? a.nnz : rg;
    a.nc = a.nnx > rg ? a.nnx : rg;
    a.nchain = 1; a.np = np; a.nsrc = nsrc; a.nrc = nrc;
    const long nprob = (long)np * nsrc, wsz = (long)a.ld * a.nc;
    const size_t nv = (size_t)(nvx + 2) * (nvz + 2) * np, nn = (size_t)a.nnx * a.nnz * np;
    a.nnx = (nvx - 1) * gdx + 1; a.nnz = (nvz - 1) * gdz + 1;   // for the LDS check below
    const bool fast = method == 0 && mct_fm2d_fast_ok(a);
    void *bufs[16] = {nullptr};
    size_t sizes[16] = {nv * 8, 2 * nn * 8, 8 * (size_t)nsrc, 8 * (size_t)nsrc, 8 * (size_t)nrc, 8 * (size_t)nrc,
                        4 * (size_t)np * nsrc * nrc, 8 * (size_t)(nprob * wsz), 8 * (size_t)(nprob * wsz),
                        8 * (size_t)(nprob * wsz), 4 * (size_t)(nprob * wsz), 4 * (size_t)(nprob * wsz),
                        4 * (size_t)(nprob * (wsz + 8)), 8 * (size_t)nprob * nrc, 4, 4 * (size_t)nprob};
    int rc = MCT_OK;
    for (int i = 0; i < 16 && !rc; i++)
        if (hipMalloc(&bufs[i], sizes[i]) != hipSuccess) rc = MCT_ERR_NOMEM;
    if (!rc && (hipMemcpy(bufs[0], velv, nv * 8, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(bufs[2], scx, 8 * (size_t)nsrc, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(bufs[3], scz, 8 * (size_t)nsrc, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(bufs[4], rcx, 8 * (size_t)nrc, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(bufs[5], rcz, 8 * (size_t)nrc, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(bufs[6], srs, 4 * (size_t)np * nsrc * nrc, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(bufs[13], ttime, 8 * (size_t)nprob * nrc, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(bufs[14], 0, 4) != hipSuccess))
        rc = MCT_ERR_HIP;
    if (!rc) {
        a.velv = (const double *)bufs[0]; a.veln = (double *)bufs[1]; a.sln = (double *)bufs[1] + nn;
        a.scx = (const double *)bufs[2]; a.scz = (const double *)bufs[3];
        a.rcx = (const double *)bufs[4]; a.rcz = (const double *)bufs[5];
        a.srs = (const int32_t *)bufs[6];
        a.ttn = (double *)bufs[7]; a.ttr = (double *)bufs[8]; a.vr = (double *)bufs[9];
        a.nsts = (int32_t *)bufs[10]; a.nstr = (int32_t *)bufs[11]; a.heap = (int32_t *)bufs[12];
        a.ttime = (double *)bufs[13]; a.status = (int32_t *)bufs[14];
        a.srs_pstride = (long)nsrc * nrc;
        rc = fast ? mct_launch_fm2d_fast(a, nprob, (int32_t *)bufs[15], 0) : mct_launch_fm2d(a, nprob, 0);
        if (!rc && hipDeviceSynchronize() != hipSuccess) rc = MCT_ERR_HIP;
    }
    int32_t st = 0;
    if (!rc && (hipMemcpy(ttime, bufs[13], 8 * (size_t)nprob * nrc, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(&st, bufs[14], 4, hipMemcpyDeviceToHost) != hipSuccess))
        rc = MCT_ERR_HIP;
    g_modrays_handover = 0;
    if (!rc && fast) {
        int32_t *pf = (int32_t *)malloc(4 * (size_t)nprob);
        if (pf && hipMemcpy(pf, bufs[15], 4 * (size_t)nprob, hipMemcpyDeviceToHost) == hipSuccess)
            for (long i = 0; i < nprob; i++) g_modrays_handover += pf[i] == 1 || pf[i] == 3;
        free(pf);
    }
    for (int i = 0; i < 16; i++) if (bufs[i]) (void)hipFree(bufs[i]);
    if (!rc && st < 0) rc = st;
    return rc;
}
