// fm2d_fast.hip -- parallel 2-D phase traveltimes with the reference's answer.
//
// modrays' narrow-band fast marching (fm2d/fm2d_ttime.f90:41-238) freezes
// nodes in time order; a node's final time is the local upwind update
// (fouds1/fouds2, :248-518) over the neighbours frozen before it, i.e. over
// the neighbours with smaller times.  That makes the final field the fixed
// point of
//     T(x) = fouds(x; alive = {neighbours y : T(y) < T(x)})
// which an order-free iteration reaches as well, with the same floating-point
// operations per node.  Three places depend on the heap's history and are
// reproduced as the reference does them:
//   * the source: the four corner times of the source cell are heap keys that
//     later pops overwrite, so the first pops run the reference's own heap
//     loop until the source cell and two rings around it are frozen (the
//     seed);
//   * the refined source grid stops when the first node on an inner edge
//     reaches the top of the heap (:143-163).  From the converged refined
//     field the stop is rebuilt: frozen = times below that node's time, close
//     = unfrozen nodes with a frozen neighbour, whose tentative time is the
//     update over the neighbours frozen when the last of them was;
//   * the coarse restart from the remapped band (:86-98, re-banded frozen
//     nodes) runs the reference's heap loop until the window of the refined
//     grid and one ring around it are frozen.
// Everything else is a Gauss-Seidel sweep over anti-diagonals (no two nodes of
// one anti-diagonal share a stencil, so the nodes of a diagonal update in
// parallel) until a whole round of four sweep directions changes nothing.
//
// k_fm2d_near: one wave per (chain, period, source), LDS-resident refined
//   grid and coarse neighbourhood of the source; writes the frozen coarse
//   nodes and a fixed mask.
// k_fm2d_sweep: one workgroup per problem, sweeps the coarse grid in global
//   memory, then picks receivers (srtimes, fm2dray_cartesian.f90:682-797).
// A problem whose heap loop would leave the LDS neighbourhood is flagged and
// solved by the sequential kernel instead.
#include "mct_internal.h"
#include <string.h>
#include <stdlib.h>

#pragma clang fp contract(off)

#define FF_BIG 1e38
#define FF_THR0 1e37
#define FF_FIN 1e30
#define FF_MARGIN 16          // coarse neighbourhood: refined window +- this many nodes
#define FF_MAX_ROUNDS 512
#define FF_SEED_RINGS 2       // the seed freezes the source cell and this many rings around it
// Why the heap-order check handed a problem to the sequential solver (bits
// of the problem's statistic, k_fm2d_verify and k_fm2d_near).
#define FF_VR_NOTIME 1      // no finite time
#define FF_VR_TIE 2         // equal times among the node and its open neighbours below it
#define FF_VR_EARLY 4       // (b) a key at or below the next neighbour's time
#define FF_VR_RISE 8        // (a) a key that rises
#define FF_VR_FINAL 16      // (e) the last key is not the time
#define FF_VR_HEAP 32       // the band at the hand-over is not a valid heap (k_fm2d_near)
#define FF_VR_NOCONV 64     // the sweep did not converge


namespace {

// A view of (part of) a grid: 1-based (iz, ix) in [1,nnz] x [1,nnx]; storage
// covers [oz+1, oz+lnz] x [ox+1, ox+lnx] with leading dimension ld (z fastest).
struct fview {
    int nnx, nnz;
    double dnx, dnz;
    double *T;
    int16_t *ns;          // LDS states (-1 far, 0 frozen, >0 heap position) or null
    int ld, ox, oz, lnx, lnz;
    __device__ __forceinline__ long at(int iz, int ix) const { return (long)(ix - ox - 1) * ld + (iz - oz - 1); }
    __device__ __forceinline__ double &t(int iz, int ix) const { return T[at(iz, ix)]; }
    __device__ __forceinline__ int16_t &n(int iz, int ix) const { return ns[at(iz, ix)]; }
    __device__ __forceinline__ bool local(int iz, int ix) const
    {
        return ix > ox && ix <= ox + lnx && iz > oz && iz <= oz + lnz;
    }
};

// ---- the reference's local updates over an "alive" predicate ---------------------
template <class AL>
__device__ __forceinline__ double ff_fouds1(const fview &V, const AL &al, int iz, int ix, double slown)
{
    int tsw1 = 0;
    double travm = 0.0;
    for (int j = ix - 1; j <= ix + 1; j += 2)
        for (int k = iz - 1; k <= iz + 1; k += 2) {
            if (!(j >= 1 && j <= V.nnx && k >= 1 && k <= V.nnz)) continue;
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0;
            if (al(iz, j)) {
                swsol = 1;
                if (al(k, ix)) {
                    double u = V.dnx, v = V.dnz;
                    double em = V.t(k, ix) - V.t(iz, j);
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = V.t(iz, j);
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (V.dnx * V.dnx));
                    tref = V.t(iz, j);
                }
            } else if (al(k, ix)) {
                swsol = 1;
                a = 1.0; b = 0.0;
                c = -((slown * V.dnz) * (slown * V.dnz));
                tref = V.t(k, ix);
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

template <class AL>
__device__ __forceinline__ double ff_fouds2(const fview &V, const AL &al, int iz, int ix, double slown)
{
    int tsw1 = 0;
    double travm = 0.0;
    for (int j = ix - 1; j <= ix + 1; j += 2) {
        if (!(j >= 1 && j <= V.nnx)) continue;
        int swj = -1, j2;
        if (j == ix - 1) {
            j2 = j - 1;
            if (j2 >= 1 && al(iz, j2)) swj = 0;
        } else {
            j2 = j + 1;
            if (j2 <= V.nnx && al(iz, j2)) swj = 0;
        }
        if (al(iz, j) && swj == 0) {
            swj = -1;
            if (V.t(iz, j) > V.t(iz, j2)) swj = 0;
        } else {
            swj = -1;
        }
        for (int k = iz - 1; k <= iz + 1; k += 2) {
            if (!(k >= 1 && k <= V.nnz)) continue;
            int swk = -1, k2;
            if (k == iz - 1) {
                k2 = k - 1;
                if (k2 >= 1 && al(k2, ix)) swk = 0;
            } else {
                k2 = k + 1;
                if (k2 <= V.nnz && al(k2, ix)) swk = 0;
            }
            if (al(k, ix) && swk == 0) {
                swk = -1;
                if (V.t(k, ix) > V.t(k2, ix)) swk = 0;
            } else {
                swk = -1;
            }
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0, tdiv = 1.0, u, v, em;
            if (swj == 0) {
                swsol = 1;
                if (swk == 0) {
                    u = 2.0 * V.dnx;
                    v = 2.0 * V.dnz;
                    em = 4.0 * V.t(iz, j) - V.t(iz, j2) - 4.0 * V.t(k, ix);
                    em = em + V.t(k2, ix);
                    a = v * v + u * u;
                    b = 2.0 * em * (u * u);
                    c = (u * u) * (em * em - (slown * slown) * (v * v));
                    tref = 4.0 * V.t(iz, j) - V.t(iz, j2);
                    tdiv = 3.0;
                } else if (al(k, ix)) {
                    u = V.dnz;
                    v = 2.0 * V.dnx;
                    em = 3.0 * V.t(k, ix) - 4.0 * V.t(iz, j) + V.t(iz, j2);
                    a = v * v + 9.0 * (u * u);
                    b = 6.0 * em * (u * u);
                    c = (u * u) * (em * em - (slown * slown) * (v * v));
                    tref = V.t(k, ix);
                    tdiv = 1.0;
                } else {
                    u = 2.0 * V.dnx;
                    a = 1.0; b = 0.0;
                    c = -((u * u) * (slown * slown));
                    tref = 4.0 * V.t(iz, j) - V.t(iz, j2);
                    tdiv = 3.0;
                }
            } else if (al(iz, j)) {
                swsol = 1;
                if (swk == 0) {
                    u = V.dnx;
                    v = 2.0 * V.dnz;
                    em = 3.0 * V.t(iz, j) - 4.0 * V.t(k, ix) + V.t(k2, ix);
                    a = v * v + 9.0 * (u * u);
                    b = 6.0 * em * (u * u);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = V.t(iz, j);
                    tdiv = 1.0;
                } else if (al(k, ix)) {
                    u = V.dnx;
                    v = V.dnz;
                    em = V.t(k, ix) - V.t(iz, j);
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = V.t(iz, j);
                    tdiv = 1.0;
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (V.dnx * V.dnx));
                    tref = V.t(iz, j);
                    tdiv = 1.0;
                }
            } else {
                if (swk == 0) {
                    swsol = 1;
                    u = 2.0 * V.dnz;
                    a = 1.0; b = 0.0;
                    c = -((u * u) * (slown * slown));
                    tref = 4.0 * V.t(k, ix) - V.t(k2, ix);
                    tdiv = 3.0;
                } else if (al(k, ix)) {
                    swsol = 1;
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (V.dnz * V.dnz));
                    tref = V.t(k, ix);
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

template <class AL>
__device__ __forceinline__ double ff_fouds(const fview &V, const AL &al, int iz, int ix, double slown, int fom)
{
    return fom == 0 ? ff_fouds1(V, al, iz, ix, slown) : ff_fouds2(V, al, iz, ix, slown);
}

// The same updates on a register copy of the node's axis stencil: [0] the
// x-1 / z-1 side, [1] the x+1 / z+1 side; t1 one node away, t2 two away;
// in* false outside the grid.  Arithmetic identical to ff_fouds1/ff_fouds2.
struct ff_sten {
    double tx1[2], tx2[2], tz1[2], tz2[2];
    bool ix1[2], ix2[2], iz1[2], iz2[2];
};

__host__ __device__ __forceinline__ bool ff_al(double t, bool in, double thr) { return in && t < thr && t < FF_FIN; }

// Alive mask of a stencil for threshold thr: bit s x-1 side / x+1 side one
// away, 2+s two away, 4+s z one away, 6+s z two away.
__host__ __device__ __forceinline__ unsigned ff_mask(const ff_sten &S, double thr)
{
    unsigned m = 0;
    for (int s = 0; s < 2; s++) {
        m |= (unsigned)ff_al(S.tx1[s], S.ix1[s], thr) << s;
        m |= (unsigned)ff_al(S.tx2[s], S.ix2[s], thr) << (2 + s);
        m |= (unsigned)ff_al(S.tz1[s], S.iz1[s], thr) << (4 + s);
        m |= (unsigned)ff_al(S.tz2[s], S.iz2[s], thr) << (6 + s);
    }
    return m;
}
#define FF_AX1(s) ((m >> (s)) & 1u)
#define FF_AX2(s) ((m >> (2 + (s))) & 1u)
#define FF_AZ1(s) ((m >> (4 + (s))) & 1u)
#define FF_AZ2(s) ((m >> (6 + (s))) & 1u)
#define FF_AX1_(m, s) (((m) >> (s)) & 1u)
#define FF_AZ1_(m, s) (((m) >> (4 + (s))) & 1u)

__host__ __device__ double ff_fouds1s(const ff_sten &S, unsigned m, double dnx, double dnz, double slown)
{
    int tsw1 = 0;
    double travm = 0.0;
    for (int sj = 0; sj < 2; sj++)
        for (int sk = 0; sk < 2; sk++) {
            if (!(S.ix1[sj] && S.iz1[sk])) continue;
            int swsol = 0;
            double a = 0, b = 0, c = 0, tref = 0;
            const double tj = S.tx1[sj], tk = S.tz1[sk];
            if (FF_AX1(sj)) {
                swsol = 1;
                if (FF_AZ1(sk)) {
                    double u = dnx, v = dnz;
                    double em = tk - tj;
                    a = u * u + v * v;
                    b = -(2.0 * (u * u) * em);
                    c = (u * u) * (em * em - (v * v) * (slown * slown));
                    tref = tj;
                } else {
                    a = 1.0; b = 0.0;
                    c = -((slown * slown) * (dnx * dnx));
                    tref = tj;
                }
            } else if (FF_AZ1(sk)) {
                swsol = 1;
                a = 1.0; b = 0.0;
                c = -((slown * dnz) * (slown * dnz));
                tref = tk;
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

__host__ __device__ double ff_fouds2s(const ff_sten &S, unsigned m, double dnx, double dnz, double slown)
{
    int tsw1 = 0;
    double travm = 0.0;
    for (int sj = 0; sj < 2; sj++) {
        if (!S.ix1[sj]) continue;
        const double tj = S.tx1[sj], tj2 = S.tx2[sj];
        int swj = -1;
        if (FF_AX2(sj)) swj = 0;
        if (FF_AX1(sj) && swj == 0) {
            swj = -1;
            if (tj > tj2) swj = 0;
        } else {
            swj = -1;
        }
        for (int sk = 0; sk < 2; sk++) {
            if (!S.iz1[sk]) continue;
            const double tk = S.tz1[sk], tk2 = S.tz2[sk];
            int swk = -1;
            if (FF_AZ2(sk)) swk = 0;
            if (FF_AZ1(sk) && swk == 0) {
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
                } else if (FF_AZ1(sk)) {
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
            } else if (FF_AX1(sj)) {
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
                } else if (FF_AZ1(sk)) {
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
                } else if (FF_AZ1(sk)) {
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

template <int FOM>
__host__ __device__ __forceinline__ double ff_cand1s(const ff_sten &S, unsigned m, double dnx, double dnz, double slown)
{
    if (!(m & 0x33u)) return FF_BIG;   // no 4-neighbour alive
    return FOM == 0 ? ff_fouds1s(S, m, dnx, dnz, slown) : ff_fouds2s(S, m, dnx, dnz, slown);
}

// The self-consistent update: the neighbours counted are those below the
// result.  An update over the same alive set is the same number, so the loop
// stops as soon as the set repeats.  The first-order update (FOM 0) reads the
// four nearest neighbours only; its stencil leaves the two-away entries out.
template <int FOM>
__host__ __device__ double ff_cands(const ff_sten &S, double dnx, double dnz, double slown)
{
    unsigned mp = ff_mask(S, FF_THR0);
    double c = ff_cand1s<FOM>(S, mp, dnx, dnz, slown);
    for (int k = 0; k < 8; k++) {
        if (c >= FF_THR0) break;
        unsigned m = ff_mask(S, c);
        if (m == mp) break;                 // the update would return c again
        double c2 = ff_cand1s<FOM>(S, m, dnx, dnz, slown);
        if (c2 == c || c2 >= FF_THR0) break;
        c = c2;
        mp = m;
    }
    return c;
}

// The first-order self-consistent update (ff_cands<0>) with each square root
// and quadratic solve done once per node.  fouds1's candidates for a mask are
// the two-neighbour solution of a quadrant whose x and z neighbours are both
// alive, else the one-neighbour time of its alive neighbour; the masks of the
// later iterations are subsets of the first one (thresholds below FF_THR0), so
// the two-neighbour solutions of the first mask cover them all.  Every value is
// computed by the same operations as in ff_fouds1s, so the result is the same
// number.
struct ff_k0 {
    double uu, vv, a4, a2;    // dnx^2, dnz^2, 4a, 2a with a = dnx^2 + dnz^2
    double dnz;
    double ia2;               // 1 / (2a), exact when a2p
    bool a2p;                 // 2a is a power of two: x / (2a) == x * (1 / (2a)) bit for bit
    bool hsame;               // dnx == dnz, a power of two: the x and z one-neighbour steps are one number
};
__host__ __device__ __forceinline__ ff_k0 ff_k0_make(double dnx, double dnz)
{
    ff_k0 K;
    K.uu = dnx * dnx;
    K.vv = dnz * dnz;
    const double a = K.uu + K.vv;
    K.a4 = 4.0 * a;
    K.a2 = 2.0 * a;
    K.dnz = dnz;
    // a correctly rounded quotient by 2^k is the correctly rounded product with
    // 2^-k (both round the same real number), so the division can go
    int e = 0;
    K.a2p = K.a2 > 0.0 && frexp(K.a2, &e) == 0.5;
    K.ia2 = K.a2p ? ldexp(1.0, 1 - e) : 0.0;
    // with dnx == dnz == 2^k, (slown*dnz)^2 rounds to (slown^2)*dnx^2 exactly
    // (scaling by a power of two commutes with rounding), so hz == hx
    int ex = 0;
    K.hsame = dnx == dnz && frexp(dnx, &ex) == 0.5 && ex > -40 && ex < 40;
    return K;
}

__host__ __device__ __forceinline__ double ff_q2(double tj, double tk, const ff_k0 &K, double ss)
{
    const double em = tk - tj;
    const double b = -(2.0 * K.uu * em);
    const double c = K.uu * (em * em - K.vv * ss);
    double rd1 = b * b - K.a4 * c;
    if (rd1 < 0.0) rd1 = 0.0;
    const double num = -b + sqrt(rd1);
    return tj + (K.a2p ? num * K.ia2 : num / K.a2);
}

__host__ __device__ double ff_cands0f(const ff_sten &S, const ff_k0 &K, double slown)
{
    unsigned mp = ff_mask(S, FF_THR0);
    if (!(mp & 0x33u)) return FF_BIG;
    const double ss = slown * slown;
    // one alive neighbour: a = 1, b = 0, so tdsh = sqrt(-4c) / 2
    const double hx = sqrt(4.0 * (ss * K.uu)) * 0.5;
    const double sdz = slown * K.dnz;
    const double hz = K.hsame ? hx : sqrt(4.0 * (sdz * sdz)) * 0.5;
    double q[2][2];
    for (int sj = 0; sj < 2; sj++)
        for (int sk = 0; sk < 2; sk++)
            q[sj][sk] = (S.ix1[sj] && S.iz1[sk] && FF_AX1_(mp, sj) && FF_AZ1_(mp, sk))
                            ? ff_q2(S.tx1[sj], S.tz1[sk], K, ss) : 0.0;
    auto sel = [&](unsigned m) -> double {
        int any = 0;
        double travm = 0.0;
        for (int sj = 0; sj < 2; sj++)
            for (int sk = 0; sk < 2; sk++) {
                if (!(S.ix1[sj] && S.iz1[sk])) continue;
                double trav;
                if (FF_AX1_(m, sj)) trav = FF_AZ1_(m, sk) ? q[sj][sk] : S.tx1[sj] + hx;
                else if (FF_AZ1_(m, sk)) trav = S.tz1[sk] + hz;
                else continue;
                travm = any ? fmin(trav, travm) : trav;
                any = 1;
            }
        return travm;
    };
    double c = sel(mp);
    for (int k = 0; k < 8; k++) {
        if (c >= FF_THR0) break;
        unsigned m = ff_mask(S, c);
        if (m == mp) break;
        double c2 = (m & 0x33u) ? sel(m) : FF_BIG;
        if (c2 == c || c2 >= FF_THR0) break;
        c = c2;
        mp = m;
    }
    return c;
}

template <int FOM>
__device__ __forceinline__ void ff_load_sten(const fview &V, int iz, int ix, ff_sten &S)
{
    S.ix1[0] = ix - 1 >= 1; S.ix1[1] = ix + 1 <= V.nnx;
    S.ix2[0] = FOM != 0 && ix - 2 >= 1; S.ix2[1] = FOM != 0 && ix + 2 <= V.nnx;
    S.iz1[0] = iz - 1 >= 1; S.iz1[1] = iz + 1 <= V.nnz;
    S.iz2[0] = FOM != 0 && iz - 2 >= 1; S.iz2[1] = FOM != 0 && iz + 2 <= V.nnz;
    S.tx1[0] = S.ix1[0] ? V.t(iz, ix - 1) : FF_BIG;
    S.tx1[1] = S.ix1[1] ? V.t(iz, ix + 1) : FF_BIG;
    S.tx2[0] = S.ix2[0] ? V.t(iz, ix - 2) : FF_BIG;
    S.tx2[1] = S.ix2[1] ? V.t(iz, ix + 2) : FF_BIG;
    S.tz1[0] = S.iz1[0] ? V.t(iz - 1, ix) : FF_BIG;
    S.tz1[1] = S.iz1[1] ? V.t(iz + 1, ix) : FF_BIG;
    S.tz2[0] = S.iz2[0] ? V.t(iz - 2, ix) : FF_BIG;
    S.tz2[1] = S.iz2[1] ? V.t(iz + 2, ix) : FF_BIG;
}

// Gauss-Seidel rounds over anti-diagonals in four directions until a round
// changes no time by more than 1e-14 relative.  fixed(z,x) nodes are kept.
// With the LDS work lists (dirty bitmap over the grid, z fastest, and a flag
// per anti-diagonal of both families, set when one of its nodes is marked)
// only nodes whose stencil changed since their last update are updated, and
// diagonals without marks since their last visit are skipped whole.  A node
// that changes marks the nodes whose stencil holds it (up to two along each
// axis); at the start the open neighbours of fixed nodes are marked.  An update
// of a node whose stencil has not changed returns its time again, so skipping
// it changes nothing: the times are those of full sweeps in the same order.
// Returns the rounds used, or -1 without convergence.
struct ff_work {
    uint32_t *bits;   // nnx*nnz bits
    int *cs, *cd;     // (nnx+nnz-1) flags: sum diagonals (x-1)+(z-1), difference diagonals (nnx-x)+(z-1)
};

// The refined grid's stop (fm2d_ttime.f90:143-163) freezes only nodes below the
// smallest time on an inner edge.  For first-order updates, which only lower a
// node's time, any inner-edge node's current time bounds that stop from above;
// a node whose update lands above the bound is left alone (its time cannot
// reach below the stop, and nodes below the stop depend only on smaller ones),
// so the sweep stops spreading into the refined grid's far corners.
struct ff_edges {
    unsigned long long *bound;       // LDS: the bound's bits (positive doubles order as integers)
    bool left, right, top, bottom;   // which edges are inner edges
};

template <int FOM, bool FAST, class FX, class VL>
__device__ __forceinline__ int ff_sweep(const fview &V, const FX &fixed, const VL &vel, int *sh_flag,
                        const ff_work *W, int *stat = nullptr, const ff_edges *E = nullptr)
{
    const bool prune = FOM == 0 && E != nullptr;
    auto on_edge = [&](int iz, int ix) -> bool {
        return (ix == 1 && E->left) || (ix == V.nnx && E->right) || (iz == 1 && E->top) || (iz == V.nnz && E->bottom);
    };
    const ff_k0 K0 = ff_k0_make(V.dnx, V.dnz);
    int n_steps = 0, n_evals = 0;
    const int nt = blockDim.x, tid = threadIdx.x;
    const int reach = FOM == 0 ? 1 : 2;
    const long nn = (long)V.nnx * V.nnz;
    const int ND = V.nnx + V.nnz - 1;
    auto bit = [&](int iz, int ix) -> long { return (long)(ix - 1) * V.nnz + (iz - 1); };
    auto mark = [&](int iz, int ix) {
        long q = bit(iz, ix);
        atomicOr(&W->bits[q >> 5], 1u << (q & 31));
        W->cs[(ix - 1) + (iz - 1)] = 1;
        W->cd[(V.nnx - ix) + (iz - 1)] = 1;
    };
    if (W) {
        for (long w = tid; w < (nn + 31) / 32; w += nt) W->bits[w] = 0u;
        for (int d = tid; d < ND; d += nt) { W->cs[d] = 0; W->cd[d] = 0; }
        __syncthreads();
        for (long q = tid; q < nn; q += nt) {
            int x = (int)(q / V.nnz) + 1, z = (int)(q % V.nnz) + 1;
            if (fixed(z, x)) continue;
            bool nb = (x > 1 && fixed(z, x - 1)) || (x < V.nnx && fixed(z, x + 1)) ||
                      (z > 1 && fixed(z - 1, x)) || (z < V.nnz && fixed(z + 1, x));
            if (nb) mark(z, x);
        }
        __syncthreads();
    }
    for (int round = 0; round < FF_MAX_ROUNDS; round++) {
        int changed = 0;
        for (int dir = 0; dir < 4; dir++) {
            for (int d = 0; d < ND; d++) {
                if (W) {
                    int *fl = dir == 0 ? &W->cs[d] : dir == 3 ? &W->cs[ND - 1 - d]
                            : dir == 1 ? &W->cd[d] : &W->cd[ND - 1 - d];
                    const int f = *fl;
                    // every wave reads the flag before it is cleared (marks of this
                    // step fall on the neighbouring diagonals, never on this one)
                    if (nt > 64) __syncthreads();
                    if (f == 0) continue;       // nothing to do on this diagonal
                    if (tid == 0) *fl = 0;
                }
                n_steps++;
                int a_lo = d - (V.nnz - 1) > 0 ? d - (V.nnz - 1) : 0;
                int a_hi = d < V.nnx - 1 ? d : V.nnx - 1;
                for (int a = a_lo + tid; a <= a_hi; a += nt) {
                    int b = d - a;
                    int x = (dir & 1) ? V.nnx - a : 1 + a;
                    int z = (dir & 2) ? V.nnz - b : 1 + b;
                    // the node's velocity is read before its work bit, so that the
                    // load is in flight during the LDS reads below (a lane whose bit
                    // is clear drops it)
                    const double vz = vel(z, x);
                    __builtin_amdgcn_sched_barrier(0);
                    if (W) {
                        long q = bit(z, x);
                        uint32_t m = 1u << (q & 31);
                        if (!(W->bits[q >> 5] & m)) continue;
                        atomicAnd(&W->bits[q >> 5], ~m);
                    } else if (fixed(z, x)) {
                        continue;   // with the work list, fixed nodes are never marked
                    }
                    n_evals++;
                    ff_sten S;
                    ff_load_sten<FOM>(V, z, x, S);
                    const double t = V.t(z, x);
                    double slown = 1.0 / vz;
                    double c = (FAST && FOM == 0) ? ff_cands0f(S, K0, slown) : ff_cands<FOM>(S, V.dnx, V.dnz, slown);
                    if (c < FF_THR0 && c != t) {
                        if (prune && c > __longlong_as_double((long long)*E->bound)) continue;
                        if (prune && on_edge(z, x)) atomicMin(E->bound, (unsigned long long)__double_as_longlong(c));
                        if (fabs(c - t) > 1e-14 * c) changed = 1;
                        V.t(z, x) = c;
                        if (W) {
                            for (int r = 1; r <= reach; r++) {
                                if (x - r >= 1 && !fixed(z, x - r)) mark(z, x - r);
                                if (x + r <= V.nnx && !fixed(z, x + r)) mark(z, x + r);
                                if (z - r >= 1 && !fixed(z - r, x)) mark(z - r, x);
                                if (z + r <= V.nnz && !fixed(z + r, x)) mark(z + r, x);
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (tid == 0) *sh_flag = 0;
        __syncthreads();
        if (changed) *sh_flag = 1;
        __syncthreads();
        int any = *sh_flag;
        __syncthreads();
        if (!any) {
            if (stat) { atomicAdd(&stat[1], tid == 0 ? n_steps : 0); atomicAdd(&stat[2], n_evals); }
            return round + 1;
        }
    }
    return -1;
}

// ---- heap on an LDS view (fm2d_ttime.f90:527-680) ---------------------------------
#define HPX(e) ((e) >> 16)
#define HPZ(e) ((e) & 0xffff)

struct ff_heap {
    int32_t *btg;    // 1-based
    int ntr;
    int cap;         // entries the LDS heap holds
    int full;        // an insertion did not fit
};

__device__ __forceinline__ void ff_addtree(const fview &V, ff_heap &H, int iz, int ix)
{
    if (H.ntr + 1 > H.cap) { H.full = 1; return; }
    H.ntr = H.ntr + 1;
    V.n(iz, ix) = (int16_t)H.ntr;
    H.btg[H.ntr] = (ix << 16) | iz;
    int tpc = H.ntr, tpp = tpc / 2;
    const double tv = V.t(iz, ix);
    while (tpp > 0) {
        int ep = H.btg[tpp];
        if (tv < V.t(HPZ(ep), HPX(ep))) {
            V.n(iz, ix) = (int16_t)tpp;
            V.n(HPZ(ep), HPX(ep)) = (int16_t)tpc;
            H.btg[tpp] = (ix << 16) | iz;      // the moving entry: the node's own
            H.btg[tpc] = ep;
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
}

__device__ __forceinline__ void ff_downtree(const fview &V, ff_heap &H)
{
    if (H.ntr == 1) { H.ntr = H.ntr - 1; return; }
    const int el = H.btg[H.ntr];
    V.n(HPZ(el), HPX(el)) = 1;
    H.btg[1] = el;
    H.ntr = H.ntr - 1;
    // the entry moving down is el at every step (each comparison is against it
    // as the parent), so its key is read once and the chosen child's entry and
    // key are the ones just read
    const double kel = V.t(HPZ(el), HPX(el));
    int tpp = 1, tpc = 2 * tpp;
    while (tpc < H.ntr) {
        const int e0 = H.btg[tpc], e1 = H.btg[tpc + 1];
        const double rd1 = V.t(HPZ(e0), HPX(e0));
        const double rd2 = V.t(HPZ(e1), HPX(e1));
        const bool right = rd1 > rd2;
        if (right) tpc = tpc + 1;
        const int ec = right ? e1 : e0;
        const double kc = right ? rd2 : rd1;
        if (kc < kel) {
            V.n(HPZ(el), HPX(el)) = (int16_t)tpc;
            V.n(HPZ(ec), HPX(ec)) = (int16_t)tpp;
            H.btg[tpc] = el;
            H.btg[tpp] = ec;
            tpp = tpc;
            tpc = 2 * tpp;
        } else {
            tpc = H.ntr + 1;
        }
    }
    if (tpc == H.ntr) {
        const int ec = H.btg[tpc];
        if (V.t(HPZ(ec), HPX(ec)) < kel) {
            V.n(HPZ(el), HPX(el)) = (int16_t)tpc;
            V.n(HPZ(ec), HPX(ec)) = (int16_t)tpp;
            H.btg[tpc] = el;
            H.btg[tpp] = ec;
        }
    }
}

__device__ __forceinline__ void ff_updtree(const fview &V, ff_heap &H, int iz, int ix)
{
    int tpc = V.n(iz, ix), tpp = tpc / 2;
    const double tv = V.t(iz, ix);
    while (tpp > 0) {
        int ep = H.btg[tpp];
        if (tv < V.t(HPZ(ep), HPX(ep))) {
            V.n(iz, ix) = (int16_t)tpp;
            V.n(HPZ(ep), HPX(ep)) = (int16_t)tpc;
            H.btg[tpp] = (ix << 16) | iz;      // the moving entry: the node's own
            H.btg[tpc] = ep;
            tpc = tpp;
            tpp = tpc / 2;
        } else {
            tpp = 0;
        }
    }
}

__device__ __forceinline__ double ff_bilinear(const double nv[2][2], double dsx, double dsz, double dnx, double dnz)
{
    double biv = 0.0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            double produ = (1.0 - fabs(((i - 1) * dnx - dsx) / dnx)) * (1.0 - fabs(((j - 1) * dnz - dsz) / dnz));
            biv = biv + nv[i - 1][j - 1] * produ;
        }
    return biv;
}

// Heap-loop stop rules.
struct ff_stop {
    int mode;                 // 0 none, 1 refined inner-edge exit, 2 block frozen, 3 two generations,
                              // 4 the exit rule, then the block rule
    int vnl, vnr, vnt, vnb;   // refined-grid bookkeeping (mode 1)
    int x0, x1, z0, z1;       // block (mode 2)
    int need, have;
    int exited;               // mode 1: exit node reached
    int overflow;             // a node outside the LDS storage would be touched
    uint32_t *gen_bits;       // mode 3: nodes of the two generations (bit per stored node)
    int gen;                  // mode 3: 0 first generation, 1 second
};

__device__ __forceinline__ void ff_gen_set(const fview &V, uint32_t *b, int iz, int ix)
{
    long q = V.at(iz, ix);
    b[q >> 5] |= 1u << (q & 31);
}
__device__ __forceinline__ bool ff_gen_get(const fview &V, const uint32_t *b, int iz, int ix)
{
    long q = V.at(iz, ix);
    return (b[q >> 5] >> (q & 31)) & 1u;
}

// The reference's heap loop over the view (travel's :137-237) with fouds and
// velocities vel(z,x); stops by rule S.  Called by the whole wave: lane 0
// keeps the heap and the stop rule (H and S are lane 0's); the updates of the
// popped node's four neighbours read frozen nodes only, which none of the four
// is, so lanes 0-3 compute them at once and lane 0 then enters them into the
// heap in the reference's order (x-1, x+1, z-1, z+1).
template <class VL>
__device__ __forceinline__ void ff_heap_loop(const fview &V, ff_heap &H, const VL &vel, int fom, ff_stop &S)
{
    const int lane = threadIdx.x;
    auto al = [&](int z, int x) -> bool { return V.n(z, x) == 0; };
    for (;;) {
        int go = 0, ix = 0, iz = 0;
        if (lane == 0) {
            do {
                if (H.ntr <= 0) break;
                if (H.full) { S.overflow = 1; break; }
                int e = H.btg[1];
                ix = HPX(e); iz = HPZ(e);
                if (S.mode == 1 || S.mode == 4) {
                    int swrg = 0;
                    if (ix == 1 && S.vnl != 1) swrg = 1;
                    if (ix == V.nnx && S.vnr != V.nnx) swrg = 1;
                    if (iz == 1 && S.vnt != 1) swrg = 1;
                    if (iz == V.nnz && S.vnb != V.nnz) swrg = 1;
                    if (swrg == 1) { V.n(iz, ix) = 0; S.exited = 1; break; }
                }
                if ((S.mode == 2 || S.mode == 4) && S.have >= S.need) break;
                if (S.mode == 3 && S.have >= S.need) {
                    // every node of the generation is frozen: the second generation is
                    // the band at this moment; after it the heap loop is a plain one
                    if (S.gen == 1 || H.ntr == 0) break;
                    for (int q = 1; q <= H.ntr; q++) ff_gen_set(V, S.gen_bits, HPZ(H.btg[q]), HPX(H.btg[q]));
                    S.need += H.ntr;
                    S.gen = 1;
                    e = H.btg[1];
                    ix = HPX(e); iz = HPZ(e);
                }
                // the updates around a popped node read up to three nodes away
                {
                    int xa = ix - 3 < 1 ? 1 : ix - 3, xb = ix + 3 > V.nnx ? V.nnx : ix + 3;
                    int za = iz - 3 < 1 ? 1 : iz - 3, zb = iz + 3 > V.nnz ? V.nnz : iz + 3;
                    if (!V.local(za, xa) || !V.local(zb, xb)) { S.overflow = 1; break; }
                }
                V.n(iz, ix) = 0;
                if ((S.mode == 2 || S.mode == 4) && ix >= S.x0 && ix <= S.x1 && iz >= S.z0 && iz <= S.z1) S.have++;
                if (S.mode == 3 && ff_gen_get(V, S.gen_bits, iz, ix)) S.have++;
                ff_downtree(V, H);
                go = 1;
            } while (0);
        }
        go = __shfl(go, 0, 64);
        if (!go) break;
        ix = __shfl(ix, 0, 64);
        iz = __shfl(iz, 0, 64);
        __syncthreads();
        int nx = ix, nz = iz;
        if (lane == 0) nx = ix - 1;
        else if (lane == 1) nx = ix + 1;
        else if (lane == 2) nz = iz - 1;
        else nz = iz + 1;
        const bool inb = lane < 4 && nx >= 1 && nx <= V.nnx && nz >= 1 && nz <= V.nnz;
        const int st = inb ? (int)V.n(nz, nx) : 0;
        double tv = 0.0;
        if (st != 0) tv = ff_fouds(V, al, nz, nx, 1.0 / vel(nz, nx), fom);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int stk = __shfl(st, k, 64);
            const double tvk = __shfl(tv, k, 64);
            if (lane == 0 && stk != 0) {
                const int xk = k == 0 ? ix - 1 : (k == 1 ? ix + 1 : ix);
                const int zk = k == 2 ? iz - 1 : (k == 3 ? iz + 1 : iz);
                V.t(zk, xk) = tvk;
                if (stk == -1) ff_addtree(V, H, zk, xk);
                else ff_updtree(V, H, zk, xk);
            }
        }
        __syncthreads();
    }
}

// travel's corner initialisation (fm2d_ttime.f90:105-126) and heap start.
template <class VL>
__device__ __forceinline__ void ff_corners(const fview &V, ff_heap &H, const VL &vel, double scx, double scz,
                           double gox, double goz, int isx, int isz)
{
    double vss[2][2];
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) vss[i - 1][j - 1] = vel(isz - 1 + j, isx - 1 + i);
    double dsx = (scx - gox) - (isx - 1) * V.dnx;
    double dsz = (scz - goz) - (isz - 1) * V.dnz;
    double vsrc = ff_bilinear(vss, dsx, dsz, V.dnx, V.dnz);
    H.ntr = 0;
    for (int i = 1; i <= 2; i++)
        for (int j = 1; j <= 2; j++) {
            double ex = dsx - (i - 1) * V.dnx, ez = dsz - (j - 1) * V.dnz;
            double ds = sqrt(ex * ex + ez * ez);
            V.t(isz - 1 + j, isx - 1 + i) = 2.0 * ds / (vss[i - 1][j - 1] + vsrc);
            ff_addtree(V, H, isz - 1 + j, isx - 1 + i);
        }
}

__device__ __forceinline__ void bspl_w4(double u, double w[4])
{
    w[0] = (1.0 - u) * (1.0 - u) * (1.0 - u) / 6.0;
    w[1] = (4.0 - 6.0 * (u * u) + 3.0 * (u * u * u)) / 6.0;
    w[2] = (1.0 + 3.0 * u + 3.0 * (u * u) - 3.0 * (u * u * u)) / 6.0;
    w[3] = (u * u * u) / 6.0;
}

// Work-group -> problem slot map that keeps neighbouring slots (the sources
// of one period, which read the same velocity grid) on one XCD: work-groups
// are dealt to the 8 XCDs round-robin, so XCD x gets a contiguous slot range.
__device__ __forceinline__ long ff_xcd_slot(long b, long nb)
{
    const long q = nb / 8, r = nb % 8, x = b % 8;
    return x * q + (x < r ? x : r) + b / 8;
}

struct near_geom {
    int s, p, c;
    long prob, slot;
    int isx, isz, vnl, vnr, vnt, vnb;
};

// Problem bookkeeping shared by both kernels; returns false to skip.
__device__ __forceinline__ bool ff_problem(const mct_fm2d_args &A, long slot, near_geom &g, bool report)
{
    g.slot = slot;
    g.prob = A.prob0 + slot;
    g.s = (int)(g.prob % A.nsrc);
    const long cp = g.prob / A.nsrc;
    g.p = (int)(cp % A.np);
    g.c = (int)(cp / A.np);
    if (A.status[g.c] > 0) return false;
    const int32_t *srs = A.srs + (long)g.p * A.srs_pstride + (long)g.s * A.nrc;
    int any = 0;
    for (int r = 0; r < A.nrc; r++) any += srs[r];
    if (any == 0 && g.s != 0) return false;   // modrays skips sources without pairs
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    const double x = A.scx[g.s], z = A.scz[g.s];
    int isx = (int)((x - A.gox) / dnx) + 1, isz = (int)((z - A.goz) / dnz) + 1;
    if (isx < 1 || isx > A.nnx || isz < 1 || isz > A.nnz) {
        if (report) atomicMin(&A.status[g.c], (int)MCT_ERR_OUTSIDE);
        return false;
    }
    if (isx == A.nnx) isx = isx - 1;
    if (isz == A.nnz) isz = isz - 1;
    g.isx = isx; g.isz = isz;
    g.vnl = isx - A.sgs > 1 ? isx - A.sgs : 1;
    g.vnr = isx + A.sgs < A.nnx ? isx + A.sgs : A.nnx;
    g.vnt = isz - A.sgs > 1 ? isz - A.sgs : 1;
    g.vnb = isz + A.sgs < A.nnz ? isz + A.sgs : A.nnz;
    return true;
}

// The coarse neighbourhood of the source held by k_fm2d_near (1-based,
// inclusive): the refined window (or the source cell) +- FF_MARGIN, clamped.
// Fixed (frozen) nodes of the 2-D problem lie in it, nowhere else.
struct ff_win {
    int x0, x1, z0, z1;
    __device__ __forceinline__ bool in(int iz, int ix) const { return ix >= x0 && ix <= x1 && iz >= z0 && iz <= z1; }
    __device__ __forceinline__ int nz() const { return z1 - z0 + 1; }
    __device__ __forceinline__ int nodes() const { return (x1 - x0 + 1) * (z1 - z0 + 1); }
};
// words of a bitmap over the largest such neighbourhood
__host__ __device__ __forceinline__ int ff_win_words(const mct_fm2d_args &A)
{
    const int ln = (A.asgr == 1 ? 2 * A.sgs + 1 : 2) + 2 * FF_MARGIN;
    const int lnx = ln < A.nnx ? ln : A.nnx, lnz = ln < A.nnz ? ln : A.nnz;
    return (lnx * lnz + 31) / 32;
}
__device__ __forceinline__ ff_win ff_window(const mct_fm2d_args &A, const near_geom &g)
{
    ff_win w;
    if (A.asgr == 1) {
        w.x0 = g.vnl - FF_MARGIN; w.x1 = g.vnr + FF_MARGIN; w.z0 = g.vnt - FF_MARGIN; w.z1 = g.vnb + FF_MARGIN;
    } else {
        w.x0 = g.isx - FF_MARGIN; w.x1 = g.isx + 1 + FF_MARGIN; w.z0 = g.isz - FF_MARGIN; w.z1 = g.isz + 1 + FF_MARGIN;
    }
    w.x0 = w.x0 < 1 ? 1 : w.x0; w.z0 = w.z0 < 1 ? 1 : w.z0;
    w.x1 = w.x1 > A.nnx ? A.nnx : w.x1; w.z1 = w.z1 > A.nnz ? A.nnz : w.z1;
    return w;
}

// wave-wide lexicographic minimum of (t, idx)
__device__ __forceinline__ void ff_wave_min(double &t, int &idx)
{
    for (int o = 32; o >= 1; o >>= 1) {
        double t2 = __shfl_xor(t, o, 64);
        int i2 = __shfl_xor(idx, o, 64);
        if (t2 < t || (t2 == t && i2 < idx)) { t = t2; idx = i2; }
    }
}

} // namespace

// srtimes (fm2dray_cartesian.f90:682-797): receiver times from the converged
// coarse field of one problem, one thread per receiver.
template <class VL>
__device__ __forceinline__ void ff_srtimes(const mct_fm2d_args &A, const near_geom &g, const fview &V, const VL &vel)
{
    const double dnx = V.dnx, dnz = V.dnz;
    const double x = A.scx[g.s], z = A.scz[g.s];
    const int32_t *srs = A.srs + (long)g.p * A.srs_pstride + (long)g.s * A.nrc;
    double *out = A.ttime + g.prob * A.nrc;
    for (int r = threadIdx.x + 1; r <= A.nrc; r += blockDim.x) {
        if (srs[r - 1] == 0) continue;
        const double rx = A.rcx[r - 1], rz = A.rcz[r - 1];
        int irx = (int)floor((rx - A.gox) / dnx) + 1;
        int irz = (int)floor((rz - A.goz) / dnz) + 1;
        if (irx < 1 || irx > A.nnx || irz < 1 || irz > A.nnz) { atomicMin(&A.status[g.c], (int)MCT_ERR_OUTSIDE); continue; }
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
                for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = vel(csz - 1 + l, csx - 1 + k);
            double drx = (x - A.gox) - (jsx - 1) * dnx, drz = (z - A.goz) - (jsz - 1) * dnz;
            double vels = ff_bilinear(vss, drx, drz, dnx, dnz);
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) vss[k - 1][l - 1] = vel(irz - 1 + l, irx - 1 + k);
            drx = (rx - A.gox) - (irx - 1) * dnx;
            drz = (rz - A.goz) - (irz - 1) * dnz;
            double velr = ff_bilinear(vss, drx, drz, dnx, dnz);
            trr = 2.0 * sred / (vels + velr);
        } else {
            double drx = (rx - A.gox) - (irx - 1) * dnx;
            double drz = (rz - A.goz) - (irz - 1) * dnz;
            trr = 0.0;
            for (int k = 1; k <= 2; k++)
                for (int l = 1; l <= 2; l++) {
                    double produ = (1.0 - fabs(((l - 1) * dnz - drz) / dnz)) * (1.0 - fabs(((k - 1) * dnx - drx) / dnx));
                    trr = trr + V.t(irz - 1 + l, irx - 1 + k) * produ;
                }
        }
        out[r - 1] = trr;
    }
}

// One wave per problem.  LDS: region (doubles + int16 states) sized for the
// larger of the refined grid and the coarse neighbourhood, the heap, and the
// remap staging of the refined window's coarse nodes.
__global__ __launch_bounds__(64) void k_fm2d_near(mct_fm2d_args A, int region_nodes, int heap_cap,
                                                  int win_nodes, int32_t *pflag)
{
    extern __shared__ double ff_lds[];
    double *RT = ff_lds;
    int16_t *RN = (int16_t *)(RT + region_nodes);
    int32_t *HB = (int32_t *)(RN + ((region_nodes + 3) & ~3));
    double *WT = (double *)(HB + ((heap_cap + 2 + 1) & ~1));
    int16_t *WN = (int16_t *)(WT + win_nodes);
    uint32_t *DB = (uint32_t *)(WN + ((win_nodes + 1) & ~1));
    int *DC = (int *)(DB + (region_nodes + 31) / 32);   // 2 x diag_cap counts
    ff_work RW;
    RW.bits = DB;
    __shared__ int sh_flag, sh_exit, sh_ovf;
    __shared__ double sh_texit;
    __shared__ int sh_eidx, sh_nbl;

    const int lane = threadIdx.x;
    near_geom g;
    const long slot = ff_xcd_slot(blockIdx.x, gridDim.x);
    if (lane == 0 && A.rounds_out) A.rounds_out[4 * slot + 3] = 0;   // set when the check hands it over
    if (!ff_problem(A, slot, g, true)) {
        if (lane == 0) pflag[slot] = 2;   // nothing to do
        return;
    }
    const long wsz = (long)A.ld * A.nc;
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    const double x = A.scx[g.s], z = A.scz[g.s];
    const long cp = g.prob / A.nsrc;
    const double *velnc = A.veln + cp * (long)A.nnx * A.nnz;
    const double *velv = A.velv + cp * (long)(A.nvz + 2) * (A.nvx + 2);
    const int ldv = A.nvz + 2;
    auto cvel = [&](int iz, int ix) -> double { return velnc[(long)(ix - 1) * A.nnz + (iz - 1)]; };
    if (lane == 0) { sh_exit = 0; sh_ovf = 0; }
    __syncthreads();

    // coarse neighbourhood of the source
    const ff_win win = ff_window(A, g);
    const int lx0 = win.x0, lx1 = win.x1, lz0 = win.z0, lz1 = win.z1;
    fview C;
    C.nnx = A.nnx; C.nnz = A.nnz; C.dnx = dnx; C.dnz = dnz;
    C.T = RT; C.ns = RN; C.ox = lx0 - 1; C.oz = lz0 - 1; C.lnx = lx1 - lx0 + 1; C.lnz = lz1 - lz0 + 1; C.ld = C.lnz;
    ff_heap H;
    H.btg = HB;
    H.ntr = 0;
    H.cap = heap_cap;
    H.full = 0;

    if (A.asgr == 1) {
        // ---- refined source grid: bsplrefine, seed, sweep, rebuilt stop ----
        fview R;
        R.nnx = (g.vnr - g.vnl) * A.sgdl + 1;
        R.nnz = (g.vnb - g.vnt) * A.sgdl + 1;
        R.dnx = A.dvx / (double)(float)(A.gdx * A.sgdl);
        R.dnz = A.dvz / (double)(float)(A.gdz * A.sgdl);
        R.T = RT; R.ns = RN; R.ox = 0; R.oz = 0; R.lnx = R.nnx; R.lnz = R.nnz; R.ld = R.nnz;
        const double gorx = A.gox + dnx * (g.vnl - 1), gorz = A.goz + dnz * (g.vnt - 1);
        double *vr = A.vr + slot * wsz;
        const int nrxr = A.gdx * A.sgdl, nrzr = A.gdz * A.sgdl;
        const int origx = (g.vnl - 1) * A.sgdl + 1, origz = (g.vnt - 1) * A.sgdl + 1;
        const int rn = R.nnx * R.nnz;
        for (int q = lane; q < rn; q += 64) {
            int idm2 = q / R.nnz + 1, idm1 = q % R.nnz + 1;
            int st1 = idm1 + origz - 1, st2 = idm2 + origx - 1;   // fine indices
            int ii = (st1 - 1) / nrzr + 1, k = (st1 - 1) % nrzr + 1;
            if (ii > A.nvz - 1) { ii = A.nvz - 1; k = nrzr + 1; }
            int jj = (st2 - 1) / nrxr + 1, l = (st2 - 1) % nrxr + 1;
            if (jj > A.nvx - 1) { jj = A.nvx - 1; l = nrxr + 1; }
            double u = nrxr; u = (l - 1) / u;
            double v = nrzr; v = (k - 1) / v;
            double ui[4], vi[4];
            bspl_w4(u, ui);
            bspl_w4(v, vi);
            double sum[4];
            for (int i1 = 1; i1 <= 4; i1++) {
                sum[i1 - 1] = 0.0;
                for (int j1 = 1; j1 <= 4; j1++)
                    sum[i1 - 1] = sum[i1 - 1] + ui[j1 - 1] * velv[(jj - 2 + j1) * ldv + (ii - 2 + i1)];
                sum[i1 - 1] = vi[i1 - 1] * sum[i1 - 1];
            }
            vr[(long)(idm2 - 1) * A.ld + (idm1 - 1)] = sum[0] + sum[1] + sum[2] + sum[3];
            RN[q] = -1;
            RT[q] = FF_BIG;
        }
        __syncthreads();
        auto rvel = [&](int iz, int ix) -> double { return vr[(long)(ix - 1) * A.ld + (iz - 1)]; };
        // seed: the reference's heap loop until the source cell's corners are frozen
        // and two rings around them (the block rule is checked after the exit rule,
        // as the reference checks its exit first)
        ff_stop S;
        memset(&S, 0, sizeof S);
        if (lane == 0) {
            int risx = (int)((x - gorx) / R.dnx) + 1, risz = (int)((z - gorz) / R.dnz) + 1;
            bool ok = !(risx < 1 || risx > R.nnx || risz < 1 || risz > R.nnz);
            if (ok) {
                if (risx == R.nnx) risx = risx - 1;
                if (risz == R.nnz) risz = risz - 1;
                ff_corners(R, H, rvel, x, z, gorx, gorz, risx, risz);
                S.mode = 4;
                S.vnl = g.vnl; S.vnr = g.vnr; S.vnt = g.vnt; S.vnb = g.vnb;
                S.x0 = risx - FF_SEED_RINGS < 1 ? 1 : risx - FF_SEED_RINGS;
                S.x1 = risx + 1 + FF_SEED_RINGS > R.nnx ? R.nnx : risx + 1 + FF_SEED_RINGS;
                S.z0 = risz - FF_SEED_RINGS < 1 ? 1 : risz - FF_SEED_RINGS;
                S.z1 = risz + 1 + FF_SEED_RINGS > R.nnz ? R.nnz : risz + 1 + FF_SEED_RINGS;
                if (A.fom != 0) {
                    // mixed order: the refined grid's answer depends on the heap's
                    // history (keys rise, nodes pop out of time order), and the
                    // heap-order check covers the coarse grid only, so the heap loop
                    // solves the whole refined grid, up to the reference's exit
                    S.x0 = 1; S.x1 = R.nnx; S.z0 = 1; S.z1 = R.nnz;
                }
                S.need = (S.x1 - S.x0 + 1) * (S.z1 - S.z0 + 1);
            } else {
                sh_ovf = 3;   // source outside the refined grid: use the sequential solver
            }
        }
        __syncthreads();
        if (!sh_ovf) {
            ff_heap_loop(R, H, rvel, A.fom, S);
            if (lane == 0) {
                sh_exit = S.exited ? 1 : 0;
                if (H.full || S.overflow) sh_ovf = 2;
            }
        }
        __syncthreads();
        if (sh_ovf) {
            if (lane == 0) pflag[slot] = 1;
            return;
        }
        if (!sh_exit) {
            // frozen seed nodes are fixed; everything else is solved by sweeping
            for (int q = lane; q < rn; q += 64) {
                if (RN[q] != 0) RT[q] = FF_BIG;
            }
            __syncthreads();
            auto rfixed = [&](int iz, int ix) -> bool { return R.n(iz, ix) == 0; };
            RW.cs = DC; RW.cd = DC + (R.nnx + R.nnz - 1);
            // inner edges as the stop below tests them (fm2d_ttime.f90:143-163)
            __shared__ unsigned long long sh_bound;
            if (lane == 0) sh_bound = (unsigned long long)__double_as_longlong(FF_BIG);
            __syncthreads();
            ff_edges RE;
            RE.bound = &sh_bound;
            RE.left = g.vnl != 1; RE.right = g.vnr != R.nnx; RE.top = g.vnt != 1; RE.bottom = g.vnb != R.nnz;
            int rounds = A.fom == 0 ? ff_sweep<0, true>(R, rfixed, rvel, &sh_flag, &RW, nullptr, &RE)
                                    : ff_sweep<1, true>(R, rfixed, rvel, &sh_flag, &RW);
            if (rounds < 0) {
                if (lane == 0) pflag[slot] = 1;
                return;
            }
            // the stop: the smallest time on an inner edge (first in the reference's scan order)
            double tm = 1e300;
            int im = 0x7fffffff;
            for (int q = lane; q < rn; q += 64) {
                int i = q / R.nnz + 1, j = q % R.nnz + 1;
                // the reference compares the coarse window bounds with the refined
                // grid's size here (fm2d_ttime.f90:143-163); kept as it is
                bool inner = (i == 1 && g.vnl != 1) || (i == R.nnx && g.vnr != R.nnx) ||
                             (j == 1 && g.vnt != 1) || (j == R.nnz && g.vnb != R.nnz);
                if (inner && (RT[q] < tm || (RT[q] == tm && q < im))) { tm = RT[q]; im = q; }
            }
            ff_wave_min(tm, im);
            if (lane == 0) { sh_texit = tm; sh_eidx = tm < 1e300 ? im : -1; }
            __syncthreads();
            const double texit = sh_texit;
            const int eidx = sh_eidx;
            for (int q = lane; q < rn; q += 64) RN[q] = (RT[q] < texit || q == eidx) ? 0 : -1;
            __syncthreads();
            // close nodes and their tentative times: the open nodes with a frozen
            // neighbour other than the exit node.  They are listed first (in the
            // heap's storage, free between the two heap loops) and solved 64 at a
            // time; a node beyond the list's room is solved where it is found.
            // Only frozen nodes' times are read and only open nodes are written,
            // so the order of the solves does not matter.
            const int ex_x = eidx >= 0 ? eidx / R.nnz + 1 : -1, ex_z = eidx >= 0 ? eidx % R.nnz + 1 : -1;
            auto close_tlast = [&](int q) -> double {
                const int i = q / R.nnz + 1, j = q % R.nnz + 1;
                double tlast = -1;
                const int nb[4][2] = {{j, i - 1}, {j, i + 1}, {j - 1, i}, {j + 1, i}};
                for (int w = 0; w < 4; w++) {
                    int zz = nb[w][0], xx = nb[w][1];
                    if (zz < 1 || zz > R.nnz || xx < 1 || xx > R.nnx) continue;
                    if (R.n(zz, xx) == 0 && !(zz == ex_z && xx == ex_x) && R.t(zz, xx) > tlast) tlast = R.t(zz, xx);
                }
                return tlast;
            };
            auto close_solve = [&](int q, double tlast) {
                const int i = q / R.nnz + 1, j = q % R.nnz + 1;
                auto alt = [&](int zz, int xx) -> bool {
                    return R.n(zz, xx) == 0 && !(zz == ex_z && xx == ex_x) && R.t(zz, xx) <= tlast;
                };
                double tv = ff_fouds(R, alt, j, i, 1.0 / rvel(j, i), A.fom);
                RT[q] = tv;
                RN[q] = 1;
            };
            int *BL = HB;
            const int blcap = (heap_cap + 2 + 1) & ~1;
            if (lane == 0) sh_nbl = 0;
            __syncthreads();
            for (int q = lane; q < rn; q += 64) {
                if (RN[q] == 0) continue;
                const double tlast = close_tlast(q);
                if (tlast >= 0) {
                    const int k = atomicAdd(&sh_nbl, 1);
                    if (k < blcap) BL[k] = q;
                    else close_solve(q, tlast);
                }
            }
            __syncthreads();
            const int nbl = sh_nbl < blcap ? sh_nbl : blcap;
            for (int k = lane; k < nbl; k += 64) {
                const int q = BL[k];
                close_solve(q, close_tlast(q));
            }
            __syncthreads();
        } else {
            for (int q = lane; q < rn; q += 64)
                if (RN[q] > 0) RN[q] = 1;
            __syncthreads();
        }
        if (A.rays) {
            // rpaths reads the refined grid after the solve (ttnr / nstsr, :395-400):
            // keep its times and states (0 frozen) in the problem's slot
            double *tr = A.ttr + slot * wsz;
            int32_t *nr = A.nstr + slot * wsz;
            for (int q = lane; q < rn; q += 64) {
                const long o = (long)(q / R.nnz) * A.ld + (q % R.nnz);
                tr[o] = RT[q];
                nr[o] = RN[q];
            }
        }
        // stage the refined values at the window's coarse nodes (remap, :354-416)
        const int wnx = g.vnr - g.vnl + 1, wnz = g.vnb - g.vnt + 1;
        for (int q = lane; q < wnx * wnz; q += 64) {
            int a = q / wnz, b = q % wnz;
            int l = a * A.sgdl + 1, k = b * A.sgdl + 1;
            WN[q] = R.n(k, l);
            WT[q] = R.t(k, l);
        }
        __syncthreads();
        // coarse neighbourhood: far everywhere, then the remapped window
        for (int q = lane; q < C.lnx * C.lnz; q += 64) { RN[q] = -1; RT[q] = FF_BIG; }
        __syncthreads();
        for (int q = lane; q < wnx * wnz; q += 64) {
            int a = q / wnz, b = q % wnz;
            int idm2 = g.vnl + a, idm1 = g.vnt + b;
            int st = WN[q];
            C.n(idm1, idm2) = (int16_t)st;
            if (st >= 0) C.t(idm1, idm2) = WT[q];
        }
        __syncthreads();
        // frozen nodes next to far nodes re-enter the band
        for (int q = lane; q < C.lnx * C.lnz; q += 64) {
            int k = C.ox + q / C.lnz + 1, l = C.oz + q % C.lnz + 1;
            if (C.n(l, k) != 0) continue;
            bool nb = (l - 1 >= 1 && C.n(l - 1, k) == -1) || (l + 1 <= A.nnz && C.n(l + 1, k) == -1) ||
                      (k - 1 >= 1 && C.n(l, k - 1) == -1) || (k + 1 <= A.nnx && C.n(l, k + 1) == -1);
            if (nb) C.n(l, k) = 1;
        }
        __syncthreads();
        // the reference's coarse heap loop over two generations: the remapped
        // band (with the re-banded frozen nodes), then the band at the moment
        // the last of those froze; after that every update is a plain one
        for (int w = lane; w < (C.lnx * C.lnz + 31) / 32; w += 64) DB[w] = 0u;
        __syncthreads();
        memset(&S, 0, sizeof S);
        if (lane == 0) {
            S.mode = 3;
            S.gen_bits = DB;
            H.ntr = 0;
            H.full = 0;
        }
        {
            // the band's nodes enter the heap in scan order (x outer, z inner): the
            // wave tests 64 nodes at a time and lane 0 adds the band nodes among
            // them in lane order (adding a node changes only the states of nodes
            // already in the heap, which stay positive, so the tests stand)
            const int cz = lz1 - lz0 + 1, tot = (lx1 - lx0 + 1) * cz;
            for (int b0 = 0; b0 < tot; b0 += 64) {
                const int q = b0 + lane;
                const int i = lx0 + q / cz, j = lz0 + q % cz;
                const bool band = q < tot && C.n(j, i) > 0;
                unsigned long long m = __ballot(band);
                while (m) {
                    const int l = __builtin_ctzll(m);
                    m &= m - 1;
                    const int il = __shfl(i, l, 64), jl = __shfl(j, l, 64);
                    if (lane == 0) {
                        ff_addtree(C, H, jl, il);
                        ff_gen_set(C, DB, jl, il);
                        S.need++;
                    }
                }
            }
        }
        __syncthreads();
        ff_heap_loop(C, H, cvel, A.fom, S);
        if (lane == 0) sh_ovf = S.overflow || H.full;
        __syncthreads();
    } else {
        // ---- no refined grid: seed on the coarse grid ----
        for (int q = lane; q < C.lnx * C.lnz; q += 64) { RN[q] = -1; RT[q] = FF_BIG; }
        __syncthreads();
        ff_stop S;
        memset(&S, 0, sizeof S);
        if (lane == 0) {
            ff_corners(C, H, cvel, x, z, A.gox, A.goz, g.isx, g.isz);
            S.mode = 2;
            S.x0 = g.isx - FF_SEED_RINGS < 1 ? 1 : g.isx - FF_SEED_RINGS;
            S.x1 = g.isx + 1 + FF_SEED_RINGS > A.nnx ? A.nnx : g.isx + 1 + FF_SEED_RINGS;
            S.z0 = g.isz - FF_SEED_RINGS < 1 ? 1 : g.isz - FF_SEED_RINGS;
            S.z1 = g.isz + 1 + FF_SEED_RINGS > A.nnz ? A.nnz : g.isz + 1 + FF_SEED_RINGS;
            S.need = (S.x1 - S.x0 + 1) * (S.z1 - S.z0 + 1);
            S.have = 0;
        }
        __syncthreads();
        ff_heap_loop(C, H, cvel, A.fom, S);
        if (lane == 0) sh_ovf = S.overflow || H.full;
        __syncthreads();
    }
    if (A.verify > 0 && lane == 0 && !sh_ovf) {
        // the band handed to the sweep must be a valid heap: fouds overwrites a
        // key and updtree only sifts up, so a raised key can leave it out of
        // order, and then the pops that follow are not in time order
        for (int q = 2; q <= H.ntr; q++) {
            const int ec = H.btg[q], ep = H.btg[q / 2];
            if (C.t(HPZ(ec), HPX(ec)) < C.t(HPZ(ep), HPX(ep))) { sh_ovf = 4; break; }
        }
    }
    __syncthreads();
    if (sh_ovf) {
        if (lane == 0) {
            pflag[slot] = 1;
            if (A.rounds_out && sh_ovf == 4) A.rounds_out[4 * slot + 3] = FF_VR_HEAP;
        }
        return;
    }
    // write the coarse grid: frozen nodes with their times, the rest open; the
    // fixed flags only in the neighbourhood (the sweeps read them nowhere else)
    double *T = A.ttn + slot * wsz;
    int32_t *F = A.nsts + slot * wsz;
    for (int k = 1; k <= A.nnx; k++) {
        double *Tc = T + (long)(k - 1) * A.ld;
        const bool wk = k > C.ox && k <= C.ox + C.lnx;
        for (int l = lane + 1; l <= A.nnz; l += 64)
            Tc[l - 1] = (wk && l > C.oz && l <= C.oz + C.lnz && C.n(l, k) == 0) ? C.t(l, k) : FF_BIG;
    }
    for (int q = lane; q < C.lnx * C.lnz; q += 64) {
        int k = C.ox + q / C.lnz + 1, l = C.oz + q % C.lnz + 1;
        F[(long)(k - 1) * A.ld + (l - 1)] = C.n(l, k) == 0 ? 1 : 0;
    }
    if (A.verify > 0) {
        // the band's keys at the hand-over, for k_fm2d_verify (the refined
        // velocities in the slot's vr are no longer read)
        double *Wb = A.vr + slot * wsz;
        for (int q = lane; q < C.lnx * C.lnz; q += 64) {
            int k = C.ox + q / C.lnz + 1, l = C.oz + q % C.lnz + 1;
            Wb[(long)(k - 1) * A.ld + (l - 1)] = C.n(l, k) > 0 ? C.t(l, k) : FF_BIG;
        }
    }
    if (lane == 0) pflag[slot] = 0;
}

// One workgroup per problem: sweep the coarse grid, then pick the receivers.
template <int FOM>
__global__ __launch_bounds__(256) void k_fm2d_sweep(mct_fm2d_args A, int32_t *pflag, int use_dirty)
{
    extern __shared__ uint32_t ff_dirty[];
    __shared__ int sh_flag;
    const long slot = ff_xcd_slot(blockIdx.x, gridDim.x);
    if (pflag[slot] != 0) return;
    near_geom g;
    if (!ff_problem(A, slot, g, false)) return;
    const long wsz = (long)A.ld * A.nc;
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    const long cp = g.prob / A.nsrc;
    const double *velnc = A.veln + cp * (long)A.nnx * A.nnz;
    fview V;
    V.nnx = A.nnx; V.nnz = A.nnz; V.dnx = dnx; V.dnz = dnz;
    V.T = A.ttn + slot * wsz; V.ns = nullptr; V.ox = 0; V.oz = 0; V.lnx = A.nnx; V.lnz = A.nnz; V.ld = A.ld;
    const int32_t *F = A.nsts + slot * wsz;
    const ff_win win = ff_window(A, g);      // fixed flags are written there only
    // the fixed mask as an LDS bitmap (z fastest) when the work lists are in LDS
    const long nn = (long)A.nnx * A.nnz;
    uint32_t *FB = use_dirty ? ff_dirty + (nn + 31) / 32 + 2 * (A.nnx + A.nnz) : nullptr;
    if (FB) {
        for (long w = threadIdx.x; w < (nn + 31) / 32; w += blockDim.x) {
            uint32_t m = 0;
            for (int b = 0; b < 32; b++) {
                long q = w * 32 + b;
                if (q < nn) {
                    int x = (int)(q / A.nnz) + 1, z = (int)(q % A.nnz) + 1;
                    if (win.in(z, x) && F[(long)(x - 1) * A.ld + (z - 1)] != 0) m |= 1u << b;
                }
            }
            FB[w] = m;
        }
        __syncthreads();
    }
    auto fixed = [&](int iz, int ix) -> bool {
        if (FB) {
            long q = (long)(ix - 1) * A.nnz + (iz - 1);
            return (FB[q >> 5] >> (q & 31)) & 1u;
        }
        return win.in(iz, ix) && F[(long)(ix - 1) * A.ld + (iz - 1)] != 0;
    };
    auto vel = [&](int iz, int ix) -> double { return velnc[(long)(ix - 1) * A.nnz + (iz - 1)]; };
    ff_work W;
    W.bits = ff_dirty;
    W.cs = (int *)(ff_dirty + ((long)A.nnx * A.nnz + 31) / 32);
    W.cd = W.cs + (A.nnx + A.nnz - 1);
    int *stat = A.rounds_out ? A.rounds_out + 4 * slot : nullptr;
    if (stat && threadIdx.x == 0) { stat[1] = 0; stat[2] = 0; }
    __syncthreads();
    int rounds = ff_sweep<FOM, false>(V, fixed, vel, &sh_flag, use_dirty ? &W : nullptr, stat);
    if (rounds < 0) {
        // the mixed-order update need not settle (it is not monotone): with the
        // heap-order check on, the sequential solver takes the problem
        if (threadIdx.x == 0) {
            if (A.verify > 0) {
                pflag[slot] = 1;
                if (A.rounds_out) A.rounds_out[4 * slot + 3] = FF_VR_NOCONV;
            } else {
                atomicMin(&A.status[g.c], (int)MCT_ERR_NOCONV);
            }
        }
        return;
    }
    if (threadIdx.x == 0 && stat) stat[0] = rounds;
    ff_srtimes(A, g, V, vel);
}

// ---- strip-pipelined sweep -------------------------------------------------------
// The Gauss-Seidel sweeps of ff_sweep in another visiting order that reads the
// same values.  A sweep crosses the grid in strips of 64 columns along its x
// direction; in a strip lane i owns column x_i and at step t updates the node
// b = t - i along the sweep's z direction.  Node (x_i, b) so comes after
// (x_i, b-1) and (x_i, b-2) (same lane, earlier steps) and after (x_{i-1}, b)
// and (x_{i-2}, b) (lanes i-1, i-2 one and two steps earlier, or an earlier
// strip), and before (x_i, b+1), (x_i, b+2), (x_{i+1}, b), (x_{i+2}, b): the
// same relations to its stencil as in the anti-diagonal order, so every update
// reads the values it reads there, and the dirty marks fall as they do there.
// Each lane keeps a window of its own column in registers (new times behind b,
// old times ahead), takes its x neighbours' times from the neighbouring lanes,
// and streams its column's old times and velocities from memory FF_PF to
// 2*FF_PF steps ahead; the strip's outer neighbour columns are staged in LDS.
// A strip without dirty nodes is skipped.
#define FF_PF 4

// Lane i takes lane i-1's value (wave_shr:1) / lane i+1's (wave_shl:1) by DPP
// moves, two per double; a lane without a source keeps its own value, as
// __shfl_up / __shfl_down do.  Called with every lane of the wave active.
__device__ __forceinline__ double ff_from_below(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double ff_from_above(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)b, (int)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(b >> 32), (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int FOM>
__global__ __launch_bounds__(64) void k_fm2d_strip(mct_fm2d_args A, int32_t *pflag)
{
    extern __shared__ double ff_sld[];
    const long slot = ff_xcd_slot(blockIdx.x, gridDim.x);
    if (pflag[slot] != 0) return;
    near_geom g;
    if (!ff_problem(A, slot, g, false)) return;
    const int lane = threadIdx.x;
    const int nnx = A.nnx, nnz = A.nnz, ld = A.ld;
    const long wsz = (long)ld * A.nc;
    const double dnx = A.dvx / A.gdx, dnz = A.dvz / A.gdz;
    const long cp = g.prob / A.nsrc;
    const double *velnc = A.veln + cp * (long)nnx * nnz;
    const double *slnc = A.sln ? A.sln + cp * (long)nnx * nnz : nullptr;
    double *T = A.ttn + slot * wsz;
    const int32_t *F = A.nsts + slot * wsz;
    const long nn = (long)nnx * nnz;
    const int nw = (int)((nn + 31) / 32);
    const int ns = (nnx + 63) / 64;
    constexpr int NSIDE = FOM == 0 ? 2 : 4;
    constexpr int reach = FOM == 0 ? 1 : 2;
    const ff_k0 K0 = ff_k0_make(dnx, dnz);
    // fixed nodes lie in the near-source neighbourhood: a bitmap of it only
    const ff_win win = ff_window(A, g);
    const int fw = ff_win_words(A);
    uint32_t *DB = (uint32_t *)ff_sld;        // dirty bits, z fastest
    uint32_t *FB = DB + nw;                   // fixed bits of the neighbourhood, z fastest
    int *cnt = (int *)(FB + fw + ((nw + fw) & 1));   // per strip of 64 columns: marked since its last visit
    double *side = (double *)(cnt + ((ns + 1) & ~1));   // staged columns L1, R1, L2, R2 (nnz each)

    for (int w = lane; w < nw; w += 64) DB[w] = 0u;
    const int wn = win.nodes(), wnz = win.nz();
    for (int w = lane; w < fw; w += 64) {
        uint32_t m = 0;
        for (int b = 0; b < 32; b++) {
            const int q = w * 32 + b;
            if (q < wn) {
                const int x = win.x0 + q / wnz, z = win.z0 + q % wnz;
                if (F[(long)(x - 1) * ld + (z - 1)] != 0) m |= 1u << b;
            }
        }
        FB[w] = m;
    }
    for (int k = lane; k < ns; k += 64) cnt[k] = 0;
    __syncthreads();
    auto fixed = [&](int z, int x) -> bool {       // 1-based
        if (!win.in(z, x)) return false;
        const int q = (x - win.x0) * wnz + (z - win.z0);
        return (FB[q >> 5] >> (q & 31)) & 1u;
    };
    auto mark = [&](int z, int x) {
        long q = (long)(x - 1) * nnz + (z - 1);
        atomicOr(&DB[q >> 5], 1u << (q & 31));
        cnt[(x - 1) >> 6] = 1;
    };
    {
        // the open neighbours of fixed nodes: within the neighbourhood and one around it
        const int ex0 = win.x0 > 1 ? win.x0 - 1 : 1, ex1 = win.x1 < nnx ? win.x1 + 1 : nnx;
        const int ez0 = win.z0 > 1 ? win.z0 - 1 : 1, ez1 = win.z1 < nnz ? win.z1 + 1 : nnz;
        const int enz = ez1 - ez0 + 1, en = (ex1 - ex0 + 1) * enz;
        for (int q = lane; q < en; q += 64) {
            const int x = ex0 + q / enz, z = ez0 + q % enz;
            if (fixed(z, x)) continue;
            bool nb = (x > 1 && fixed(z, x - 1)) || (x < nnx && fixed(z, x + 1)) ||
                      (z > 1 && fixed(z - 1, x)) || (z < nnz && fixed(z + 1, x));
            if (nb) mark(z, x);
        }
    }
    __syncthreads();

    int n_steps = 0, n_evals = 0, rounds = -1;
    for (int round = 0; round < FF_MAX_ROUNDS && rounds < 0; round++) {
        int changed = 0;
        for (int dir = 0; dir < 4; dir++) {
            const int xs = (dir & 1) ? -1 : 1, zs = (dir & 2) ? -1 : 1;
            for (int si = 0; si < ns; si++) {
                const int s = xs > 0 ? si : ns - 1 - si;
                __syncthreads();
                if (cnt[s] == 0) continue;
                cnt[s] = 0;     // marks of this strip's steps set it again
                const int xl0 = xs > 0 ? 64 * s + 1 : 64 * s + 64;      // lane 0's column
                const int xl63 = xs > 0 ? xl0 + 63 : xl0 - 63;
                const int x = xs > 0 ? xl0 + lane : xl0 - lane;
                const bool xin = x >= 1 && x <= nnx;
                const long colT = (long)(x - 1) * ld, colV = (long)(x - 1) * nnz;
                // the strip's marked nodes as pipeline steps t = b + lane: the visit runs
                // from the first to the last of them, and on while its changes mark
                // nodes further on
                int tl = 0x7fffffff, th = -1;
                if (xin) {
                    int zlo = -1, zhi = -1;
                    const long q0 = colV, q1 = colV + nnz;
                    for (long w = q0 >> 5; w <= (q1 - 1) >> 5; w++) {
                        uint32_t v = DB[w];
                        if (w == (q0 >> 5)) v &= ~0u << (q0 & 31);
                        if (w == ((q1 - 1) >> 5) && (q1 & 31)) v &= ~(~0u << (q1 & 31));
                        if (v) {
                            const int lo = (int)(w * 32 - q0) + __ffs((int)v) - 1;
                            const int hi = (int)(w * 32 - q0) + 31 - __clz((int)v);
                            if (zlo < 0) zlo = lo;
                            zhi = hi;
                        }
                    }
                    if (zlo >= 0) {
                        tl = (zs > 0 ? zlo : nnz - 1 - zhi) + lane;
                        th = (zs > 0 ? zhi : nnz - 1 - zlo) + lane;
                    }
                }
                for (int o = 32; o >= 1; o >>= 1) {
                    const int tl2 = __shfl_xor(tl, o, 64), th2 = __shfl_xor(th, o, 64);
                    tl = tl2 < tl ? tl2 : tl;
                    th = th2 > th ? th2 : th;
                }
                if (th < 0) continue;       // no marked node left in the strip
                {
                    const int cx[4] = {xl0 - xs, xl63 + xs, xl0 - 2 * xs, xl63 + 2 * xs};
                    for (int c = 0; c < NSIDE; c++) {
                        const bool ok = cx[c] >= 1 && cx[c] <= nnx;
                        for (int k = lane; k < nnz; k += 64)
                            side[c * nnz + k] = ok ? T[(long)(cx[c] - 1) * ld + k] : FF_BIG;
                    }
                }
                __syncthreads();
                const int first = xs > 0 ? 0 : (xl0 > nnx ? xl0 - nnx : 0);
                const int last = xs > 0 ? (nnx - xl0 < 63 ? nnx - xl0 : 63) : 63;
                const int tmax = nnz - 1 + last;
                const int tstart = tl > first ? tl : first;
                int tend = th < tmax ? th : tmax;
                // The stream's loads have no condition: a lane whose node is outside
                // the grid reads the slot's first element (one line that every such
                // lane of the instruction shares), and the value is replaced where it
                // is used, so the loads stay in flight until then (a load under a
                // condition would be waited for at once).
                auto inb = [&](int bb) -> bool { return xin && bb >= 0 && bb < nnz; };
                auto zrow = [&](int bb) -> int { return zs > 0 ? bb : nnz - 1 - bb; };
                auto tload = [&](int bb) -> double {
                    double v = 0.0;
                    if (inb(bb)) v = T[colT + zrow(bb)];
                    return v;
                };
                auto told = [&](int bb) -> double {
                    const double v = tload(bb);
                    return inb(bb) ? v : FF_BIG;
                };
                // the node's slowness 1.0 / veln (fouds' slown); used only where the
                // node is in the column
                auto vload = [&](int bb) -> double {
                    double v = 1.0;
                    if (inb(bb)) v = slnc[colV + zrow(bb)];
                    return v;
                };
                constexpr int D = 2 + FOM;      // the stream delivers the old time D nodes ahead
                const int b0 = tstart - lane;
                // behind the start nothing has changed in this visit: the stored times
                double pn1 = told(b0 - 1), pn2 = FOM ? told(b0 - 2) : FF_BIG;
                double c0 = told(b0), o1 = told(b0 + 1), o2 = FOM ? told(b0 + 2) : 0.0;
                double sa[FF_PF], sb[FF_PF], va[FF_PF], vb[FF_PF];
#pragma unroll
                for (int j = 0; j < FF_PF; j++) { sa[j] = tload(b0 + j + D); va[j] = vload(b0 + j); }
                // the window and the first chunk in registers before the step loop, so the
                // loop's waits are for the chunk in flight only
                __builtin_amdgcn_s_waitcnt(0);
                auto step = [&](int t, double snext_raw, double scur) {
                    const int b = t - lane;
                    const double snext = inb(b + D) ? snext_raw : FF_BIG;
                    const bool act = xin && b >= 0 && b < nnz;
                    const int z0 = zs > 0 ? b : nnz - 1 - b;              // 0-based z
                    // lanes 0/63 (and 1/62 for the second neighbours) take these from LDS below
                    double xm1 = ff_from_below(pn1), xp1 = ff_from_above(o1);
                    double xm2 = 0.0, xp2 = 0.0;
                    if (FOM) { xm2 = ff_from_below(ff_from_below(pn2)); xp2 = ff_from_above(ff_from_above(o2)); }
                    double newv = c0;
                    if (act) {
                        const long q = colV + z0;
                        const uint32_t m = 1u << (q & 31);
                        if (DB[q >> 5] & m) {        // fixed nodes are never marked
                            atomicAnd(&DB[q >> 5], ~m);
                            n_evals++;
                            if (lane == 0) xm1 = side[z0];
                            if (lane == 63) xp1 = side[nnz + z0];
                            if (FOM) {
                                if (lane == 0) xm2 = side[2 * nnz + z0];
                                if (lane == 1) xm2 = side[z0];
                                if (lane == 63) xp2 = side[3 * nnz + z0];
                                if (lane == 62) xp2 = side[nnz + z0];
                            }
                            // the stencil in grid orientation: [0] the x-1 / z-1 side
                            ff_sten S;
                            S.ix1[0] = x - 1 >= 1; S.ix1[1] = x + 1 <= nnx;
                            S.iz1[0] = z0 >= 1; S.iz1[1] = z0 + 1 < nnz;
                            S.ix2[0] = FOM != 0 && x - 2 >= 1; S.ix2[1] = FOM != 0 && x + 2 <= nnx;
                            S.iz2[0] = FOM != 0 && z0 >= 2; S.iz2[1] = FOM != 0 && z0 + 2 < nnz;
                            const double lx1 = xs > 0 ? xm1 : xp1, hx1 = xs > 0 ? xp1 : xm1;
                            const double lz1 = zs > 0 ? pn1 : o1, hz1 = zs > 0 ? o1 : pn1;
                            S.tx1[0] = S.ix1[0] ? lx1 : FF_BIG; S.tx1[1] = S.ix1[1] ? hx1 : FF_BIG;
                            S.tz1[0] = S.iz1[0] ? lz1 : FF_BIG; S.tz1[1] = S.iz1[1] ? hz1 : FF_BIG;
                            if (FOM) {
                                const double lx2 = xs > 0 ? xm2 : xp2, hx2 = xs > 0 ? xp2 : xm2;
                                const double lz2 = zs > 0 ? pn2 : o2, hz2 = zs > 0 ? o2 : pn2;
                                S.tx2[0] = S.ix2[0] ? lx2 : FF_BIG; S.tx2[1] = S.ix2[1] ? hx2 : FF_BIG;
                                S.tz2[0] = S.iz2[0] ? lz2 : FF_BIG; S.tz2[1] = S.iz2[1] ? hz2 : FF_BIG;
                            } else {
                                S.tx2[0] = S.tx2[1] = S.tz2[0] = S.tz2[1] = FF_BIG;
                            }
                            const double c = FOM == 0 ? ff_cands0f(S, K0, scur)
                                                      : ff_cands<FOM>(S, dnx, dnz, scur);
                            if (c < FF_THR0 && c != c0) {
                                if (fabs(c - c0) > 1e-14 * c) changed = 1;
                                newv = c;
                                T[colT + z0] = c;
                                const int z = z0 + 1;
                                for (int r = 1; r <= reach; r++) {
                                    if (x - r >= 1 && !fixed(z, x - r)) mark(z, x - r);
                                    if (x + r <= nnx && !fixed(z, x + r)) mark(z, x + r);
                                    if (z - r >= 1 && !fixed(z - r, x)) mark(z - r, x);
                                    if (z + r <= nnz && !fixed(z + r, x)) mark(z + r, x);
                                }
                            }
                        }
                    }
                    if (__ballot(newv != c0) != 0ull) {
                        const int e = t + reach < tmax ? t + reach : tmax;
                        tend = e > tend ? e : tend;
                    }
                    pn2 = pn1; pn1 = newv; c0 = o1;
                    if (FOM) { o1 = o2; o2 = snext; } else { o1 = snext; }
                };
                // each pass issues the loads of the next FF_PF steps, then runs its
                // own FF_PF steps on the values the previous pass loaded (static
                // register indices, so the waits count only the loads it needs)
                for (int t0 = tstart; t0 <= tend; t0 += FF_PF) {
#pragma unroll
                    for (int j = 0; j < FF_PF; j++) {
                        sb[j] = tload(t0 + FF_PF + j - lane + D);
                        vb[j] = vload(t0 + FF_PF + j - lane);
                    }
#pragma unroll
                    for (int j = 0; j < FF_PF; j++)
                        if (t0 + j <= tend) step(t0 + j, sa[j], va[j]);
#pragma unroll
                    for (int j = 0; j < FF_PF; j++) { sa[j] = sb[j]; va[j] = vb[j]; }
                }
                if (lane == 0) n_steps += tend - tstart + 1;
                // the flag stays set only if the strip still holds marked nodes (marks
                // of its own steps ahead of the pipeline were taken in this visit)
                bool left = false;
                if (xin) {
                    const long q0 = colV, q1 = colV + nnz;
                    for (long w = q0 >> 5; w <= (q1 - 1) >> 5; w++) {
                        uint32_t v = DB[w];
                        if (w == (q0 >> 5)) v &= ~0u << (q0 & 31);
                        if (w == ((q1 - 1) >> 5) && (q1 & 31)) v &= ~(~0u << (q1 & 31));
                        left |= v != 0u;
                    }
                }
                if (__ballot(left) == 0ull) cnt[s] = 0;
            }
        }
        if (__ballot(changed) == 0ull) rounds = round + 1;
    }
    __syncthreads();
    int *stat = A.rounds_out ? A.rounds_out + 4 * slot : nullptr;
    if (stat) {
        if (lane == 0) { stat[0] = rounds; stat[1] = n_steps; }
        int tot = n_evals;
        for (int o = 32; o >= 1; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (lane == 0) stat[2] = tot;
    }
    if (rounds < 0) {
        if (lane == 0) {
            if (A.verify > 0) {
                pflag[slot] = 1;
                if (A.rounds_out) A.rounds_out[4 * slot + 3] = FF_VR_NOCONV;
            } else {
                atomicMin(&A.status[g.c], (int)MCT_ERR_NOCONV);
            }
        }
        return;
    }
    fview V;
    V.nnx = nnx; V.nnz = nnz; V.dnx = dnx; V.dnz = dnz;
    V.T = T; V.ns = nullptr; V.ox = 0; V.oz = 0; V.lnx = nnx; V.lnz = nnz; V.ld = ld;
    auto vel = [&](int iz, int ix) -> double { return velnc[(long)(ix - 1) * nnz + (iz - 1)]; };
    ff_srtimes(A, g, V, vel);
}

// ---- the heap loop's answer, checked -------------------------------------------
// The sweep's field T satisfies T(n) = fouds(n; alive = {T < T(n)}).  The
// reference's heap loop gives the same field if, from the hand-over on, it
// freezes the open nodes in the order of T.  Its key for an open node n is
// the band key W(n) at the hand-over (or none), then the update after each
// of n's open neighbours d_1 < d_2 < ... < d_r (the ones with T below T(n))
// freezes, over the alive set then: the fixed nodes and the open nodes with
// T <= T(d_i).  Call those keys V_0 = W(n), V_1, ..., V_r.  If for every open
// node
//   (a) V_i <= V_{i-1}: keys only fall, so updtree's sift-up keeps the heap
//       valid (fouds overwrites a key, it may rise);
//   (b) V_{i-1} > T(d_i): n is not popped before the neighbour that updates
//       it next;
//   (c) T(n) differs from every neighbour's T, and the T(d_i) from each
//       other: the freeze order is not left to the heap's tie positions;
//   (e) V_r = T(n);
// and the band at the hand-over is a valid heap (checked by k_fm2d_near),
// then by induction over T the heap loop pops the open nodes in the order of
// T, each with key T(n): the node with the smallest T of the unfrozen ones
// holds key T(n) by (e), every other one a larger key by (b) and (c).  So a
// problem that passes has the reference's times; one that fails goes to the
// sequential solver.  The sweep stops at 1e-14 relative change, so (a) and
// (e) compare within 1e-12 relative.
template <int FOM, class FX>
__device__ __forceinline__ int ff_verify_node(const fview &V, const FX &fixed, double w0, int z, int x,
                                              double slown)
{
    const double tn = V.t(z, x);
    if (!(tn < FF_FIN)) return FF_VR_NOTIME;
    double d[4];
    int r = 0;
    const int nz_[4] = {z, z, z - 1, z + 1}, nx_[4] = {x - 1, x + 1, x, x};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int zz = nz_[k], xx = nx_[k];
        if (zz < 1 || zz > V.nnz || xx < 1 || xx > V.nnx) continue;
        const double t = V.t(zz, xx);
        if (t == tn) return FF_VR_TIE;
        if (t < tn && !fixed(zz, xx)) {
            int i = r++;
            while (i > 0 && d[i - 1] > t) { d[i] = d[i - 1]; i--; }
            d[i] = t;
        }
    }
    for (int i = 1; i < r; i++)
        if (d[i] == d[i - 1]) return FF_VR_TIE;
    double v = w0;
    for (int i = 0; i < r; i++) {
        if (!(v > d[i])) return FF_VR_EARLY;
        const double thr = d[i];
        auto al = [&](int zz, int xx) -> bool { return fixed(zz, xx) || V.t(zz, xx) <= thr; };
        const double vn = FOM == 0 ? ff_fouds1(V, al, z, x, slown) : ff_fouds2(V, al, z, x, slown);
        if (vn > v && vn - v > 1e-12 * vn) return FF_VR_RISE;
        v = vn;
    }
    return (v == tn || fabs(v - tn) <= 1e-12 * tn) ? 0 : FF_VR_FINAL;
}

// One workgroup per problem over the converged coarse field; a problem that
// fails the check is flagged for the sequential solver (pflag 1).
template <int FOM>
__global__ __launch_bounds__(256) void k_fm2d_verify(mct_fm2d_args A, int32_t *pflag)
{
    extern __shared__ uint32_t ff_vfb[];
    __shared__ int bad;
    const long slot = ff_xcd_slot(blockIdx.x, gridDim.x);
    if (pflag[slot] != 0) return;
    near_geom g;
    if (!ff_problem(A, slot, g, false)) return;
    const int nnx = A.nnx, nnz = A.nnz, ld = A.ld;
    const long wsz = (long)ld * A.nc;
    const long cp = g.prob / A.nsrc;
    const double *velnc = A.veln + cp * (long)nnx * nnz;
    fview V;
    V.nnx = nnx; V.nnz = nnz; V.dnx = A.dvx / A.gdx; V.dnz = A.dvz / A.gdz;
    V.T = A.ttn + slot * wsz; V.ns = nullptr; V.ox = 0; V.oz = 0; V.lnx = nnx; V.lnz = nnz; V.ld = ld;
    const int32_t *F = A.nsts + slot * wsz;
    const double *Wb = A.vr + slot * wsz;
    const ff_win win = ff_window(A, g);
    const int wn = win.nodes(), wnz = win.nz(), fw = (wn + 31) / 32;
    for (int w = threadIdx.x; w < fw; w += blockDim.x) {
        uint32_t m = 0;
        for (int b = 0; b < 32; b++) {
            const int q = w * 32 + b;
            if (q < wn) {
                const int x = win.x0 + q / wnz, z = win.z0 + q % wnz;
                if (F[(long)(x - 1) * ld + (z - 1)] != 0) m |= 1u << b;
            }
        }
        ff_vfb[w] = m;
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    auto fixed = [&](int z, int x) -> bool {
        if (!win.in(z, x)) return false;
        const int q = (x - win.x0) * wnz + (z - win.z0);
        return (ff_vfb[q >> 5] >> (q & 31)) & 1u;
    };
    const long nn = (long)nnx * nnz;
    for (long q = threadIdx.x; q < nn; q += blockDim.x) {
        if (*(volatile int *)&bad) break;
        const int x = (int)(q / nnz) + 1, z = (int)(q % nnz) + 1;
        if (fixed(z, x)) continue;
        const double w0 = win.in(z, x) ? Wb[(long)(x - 1) * ld + (z - 1)] : FF_BIG;
        const int why = ff_verify_node<FOM>(V, fixed, w0, z, x, 1.0 / velnc[(long)(x - 1) * nnz + (z - 1)]);
        if (why) atomicOr(&bad, why);
    }
    __syncthreads();
    if (threadIdx.x == 0 && bad) {
        pflag[slot] = 1;
        if (A.rounds_out) A.rounds_out[4 * slot + 3] = bad;
    }
}

// LDS bytes of k_fm2d_strip, or 0 when the grid needs more than the limit.
static size_t ff_strip_lds(const mct_fm2d_args &a)
{
    const long nn = (long)a.nnx * a.nnz;
    const long nw = (nn + 31) / 32;
    const long ns = (a.nnx + 63) / 64;
    const long fw = ff_win_words(a);
    const size_t b = 4 * (size_t)(nw + fw + ((nw + fw) & 1) + ((ns + 1) & ~1L)) +
                     8 * (size_t)(a.fom == 0 ? 2 : 4) * a.nnz;
    return b <= 48 * 1024 ? b : 0;
}

// LDS bytes of k_fm2d_near for the geometry in a.
static size_t ff_near_lds(const mct_fm2d_args &a, int &region, int &heap_cap, int &win)
{
    const int rg = 2 * a.sgs * a.sgdl + 1;
    const int wn = 2 * a.sgs + 1;
    const int ln = (a.asgr == 1 ? wn : 2) + 2 * FF_MARGIN;
    const int lnx = ln < a.nnx ? ln : a.nnx, lnz = ln < a.nnz ? ln : a.nnz;
    const int rgx = rg < (a.nnx - 1) * a.sgdl + 1 ? rg : (a.nnx - 1) * a.sgdl + 1;
    const int rgz = rg < (a.nnz - 1) * a.sgdl + 1 ? rg : (a.nnz - 1) * a.sgdl + 1;
    const int rnodes = a.asgr == 1 ? rgx * rgz : 0;
    region = rnodes > lnx * lnz ? rnodes : lnx * lnz;
    // the heap holds a narrow band only; a problem whose band outgrows it goes
    // to the sequential solver
    heap_cap = region < 512 ? region : 512;
    win = a.asgr == 1 ? wn * wn : 1;
    // times and states of the region, heap, remap staging, bit set, diagonal flags
    size_t b = 8 * (size_t)region + 2 * (size_t)((region + 3) & ~3) + 4 * (size_t)((heap_cap + 2 + 1) & ~1) +
               8 * (size_t)win + 2 * (size_t)((win + 1) & ~1) + 4 * (size_t)((region + 31) / 32) +
               4 * 2 * (size_t)(rgx + rgz) + 64;
    return b;
}

int mct_fm2d_fast_ok(const mct_fm2d_args &a)
{
    int region, heap_cap, win;
    size_t b = ff_near_lds(a, region, heap_cap, win);
    // heap positions and coordinates are 16-bit; LDS within the device's limit per workgroup
    static int lds_max = 0;
    if (!lds_max) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 64 * 1024;
        lds_max = v;
    }
    return b <= (size_t)lds_max && region < 32767 && a.nnx < 32767 && a.nnz < 32767;
}

int mct_launch_fm2d_fast(const mct_fm2d_args &a, long chunk, int32_t *pflag, hipStream_t s)
{
    if (!a.sln) return MCT_ERR_ARG;       // the sweeps read the gridder's slowness
    int region, heap_cap, win;
    size_t lds = ff_near_lds(a, region, heap_cap, win);
    int rc = mct_launch_fm2d_gridder(a, s);
    if (rc) return rc;
    if (lds > 64 * 1024)
        MCT_CHECK(hipFuncSetAttribute((const void *)k_fm2d_near, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // the sweep's dirty bitmap in LDS when it fits
    const size_t dirty_bytes = 2 * 4 * (((size_t)a.nnx * a.nnz + 31) / 32) + 4 * 2 * (size_t)(a.nnx + a.nnz);
    const int use_dirty = dirty_bytes <= 48 * 1024;
    const size_t dirty_lds = use_dirty ? dirty_bytes : 0;
    // threads per 2-D sweep (one problem per workgroup; one wave measured fastest):
    // environment MCT_SWEEP_THREADS
    static int env_threads = 0;
    if (!env_threads) {
        const char *e = getenv("MCT_SWEEP_THREADS");
        int v = e ? atoi(e) : 64;
        env_threads = (v == 64 || v == 128 || v == 256) ? v : 64;
    }
    const int sweep_threads = a.sweep_threads ? a.sweep_threads : env_threads;
    // the strip-pipelined sweep unless disabled (MCT_SWEEP_KIND=diag, or the
    // context's setting) or too large for LDS
    const char *sk = getenv("MCT_SWEEP_KIND");
    const int sweep_kind = a.sweep_kind ? (a.sweep_kind == 1 ? 1 : 0) : ((sk && strcmp(sk, "diag") == 0) ? 0 : 1);
    const size_t strip_lds = sweep_kind ? ff_strip_lds(a) : 0;
    const long nprob = (long)a.nchain * a.np * a.nsrc;
    if (chunk < 1) return MCT_ERR_ARG;
    // the heap-order check (k_fm2d_verify): on by default for the mixed-order update
    int verify = a.verify;
    if (verify == 0) {
        const char *ev = getenv("MCT_FM2D_VERIFY");
        verify = ev ? (atoi(ev) ? 1 : -1) : (a.fom == 1 ? 1 : -1);
    }
    // the wave-per-problem sequential solver for the handed-over problems unless
    // the environment sets MCT_FM2D_WAVE=0 (one lane per problem for all of them)
    const char *ew = getenv("MCT_FM2D_WAVE");
    const int wave_off = ew && atoi(ew) == 0;
    // Mixed order (fom 1): the heap loop itself, one wave per problem, for every
    // problem.  The sweeps reach the heap loop's answer only where the check
    // above passes (a fifth of the problems at 201x201), and sweeping and
    // checking all of them cost more than solving the rest sequentially:
    // 1.9 s instead of 4.1 s per 64-chain batch.  MCT_FM2D_ORDER1=check keeps
    // the sweeps and the check.
    const char *eo = getenv("MCT_FM2D_ORDER1");
    const int order1_check = eo && strcmp(eo, "check") == 0;
    if (a.fom == 1 && !order1_check && !wave_off) {
        for (long p0 = 0; p0 < nprob; p0 += chunk) {
            mct_fm2d_args b = a;
            b.prob0 = p0;
            b.nrun = nprob - p0 < chunk ? nprob - p0 : chunk;
            if (a.ev_sweep && p0 == 0) MCT_CHECK(hipEventRecord(a.ev_sweep, s));
            if ((rc = mct_fill_i32(pflag, b.nrun, 1, s))) return rc;
            if (a.rounds_out && (rc = mct_fill_i32(a.rounds_out, 4 * b.nrun, 0, s))) return rc;
            if ((rc = mct_launch_fm2d_wave(b, pflag, s))) return rc;
            b.only_flag = pflag;
            if ((rc = mct_launch_fm2d_solve(b, s))) return rc;
            if ((rc = mct_launch_fm2d_rays(b, s))) return rc;
        }
        return MCT_OK;
    }
    for (long p0 = 0; p0 < nprob; p0 += chunk) {
        mct_fm2d_args b = a;
        b.verify = verify > 0 ? 1 : -1;
        b.prob0 = p0;
        b.nrun = nprob - p0 < chunk ? nprob - p0 : chunk;
        hipLaunchKernelGGL(k_fm2d_near, dim3((unsigned)b.nrun), dim3(64), lds, s, b, region, heap_cap, win, pflag);
        MCT_CHECK(hipGetLastError());
        if (a.ev_sweep && p0 == 0) MCT_CHECK(hipEventRecord(a.ev_sweep, s));
        if (strip_lds && a.fom == 0)
            hipLaunchKernelGGL(k_fm2d_strip<0>, dim3((unsigned)b.nrun), dim3(64), strip_lds, s, b, pflag);
        else if (strip_lds)
            hipLaunchKernelGGL(k_fm2d_strip<1>, dim3((unsigned)b.nrun), dim3(64), strip_lds, s, b, pflag);
        else if (a.fom == 0)
            hipLaunchKernelGGL(k_fm2d_sweep<0>, dim3((unsigned)b.nrun), dim3(sweep_threads), dirty_lds, s, b,
                               pflag, use_dirty);
        else
            hipLaunchKernelGGL(k_fm2d_sweep<1>, dim3((unsigned)b.nrun), dim3(sweep_threads), dirty_lds, s, b,
                               pflag, use_dirty);
        MCT_CHECK(hipGetLastError());
        if (verify > 0) {
            const size_t vlds = 4 * (size_t)ff_win_words(a);
            if (a.fom == 0)
                hipLaunchKernelGGL(k_fm2d_verify<0>, dim3((unsigned)b.nrun), dim3(256), vlds, s, b, pflag);
            else
                hipLaunchKernelGGL(k_fm2d_verify<1>, dim3((unsigned)b.nrun), dim3(256), vlds, s, b, pflag);
            MCT_CHECK(hipGetLastError());
        }
        // problems whose heap loop left the LDS neighbourhood, or that the check
        // above failed: the sequential heap loop, one wave per problem, and one
        // lane per problem for the few whose band outgrows that wave's LDS heap
        if (!wave_off && (rc = mct_launch_fm2d_wave(b, pflag, s))) return rc;
        b.only_flag = pflag;
        if ((rc = mct_launch_fm2d_solve(b, s))) return rc;
        if ((rc = mct_launch_fm2d_rays(b, s))) return rc;
    }
    return MCT_OK;
}
