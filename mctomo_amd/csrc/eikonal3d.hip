// eikonal3d.hip -- 3-D first-arrival traveltimes on gfx950.
//
// Replaces the SFMM solve inside fastMarching3D (fastmarching/fastMarching.cpp
// :26-129, fmm/fmm.hpp:92-158).  Instead of a serial heap, the grid is cut into
// 4x4x4 tiles, one wave64 per tile, one node per lane.  A round processes every
// active tile of every problem (problem = chain x wave type x source): the wave
// stages its tile and a 2-deep face halo in LDS, relaxes the local upwind
// update to convergence (a wave ballot ends the loop), writes back, and
// activates the face neighbours whose halo it changed.  Rounds repeat until no
// tile is active (block fast-iterative method).
//
// The local update is the reference's: solveEikonal (eikonalsolver.hpp:50-79)
// for order 0, highAccuracySolveEikonal (:83-145) for order 1, with the cell's
// current value as the filter and a new value kept only when it improves
// (fmm.hpp:141-150).  For order 0 the FMM answer is the unique fixed point of
// that update, so the tiles converge to it; arithmetic is FP32 with each
// quadratic written relative to its smallest neighbour time, which removes the
// cancellation of the reference's b^2-4ac form.
#include "mct_internal.h"
#include <float.h>
#include <stdlib.h>

#pragma clang fp contract(fast)

#define EIK_INF __builtin_huge_valf()
template <class TT> __device__ __forceinline__ TT tinf();
template <> __device__ __forceinline__ float tinf<float>() { return __builtin_huge_valf(); }
template <> __device__ __forceinline__ double tinf<double>() { return __builtin_huge_val(); }
// LDS box: the 4^3 tile at offset 2 inside an 8^3 box (2-deep face halo).
#define BX(x, y, z) ((((z) + 2) << 6) | (((y) + 2) << 3) | ((x) + 2))

// The local solves below run in R = float (velocities and increments in FP32)
// or R = double (the FP64 mode: the reference's own precision).
__device__ __forceinline__ float lsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double lsqrt(double x) { return sqrt(x); }

// Compare-and-swap of (time, 1/h^2, h) triples by time; the swap keeps the
// earlier dimension first on equal times (std::sort on <= 16 elements is an
// insertion sort, which is stable).
template <class R>
__device__ __forceinline__ void cas3(R &t0, R &a0, R &h0, R &t1, R &a1, R &h1)
{
    bool sw = t1 < t0;
    R tt = sw ? t1 : t0, aa = sw ? a1 : a0, hh = sw ? h1 : h0;
    t1 = sw ? t0 : t1; a1 = sw ? a0 : a1; h1 = sw ? h0 : h1;
    t0 = tt; a0 = aa; h0 = hh;
}

template <class R>
__device__ __forceinline__ void cas4(R &t0, R &u0, R &a0, R &h0, R &t1, R &u1, R &a1, R &h1)
{
    bool sw = t1 < t0;
    R tt = sw ? t1 : t0, uu = sw ? u1 : u0, aa = sw ? a1 : a0, hh = sw ? h1 : h0;
    t1 = sw ? t0 : t1; u1 = sw ? u0 : u1; a1 = sw ? a0 : a1; h1 = sw ? h0 : h1;
    t0 = tt; u0 = uu; a0 = aa; h0 = hh;
}

// root of sum_i al_i (tau - dl_i)^2 = s2 (larger root), INF if none
template <class R>
__device__ __forceinline__ R root2(R a0, R d0, R a1, R d1, R s2)
{
    R A = a0 + a1, B = a0 * d0 + a1 * d1, C = a0 * d0 * d0 + a1 * d1 * d1 - s2;
    R disc = B * B - A * C;
    return disc < R(0) ? tinf<R>() : (B + lsqrt(disc)) / A;
}

template <class R>
__device__ __forceinline__ R root3(R a0, R d0, R a1, R d1, R a2, R d2, R s2)
{
    R A = a0 + a1 + a2, B = a0 * d0 + a1 * d1 + a2 * d2;
    R C = a0 * d0 * d0 + a1 * d1 * d1 + a2 * d2 * d2 - s2;
    R disc = B * B - A * C;
    return disc < R(0) ? tinf<R>() : (B + lsqrt(disc)) / A;
}

// solveEikonal (first order): dimensions whose smaller neighbour beats the
// cell's own value, sorted, grown while the next one is below the solution.
// Written relative to the smallest neighbour time t0.
template <class R>
__device__ __forceinline__ R update_order0(const R m[3], R tself, const R h[3], const R ih2[3], R inv_v, R s2)
{
    const R INF = tinf<R>();
    R t0 = m[0] < tself ? m[0] : INF, t1 = m[1] < tself ? m[1] : INF, t2 = m[2] < tself ? m[2] : INF;
    R a0 = ih2[0], a1 = ih2[1], a2 = ih2[2], h0 = h[0], h1 = h[1], h2 = h[2];
    cas3(t0, a0, h0, t1, a1, h1);
    cas3(t1, a1, h1, t2, a2, h2);
    cas3(t0, a0, h0, t1, a1, h1);
    if (t0 == INF) return INF;
    R u1 = t0 + h0 * inv_v;
    if (!(u1 > t1)) return u1;                     // one dimension, or the next is not below
    R d1 = t1 - t0;
    R tau2 = root2(a0, R(0), a1, d1, s2);
    if (!(tau2 + t0 > t2)) return t0 + tau2;
    R tau3 = root3(a0, R(0), a1, d1, a2, t2 - t0, s2);
    return t0 + tau3;
}

// highAccuracySolveEikonal: per dimension the smaller one-away time t and, when
// the two-away time on that side is upwind (<= t), the second-order value
// (4t - t2)/3 with weight 9/4; try all included dimensions first and drop the
// largest while the root does not exceed that dimension's time.
template <class R>
__device__ __forceinline__ R update_order1(const R m1[3], const R m2[3], R tself, const R h[3], const R ih2[3],
                                           R inv_v, R s2)
{
    const R INF = tinf<R>();
    R t[3], u[3], al[3], hh[3];
    for (int d = 0; d < 3; d++) {
        bool inc = m1[d] < tself;
        bool two = m2[d] < INF;
        t[d] = inc ? m1[d] : INF;
        u[d] = two ? m2[d] : INF;
        al[d] = two ? R(2.25) * ih2[d] : ih2[d];
        hh[d] = h[d];
    }
    cas4(t[0], u[0], al[0], hh[0], t[1], u[1], al[1], hh[1]);
    cas4(t[1], u[1], al[1], hh[1], t[2], u[2], al[2], hh[2]);
    cas4(t[0], u[0], al[0], hh[0], t[1], u[1], al[1], hh[1]);
    if (t[0] == INF) return tself;
    const R t0 = t[0];
    R dl[3];
    for (int d = 0; d < 3; d++)
        dl[d] = u[d] < INF ? (R(4) * (t[d] - t0) - (u[d] - t0)) * (R(1) / R(3)) : t[d] - t0;
    if (t[2] < INF) {
        R tau = root3(al[0], dl[0], al[1], dl[1], al[2], dl[2], s2);
        if (tau < INF && tau > t[2] - t0) return t0 + tau;
    }
    if (t[1] < INF) {
        R tau = root2(al[0], dl[0], al[1], dl[1], s2);
        if (tau < INF && tau > t[1] - t0) return t0 + tau;
    }
    return t0 + hh[0] * inv_v;
}

// ---- initial points --------------------------------------------------------------
// One thread per (problem, corner): Solver::setInitialPoints (solver.hpp:72-92)
// with getVel / getNeighbors(point) (ndgridmap.hpp:158-191,433-460).
__device__ __forceinline__ void pt_coord(const mct_geom &G, const double p[3], const double leaf[3],
                                         int c[3])
{
    int dims[3] = {G.nx, G.ny, G.nz};
    for (int d = 0; d < 3; d++) {
        // int(p/h) truncates toward zero, then clamps (getPointCoord);
        // done in double so out-of-range inputs never reach the conversion
        double qd = p[d] / leaf[d];
        int ci = !(qd >= 1.0) ? 0 : (qd >= (double)dims[d] ? dims[d] - 1 : (int)qd);
        c[d] = ci;
    }
}

__device__ __forceinline__ long tl_index(const mct_geom &G, int x, int y, int z)
{
    int tile = ((z >> 2) * G.nty + (y >> 2)) * G.ntx + (x >> 2);
    int lane = ((z & 3) << 4) | ((y & 3) << 2) | (x & 3);
    return ((long)tile << 6) | lane;
}

template <class TT>
__global__ void k_eik_clear(TT *T, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        T[i] = tinf<TT>();
}

template <class TT>
__global__ void k_eik_init(mct_eik_problem_set P, const double *src_pts, int src_stride,
                           long src_chain_stride, double xmin, double ymin, double zmin, double lx,
                           double ly, double lz)
{
    int prob = blockIdx.x * blockDim.x + threadIdx.x;
    if (prob >= P.nprob) return;
    const mct_geom &G = P.G;
    int s = prob % P.nsrc_eff, cw = prob / P.nsrc_eff;          // cw = chain*nwave + wave
    int c = cw / P.nwave;
    const double *sp0 = src_pts + (long)c * src_chain_stride + (long)s * src_stride;
    double p[3] = {sp0[0] - xmin, sp0[1] - ymin, sp0[2] - zmin};
    double leaf[3] = {lx, ly, lz};
    int dims[3] = {G.nx, G.ny, G.nz};
    int pc[3];
    pt_coord(G, p, leaf, pc);
    const float *V = P.vel ? P.vel + (long)cw * G.ntiles * MCT_TILE_N : nullptr;
    const double *V64 = P.vel64 ? P.vel64 + (long)cw * G.ntiles * MCT_TILE_N : nullptr;
    TT *T = (TT *)P.T + (long)prob * G.ntiles * MCT_TILE_N;
    int cx[8], cy[8], cz[8];
    double vn[8];
    for (int v = 0; v < 8; v++) {
        int vc[3];
        for (int d = 0; d < 3; d++) {
            vc[d] = pc[d] + ((v >> (2 - d)) & 1);
            if (vc[d] >= dims[d]) vc[d] = dims[d] - 1;
        }
        cx[v] = vc[0]; cy[v] = vc[1]; cz[v] = vc[2];
        const long q = tl_index(G, vc[0], vc[1], vc[2]);
        vn[v] = V64 ? V64[q] : (double)V[q];
    }
    // getVel: trilinear weights with clamped corners
    double vs = 0.0;
    for (int v = 0; v < 8; v++) {
        int cc[3] = {cx[v], cy[v], cz[v]};
        double w = 1.0;
        for (int d = 0; d < 3; d++) {
            double sd = p[d] - leaf[d] * pc[d];
            double wj = fabs((double)(cc[d] - pc[d]) * leaf[d] - leaf[d] + sd);
            w = w * wj / leaf[d];
        }
        vs = vs + w * vn[v];
    }
    for (int v = 0; v < 8; v++) {
        double q[3] = {cx[v] * leaf[0], cy[v] * leaf[1], cz[v] * leaf[2]};
        double dist = 0.0;
        for (int d = 0; d < 3; d++) dist = dist + (q[d] - p[d]) * (q[d] - p[d]);
        dist = sqrt(dist);
        T[tl_index(G, cx[v], cy[v], cz[v])] = (TT)(dist * 2 / (vs + vn[v]));
    }
    // seed: the 3x3x3 block of tiles around each corner's tile
    P.any[0] = 1u;
    for (int v = 0; v < 8; v++) {
        int tx = cx[v] >> 2, ty = cy[v] >> 2, tz = cz[v] >> 2;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int ax = tx + dx, ay = ty + dy, az = tz + dz;
                    if (ax < 0 || ay < 0 || az < 0 || ax >= G.ntx || ay >= G.nty || az >= G.ntz) continue;
                    long key = (long)prob * G.ntiles + (long)((az * G.nty + ay) * G.ntx + ax);
                    P.flags[0][key] = 1;
                    P.key[0][key] = 0u;
                }
    }
}

// ---- one relaxation round ----------------------------------------------------------
__device__ __forceinline__ float wave_min(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }


// One relaxation round.  Active tiles are byte flags (no lists, no counters);
// each carries a key, the smallest arrival time that activated it.  A tile is
// relaxed only once the round's band time tau has passed its key, so the
// front is swept roughly in arrival order (a tile sees its upwind halo nearly
// final); tiles not yet due are carried to the next round unchanged.
//
// TT is the storage type of arrival times.  The local quadratic is always
// solved in FP32 on times relative to the smallest upwind neighbour t0 (an
// increment of about h/v), then added to t0 in TT: with TT = double the
// rounding of each step is that of the small increment, so errors do not
// accumulate along a ray at the FP32 level.
template <int ORDER, class TT, class CT>
__global__ __launch_bounds__(64) void k_eik_round(mct_eik_problem_set P, int round, double hxd, double hyd,
                                                  double hzd)
{
    __shared__ TT box[512];
    const mct_geom &G = P.G;
    const TT TINF = tinf<TT>();
    const CT CINF = tinf<CT>();
    const int cur = round & 1, nxt = cur ^ 1;
    const int lane = threadIdx.x;
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const CT h[3] = {(CT)hxd, (CT)hyd, (CT)hzd};
    const CT ih2[3] = {CT(1) / (h[0] * h[0]), CT(1) / (h[1] * h[1]), CT(1) / (h[2] * h[2])};
    if (blockIdx.x == 0 && lane == 0) P.any[(round + 2) % 3] = 0u;   // flag of round r+2
    const uint32_t ntiles = (uint32_t)G.ntiles;
    const uint32_t nkeys = (uint32_t)P.nprob * ntiles;
    const uint32_t stride = gridDim.x;
    uint8_t *fcur = P.flags[cur], *fnxt = P.flags[nxt];
    uint32_t *kcur = P.key[cur], *knxt = P.key[nxt];
    unsigned int *any_next = &P.any[(round + 1) % 3];
    const bool banded = round < P.band_rounds;
    // Relative change that counts.  In the FP64 mode every decrease counts, so
    // the rounds end only at the update's fixed point, whatever order the tiles
    // ran in (a batch gives each chain the times its own call gives); with FP32
    // increments a change is at least an FP32 step of the increment anyway.
    const TT thr = sizeof(CT) == 8 ? (TT)0 : sizeof(TT) == 8 ? (TT)1e-9 : (TT)1e-7;
    TT *Tall = (TT *)P.T;
    // face lane masks (bit l = lane l): the layers a neighbour reads as halo
    const int deep = ORDER == 0 ? 0 : 1;
    uint64_t fm_mine = 0;
    for (int l = 0; l < 64; l++) {
        int ax = l & 3, ay = (l >> 2) & 3, az = l >> 4, f = lane;
        bool in = (f == 0 && ax <= deep) || (f == 1 && ax >= 3 - deep) || (f == 2 && ay <= deep) ||
                  (f == 3 && ay >= 3 - deep) || (f == 4 && az <= deep) || (f == 5 && az >= 3 - deep);
        if (in) fm_mine |= 1ull << l;
    }

    // Keys are visited lane-strided: block b, pass i covers keys
    // b + stride*(64*i + lane), so every key has exactly one owner.
    for (uint32_t base = blockIdx.x; base < nkeys; base += stride * 64u) {
        const uint32_t mykey = base + stride * (uint32_t)lane;
        uint8_t f = mykey < nkeys ? fcur[mykey] : 0;
        bool due = false;
        if (f) {
            fcur[mykey] = 0;                       // consumed
            uint32_t kc = kcur[mykey];
            kcur[mykey] = 0xffffffffu;
            const uint32_t prob = mykey / ntiles;
            const float tau = banded ? (float)(round + 1) * P.band_dt[prob / (uint32_t)P.nsrc_eff]
                                     : __builtin_huge_valf();
            if (__uint_as_float(kc) <= tau || kc == 0xffffffffu) {
                due = true;
            } else {                               // not yet: carry to the next round
                fnxt[mykey] = 1;
                atomicMin(&knxt[mykey], kc);
                *any_next = 1u;
            }
        }
        uint64_t todo = __ballot(due);
        while (todo) {
            const int src_lane = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t key = base + stride * (uint32_t)src_lane;
            const uint32_t prob = key / ntiles;
            const uint32_t tile = key - prob * ntiles;
            const int tx = (int)(tile % (uint32_t)G.ntx);
            const int tyz = (int)(tile / (uint32_t)G.ntx);
            const int ty = tyz % G.nty, tz = tyz / G.nty;
            const int gx = tx * 4 + lx, gy = ty * 4 + ly, gz = tz * 4 + lz;
            const bool inside = gx < G.nx && gy < G.ny && gz < G.nz;
            TT *T = Tall + (size_t)prob * ((size_t)ntiles << 6);
            const uint32_t cw = prob / (uint32_t)P.nsrc_eff;
            const size_t vi = (size_t)cw * ((size_t)ntiles << 6) + ((size_t)tile << 6) + lane;
            const CT v = sizeof(CT) == 8 ? (CT)P.vel64[vi] : (CT)P.vel[vi];
            TT t = inside ? T[((size_t)tile << 6) + lane] : TINF;
            const TT t_in = t;
            box[BX(lx, ly, lz)] = t;
            // face halo: 6 faces x 2 layers x 16 nodes = 192 values, 3 per lane
            for (int r = 0; r < 3; r++) {
                int q = lane + 64 * r;
                int face = q >> 5, layer = (q >> 4) & 1, a = q & 3, b = (q >> 2) & 3;
                int ox, oy, oz;
                int dep = -1 - layer;
                switch (face) {
                    case 0: ox = dep; oy = a; oz = b; break;          // -x
                    case 1: ox = 3 - dep; oy = a; oz = b; break;      // +x
                    case 2: ox = a; oy = dep; oz = b; break;          // -y
                    case 3: ox = a; oy = 3 - dep; oz = b; break;      // +y
                    case 4: ox = a; oy = b; oz = dep; break;          // -z
                    default: ox = a; oy = b; oz = 3 - dep; break;     // +z
                }
                int nxg = tx * 4 + ox, nyg = ty * 4 + oy, nzg = tz * 4 + oz;
                TT hv = TINF;
                if ((ORDER != 0 || layer == 0) && nxg >= 0 && nyg >= 0 && nzg >= 0 && nxg < G.nx &&
                    nyg < G.ny && nzg < G.nz)
                    hv = T[tl_index(G, nxg, nyg, nzg)];
                box[BX(ox, oy, oz)] = hv;
            }
            __syncthreads();
            const CT inv_v = CT(1) / v;
            const CT s2 = inv_v * inv_v;
            const bool solve = inside && v >= CT(2.220446e-11);   // isOccupied(): velocity < COMP_MARGIN
            const int me = BX(lx, ly, lz);
            const int st[3] = {1, 8, 64};
            bool capped = true;
            for (int it = 0; it < 48; it++) {
                TT n1[3], n2[3];
                TT t0 = TINF;
                for (int d = 0; d < 3; d++) {
                    TT a1 = box[me - st[d]], b1 = box[me + st[d]];
                    bool useb = b1 < a1;
                    n1[d] = useb ? b1 : a1;
                    if (ORDER != 0) {
                        TT two = useb ? box[me + 2 * st[d]] : box[me - 2 * st[d]];
                        n2[d] = (n1[d] < TINF && two <= n1[d]) ? two : TINF;
                    }
                    t0 = n1[d] < t0 ? n1[d] : t0;
                }
                TT u = TINF;
                if (solve && t0 < t) {
                    // times relative to t0 in CT; excluded/absent neighbours stay INF
                    CT m1[3], m2[3];
                    const CT fself = t < TINF ? (CT)(t - t0) : CINF;
                    for (int d = 0; d < 3; d++) {
                        m1[d] = n1[d] < TINF ? (CT)(n1[d] - t0) : CINF;
                        if (ORDER != 0) m2[d] = n2[d] < TINF ? (CT)(n2[d] - t0) : CINF;
                    }
                    CT du = ORDER == 0 ? update_order0<CT>(m1, fself, h, ih2, inv_v, s2)
                                       : update_order1<CT>(m1, m2, fself, h, ih2, inv_v, s2);
                    if (du < CINF) u = t0 + (TT)du;
                }
                bool better = u < t;
                bool sig = better && (u < t - thr * t || t == TINF);
                if (better) t = u;
                __syncthreads();
                box[me] = t;
                __syncthreads();
                if (__ballot(sig) == 0ull) { capped = false; break; }
            }
            const bool changed = t < t_in && (t < t_in - thr * t_in || t_in == TINF);
            if (inside && t < t_in) T[((size_t)tile << 6) + lane] = t;
            const uint64_t bal = __ballot(changed);
            if (bal || capped) {
                // key of the tiles this visit activates: the smallest changed time
                const float kmin = wave_min(changed ? (float)t : __builtin_huge_valf());
                const float kown = wave_min(inside ? (float)t : __builtin_huge_valf());
                bool act = false;
                uint32_t nkey = 0;
                uint32_t kbits = fbits(kmin);
                if (lane < 6 && (bal & fm_mine)) {
                    int ax = tx + (lane == 0 ? -1 : lane == 1 ? 1 : 0);
                    int ay = ty + (lane == 2 ? -1 : lane == 3 ? 1 : 0);
                    int az = tz + (lane == 4 ? -1 : lane == 5 ? 1 : 0);
                    if (ax >= 0 && ay >= 0 && az >= 0 && ax < G.ntx && ay < G.nty && az < G.ntz) {
                        nkey = prob * ntiles + (uint32_t)((az * G.nty + ay) * G.ntx + ax);
                        act = true;
                    }
                }
                if (lane == 6 && capped) { nkey = key; act = true; kbits = fbits(kown); }
                if (act) {
                    fnxt[nkey] = 1;
                    atomicMin(&knxt[nkey], kbits);
                }
                if (__ballot(act) != 0ull && lane == 0) *any_next = 1u;
            }
            __syncthreads();
        }
    }
}

int mct_launch_eik_init(const mct_eik_problem_set &P, const double *src_pts, int src_stride,
                        long src_chain_stride, double xmin, double ymin, double zmin,
                        double lx, double ly, double lz, hipStream_t s)
{
    long n = (long)P.nprob * P.G.ntiles * MCT_TILE_N;
    long nf = (long)P.nprob * P.G.ntiles;
    if (P.t64) hipLaunchKernelGGL(k_eik_clear<double>, dim3(4096), dim3(256), 0, s, (double *)P.T, n);
    else hipLaunchKernelGGL(k_eik_clear<float>, dim3(4096), dim3(256), 0, s, (float *)P.T, n);
    MCT_CHECK(hipGetLastError());
    MCT_CHECK(hipMemsetAsync(P.flags[0], 0, (size_t)((nf + 3) & ~3L), s));
    MCT_CHECK(hipMemsetAsync(P.flags[1], 0, (size_t)((nf + 3) & ~3L), s));
    MCT_CHECK(hipMemsetAsync(P.any, 0, 3 * sizeof(uint32_t), s));
    MCT_CHECK(hipMemsetAsync(P.key[0], 0xff, sizeof(uint32_t) * nf, s));
    MCT_CHECK(hipMemsetAsync(P.key[1], 0xff, sizeof(uint32_t) * nf, s));
    if (P.t64)
        hipLaunchKernelGGL(k_eik_init<double>, dim3((P.nprob + 63) / 64), dim3(64), 0, s, P, src_pts,
                           src_stride, src_chain_stride, xmin, ymin, zmin, lx, ly, lz);
    else
        hipLaunchKernelGGL(k_eik_init<float>, dim3((P.nprob + 63) / 64), dim3(64), 0, s, P, src_pts,
                           src_stride, src_chain_stride, xmin, ymin, zmin, lx, ly, lz);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_eik_round(const mct_eik_problem_set &P, int round, double hx, double hy, double hz,
                         int grid_blocks, hipStream_t s)
{
    dim3 g(grid_blocks), b(64);
    if (P.vel64) {   // FP64 velocities and local solves (arrival times FP64 too)
        if (!P.t64) return MCT_ERR_ARG;
        if (P.order == 1) hipLaunchKernelGGL((k_eik_round<1, double, double>), g, b, 0, s, P, round, hx, hy, hz);
        else hipLaunchKernelGGL((k_eik_round<0, double, double>), g, b, 0, s, P, round, hx, hy, hz);
    } else if (P.t64) {
        if (P.order == 1) hipLaunchKernelGGL((k_eik_round<1, double, float>), g, b, 0, s, P, round, hx, hy, hz);
        else hipLaunchKernelGGL((k_eik_round<0, double, float>), g, b, 0, s, P, round, hx, hy, hz);
    } else {
        if (P.order == 1) hipLaunchKernelGGL((k_eik_round<1, float, float>), g, b, 0, s, P, round, hx, hy, hz);
        else hipLaunchKernelGGL((k_eik_round<0, float, float>), g, b, 0, s, P, round, hx, hy, hz);
    }
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

// Band time step per (chain, wave): the time the front needs to cross
// `tiles_per_round` tiles at the fastest velocity of that model.
template <class VT>
__global__ __launch_bounds__(256) void k_band_dt(const VT *vel, long tsz, float hmin, float tpr, float *band_dt)
{
    __shared__ float sm[4];
    const int cw = blockIdx.x;
    const VT *V = vel + (long)cw * tsz;
    float m = 0.f;
    for (long i = threadIdx.x; i < tsz; i += blockDim.x) m = fmaxf(m, (float)V[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float mm = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
        band_dt[cw] = mm > 0.f ? tpr * 4.f * hmin / mm : EIK_INF;
    }
}

int mct_launch_band_dt(const float *vel, const double *vel64, long tsz, int ncw, float hmin, float tiles_per_round,
                       float *band_dt, hipStream_t s)
{
    if (vel64)
        hipLaunchKernelGGL(k_band_dt<double>, dim3(ncw), dim3(256), 0, s, vel64, tsz, hmin, tiles_per_round, band_dt);
    else
        hipLaunchKernelGGL(k_band_dt<float>, dim3(ncw), dim3(256), 0, s, vel, tsz, hmin, tiles_per_round, band_dt);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

float mct_default_band_tiles()
{
    const char *e = getenv("MCT_BAND_TILES");
    float v = e ? (float)atof(e) : 0.5f;
    return v > 0.f ? v : 0.5f;
}

int mct_default_f64()
{
    const char *e = getenv("MCT_EIK_F64");
    return e ? (atoi(e) != 0) : 1;
}

int mct_default_t64()
{
    const char *e = getenv("MCT_EIK_F32");
    return (e && atoi(e) == 1) ? 0 : 1;
}
