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

#pragma clang fp contract(fast)

#define EIK_INF __builtin_huge_valf()
// LDS box: the 4^3 tile at offset 2 inside an 8^3 box (2-deep face halo).
#define BX(x, y, z) ((((z) + 2) << 6) | (((y) + 2) << 3) | ((x) + 2))

struct eik_term { float t1, t2; int dim; };

// Solve sum_i alpha_i (U - tt_i)^2 = s2 for U, shifted by t0 (<= every tt_i).
__device__ __forceinline__ float quad_root(const float *alpha, const float *dlt, int n, float s2)
{
    float A = 0.f, B = 0.f, C = -s2;
    for (int i = 0; i < n; i++) {
        A += alpha[i];
        B += alpha[i] * dlt[i];
        C += alpha[i] * dlt[i] * dlt[i];
    }
    float disc = B * B - A * C;
    if (disc < 0.f) return EIK_INF;
    return (B + sqrtf(disc)) / A;
}

// solveEikonal: first order, grow the dimension count while it improves.
__device__ __forceinline__ float update_order0(const float m[3], float tself, const float h[3],
                                               const float ih2[3], float inv_v, float s2)
{
    eik_term T[3];
    int a = 0;
    for (int d = 0; d < 3; d++)
        if (m[d] < tself) { T[a].t1 = m[d]; T[a].dim = d; a++; }   // m is INF when absent
    if (a == 0) return EIK_INF;
    // stable insertion sort by t1 (std::sort on <= 16 elements)
    for (int i = 1; i < a; i++) {
        eik_term x = T[i];
        int j = i - 1;
        while (j >= 0 && x.t1 < T[j].t1) { T[j + 1] = T[j]; j--; }
        T[j + 1] = x;
    }
    float t0 = T[0].t1;
    float up = t0 + h[T[0].dim] * inv_v;
    if (a == 1 || up <= T[1].t1) return up;
    float al[3], dl[3];
    for (int i = 0; i < a; i++) { al[i] = ih2[T[i].dim]; dl[i] = T[i].t1 - t0; }
    float tau = quad_root(al, dl, 2, s2);
    if (a == 2 || tau <= dl[2]) return t0 + tau;
    tau = quad_root(al, dl, 3, s2);
    return t0 + tau;
}

// highAccuracySolveEikonal: second order where the two-away value is upwind.
__device__ __forceinline__ float update_order1(const float m1[3], const float m2[3], float tself,
                                               const float h[3], const float ih2[3], float inv_v,
                                               float s2)
{
    eik_term T[3];
    int a = 0;
    for (int d = 0; d < 3; d++)
        if (m1[d] < tself) { T[a].t1 = m1[d]; T[a].t2 = m2[d]; T[a].dim = d; a++; }
    if (a == 0) return tself;
    for (int i = 1; i < a; i++) {
        eik_term x = T[i];
        int j = i - 1;
        while (j >= 0 && x.t1 < T[j].t1) { T[j + 1] = T[j]; j--; }
        T[j + 1] = x;
    }
    float t0 = T[0].t1;
    float al[3], dl[3];
    for (int i = 0; i < a; i++) {
        float hh = ih2[T[i].dim];
        if (T[i].t2 < EIK_INF) {
            al[i] = 2.25f * hh;
            dl[i] = (4.f * (T[i].t1 - t0) - (T[i].t2 - t0)) * (1.f / 3.f);
        } else {
            al[i] = hh;
            dl[i] = T[i].t1 - t0;
        }
    }
    for (int i = a; i >= 2; --i) {
        float tau = quad_root(al, dl, i, s2);
        if (tau < EIK_INF && tau > T[i - 1].t1 - t0) return t0 + tau;
    }
    return t0 + h[T[0].dim] * inv_v;
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

__device__ __forceinline__ bool flag_set(uint8_t *flags, uint32_t key)
{
    unsigned int *wordp = (unsigned int *)((uintptr_t)(flags + key) & ~(uintptr_t)3);
    unsigned int sh = (unsigned int)(((uintptr_t)(flags + key) & 3u) * 8u);
    unsigned int old = atomicOr(wordp, 1u << sh);
    return !((old >> sh) & 1u);
}

__global__ void k_eik_clear(float *T, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        T[i] = EIK_INF;
}

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
    const float *V = P.vel + (long)cw * G.ntiles * MCT_TILE_N;
    float *T = P.T + (long)prob * G.ntiles * MCT_TILE_N;
    int cx[8], cy[8], cz[8];
    double vn[8];
    for (int v = 0; v < 8; v++) {
        int vc[3];
        for (int d = 0; d < 3; d++) {
            vc[d] = pc[d] + ((v >> (2 - d)) & 1);
            if (vc[d] >= dims[d]) vc[d] = dims[d] - 1;
        }
        cx[v] = vc[0]; cy[v] = vc[1]; cz[v] = vc[2];
        vn[v] = (double)V[tl_index(G, vc[0], vc[1], vc[2])];
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
        T[tl_index(G, cx[v], cy[v], cz[v])] = (float)(dist * 2 / (vs + vn[v]));
    }
    // seed: the 3x3x3 block of tiles around each corner's tile
    for (int v = 0; v < 8; v++) {
        int tx = cx[v] >> 2, ty = cy[v] >> 2, tz = cz[v] >> 2;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int ax = tx + dx, ay = ty + dy, az = tz + dz;
                    if (ax < 0 || ay < 0 || az < 0 || ax >= G.ntx || ay >= G.nty || az >= G.ntz) continue;
                    uint32_t key = (uint32_t)prob * (uint32_t)G.ntiles + (uint32_t)((az * G.nty + ay) * G.ntx + ax);
                    if (flag_set(P.flags[0], key)) {
                        uint32_t pos = atomicAdd(&P.counts[0], 1u);
                        P.list[0][pos] = key;
                    }
                }
    }
}

// ---- one relaxation round ----------------------------------------------------------
template <int ORDER>
__global__ __launch_bounds__(64) void k_eik_round(mct_eik_problem_set P, int round, float hx, float hy, float hz)
{
    __shared__ float box[512];
    const mct_geom &G = P.G;
    const int cur = round & 1, nxt = cur ^ 1;
    const uint32_t count = P.counts[round % 3];
    const int lane = threadIdx.x;
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const float h[3] = {hx, hy, hz};
    const float ih2[3] = {1.f / (hx * hx), 1.f / (hy * hy), 1.f / (hz * hz)};
    if (blockIdx.x == 0 && lane == 0) P.counts[(round + 2) % 3] = 0u;   // next round's append target
    const long tsz = (long)G.ntiles * MCT_TILE_N;

    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const uint32_t key = P.list[cur][e];
        const int prob = (int)(key / (uint32_t)G.ntiles);
        const int tile = (int)(key % (uint32_t)G.ntiles);
        // clear this round's membership so a neighbour can queue the tile again
        if (lane == 0) P.flags[cur][key] = 0;
        const int tx = tile % G.ntx, ty = (tile / G.ntx) % G.nty, tz = tile / (G.ntx * G.nty);
        const int gx = tx * 4 + lx, gy = ty * 4 + ly, gz = tz * 4 + lz;
        const bool inside = gx < G.nx && gy < G.ny && gz < G.nz;
        float *T = P.T + (long)prob * tsz;
        const int cw = prob / P.nsrc_eff;
        const float v = P.vel[(long)cw * tsz + ((long)tile << 6) + lane];
        float t = inside ? T[((long)tile << 6) + lane] : EIK_INF;
        const float t_in = t;
        box[BX(lx, ly, lz)] = t;
        // face halo: 6 faces x 2 layers x 16 nodes = 192 values, 3 per lane
        for (int r = 0; r < 3; r++) {
            int q = lane + 64 * r;
            int face = q >> 5, layer = (q >> 4) & 1, a = q & 3, b = (q >> 2) & 3;
            int ox, oy, oz;
            int dep = -1 - layer;
            switch (face) {
                case 0: ox = dep; oy = a; oz = b; break;          // -x
                case 1: ox = 3 - dep; oy = a; oz = b; break;      // +x  (4,5)
                case 2: ox = a; oy = dep; oz = b; break;          // -y
                case 3: ox = a; oy = 3 - dep; oz = b; break;      // +y
                case 4: ox = a; oy = b; oz = dep; break;          // -z
                default: ox = a; oy = b; oz = 3 - dep; break;     // +z
            }
            int nxg = tx * 4 + ox, nyg = ty * 4 + oy, nzg = tz * 4 + oz;
            float hv = EIK_INF;
            if (nxg >= 0 && nyg >= 0 && nzg >= 0 && nxg < G.nx && nyg < G.ny && nzg < G.nz)
                hv = T[tl_index(G, nxg, nyg, nzg)];
            box[BX(ox, oy, oz)] = hv;
        }
        __syncthreads();
        const float inv_v = 1.f / v;
        const float s2 = inv_v * inv_v;
        const bool solve = inside && v >= 2.220446e-11f;   // isOccupied(): velocity < COMP_MARGIN
        const int me = BX(lx, ly, lz);
        const int st[3] = {1, 8, 64};
        bool capped = true;
        for (int it = 0; it < 48; it++) {
            float m1[3], m2[3];
            for (int d = 0; d < 3; d++) {
                float a1 = box[me - st[d]], b1 = box[me + st[d]];
                if (ORDER == 0) {
                    m1[d] = fminf(a1, b1);
                } else {
                    float a2 = box[me - 2 * st[d]], b2 = box[me + 2 * st[d]];
                    // minT over the two sides, minT2 from the winning side if upwind
                    float mt = EIK_INF, mt2 = EIK_INF;
                    if (a1 < mt) { mt = a1; mt2 = (a2 <= a1) ? a2 : EIK_INF; }
                    if (b1 < mt) { mt = b1; mt2 = (b2 <= b1) ? b2 : EIK_INF; }
                    m1[d] = mt; m2[d] = mt2;
                }
            }
            float u = EIK_INF;
            if (solve) u = ORDER == 0 ? update_order0(m1, t, h, ih2, inv_v, s2)
                                      : update_order1(m1, m2, t, h, ih2, inv_v, s2);
            bool better = u < t;
            bool sig = better && (u < t - 1e-7f * t || t == EIK_INF);
            if (better) t = u;
            __syncthreads();
            box[me] = t;
            __syncthreads();
            if (__ballot(sig) == 0ull) { capped = false; break; }
        }
        const bool changed = t < t_in && (t < t_in - 1e-7f * t_in || t_in == EIK_INF);
        if (inside && t < t_in) T[((long)tile << 6) + lane] = t;
        // neighbours read a 1-deep (order 0) or 2-deep (order 1) halo
        const int deep = ORDER == 0 ? 0 : 1;
        unsigned long long bal = __ballot(changed);
        if (capped && lane == 0 && flag_set(P.flags[nxt], key)) {
            // local relaxation hit its cap: queue the tile again
            uint32_t pos = atomicAdd(&P.counts[(round + 1) % 3], 1u);
            P.list[nxt][pos] = key;
        }
        if (bal) {
            uint64_t fm[6] = {0, 0, 0, 0, 0, 0};
            // lane masks of the faces (bit l = lane l)
            for (int l = 0; l < 64; l++) {
                int ax = l & 3, ay = (l >> 2) & 3, az = l >> 4;
                if (ax <= deep) fm[0] |= 1ull << l;
                if (ax >= 3 - deep) fm[1] |= 1ull << l;
                if (ay <= deep) fm[2] |= 1ull << l;
                if (ay >= 3 - deep) fm[3] |= 1ull << l;
                if (az <= deep) fm[4] |= 1ull << l;
                if (az >= 3 - deep) fm[5] |= 1ull << l;
            }
            if (lane < 6 && (bal & fm[lane])) {
                int ax = tx + (lane == 0 ? -1 : lane == 1 ? 1 : 0);
                int ay = ty + (lane == 2 ? -1 : lane == 3 ? 1 : 0);
                int az = tz + (lane == 4 ? -1 : lane == 5 ? 1 : 0);
                if (ax >= 0 && ay >= 0 && az >= 0 && ax < G.ntx && ay < G.nty && az < G.ntz) {
                    uint32_t nkey = (uint32_t)prob * (uint32_t)G.ntiles + (uint32_t)((az * G.nty + ay) * G.ntx + ax);
                    if (flag_set(P.flags[nxt], nkey)) {
                        uint32_t pos = atomicAdd(&P.counts[(round + 1) % 3], 1u);
                        P.list[nxt][pos] = nkey;
                    }
                }
            }
        }
        __syncthreads();
    }
}

int mct_launch_eik_init(const mct_eik_problem_set &P, const double *src_pts, int src_stride,
                        long src_chain_stride, double xmin, double ymin, double zmin,
                        double lx, double ly, double lz, hipStream_t s)
{
    long n = (long)P.nprob * P.G.ntiles * MCT_TILE_N;
    long nf = (long)P.nprob * P.G.ntiles;
    hipLaunchKernelGGL(k_eik_clear, dim3(4096), dim3(256), 0, s, P.T, n);
    MCT_CHECK(hipGetLastError());
    MCT_CHECK(hipMemsetAsync(P.flags[0], 0, (size_t)((nf + 3) & ~3L), s));
    MCT_CHECK(hipMemsetAsync(P.flags[1], 0, (size_t)((nf + 3) & ~3L), s));
    MCT_CHECK(hipMemsetAsync(P.counts, 0, 3 * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_eik_init, dim3((P.nprob + 63) / 64), dim3(64), 0, s, P, src_pts, src_stride,
                       src_chain_stride, xmin, ymin, zmin, lx, ly, lz);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_launch_eik_round(const mct_eik_problem_set &P, int round, float hx, float hy, float hz,
                         int grid_blocks, hipStream_t s)
{
    if (P.order == 1)
        hipLaunchKernelGGL(k_eik_round<1>, dim3(grid_blocks), dim3(64), 0, s, P, round, hx, hy, hz);
    else
        hipLaunchKernelGGL(k_eik_round<0>, dim3(grid_blocks), dim3(64), 0, s, P, round, hx, hy, hz);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}
