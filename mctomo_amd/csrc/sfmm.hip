// sfmm.hip -- exact 3-D first-arrival traveltimes: the SFMM heap loop of
// fastMarching3D replayed step for step, one wave per problem.
//
// fastMarching3D (fastmarching/fastMarching.cpp:26-129) runs
// FMM::computeInternal (fmm/fmm.hpp:92-158) over SFMM's queue (fmm/sfmm.hpp
// :44-52, fmm/fmdata/fmpriorityqueue.hpp:27-86): a binary heap of cells whose
// keys are read live from the cells (FMCompare, fmm/fmdata/fmcompare.hpp
// :29-32); a cell is pushed again on every improvement and its older entries
// stay in the heap.  With the mixed-order update (highAccuracySolveEikonal,
// fmm/eikonalsolver.hpp:83-145) a cell's time depends on the values its
// neighbours held when each of them was popped, so the times are a function
// of the pop order, not of the field alone.  The only way to reproduce them is
// to replay the heap: the same pushes, the same sift steps (GCC libstdc++
// push_heap / pop_heap: the hole is moved down to a leaf, then val sifts up),
// and the same local solves in FP64 in the reference's operation order
// (contraction off).
//
// One wave runs one problem: the replay stays sequential, the lanes carry the
// memory traffic.  A pop reads the five-level subtree under the hole (62
// entries, then their live keys) in two round trips and walks it with lane
// reads; the val sift-up that ends a pop needs no loads (its comparisons are
// against the entries just walked).  A push reads every ancestor of its slot
// at once.  The six neighbours of a popped cell are solved by six lanes; only
// the plus neighbour of a dimension reads a value the minus neighbour may
// have just changed (its two-away cell), so the plus lanes solve again after
// the minus lanes' updates.  Frozen cells keep their time with the sign bit
// set (times are never negative): a neighbour's state and value arrive in one
// load, the keys are |t|, and a last pass clears the bit.
#include "mct_internal.h"
#include <float.h>
#include <stdlib.h>
#include <string.h>

#pragma clang fp contract(off)

#define SF_MARGIN (DBL_EPSILON * 1e5)     // utils::COMP_MARGIN (utils/utils.h:27)
#define SF_LEVELS 5                       // subtree levels read per sift-down block

enum { SF_ERR_HEAP = 1u, SF_ERR_POPS = 2u, SF_ERR_OPEN_INF = 4u };

namespace {

__device__ __forceinline__ double sf_inf() { return __builtin_huge_val(); }

__device__ __forceinline__ double rl_d(double v, int l)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ uint32_t rl_u(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

__device__ __forceinline__ uint32_t first_u(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

__device__ __forceinline__ double first_d(double v) { return rl_d(v, 0); }

// Lanes of one wave read what other lanes of it wrote: order the accesses.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ long sf_index(const mct_geom &G, int x, int y, int z)
{
    const int tile = ((z >> 2) * G.nty + (y >> 2)) * G.ntx + (x >> 2);
    const int ln = ((z & 3) << 4) | ((y & 3) << 2) | (x & 3);
    return ((long)tile << 6) | ln;
}

// The cell, its velocity and, per dimension, the times one and two cells away
// on the minus and plus sides (|t|; `h*` says the cell exists).
struct sf_stencil {
    double tself, vel;
    double m1[3], m2[3], p1[3], p2[3];
    bool hm1[3], hm2[3], hp1[3], hp2[3];
};

// Stable sort of three (t1, t2, h) triples by t1 (std::sort on <= 16
// elements is an insertion sort: equal keys keep their order).  Absent
// dimensions carry t1 = +inf and end up last.
__device__ __forceinline__ void sf_cas(double &a1, double &a2, double &ah, double &b1, double &b2, double &bh)
{
    const bool sw = b1 < a1;
    const double x1 = sw ? b1 : a1, x2 = sw ? b2 : a2, xh = sw ? bh : ah;
    b1 = sw ? a1 : b1; b2 = sw ? a2 : b2; bh = sw ? ah : bh;
    a1 = x1; a2 = x2; ah = xh;
}

__device__ __forceinline__ void sf_sort3(double t1[3], double t2[3], double hh[3])
{
    sf_cas(t1[0], t2[0], hh[0], t1[1], t2[1], hh[1]);
    sf_cas(t1[1], t2[1], hh[1], t1[2], t2[2], hh[2]);
    sf_cas(t1[0], t2[0], hh[0], t1[1], t2[1], hh[1]);
}

// solveEikonalNDims (eikonalsolver.hpp:292-322) over the first `dims` entries
__device__ __forceinline__ double sf_nd1(const double t1[3], const double hh[3], int dims, double vel)
{
    if (dims == 1) return t1[0] + hh[0] / vel;
    double a = 0.0, b = 0.0, c = -1.0 / (vel * vel);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i < dims) {
            const double alpha = 1.0 / (hh[i] * hh[i]);
            c += t1[i] * t1[i] * alpha;
            b += -2.0 * t1[i] * alpha;
            a += alpha;
        }
    }
    const double q = b * b - 4 * a * c;
    if (q < 0) return sf_inf();
    return (-b + sqrt(q)) / (2 * a);
}

// highAccuracySolveEikonalNDims (eikonalsolver.hpp:326-368)
__device__ __forceinline__ double sf_nd2(const double t1[3], const double t2[3], const double hh[3], int dims,
                                         double vel)
{
    if (dims == 1) return t1[0] + hh[0] / vel;
    double a = 0.0, b = 0.0, c = -1.0 / (vel * vel);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i < dims) {
            if (!__builtin_isinf(t2[i])) {
                const double alpha = 9.0 / 4.0 / (hh[i] * hh[i]);
                const double t = (4.0 * t1[i] - t2[i]) / 3.0;
                c += t * t * alpha;
                b += -2.0 * t * alpha;
                a += alpha;
            } else if (t1[i] < DBL_MAX) {
                const double alpha = 1.0 / (hh[i] * hh[i]);
                c += t1[i] * t1[i] * alpha;
                b += -2.0 * t1[i] * alpha;
                a += alpha;
            }
        }
    }
    const double q = b * b - 4 * a * c;
    if (q < 0) return sf_inf();
    return (-b + sqrt(q)) / (2 * a);
}

// solveEikonal (eikonalsolver.hpp:50-79): per dimension the smaller of the
// two neighbours, kept when finite and below the cell's own time; dimensions
// are added in time order until one more would not lower the result.
__device__ __forceinline__ double sf_solve1(const sf_stencil &S, const double leaf[3])
{
    double t1[3], t2[3], hh[3];
    int a = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double m = (S.hm1[d] && S.hp1[d]) ? (S.m1[d] < S.p1[d] ? S.m1[d] : S.p1[d])
                                                : (S.hm1[d] ? S.m1[d] : S.p1[d]);
        const bool in = !__builtin_isinf(m) && m < S.tself;
        t1[d] = in ? m : sf_inf();
        t2[d] = sf_inf();
        hh[d] = leaf[d];
        a += in ? 1 : 0;
    }
    if (a == 0) return sf_inf();
    sf_sort3(t1, t2, hh);
    double up = sf_inf();
#pragma unroll
    for (int i = 1; i <= 3; i++) {
        if (i <= a) {
            up = sf_nd1(t1, hh, i, S.vel);
            if (i == a || (up - t1[i < 3 ? i : 2]) < SF_MARGIN) break;
        }
    }
    return up;
}

// highAccuracySolveEikonal (eikonalsolver.hpp:83-145): per dimension the
// smaller one-away time, with the two-away time behind it when that is
// finite and not larger; the widest stencil whose root lies above its
// largest neighbour time (by COMP_MARGIN) wins.
__device__ __forceinline__ double sf_solve2(const sf_stencil &S, const double leaf[3])
{
    double t1[3], t2[3], hh[3];
    int a = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double minT = DBL_MAX, minT2 = DBL_MAX;
        if (S.hm1[d]) {
            const double tv = S.m1[d];
            if (!__builtin_isinf(tv) && tv < minT) {
                minT = tv;
                minT2 = DBL_MAX;
                if (S.hm2[d]) {
                    const double tv2 = S.m2[d];
                    if (!__builtin_isinf(tv2) && tv2 <= tv) minT2 = tv2;
                }
            }
        }
        if (S.hp1[d]) {
            const double tv = S.p1[d];
            if (!__builtin_isinf(tv) && tv < minT) {
                minT = tv;
                minT2 = DBL_MAX;
                if (S.hp2[d]) {
                    const double tv2 = S.p2[d];
                    if (!__builtin_isinf(tv2) && tv2 <= tv) minT2 = tv2;
                }
            }
        }
        const bool in = minT < DBL_MAX && minT < S.tself;
        t1[d] = in ? minT : sf_inf();
        t2[d] = (in && minT2 < DBL_MAX) ? minT2 : sf_inf();
        hh[d] = leaf[d];
        a += in ? 1 : 0;
    }
    if (a == 0) return S.tself;
    sf_sort3(t1, t2, hh);
#pragma unroll
    for (int i = 3; i >= 1; i--) {
        if (i <= a) {
            const double cur = sf_nd2(t1, t2, hh, i, S.vel);
            if (!__builtin_isinf(cur) && (cur - t1[i - 1]) > SF_MARGIN) return cur;
        }
    }
    return sf_inf();
}

template <int HIGH>
__global__ __launch_bounds__(64) void k_sfmm(mct_sfmm_args A)
{
    const int lane = threadIdx.x;
    const int prob = blockIdx.x;
    const mct_geom G = A.G;
    const long tsz = (long)G.ntiles * MCT_TILE_N;
    const int si = prob % A.nsrc_eff, cw = prob / A.nsrc_eff, ch = cw / A.nwave;
    double *T = A.T + (long)prob * tsz;
    const double *V = A.V + (long)cw * tsz;
    uint32_t *H = A.heap + (long)prob * A.hcap;
    const double leaf[3] = {A.lx, A.ly, A.lz};
    const int dims[3] = {G.nx, G.ny, G.nz};
    uint32_t err = 0;

    // ---- Solver::setInitialPoints (fmm/solver.hpp:72-92): the eight corner
    // cells of the source's grid cell get dist * 2 / (v_source + v_corner),
    // v_source trilinear with clamped corners (ndgridmap.hpp:158-191,433-489),
    // and are pushed in corner order.
    const double *sp0 = A.src + (long)ch * A.src_chain_stride + (long)si * A.src_stride;
    const double p[3] = {sp0[0] - A.xmin, sp0[1] - A.ymin, sp0[2] - A.zmin};
    int pc[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double qd = p[d] / leaf[d];
        pc[d] = !(qd >= 1.0) ? 0 : (qd >= (double)dims[d] ? dims[d] - 1 : (int)qd);
    }
    uint32_t cq[8];
    double vn[8];
    int cc[8][3];
#pragma unroll
    for (int v = 0; v < 8; v++) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            int c = pc[d] + ((v >> (2 - d)) & 1);
            cc[v][d] = c >= dims[d] ? dims[d] - 1 : c;
        }
        cq[v] = (uint32_t)sf_index(G, cc[v][0], cc[v][1], cc[v][2]);
        vn[v] = V[cq[v]];
    }
    double vs = 0.0;
#pragma unroll
    for (int v = 0; v < 8; v++) {
        double w = 1.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double sd = p[d] - leaf[d] * pc[d];
            const double wj = fabs((double)(cc[v][d] - pc[d]) * leaf[d] - leaf[d] + sd);
            w = w * wj / leaf[d];
        }
        vs = vs + w * vn[v];
    }
    if (lane == 0) {
#pragma unroll
        for (int v = 0; v < 8; v++) {
            double dist = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double xi = (cc[v][d] * leaf[d] - p[d]) * (cc[v][d] * leaf[d] - p[d]);
                dist = dist + xi;
            }
            dist = sqrt(dist);
            T[cq[v]] = dist * 2 / (vs + vn[v]);
        }
    }
    wave_sync();

    long hs = 0;
    // push_heap: the new entry sifts up from slot hs.  Lane l (1..m) holds the
    // ancestor l levels up and its live key; the entry stops below the first
    // ancestor whose key is not larger than its own.  Loads are unconditional
    // (inactive lanes read slot 0 and cell 0) so they issue together.
    auto push = [&](uint32_t cell, double kc) -> bool {
        if (hs >= A.hcap) return false;
        const unsigned long long k1 = (unsigned long long)hs + 1ull;
        hs = hs + 1;
        const int m = 63 - __builtin_clzll(k1);
        const bool act = lane >= 1 && lane <= m;
        const unsigned long long slot = act ? (k1 >> lane) - 1ull : 0ull;
        const uint32_t e = act ? H[slot] : 0u;
        const double ke = fabs(T[e]);
        const unsigned long long stop = __ballot(act && !(ke > kc));
        const int ls = stop ? __builtin_ctzll(stop) : m + 1;
        if (lane >= 1 && lane < ls) H[(k1 >> (lane - 1)) - 1] = e;
        if (lane == 0) H[(k1 >> (ls - 1)) - 1] = cell;
        wave_sync();
        return true;
    };

    for (int v = 0; v < 8 && !err; v++) {
        const double kc = first_d(fabs(T[cq[v]]));
        if (!push(cq[v], kc)) err |= SF_ERR_HEAP;
    }

    // lanes 0..61 of a sift-down block: level lv (1..5) and offset under the hole
    const int lv = 31 - __builtin_clz((unsigned)lane + 2u);
    const int off = lane + 2 - (1 << lv);
    // lanes 0..5: neighbour s = 2*dim + side of a popped cell (fmm.hpp:117-120,
    // ndgridmap.hpp:158-191: x-, x+, y-, y+, z-, z+)
    const int ndim = lane >> 1;
    const int nstep = (lane & 1) ? 1 : -1;

    const long max_pops = 64L * (long)G.nx * G.ny * G.nz + 64;
    long pops = 0;
    while (hs > 0 && !err) {
        if (++pops > max_pops) { err |= SF_ERR_POPS; break; }
        // ---- pop_heap (libstdc++ __pop_heap / __adjust_heap).  First round
        // trip: the root, the last entry and the subtree block under the root.
        const long n = hs - 1;
        hs = n;
        long idx = (1L << lv) - 1 + off;
        bool ok = lane < 62 && idx < n;
        const uint32_t h0 = H[0];
        const uint32_t hn = H[n];
        uint32_t e = H[ok ? idx : 0];
        __builtin_amdgcn_sched_barrier(0);
        const uint32_t top = first_u(h0);
        const uint32_t val = first_u(hn);
        // Second round trip: the block's live keys, the popped cell, val's key
        // and the neighbours' stencils (cells off the grid read the cell itself),
        // issued in that order so that the walk waits for the keys alone.
        const double k0 = T[e];
        const double ttop_v = T[top];
        const double kval_v = T[val];
        __builtin_amdgcn_sched_barrier(0);
        const int tile = (int)(top >> 6), tl = (int)(top & 63u);
        const int tx = tile % G.ntx, tyz = tile / G.ntx;
        const int ty = tyz % G.nty, tz = tyz / G.nty;
        const int cx = tx * 4 + (tl & 3), cy = ty * 4 + ((tl >> 2) & 3), cz = tz * 4 + (tl >> 4);
        int jc[3] = {cx, cy, cz};
#pragma unroll
        for (int d = 0; d < 3; d++) jc[d] += (d == ndim) ? nstep : 0;
        const int jn = ndim == 0 ? jc[0] : (ndim == 1 ? jc[1] : jc[2]);
        const int dn = ndim == 0 ? dims[0] : (ndim == 1 ? dims[1] : dims[2]);
        const bool nvalid = lane < 6 && jn >= 0 && jn < dn;
        if (!nvalid) { jc[0] = cx; jc[1] = cy; jc[2] = cz; }
        sf_stencil S;
        const uint32_t jq = (uint32_t)sf_index(G, jc[0], jc[1], jc[2]);
        const double traw = T[jq];
        S.vel = V[jq];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            S.hm1[d] = jc[d] >= 1; S.hm2[d] = jc[d] >= 2;
            S.hp1[d] = jc[d] <= dims[d] - 2; S.hp2[d] = jc[d] <= dims[d] - 3;
            int c[3] = {jc[0], jc[1], jc[2]};
            c[d] = S.hm1[d] ? jc[d] - 1 : jc[d]; S.m1[d] = T[sf_index(G, c[0], c[1], c[2])];
            c[d] = S.hm2[d] ? jc[d] - 2 : jc[d]; S.m2[d] = T[sf_index(G, c[0], c[1], c[2])];
            c[d] = S.hp1[d] ? jc[d] + 1 : jc[d]; S.p1[d] = T[sf_index(G, c[0], c[1], c[2])];
            c[d] = S.hp2[d] ? jc[d] + 2 : jc[d]; S.p2[d] = T[sf_index(G, c[0], c[1], c[2])];
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            // (with n = 0 the walk is empty and val, the popped entry, is
            // written back to the dead slot 0)
            const double kval = first_d(fabs(kval_v));
            double k = fabs(k0);
            const long lim = n > 0 ? (n - 1) / 2 : 0;
            long hole = 0;
            int m = 0;                         // path levels walked
            uint32_t pe = 0;                   // lane l: entry moved up from level l
            double pk = 0.0;                   //          its key
            long pp = 0;                       //          its position
            while (hole < lim) {
                if (!ok) k = 0.0;
                int r = 1;                     // subtree node (root 1, children 2r and 2r+1)
                for (int st = 0; st < SF_LEVELS && hole < lim; st++) {
                    const double kr = rl_d(k, 2 * r - 1), kl = rl_d(k, 2 * r - 2);
                    const bool left = kr > kl;
                    const int pick = left ? 2 * r - 2 : 2 * r - 1;
                    const uint32_t ec = rl_u(e, pick);
                    const long child = 2 * (hole + 1) - (left ? 1 : 0);
                    m++;
                    if (lane == m) { pe = ec; pk = left ? kl : kr; pp = child; }
                    hole = child;
                    r = pick + 2;
                }
                if (hole < lim) {              // the next block, under the new hole
                    idx = ((hole + 1) << lv) - 1 + off;
                    ok = lane < 62 && idx < n;
                    e = H[ok ? idx : 0];
                    k = fabs(T[e]);
                }
            }
            if ((n & 1) == 0 && hole == (n - 2) / 2) {
                const long child = 2 * (hole + 1);
                const uint32_t ec = first_u(H[child - 1]);
                const double kc = first_d(fabs(T[ec]));
                m++;
                if (lane == m) { pe = ec; pk = kc; pp = child - 1; }
                hole = child - 1;
            }
            // __push_heap of val from the leaf hole: it climbs past the path
            // entries whose keys are larger and stops below the first that is
            // not; the entries above that point stay one level up.
            const unsigned long long keep = __ballot(lane >= 1 && lane <= m && !(pk > kval));
            const int js = keep ? 63 - __builtin_clzll(keep) : 0;
            const uint32_t pe_below = (uint32_t)__shfl_down((int)pe, 1, 64);
            if (lane < js) H[pp] = pe_below;
            if (lane == js) H[pp] = val;
        }
        // the popped cell is frozen (fmm.hpp:123)
        if (lane == 0) T[top] = -fabs(ttop_v);
        wave_sync();
        // the stencil values are |t| (frozen cells carry the sign bit)
        S.tself = fabs(traw);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            S.m1[d] = fabs(S.m1[d]); S.m2[d] = fabs(S.m2[d]);
            S.p1[d] = fabs(S.p1[d]); S.p2[d] = fabs(S.p2[d]);
        }

        // ---- neighbour updates (fmm.hpp:124-154)
        const bool live = nvalid && !signbit(traw) && !(S.vel < SF_MARGIN);
        double nt = 0.0;
        bool upd = false;
        if (live) {
            nt = HIGH ? sf_solve2(S, leaf) : sf_solve1(S, leaf);
            upd = nt + SF_MARGIN < S.tself;
        }
        if (HIGH) {
            // the plus neighbour's two-away cell on its minus side is the
            // minus neighbour: solve again with its value after the update
            // (only when that neighbour was updated: otherwise the value the
            // plus lane loaded is already the one the reference reads)
            const double after = upd ? nt : S.tself;
            const double from_minus = __shfl(after, lane - 1 < 0 ? 0 : lane - 1, 64);
            const bool minus_upd = __shfl(upd ? 1 : 0, lane - 1 < 0 ? 0 : lane - 1, 64) != 0;
            const bool hm2n = ndim == 0 ? S.hm2[0] : (ndim == 1 ? S.hm2[1] : S.hm2[2]);
            if (live && (lane & 1) && hm2n && minus_upd) {
#pragma unroll
                for (int d = 0; d < 3; d++)
                    if (d == ndim) S.m2[d] = from_minus;
                nt = sf_solve2(S, leaf);
                upd = nt + SF_MARGIN < S.tself;
            }
        }
        if (__ballot(live && __builtin_isinf(S.tself) && __builtin_isinf(nt))) err |= SF_ERR_OPEN_INF;
        unsigned long long todo = __ballot(upd);
        while (todo && !err) {
            const int s = __builtin_ctzll(todo);
            todo &= todo - 1;
            const uint32_t cell = rl_u(jq, s);
            const double tn = rl_d(nt, s);
            if (lane == 0) T[cell] = tn;
            wave_sync();
            if (!push(cell, tn)) err |= SF_ERR_HEAP;
        }
    }
    err = first_u(err);
    if (A.pops && lane == 0) A.pops[prob] = (uint32_t)(pops > 0xffffffffL ? 0xffffffffL : pops);
    if (err && lane == 0) atomicOr(A.err, err);
}

// Frozen cells carry the sign bit: clear it (cells never reached keep +inf).
__global__ void k_sfmm_thaw(double *T, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        T[i] = fabs(T[i]);
}

} // namespace

long mct_sfmm_heap_cap(const mct_geom &G)
{
    const long cells = (long)G.nx * G.ny * G.nz;
    long cap = cells / 8;
    if (cap < 65536) cap = 65536;
    return cap;
}

int mct_launch_sfmm(const mct_sfmm_args &a, hipStream_t s)
{
    if (a.nprob < 1 || !a.T || !a.V || !a.src || !a.heap || !a.err || a.hcap < 16 ||
        a.nsrc_eff < 1 || a.nwave < 1)
        return MCT_ERR_ARG;
    if ((long)a.G.ntiles * MCT_TILE_N > 0xffffffffL) return MCT_ERR_ARG;
    MCT_CHECK(hipMemsetAsync(a.err, 0, sizeof(uint32_t), s));
    if (a.order == 1)
        hipLaunchKernelGGL(k_sfmm<1>, dim3((unsigned)a.nprob), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(k_sfmm<0>, dim3((unsigned)a.nprob), dim3(64), 0, s, a);
    MCT_CHECK(hipGetLastError());
    const long n = (long)a.nprob * a.G.ntiles * MCT_TILE_N;
    hipLaunchKernelGGL(k_sfmm_thaw, dim3(4096), dim3(256), 0, s, a.T, n);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}

int mct_sfmm_status(uint32_t err_bits)
{
    if (!err_bits) return MCT_OK;
    return (err_bits & SF_ERR_HEAP) ? MCT_ERR_NOMEM : MCT_ERR_NOCONV;
}

int mct_default_heap_order1()
{
    const char *e = getenv("MCT_EIK_ORDER1");
    return !(e && strcmp(e, "parallel") == 0);
}
