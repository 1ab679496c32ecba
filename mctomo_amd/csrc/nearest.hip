// nearest.hip -- Voronoi -> grid nearest-site rasterisation on gfx950.
//
// Replaces kdtree_to_grid (src/mcmc_loc2.f90:2002-2082), which builds a
// kdtree2 on the Voronoi sites and asks it for the 1-NN of every grid node.
//
// GPU design: a brute-force FP64 scan with the sites staged in LDS (every lane
// reads the same site: LDS broadcast), one grid node per lane, nodes visited in
// eikonal-tile order so the fused output -- the FP32 velocity tiles the eikonal
// solver reads -- is written coalesced.  The squared distance is summed in the
// reference's order ((dx^2 + dy^2) + dz^2, src/kdtree2.f90:1536-1539) with FP
// contraction off, so the argmin is bit-identical whenever it is unique.
//
// kdtree2 answers exact and near ties by its visiting order and box pruning
// (last equal-distance point wins, :1537 and :398-433; pruning by :1419-1438).
// Nodes whose best and second-best distances are within 1e-14 relative are
// flagged and replayed by a second kernel that rebuilds the reference's tree on
// the device (build_tree_for_range, :704-840) and runs its exact search.  For
// generic sites that list is empty and the replay kernel exits at once.
#include "mct_internal.h"
#include <float.h>

#pragma clang fp contract(off)

#define NEAR_BLOCK 256
#define NEAR_CHUNK 1024        // sites staged in LDS per pass (24 KiB + 8 KiB of lists)

__device__ __forceinline__ void node_of_tile_lane(const mct_geom &G, long e, int &x, int &y, int &z)
{
    int tile = (int)(e >> 6), lane = (int)(e & 63);
    int tx = tile % G.ntx, r = tile / G.ntx;
    int ty = r % G.nty, tz = r / G.nty;
    x = tx * MCT_TILE + (lane & 3);
    y = ty * MCT_TILE + ((lane >> 2) & 3);
    z = tz * MCT_TILE + (lane >> 4);
}

__device__ __forceinline__ long tile_lane_of_node(const mct_geom &G, int x, int y, int z)
{
    int tile = ((z >> 2) * G.nty + (y >> 2)) * G.ntx + (x >> 2);
    int lane = ((z & 3) << 4) | ((y & 3) << 2) | (x & 3);
    return ((long)tile << 6) | lane;
}

__device__ __forceinline__ void write_node(const mct_nearest_args &a, int c, int x, int y, int z, int id)
{
    const mct_geom &G = a.G;
    long gsz = (long)G.nx * G.ny * G.nz;
    long g = (long)z + (long)G.nz * ((long)y + (long)G.ny * x);   // Fortran (nz,ny,nx)
    const double *pr = a.params + ((long)c * a.max_cells + (id - 1)) * 3;
    if (a.sites_id) a.sites_id[c * gsz + g] = id;
    if (a.vp) a.vp[c * gsz + g] = pr[0];
    if (a.vs) a.vs[c * gsz + g] = pr[1];
    if (a.rho) a.rho[c * gsz + g] = pr[2];
    if (a.vel_tiles) {
        long tl = tile_lane_of_node(G, x, y, z);
        long wsz = (long)G.ntiles * MCT_TILE_N;
        a.vel_tiles[((long)c * a.nwave + 0) * wsz + tl] = (float)pr[0];
        if (a.nwave > 1) a.vel_tiles[((long)c * a.nwave + 1) * wsz + tl] = (float)pr[1];
    }
    if (a.vel_tiles64) {
        long tl = tile_lane_of_node(G, x, y, z);
        long wsz = (long)G.ntiles * MCT_TILE_N;
        a.vel_tiles64[((long)c * a.nwave + 0) * wsz + tl] = pr[0];
        if (a.nwave > 1) a.vel_tiles64[((long)c * a.nwave + 1) * wsz + tl] = pr[1];
    }
}

__device__ __forceinline__ bool in_box(const mct_nearest_args &a, int x, int y, int z)
{
    return x + 1 >= a.ix0 && x + 1 <= a.ix1 && y + 1 >= a.iy0 && y + 1 <= a.iy1 &&
           z + 1 >= a.iz0 && z + 1 <= a.iz1;
}

__device__ __forceinline__ bool pm_selects(const mct_nearest_args &a, int c, int x, int y, int z)
{
    if (!a.pm) return true;
    const mct_geom &G = a.G;
    long gsz = (long)G.nx * G.ny * G.nz;
    long g = (long)z + (long)G.nz * ((long)y + (long)G.ny * x);
    const double eps = (double)1e-8f;   // eps = 1e-8 default-real literal (src/mcmc_loc2.f90:51)
    return fabs(a.vs[c * gsz + g] - a.pm[1]) < eps && fabs(a.vp[c * gsz + g] - a.pm[0]) < eps;
}

__global__ __launch_bounds__(NEAR_BLOCK) void k_nearest_brute(mct_nearest_args a)
{
    __shared__ double sx[NEAR_CHUNK], sy[NEAR_CHUNK], sz[NEAR_CHUNK];
    __shared__ unsigned short cand[NEAR_BLOCK / 64][NEAR_CHUNK];
    const mct_geom &G = a.G;
    const int c = blockIdx.y;
    const int S = a.ncells ? a.ncells[c] : a.ncells_const;
    const long e = (long)blockIdx.x * NEAR_BLOCK + threadIdx.x;
    const long ntl = (long)G.ntiles * MCT_TILE_N;
    int x = 0, y = 0, z = 0;
    bool live = false;
    if (e < ntl) {
        node_of_tile_lane(G, e, x, y, z);
        if (x < G.nx && y < G.ny && z < G.nz) {
            live = in_box(a, x, y, z) && pm_selects(a, c, x, y, z);
        } else {
            // padding lanes of edge tiles: never solved; zero keeps them out of max(v)
            long wsz = (long)G.ntiles * MCT_TILE_N;
            if (a.vel_tiles)
                for (int w = 0; w < a.nwave; w++) a.vel_tiles[((long)c * a.nwave + w) * wsz + e] = 0.0f;
            if (a.vel_tiles64)
                for (int w = 0; w < a.nwave; w++) a.vel_tiles64[((long)c * a.nwave + w) * wsz + e] = 0.0;
        }
    }
    const double qx = G.xmin + (double)x * G.dx;
    const double qy = G.ymin + (double)y * G.dy;
    const double qz = G.zmin + (double)z * G.dz;
    double best = DBL_MAX, second = DBL_MAX;
    int bi = 1;
    const double *P = a.points + (long)c * a.max_cells * 3;
    // Site pruning per wave: the wave's nodes form one 4x4x4 tile, whose node
    // coordinates lie in the box [lo, hi].  ub = min over sites of the squared
    // distance to the box's farthest point bounds every node's best distance;
    // a site whose squared distance to the box's nearest point exceeds
    // ub (1 + 1e-12) is farther than every node's best by more than the
    // near-tie tolerance, so it can be neither the nearest site nor a tie.
    // Rounding is monotone in each operand, so the bounds computed in FP64 hold
    // for the distances computed in FP64 (same order of operations).
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    const long tile = ((long)blockIdx.x * NEAR_BLOCK + wv * 64) >> 6;
    const bool wlive = tile < G.ntiles && __ballot(live) != 0ull;      // wave-uniform
    double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};
    if (wlive) {
        const int t = (int)tile;
        const int tx = t % G.ntx, r = t / G.ntx, ty = r % G.nty, tz = r / G.nty;
        const int x0 = tx * MCT_TILE, y0 = ty * MCT_TILE, z0 = tz * MCT_TILE;
        lo[0] = G.xmin + (double)x0 * G.dx; hi[0] = G.xmin + (double)min(x0 + MCT_TILE - 1, G.nx - 1) * G.dx;
        lo[1] = G.ymin + (double)y0 * G.dy; hi[1] = G.ymin + (double)min(y0 + MCT_TILE - 1, G.ny - 1) * G.dy;
        lo[2] = G.zmin + (double)z0 * G.dz; hi[2] = G.zmin + (double)min(z0 + MCT_TILE - 1, G.nz - 1) * G.dz;
    }
    double ub = DBL_MAX;
    for (int base = 0; base < S; base += NEAR_CHUNK) {
        int n = min(NEAR_CHUNK, S - base);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += NEAR_BLOCK) {
            sx[i] = P[3 * (base + i) + 0];
            sy[i] = P[3 * (base + i) + 1];
            sz[i] = P[3 * (base + i) + 2];
        }
        __syncthreads();
        int ncand = 0;
        if (wlive) {
            double m = DBL_MAX;
            for (int i = ln; i < n; i += 64) {
                double f0 = fmax(fabs(sx[i] - lo[0]), fabs(sx[i] - hi[0]));
                double f1 = fmax(fabs(sy[i] - lo[1]), fabs(sy[i] - hi[1]));
                double f2 = fmax(fabs(sz[i] - lo[2]), fabs(sz[i] - hi[2]));
                double u = f0 * f0;
                u = u + f1 * f1;
                u = u + f2 * f2;
                m = fmin(m, u);
            }
            for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
            ub = fmin(ub, m);
            const double lim = ub + ub * 1e-12;
            for (int i0 = 0; i0 < n; i0 += 64) {
                const int i = i0 + ln;
                bool take = false;
                if (i < n) {
                    double g0 = fmax(fmax(lo[0] - sx[i], sx[i] - hi[0]), 0.0);
                    double g1 = fmax(fmax(lo[1] - sy[i], sy[i] - hi[1]), 0.0);
                    double g2 = fmax(fmax(lo[2] - sz[i], sz[i] - hi[2]), 0.0);
                    double l = g0 * g0;
                    l = l + g1 * g1;
                    l = l + g2 * g2;
                    take = l <= lim;
                }
                const unsigned long long bal = __ballot(take);
                if (take) cand[wv][ncand + __popcll(bal & ((1ull << ln) - 1ull))] = (unsigned short)i;
                ncand += __popcll(bal);
            }
        }
        __syncthreads();            // the lists are read by every lane of the wave
        if (live) {
            for (int j = 0; j < ncand; j++) {
                const int i = cand[wv][j];
                double d0 = sx[i] - qx;
                double sd = d0 * d0;
                double d1 = sy[i] - qy;
                sd = sd + d1 * d1;
                double d2 = sz[i] - qz;
                sd = sd + d2 * d2;
                if (sd < best) { second = best; best = sd; bi = base + i + 1; }
                else if (sd < second) second = sd;
            }
        }
    }
    if (!live) return;
    write_node(a, c, x, y, z, bi);
    if (second - best <= best * 1e-14) {
        int pos = atomicAdd(&a.flag_count[c], 1);
        if (pos < MCT_MAX_FLAGGED)
            a.flag_nodes[(long)c * MCT_MAX_FLAGGED + pos] =
                (int32_t)((long)z + (long)G.nz * ((long)y + (long)G.ny * x));
    }
}

// ---- exact kdtree2 replay ------------------------------------------------------
struct kd_build_frame { int l, u, parent, me, m, stage; };
#define KD_STACK 96   // recursion depth bound of the tree build / search

__device__ static void kd_spread(const double *data, const int32_t *ind, int c, int l, int u,
                                 double *lo, double *hi)
{   // spread_in_coordinate, src/kdtree2.f90:906-947
    double smin = data[3 * (ind[l] - 1) + c], smax = smin;
    int i;
    for (i = l + 2; i <= u; i += 2) {
        double lmin = data[3 * (ind[i - 1] - 1) + c], lmax = data[3 * (ind[i] - 1) + c];
        if (lmin > lmax) { double t = lmin; lmin = lmax; lmax = t; }
        if (smin > lmin) smin = lmin;
        if (smax < lmax) smax = lmax;
    }
    if (i == u + 1) {
        double last = data[3 * (ind[u] - 1) + c];
        if (smin > last) smin = last;
        if (smax < last) smax = last;
    }
    *lo = smin; *hi = smax;
}

// kdtree2_create(sort=.false., rearrange=.true.) on one thread, the recursion
// of build_tree_for_range unrolled onto an explicit stack.  Returns the root.
__device__ static int kd_build(const double *data, int n, int32_t *ind, mct_kd_node *nodes,
                               double *rdata)
{
    kd_build_frame st[KD_STACK];
    int sp = 0, nn = 0, root = -1;
    for (int j = 1; j <= n; j++) ind[j] = j;
    st[sp++] = {1, n, -1, -1, 0, 0};
    int ret = -1;            // id returned by the frame that just finished
    while (sp > 0) {
        kd_build_frame &f = st[sp - 1];
        if (f.stage == 0) {
            if (f.u < f.l) { ret = -1; sp--; goto deliver; }
            {
                int me = nn++;
                f.me = me;
                mct_kd_node &r = nodes[me];
                r.left = r.right = -1;
                r.cut_val_left = r.cut_val_right = 0.0;
                if (f.u - f.l <= MCT_KD_BUCKET) {
                    for (int c = 0; c < 3; c++) kd_spread(data, ind, c, f.l, f.u, &r.lo[c], &r.hi[c]);
                    r.cut_dim = 0; r.cut_val = 0.0; r.l = f.l; r.u = f.u;
                    ret = me; sp--; goto deliver;
                }
                for (int c = 0; c < 3; c++) {
                    bool recompute = true;
                    if (f.parent >= 0 && (c + 1) != nodes[f.parent].cut_dim) recompute = false;
                    if (recompute) kd_spread(data, ind, c, f.l, f.u, &r.lo[c], &r.hi[c]);
                    else { r.lo[c] = nodes[f.parent].lo[c]; r.hi[c] = nodes[f.parent].hi[c]; }
                }
                int cd = 0;
                double bs = r.hi[0] - r.lo[0];
                for (int c = 1; c < 3; c++) { double s = r.hi[c] - r.lo[c]; if (s > bs) { bs = s; cd = c; } }
                double sum = 0.0;
                for (int i = f.l; i <= f.u; i++) sum += data[3 * (ind[i] - 1) + cd];
                double average = sum / (double)(f.u - f.l + 1);
                r.cut_val = average;
                // select_on_coordinate_value, :850-878
                int lb = f.l, rb = f.u;
                while (lb < rb) {
                    if (data[3 * (ind[lb] - 1) + cd] <= average) lb++;
                    else { int t = ind[lb]; ind[lb] = ind[rb]; ind[rb] = t; rb--; }
                }
                f.m = (data[3 * (ind[lb] - 1) + cd] <= average) ? lb : lb - 1;
                r.cut_dim = cd + 1; r.l = f.l; r.u = f.u;
                f.stage = 1;
                if (sp >= KD_STACK) return -1;      // tree deeper than the stack: give up
                st[sp++] = {f.l, f.m, me, -1, 0, 0};
                continue;
            }
        } else if (f.stage == 1) {
            f.stage = 2;
            if (sp >= KD_STACK) return -1;
            st[sp++] = {f.m + 1, f.u, f.me, -1, 0, 0};
            continue;
        } else {
            mct_kd_node &r = nodes[f.me];
            int c = r.cut_dim - 1;
            if (r.right < 0) {
                mct_kd_node &L = nodes[r.left];
                for (int d = 0; d < 3; d++) { r.lo[d] = L.lo[d]; r.hi[d] = L.hi[d]; }
                r.cut_val_left = L.hi[c];
                r.cut_val = r.cut_val_left;
            } else if (r.left < 0) {
                mct_kd_node &R = nodes[r.right];
                for (int d = 0; d < 3; d++) { r.lo[d] = R.lo[d]; r.hi[d] = R.hi[d]; }
                r.cut_val_right = R.lo[c];
                r.cut_val = r.cut_val_right;
            } else {
                mct_kd_node &L = nodes[r.left], &R = nodes[r.right];
                r.cut_val_right = R.lo[c];
                r.cut_val_left = L.hi[c];
                r.cut_val = (r.cut_val_left + r.cut_val_right) / 2;
                for (int d = 0; d < 3; d++) { r.hi[d] = fmax(L.hi[d], R.hi[d]); r.lo[d] = fmin(L.lo[d], R.lo[d]); }
            }
            ret = f.me; sp--;
        }
    deliver:
        // hand `ret` to the parent frame as its left (stage 1) or right (stage 2) child
        if (sp == 0) { root = ret; break; }
        {
            kd_build_frame &p = st[sp - 1];
            if (p.stage == 1) nodes[p.me].left = ret;
            else if (p.stage == 2) nodes[p.me].right = ret;
        }
    }
    for (int i = 1; i <= n; i++)
        for (int d = 0; d < 3; d++) rdata[3 * (i - 1) + d] = data[3 * (ind[i] - 1) + d];
    return root;
}

__device__ __forceinline__ double kd_dis2(double x, double amin, double amax)
{   // dis2_from_bnd, :1446-1462
    if (x > amax) return (x - amax) * (x - amax);
    if (x < amin) return (amin - x) * (amin - x);
    return 0.0;
}

// kdtree2_n_nearest with nn=1 (:1028-1069), search (:1369-1443) unrolled.
__device__ static int kd_search(const mct_kd_node *nodes, int root, const int32_t *ind,
                                const double *rdata, const double q[3])
{
    struct frame { int node, stage, farther; double dis; };
    frame st[KD_STACK];
    int sp = 0, idx = 0;
    double ballsize = (double)FLT_MAX;   // huge(1.0), :1038
    st[sp++] = {root, 0, -1, 0.0};
    while (sp > 0) {
        frame &f = st[sp - 1];
        const mct_kd_node &nd = nodes[f.node];
        if (f.stage == 0) {
            if (!(nd.left >= 0 && nd.right >= 0)) {
                // process_terminal_node, :1496-1599 (nn=1: ballsize = last accepted)
                for (int i = nd.l; i <= nd.u; i++) {
                    double sd = 0.0;
                    bool skip = false;
                    for (int k = 0; k < 3; k++) {
                        double d = rdata[3 * (i - 1) + k] - q[k];
                        sd = sd + d * d;
                        if (sd > ballsize) { skip = true; break; }
                    }
                    if (skip) continue;
                    idx = ind[i];
                    ballsize = sd;
                }
                sp--;
                continue;
            }
            int cd = nd.cut_dim - 1;
            double qval = q[cd];
            int closer;
            if (qval < nd.cut_val) {
                closer = nd.left; f.farther = nd.right;
                f.dis = (nd.cut_val_right - qval) * (nd.cut_val_right - qval);
            } else {
                closer = nd.right; f.farther = nd.left;
                f.dis = (nd.cut_val_left - qval) * (nd.cut_val_left - qval);
            }
            f.stage = 1;
            if (closer >= 0) {
                if (sp >= KD_STACK) return -1;
                st[sp++] = {closer, 0, -1, 0.0};
            }
            continue;
        }
        // closer subtree done
        int farther = f.farther, cd = nd.cut_dim - 1;
        double dis = f.dis;
        sp--;
        if (farther < 0 || !(dis <= ballsize)) continue;
        bool prune = false;
        for (int i = 0; i < 3; i++) {
            if (i != cd) {
                dis = dis + kd_dis2(q[i], nd.lo[i], nd.hi[i]);
                if (dis > ballsize) { prune = true; break; }
            }
        }
        if (!prune) {
            if (sp >= KD_STACK) return -1;
            st[sp++] = {farther, 0, -1, 0.0};
        }
    }
    return idx;
}

__global__ __launch_bounds__(256) void k_nearest_replay(mct_nearest_args a)
{
    const int c = blockIdx.x;
    const int cnt = a.flag_count[c];
    if (cnt == 0) return;
    const mct_geom &G = a.G;
    const int S = a.ncells ? a.ncells[c] : a.ncells_const;
    const double *P = a.points + (long)c * a.max_cells * 3;
    mct_kd_node *nodes = a.kd_nodes + (long)c * (2 * a.max_cells + 2);
    int32_t *ind = a.kd_ind + (long)c * (a.max_cells + 1);
    double *rdata = a.kd_rdata + (long)c * 3 * a.max_cells;
    __shared__ int root;
    if (threadIdx.x == 0) root = kd_build(P, S, ind, nodes, rdata);
    __syncthreads();
    __threadfence_block();
    if (root < 0) {   // degenerate tree: keep the brute-force answer, report it
        if (threadIdx.x == 0) a.flag_count[c] = -cnt;
        return;
    }
    const bool full = cnt > MCT_MAX_FLAGGED;
    const long total = full ? (long)G.nx * G.ny * G.nz : cnt;
    for (long t = threadIdx.x; t < total; t += blockDim.x) {
        long g = full ? t : a.flag_nodes[(long)c * MCT_MAX_FLAGGED + t];
        int z = (int)(g % G.nz), y = (int)((g / G.nz) % G.ny), x = (int)(g / ((long)G.nz * G.ny));
        if (full && !(in_box(a, x, y, z) && pm_selects(a, c, x, y, z))) continue;
        double q[3] = {G.xmin + (double)x * G.dx, G.ymin + (double)y * G.dy, G.zmin + (double)z * G.dz};
        int id = kd_search(nodes, root, ind, rdata, q);
        if (id > 0) write_node(a, c, x, y, z, id);
    }
}

int mct_launch_nearest(const mct_nearest_args &a, hipStream_t s)
{
    long ntl = (long)a.G.ntiles * MCT_TILE_N;
    dim3 grid((unsigned)((ntl + NEAR_BLOCK - 1) / NEAR_BLOCK), (unsigned)a.nchain);
    int rc = mct_fill_i32(a.flag_count, a.nchain, 0, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_nearest_brute, grid, dim3(NEAR_BLOCK), 0, s, a);
    MCT_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_nearest_replay, dim3(a.nchain), dim3(256), 0, s, a);
    MCT_CHECK(hipGetLastError());
    return MCT_OK;
}
