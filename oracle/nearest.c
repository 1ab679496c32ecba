/*
 * nearest.c -- oracle for Voronoi -> grid nearest-site assignment.
 * TEST INFRASTRUCTURE ONLY (see mct_oracle.h).
 *
 * Restates the kd-tree of src/kdtree2.f90 (Kennel's kdtree2, double
 * precision, bucket_size = 12, sort=.false., rearrange=.true.) and the
 * caller kdtree_to_grid (src/mcmc_loc2.f90:2002-2082).  The tree is needed,
 * not only a brute-force argmin, because the reference's answer on exact and
 * near ties depends on the visiting order and on the box pruning.
 */
#include "mct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define KD_BUCKET 12 /* src/kdtree2.f90:516 */

typedef struct {
    int cut_dim;               /* 1-based, 0 for a bucket */
    double cut_val, cut_val_left, cut_val_right;
    int l, u;                  /* 1-based range into ind[] */
    int left, right;           /* child node ids, -1 = null */
    double lo[3], hi[3];       /* bounding box */
} kd_node;

typedef struct {
    const double *data;        /* (3,n) original data */
    int n;
    int *ind;                  /* 1-based permutation, ind[1..n] */
    double *rdata;             /* rearranged (3,n): rdata(:,i) = data(:,ind(i)) */
    kd_node *nodes;
    int nnodes, cap;
    int root;
} kd_tree;

static double V(const kd_tree *t, int c, int idx) { return t->data[3 * (idx - 1) + (c - 1)]; }

/* spread_in_coordinate, src/kdtree2.f90:906-947 */
static void spread(const kd_tree *t, int c, int l, int u, double *lo, double *hi)
{
    double smin = V(t, c, t->ind[l]), smax = smin;
    int i;
    for (i = l + 2; i <= u; i += 2) {
        double lmin = V(t, c, t->ind[i - 1]), lmax = V(t, c, t->ind[i]);
        if (lmin > lmax) { double tt = lmin; lmin = lmax; lmax = tt; }
        if (smin > lmin) smin = lmin;
        if (smax < lmax) smax = lmax;
    }
    if (i == u + 1) {
        double last = V(t, c, t->ind[u]);
        if (smin > last) smin = last;
        if (smax < last) smax = last;
    }
    *lo = smin; *hi = smax;
}

/* select_on_coordinate_value, src/kdtree2.f90:850-878 */
static int select_value(kd_tree *t, int c, double alpha, int li, int ui)
{
    int lb = li, rb = ui;
    while (lb < rb) {
        if (V(t, c, t->ind[lb]) <= alpha) lb++;
        else { int tmp = t->ind[lb]; t->ind[lb] = t->ind[rb]; t->ind[rb] = tmp; rb--; }
    }
    return (V(t, c, t->ind[lb]) <= alpha) ? lb : lb - 1;
}

static int new_node(kd_tree *t)
{
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 64;
        t->nodes = (kd_node *)realloc(t->nodes, (size_t)t->cap * sizeof(kd_node));
    }
    memset(&t->nodes[t->nnodes], 0, sizeof(kd_node));
    return t->nnodes++;
}

/* build_tree_for_range, src/kdtree2.f90:704-840 */
static int build_range(kd_tree *t, int l, int u, int parent)
{
    if (u < l) return -1;
    int me = new_node(t);
    kd_node *r = &t->nodes[me];
    if (u - l <= KD_BUCKET) {
        for (int c = 1; c <= 3; c++) spread(t, c, l, u, &r->lo[c - 1], &r->hi[c - 1]);
        r->cut_dim = 0; r->cut_val = 0.0; r->l = l; r->u = u; r->left = r->right = -1;
        return me;
    }
    for (int c = 1; c <= 3; c++) {
        int recompute = 1;
        if (parent >= 0 && c != t->nodes[parent].cut_dim) recompute = 0;
        if (recompute) spread(t, c, l, u, &r->lo[c - 1], &r->hi[c - 1]);
        else { r->lo[c - 1] = t->nodes[parent].lo[c - 1]; r->hi[c - 1] = t->nodes[parent].hi[c - 1]; }
    }
    int c = 1; /* maxloc: first maximal spread */
    double best = r->hi[0] - r->lo[0];
    for (int d = 2; d <= 3; d++) {
        double s = r->hi[d - 1] - r->lo[d - 1];
        if (s > best) { best = s; c = d; }
    }
    double sum = 0.0;
    for (int i = l; i <= u; i++) sum += V(t, c, t->ind[i]);
    double average = sum / (double)(u - l + 1);
    r->cut_val = average;
    int m = select_value(t, c, average, l, u);
    r->cut_dim = c; r->l = l; r->u = u;
    int left = build_range(t, l, m, me);
    int right = build_range(t, m + 1, u, me);
    r = &t->nodes[me]; /* realloc may have moved the array */
    r->left = left; r->right = right;
    if (right < 0) {
        kd_node *L = &t->nodes[left];
        memcpy(r->lo, L->lo, sizeof r->lo); memcpy(r->hi, L->hi, sizeof r->hi);
        r->cut_val_left = L->hi[c - 1];
        r->cut_val = r->cut_val_left;
    } else if (left < 0) {
        kd_node *R = &t->nodes[right];
        memcpy(r->lo, R->lo, sizeof r->lo); memcpy(r->hi, R->hi, sizeof r->hi);
        r->cut_val_right = R->lo[c - 1];
        r->cut_val = r->cut_val_right;
    } else {
        kd_node *L = &t->nodes[left], *R = &t->nodes[right];
        r->cut_val_right = R->lo[c - 1];
        r->cut_val_left = L->hi[c - 1];
        r->cut_val = (r->cut_val_left + r->cut_val_right) / 2;
        for (int d = 0; d < 3; d++) {
            r->hi[d] = fmax(L->hi[d], R->hi[d]);
            r->lo[d] = fmin(L->lo[d], R->lo[d]);
        }
    }
    return me;
}

/* kdtree2_create, src/kdtree2.f90:609-689 */
static void kd_create(kd_tree *t, const double *pts, int n)
{
    memset(t, 0, sizeof *t);
    t->data = pts; t->n = n;
    t->ind = (int *)malloc((size_t)(n + 1) * sizeof(int));
    for (int j = 1; j <= n; j++) t->ind[j] = j;
    t->root = build_range(t, 1, n, -1);
    t->rdata = (double *)malloc((size_t)3 * (n + 1) * sizeof(double));
    for (int i = 1; i <= n; i++)
        for (int d = 0; d < 3; d++) t->rdata[3 * (i - 1) + d] = pts[3 * (t->ind[i] - 1) + d];
}

static void kd_free(kd_tree *t) { free(t->ind); free(t->rdata); free(t->nodes); }

typedef struct { double ballsize; int nfound; int idx; const double *qv; } kd_search;

static double dis2_from_bnd(double x, double amin, double amax) /* :1446-1462 */
{
    if (x > amax) return (x - amax) * (x - amax);
    if (x < amin) return (amin - x) * (amin - x);
    return 0.0;
}

/* process_terminal_node, src/kdtree2.f90:1496-1599, specialised to nn=1:
 * pq_insert / pq_replace_max with one element make ballsize the distance of
 * the last accepted point (:296-338, :381-450). */
static void terminal(const kd_tree *t, const kd_node *nd, kd_search *s)
{
    double ballsize = s->ballsize;
    for (int i = nd->l; i <= nd->u; i++) {
        double sd = 0.0;
        int skip = 0;
        for (int k = 0; k < 3; k++) {
            double d = t->rdata[3 * (i - 1) + k] - s->qv[k];
            sd = sd + d * d;
            if (sd > ballsize) { skip = 1; break; }
        }
        if (skip) continue;
        s->idx = t->ind[i];
        s->nfound = 1;
        ballsize = sd;
    }
    s->ballsize = ballsize;
}

/* search, src/kdtree2.f90:1369-1443 */
static void search(const kd_tree *t, int node, kd_search *s)
{
    const kd_node *nd = &t->nodes[node];
    if (!(nd->left >= 0 && nd->right >= 0)) { terminal(t, nd, s); return; }
    int cd = nd->cut_dim;
    double qval = s->qv[cd - 1], dis;
    int closer, farther;
    if (qval < nd->cut_val) {
        closer = nd->left; farther = nd->right;
        dis = (nd->cut_val_right - qval) * (nd->cut_val_right - qval);
    } else {
        closer = nd->right; farther = nd->left;
        dis = (nd->cut_val_left - qval) * (nd->cut_val_left - qval);
    }
    if (closer >= 0) search(t, closer, s);
    if (farther >= 0) {
        double ballsize = s->ballsize;
        if (dis <= ballsize) {
            for (int i = 1; i <= 3; i++) {
                if (i != cd) {
                    dis = dis + dis2_from_bnd(s->qv[i - 1], nd->lo[i - 1], nd->hi[i - 1]);
                    if (dis > ballsize) return;
                }
            }
            search(t, farther, s);
        }
    }
}

static int kd_nearest(const kd_tree *t, const double *q)
{
    kd_search s;
    s.ballsize = (double)FLT_MAX; /* huge(1.0): default-real huge, :1038 */
    s.nfound = 0; s.idx = 0; s.qv = q;
    search(t, t->root, &s);
    return s.idx;
}

int orc_kdtree_nearest(const double *points, int ncells, const double *q)
{
    kd_tree t;
    kd_create(&t, points, ncells);
    int r = kd_nearest(&t, q);
    kd_free(&t);
    return r;
}

int orc_brute_nearest(const double *points, int ncells, const double *q,
                      double *best_sd, int *n_at_best)
{
    double best = INFINITY;
    int bi = 0, cnt = 0;
    for (int c = 0; c < ncells; c++) {
        double sd = 0.0;
        for (int k = 0; k < 3; k++) { double d = points[3 * c + k] - q[k]; sd = sd + d * d; }
        if (sd < best) { best = sd; bi = c + 1; cnt = 1; }
        else if (sd == best) cnt++;
    }
    if (best_sd) *best_sd = best;
    if (n_at_best) *n_at_best = cnt;
    return bi;
}

/* kdtree_to_grid, src/mcmc_loc2.f90:2002-2082 */
int orc_kdtree_to_grid(int nx, int ny, int nz,
                       double xmin, double ymin, double zmin,
                       double dx, double dy, double dz,
                       const double *points, const double *params, int ncells,
                       const double *box, const double *pm,
                       int32_t *sites_id, double *vp, double *vs, double *rho)
{
    const double eps = (double)1e-8f; /* real(ii10), parameter :: eps = 1e-8 (default-real literal), :51 */
    kd_tree t;
    kd_create(&t, points, ncells);
    int ix0 = 1, ix1 = nx, iy0 = 1, iy1 = ny, iz0 = 1, iz1 = nz;
    if (box) {
        ix0 = (int)floor((box[0] - xmin) / dx) + 1; ix1 = (int)floor((box[3] - xmin) / dx) + 1;
        iy0 = (int)floor((box[1] - ymin) / dy) + 1; iy1 = (int)floor((box[4] - ymin) / dy) + 1;
        iz0 = (int)floor((box[2] - zmin) / dz) + 1; iz1 = (int)floor((box[5] - zmin) / dz) + 1;
        if (ix0 < 1) ix0 = 1;
        if (iy0 < 1) iy0 = 1;
        if (iz0 < 1) iz0 = 1;
        if (ix1 > nx) ix1 = nx;
        if (iy1 > ny) iy1 = ny;
        if (iz1 > nz) iz1 = nz;
    }
    #pragma omp parallel for schedule(static)
    for (int i = ix0; i <= ix1; i++)
        for (int j = iy0; j <= iy1; j++)
            for (int k = iz0; k <= iz1; k++) {
                size_t g = (size_t)(k - 1) + (size_t)nz * ((size_t)(j - 1) + (size_t)ny * (size_t)(i - 1));
                double qv[3] = { xmin + (double)(i - 1) * dx, ymin + (double)(j - 1) * dy,
                                 zmin + (double)(k - 1) * dz };
                if (pm && !(fabs(vs[g] - pm[1]) < eps && fabs(vp[g] - pm[0]) < eps)) continue;
                int id = kd_nearest(&t, qv);
                sites_id[g] = id;
                vp[g] = params[3 * (id - 1) + 0];
                vs[g] = params[3 * (id - 1) + 1];
                rho[g] = params[3 * (id - 1) + 2];
            }
    kd_free(&t);
    return 0;
}
