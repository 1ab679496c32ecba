/*
 * fmm3d.c -- oracle for the 3-D first-arrival traveltime solve.
 * TEST INFRASTRUCTURE ONLY (see mct_oracle.h).
 *
 * Restates the reference's C ABI fastMarching3D (fastmarching/fastMarching.cpp
 * :26-129) with the SFMM solver it instantiates: FMM over a priority queue that
 * re-pushes a cell on every improvement and never decreases keys
 * (fmm/sfmm.hpp:44-52, fmm/fmdata/fmpriorityqueue.hpp:27-86).
 *
 * The queue is boost::heap::priority_queue<const FMCell*, compare<FMCompare>>,
 * a binary heap kept by the C++ library's push_heap/pop_heap over a vector of
 * cell pointers.  Keys are read live from the cells (FMCompare,
 * fmm/fmdata/fmcompare.hpp:29-32), so a stale entry whose cell improved carries
 * the new key.  The sift steps below follow GCC libstdc++'s push_heap /
 * pop_heap (hole-based sift-down to a leaf, then sift-up), which decide the
 * pop order between equal keys.  Boost itself is absent from this image, so the
 * equal-key order against a real Boost build is unpinned.
 *
 * Grid index: x fastest, idx = i + nx*(j + ny*k) (fastMarching.cpp:59-69).
 */
#include "mct_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <stdint.h>
#include <omp.h>

#define COMP_MARGIN (DBL_EPSILON * 1e5)   /* utils/utils.h:27 */
enum { ST_OPEN = 0, ST_NARROW = 1, ST_FROZEN = 2 };
#define NOIDX 0xffffffffu

typedef struct {
    unsigned dims[3];
    unsigned d[3];            /* d_[i] = prod(dims[0..i]) (ndgridmap.hpp resize) */
    unsigned n;
    double leaf[3];
    double *t;                /* arrival times */
    double *v;                /* velocity (occupancy) */
    uint8_t *st;
    unsigned *heap;           /* heap of cell indices */
    size_t hsize, hcap;
} fm_grid;

/* ---- heap (libstdc++ std::push_heap / std::pop_heap, comp = t[a] > t[b]) -- */
static int hcomp(const fm_grid *g, unsigned a, unsigned b) { return g->t[a] > g->t[b]; }

static void sift_up(fm_grid *g, size_t hole, size_t top, unsigned val)
{
    size_t parent = (hole - 1) / 2;
    while (hole > top && hcomp(g, g->heap[parent], val)) {
        g->heap[hole] = g->heap[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    g->heap[hole] = val;
}

static void heap_push(fm_grid *g, unsigned c)
{
    if (g->hsize == g->hcap) {
        g->hcap = g->hcap ? 2 * g->hcap : 1024;
        g->heap = (unsigned *)realloc(g->heap, g->hcap * sizeof(unsigned));
    }
    g->heap[g->hsize++] = c;
    sift_up(g, g->hsize - 1, 0, c);
}

static unsigned heap_pop(fm_grid *g)
{
    unsigned top = g->heap[0];
    size_t len = g->hsize;
    if (len > 1) {
        size_t last = len - 1;           /* element moved into the hole */
        unsigned val = g->heap[last];
        g->heap[last] = g->heap[0];
        size_t n = last;                 /* heap length during adjust */
        size_t hole = 0, child = 0;
        while (child < (n - 1) / 2) {
            child = 2 * (child + 1);
            if (hcomp(g, g->heap[child], g->heap[child - 1])) child--;
            g->heap[hole] = g->heap[child];
            hole = child;
        }
        if ((n & 1) == 0 && child == (n - 2) / 2) {
            child = 2 * (child + 1);
            g->heap[hole] = g->heap[child - 1];
            hole = child - 1;
        }
        sift_up(g, hole, 0, val);
    }
    g->hsize--;
    return top;
}

/* ---- neighbours (ndgridmap/ndgridmap.hpp:158-191, 242-275) ---------------- */
static unsigned stride(const fm_grid *g, int dim) { return dim == 0 ? 1u : g->d[dim - 1]; }

static int nbrs(const fm_grid *g, unsigned idx, unsigned out[6])
{
    int n = 0;
    for (int dim = 0; dim < 3; dim++) {
        unsigned s = stride(g, dim);
        unsigned c1 = idx - s, c2 = idx + s;
        if (c1 / g->d[dim] == idx / g->d[dim]) out[n++] = c1;
        if (c2 / g->d[dim] == idx / g->d[dim]) out[n++] = c2;
    }
    return n;
}

/* pairs (one-away, two-away) on the minus and plus side of `dim` */
static void nbr_pairs(const fm_grid *g, unsigned idx, int dim, unsigned p[4])
{
    unsigned s = stride(g, dim), D = g->d[dim], me = idx / D;
    unsigned c1a = idx - s, c1b = c1a - s, c2a = idx + s, c2b = c2a + s;
    if (c1a / D != me) c1a = NOIDX;
    if (c1b / D != me) c1b = NOIDX;
    if (c2a >= g->n || c2a / D != me) c2a = NOIDX;
    if (c2b >= g->n || c2b / D != me) c2b = NOIDX;
    p[0] = c1a; p[1] = c1b; p[2] = c2a; p[3] = c2b;
}

/* ---- local solvers (fmm/eikonalsolver.hpp) -------------------------------- */
typedef struct { double t1, t2; int dim; } tv;

static void isort(tv *a, int n) /* stable insertion sort by t1 (std::sort, n<=16) */
{
    for (int i = 1; i < n; i++) {
        tv x = a[i];
        int j = i - 1;
        while (j >= 0 && x.t1 < a[j].t1) { a[j + 1] = a[j]; j--; }
        a[j + 1] = x;
    }
}

/* solveEikonalNDims, eikonalsolver.hpp:292-322 */
static double solve_nd1(const fm_grid *g, unsigned idx, const tv *T, int dims)
{
    double vel = g->v[idx];
    if (dims == 1) return T[0].t1 + g->leaf[T[0].dim] / vel;
    double a = 0.0, b = 0.0, c = -1.0 / (vel * vel);
    for (int i = 0; i < dims; i++) {
        double h = g->leaf[T[i].dim], tval = T[i].t1;
        double alpha = 1.0 / (h * h);
        c += tval * tval * alpha;
        b += -2.0 * tval * alpha;
        a += alpha;
    }
    double q = b * b - 4 * a * c;
    if (q < 0) return INFINITY;
    return (-b + sqrt(q)) / (2 * a);
}

/* solveEikonal, eikonalsolver.hpp:50-79 */
static double solve1(const fm_grid *g, unsigned idx)
{
    tv T[3];
    int a = 0;
    for (int dim = 0; dim < 3; dim++) {
        unsigned s = stride(g, dim), D = g->d[dim], me = idx / D;
        unsigned c1 = idx - s, c2 = idx + s;
        double m;
        int have1 = (c1 / D == me), have2 = (c2 / D == me);
        if (have1 && have2) m = (g->t[c1] < g->t[c2]) ? g->t[c1] : g->t[c2];
        else if (have1) m = g->t[c1];
        else m = g->t[c2];
        if (!isinf(m) && m < g->t[idx]) { T[a].t1 = m; T[a].t2 = INFINITY; T[a].dim = dim; a++; }
    }
    if (a == 0) return INFINITY;
    isort(T, a);
    double up = INFINITY;
    for (int i = 1; i <= a; i++) {
        up = solve_nd1(g, idx, T, i);
        if (i == a || (up - T[i].t1) < COMP_MARGIN) break;
    }
    return up;
}

/* highAccuracySolveEikonalNDims, eikonalsolver.hpp:326-368 */
static double solve_nd2(const fm_grid *g, unsigned idx, const tv *T, int dims)
{
    double vel = g->v[idx];
    if (dims == 1) return T[0].t1 + g->leaf[T[0].dim] / vel;
    double a = 0.0, b = 0.0, c = -1.0 / (vel * vel);
    for (int i = 0; i < dims; i++) {
        double h = g->leaf[T[i].dim], tval = T[i].t1, tval2 = T[i].t2;
        if (!isinf(tval2)) {
            double alpha = 9.0 / 4.0 / (h * h);
            double t = (4.0 * tval - tval2) / 3.0;
            c += t * t * alpha;
            b += -2.0 * t * alpha;
            a += alpha;
        } else if (tval < DBL_MAX) {
            double alpha = 1.0 / (h * h);
            c += tval * tval * alpha;
            b += -2.0 * tval * alpha;
            a += alpha;
        }
    }
    double q = b * b - 4 * a * c;
    if (q < 0) return INFINITY;
    return (-b + sqrt(q)) / (2 * a);
}

/* highAccuracySolveEikonal, eikonalsolver.hpp:83-145 */
static double solve2(const fm_grid *g, unsigned idx)
{
    tv T[3];
    int a = 0;
    double tself = g->t[idx];
    for (int dim = 0; dim < 3; dim++) {
        unsigned p[4];
        nbr_pairs(g, idx, dim, p);
        double minT = DBL_MAX, minT2 = DBL_MAX;
        for (int side = 0; side < 2; side++) {
            unsigned first = p[2 * side], second = p[2 * side + 1];
            if (first != NOIDX) {
                double tvalue = g->t[first];
                if (!isinf(tvalue) && tvalue < minT) {
                    minT = tvalue;
                    minT2 = DBL_MAX;
                    if (second != NOIDX) {
                        double tvalue2 = g->t[second];
                        if (!isinf(tvalue2) && tvalue2 <= tvalue) minT2 = tvalue2;
                    }
                }
            }
        }
        if (minT2 < DBL_MAX && minT < tself) { T[a].t1 = minT; T[a].t2 = minT2; T[a].dim = dim; a++; }
        else if (minT < DBL_MAX && minT < tself) { T[a].t1 = minT; T[a].t2 = INFINITY; T[a].dim = dim; a++; }
    }
    if (a == 0) return tself;
    isort(T, a);
    double up = INFINITY;
    for (int i = a; i >= 1; --i) {
        double cur = solve_nd2(g, idx, T, i);
        if (!isinf(cur) && (cur - T[i - 1].t1) > COMP_MARGIN) { up = cur; break; }
    }
    return up;
}

/* ---- point helpers (ndgridmap.hpp:194-206, 433-489) ----------------------- */
static void point_coord(const fm_grid *g, const double *p, unsigned *c)
{
    for (int dim = 0; dim < 3; dim++) {
        int ci = (int)(p[dim] / g->leaf[dim]);
        if (ci < 0) ci = 0;
        else if ((unsigned)ci >= g->dims[dim]) ci = (int)g->dims[dim] - 1;
        c[dim] = (unsigned)ci;
    }
}

static void corner_nodes(const fm_grid *g, const double *p, unsigned out[8])
{
    unsigned c[3];
    point_coord(g, p, c);
    for (int v = 0; v < 8; v++) {
        unsigned vc[3];
        for (int dim = 0; dim < 3; dim++) {
            unsigned bit = (unsigned)(v >> (3 - dim - 1)) & 1u;
            vc[dim] = c[dim] + bit;
            if (vc[dim] >= g->dims[dim]) vc[dim] = g->dims[dim] - 1;
        }
        out[v] = vc[0] + vc[1] * g->d[0] + vc[2] * g->d[1];
    }
}

static void idx2coord(const fm_grid *g, unsigned idx, unsigned *c)
{
    c[2] = idx / g->d[1];
    unsigned aux = idx - c[2] * g->d[1];
    c[1] = aux / g->d[0];
    aux -= c[1] * g->d[0];
    c[0] = aux;
}

/* getVel / getTime: trilinear with clamped corners (ndgridmap.hpp:433-489) */
static double interp(const fm_grid *g, const double *p, const double *field)
{
    unsigned pc[3], nb[8];
    point_coord(g, p, pc);
    corner_nodes(g, p, nb);
    double acc = 0;
    for (int i = 0; i < 8; i++) {
        unsigned c[3];
        idx2coord(g, nb[i], c);
        double w = 1.0;
        for (int j = 0; j < 3; j++) {
            unsigned di = c[j] - pc[j];
            double s = p[j] - g->leaf[j] * pc[j];
            double wj = fabs((double)di * g->leaf[j] - g->leaf[j] + s);
            w = w * wj / g->leaf[j];
        }
        acc = acc + w * field[nb[i]];
    }
    return acc;
}

/* FMM::computeInternal, fmm/fmm.hpp:92-158 */
static void fmm_run(fm_grid *g, int high)
{
    while (g->hsize > 0) {
        unsigned idxMin = heap_pop(g);
        unsigned nb[6];
        int nn = nbrs(g, idxMin, nb);
        g->st[idxMin] = ST_FROZEN;
        for (int s = 0; s < nn; s++) {
            unsigned j = nb[s];
            if (g->st[j] == ST_FROZEN || g->v[j] < COMP_MARGIN) continue;
            double nt = high ? solve2(g, j) : solve1(g, j);
            if (g->st[j] == ST_NARROW) {
                if (nt + COMP_MARGIN < g->t[j]) { g->t[j] = nt; heap_push(g, j); }
            } else {
                g->st[j] = ST_NARROW;
                g->t[j] = nt;
                heap_push(g, j);
            }
        }
    }
}

int orc_fastMarching3D(int nx, int ny, int nz, double dx, double dy, double dz,
                       double xmin, double ymin, double zmin,
                       const double *vel, const double *src, const double *rev,
                       int nrev, int order, double *ttime, double *tfield)
{
    fm_grid g;
    memset(&g, 0, sizeof g);
    g.dims[0] = (unsigned)nx; g.dims[1] = (unsigned)ny; g.dims[2] = (unsigned)nz;
    g.d[0] = g.dims[0]; g.d[1] = g.d[0] * g.dims[1]; g.d[2] = g.d[1] * g.dims[2];
    g.n = g.d[2];
    g.leaf[0] = dx; g.leaf[1] = dy; g.leaf[2] = dz;
    g.t = (double *)malloc((size_t)g.n * sizeof(double));
    g.v = (double *)malloc((size_t)g.n * sizeof(double));
    g.st = (uint8_t *)calloc(g.n, 1);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                unsigned ind = (unsigned)i + (unsigned)j * nx + (unsigned)k * nx * ny;
                g.v[ind] = vel[((size_t)i * ny + j) * nz + k];
            }
    for (unsigned i = 0; i < g.n; i++) g.t[i] = INFINITY;

    double sp[3] = { src[0] - xmin, src[1] - ymin, src[2] - zmin };
    /* Solver::setInitialPoints(point), fmm/solver.hpp:72-92 */
    double vs = interp(&g, sp, g.v);
    unsigned nb[8];
    corner_nodes(&g, sp, nb);
    double t0[8];
    for (int i = 0; i < 8; i++) {
        unsigned c[3];
        idx2coord(&g, nb[i], c);
        double dist = 0;
        for (int d = 0; d < 3; d++) {
            double xi = (c[d] * g.leaf[d] - sp[d]) * (c[d] * g.leaf[d] - sp[d]);
            dist = dist + xi;
        }
        dist = sqrt(dist);
        t0[i] = dist * 2 / (vs + g.v[nb[i]]);
    }
    for (int i = 0; i < 8; i++) {
        g.t[nb[i]] = t0[i];
        g.st[nb[i]] = ST_NARROW;
        heap_push(&g, nb[i]);
    }
    fmm_run(&g, order == 1);

    for (int r = 0; r < nrev; r++) {
        double p[3] = { rev[3 * r] - xmin, rev[3 * r + 1] - ymin, rev[3 * r + 2] - zmin };
        ttime[r] = interp(&g, p, g.t);
    }
    if (tfield) {
        for (int i = 0; i < nx; i++)
            for (int j = 0; j < ny; j++)
                for (int k = 0; k < nz; k++)
                    tfield[((size_t)i * ny + j) * nz + k] = g.t[(unsigned)i + (unsigned)j * nx + (unsigned)k * nx * ny];
    }
    free(g.t); free(g.v); free(g.st); free(g.heap);
    return 0;
}

int orc_body_traveltimes(int nx, int ny, int nz, double dx, double dy, double dz,
                         double xmin, double ymin, double zmin, double scaling,
                         const double *vel, int nsrc, const double *locations,
                         int nrev, const double *rev, int order,
                         double *phase_time, int nthreads)
{
    /* likelihood_body.F90:204-224: dz and zmin are divided by grid%scaling */
    double dzs = dz / scaling, zmins = zmin / scaling;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
    for (int i = 0; i < nsrc; i++)
        orc_fastMarching3D(nx, ny, nz, dx, dy, dzs, xmin, ymin, zmins, vel, &locations[4 * i],
                           rev, nrev, order, &phase_time[(size_t)i * nrev], NULL);
    return 0;
}

/* Test hook: replay push / key-change / pop operations on the oracle's heap.
 * ops (nops,2) = (kind, cell): kind 0 push cell, 1 keys[cell] = vals[i],
 * 2 pop (popped cell appended to `popped`).  Returns the number of pops. */
int orc_heap_replay(int ncell, double *keys, int nops, const int *ops, const double *vals,
                    int *popped)
{
    fm_grid g;
    memset(&g, 0, sizeof g);
    g.t = keys;
    g.n = (unsigned)ncell;
    int np = 0;
    for (int i = 0; i < nops; i++) {
        int kind = ops[2 * i], c = ops[2 * i + 1];
        if (kind == 0) heap_push(&g, (unsigned)c);
        else if (kind == 1) keys[c] = vals[i];
        else if (g.hsize > 0) popped[np++] = (int)heap_pop(&g);
    }
    free(g.heap);
    return np;
}
