"""Pin the oracle's SFMM (oracle/fmm3d.c) with an independent restatement.

The reference's fastMarching3D (fastmarching/fastMarching.cpp:26-129) needs
Boost, which this image lacks, so it is not compiled here.  This file restates
the whole solve a second time, separately and in Python, from the reference's
text: the corner initialisation (fmm/solver.hpp:72-107, getVel
ndgridmap/ndgridmap.hpp:442-463), FMM::computeInternal (fmm/fmm.hpp:92-158)
over boost::heap::priority_queue's algorithm (std::push_heap / std::pop_heap
on a vector of cells compared by their live arrival times,
fmm/fmdata/fmpriorityqueue.hpp:27-86, fmm/fmdata/fmcompare.hpp:29-32),
solveEikonal (fmm/eikonalsolver.hpp:50-79, 292-322) and
highAccuracySolveEikonal (:83-145, 326-368: the 9/4 and (4t - t2)/3
coefficients, the two-away time taken only when t2 <= t, dimensions tried from
a down to 1 with the first result more than COMP_MARGIN above its largest
neighbour accepted, COMP_MARGIN = 1e5 * DBL_EPSILON, utils/utils.h:27-34).

On tiny grids with random velocities the oracle must equal it bit for bit, the
whole field and the trilinear receiver picks, for both orders.  Each rule of
the second-order solve is then broken on purpose in this restatement only (on
the CPU); every planted mistake must make it disagree with the oracle on some
case, so the comparison above would see that mistake in the oracle.
"""
import math
import sys

import numpy as np
import pytest

from oracle import pyoracle as po

EPS = sys.float_info.epsilon
MARGIN = EPS * 1e5
UMAX = 0xFFFFFFFF
OPEN, NARROW, FROZEN = 0, 1, 2


class Sfmm:
    """One fastMarching3D solve; `bug` names a rule to break (or None)."""

    def __init__(self, vel, leaf, bug=None):
        self.nx, self.ny, self.nz = vel.shape
        self.d = [self.nx, self.nx * self.ny, self.nx * self.ny * self.nz]
        self.n = self.d[2]
        self.leaf = leaf
        self.v = [0.0] * self.n
        for i in range(self.nx):
            for j in range(self.ny):
                for k in range(self.nz):
                    self.v[i + j * self.nx + k * self.nx * self.ny] = float(vel[i, j, k])
        self.t = [math.inf] * self.n
        self.state = [OPEN] * self.n
        self.heap = []
        self.bug = bug

    # --- boost::heap::priority_queue = std::push_heap / std::pop_heap, min on live time
    def _less_prio(self, a, b):           # FMCompare: a after b when a's time is larger
        return self.t[a] > self.t[b]

    def _sift_up(self, hole, top, val):
        h = self.heap
        parent = (hole - 1) // 2
        while hole > top and self._less_prio(h[parent], val):
            h[hole] = h[parent]
            hole = parent
            parent = (hole - 1) // 2
        h[hole] = val

    def push(self, c):
        self.heap.append(c)
        self._sift_up(len(self.heap) - 1, 0, c)

    def pop(self):
        h = self.heap
        top = h[0]
        n = len(h) - 1
        if n > 0:
            val = h[n]
            h[n] = h[0]
            hole, child = 0, 0
            while child < (n - 1) // 2:
                child = 2 * (child + 1)
                if self._less_prio(h[child], h[child - 1]):
                    child -= 1
                h[hole] = h[child]
                hole = child
            if n % 2 == 0 and child == (n - 2) // 2:
                child = 2 * (child + 1)
                h[hole] = h[child - 1]
                hole = child - 1
            self._sift_up(hole, 0, val)
        h.pop()
        return top

    # --- grid helpers (ndgridmap.hpp) -------------------------------------------------
    def _step(self, dim):
        return 1 if dim == 0 else self.d[dim - 1]

    def _same(self, c, idx, dim):
        return 0 <= c < self.n and c // self.d[dim] == idx // self.d[dim]

    def neighbours(self, idx):
        out = []
        for dim in range(3):
            s = self._step(dim)
            for c in (idx - s, idx + s):
                if self._same(c, idx, dim):
                    out.append(c)
        return out

    def coord(self, idx):
        k = idx // self.d[1]
        r = idx - k * self.d[1]
        j = r // self.d[0]
        return (r - j * self.d[0], j, k)

    def cell_of(self, p):
        dims = (self.nx, self.ny, self.nz)
        c = []
        for dim in range(3):
            q = int(p[dim] / self.leaf[dim])
            c.append(0 if q < 0 else (dims[dim] - 1 if q >= dims[dim] else q))
        return c

    def corners(self, p):
        dims = (self.nx, self.ny, self.nz)
        c = self.cell_of(p)
        out = []
        for v in range(8):
            vc = [min(c[dim] + ((v >> (2 - dim)) & 1), dims[dim] - 1) for dim in range(3)]
            out.append(vc[0] + vc[1] * self.d[0] + vc[2] * self.d[1])
        return out

    def trilinear(self, p, field):
        pc = self.cell_of(p)
        acc = 0.0
        for nb in self.corners(p):
            c = self.coord(nb)
            w = 1.0
            for j in range(3):
                di = c[j] - pc[j]
                s = p[j] - self.leaf[j] * pc[j]
                w = w * abs(di * self.leaf[j] - self.leaf[j] + s) / self.leaf[j]
            acc = acc + w * field[nb]
        return acc

    # --- local solves (eikonalsolver.hpp) ------------------------------------------------
    def solve1(self, idx):
        vals = []
        for dim in range(3):
            s = self._step(dim)
            nb = [c for c in (idx - s, idx + s) if self._same(c, idx, dim)]
            m = self.t[nb[0]] if len(nb) == 1 else (self.t[nb[0]] if self.t[nb[0]] < self.t[nb[1]] else self.t[nb[1]])
            if not math.isinf(m) and m < self.t[idx]:
                vals.append((m, dim))
        if not vals:
            return math.inf
        vals.sort(key=lambda e: e[0])
        vel = self.v[idx]
        for i in range(1, len(vals) + 1):
            if i == 1:
                up = vals[0][0] + self.leaf[vals[0][1]] / vel
            else:
                a, b, c = 0.0, 0.0, -1.0 / (vel * vel)
                for tval, dim in vals[:i]:
                    al = 1.0 / (self.leaf[dim] * self.leaf[dim])
                    c += tval * tval * al
                    b += -2.0 * tval * al
                    a += al
                q = b * b - 4 * a * c
                up = math.inf if q < 0 else (-b + math.sqrt(q)) / (2 * a)
            if i == len(vals) or (up - vals[i][0]) < MARGIN:
                break
        return up

    def solve2(self, idx):
        vals = []
        bug = self.bug
        tself = self.t[idx]
        for dim in range(3):
            s = self._step(dim)
            sides = []
            for one, two in ((idx - s, idx - 2 * s), (idx + s, idx + 2 * s)):
                sides.append((one if self._same(one, idx, dim) else UMAX,
                              two if self._same(two, idx, dim) else UMAX))
            mt, mt2 = sys.float_info.max, sys.float_info.max
            for one, two in sides:
                if one != UMAX:
                    tv = self.t[one]
                    if not math.isinf(tv) and tv < mt:
                        mt = tv
                        mt2 = sys.float_info.max
                        if two != UMAX:
                            tv2 = self.t[two]
                            ok = (tv2 < tv) if bug == "t2_strict" else (tv2 <= tv)
                            if not math.isinf(tv2) and ok:
                                mt2 = tv2
            if mt2 < sys.float_info.max and mt < tself:
                vals.append((mt, mt2, dim))
            elif mt < sys.float_info.max and mt < tself:
                vals.append((mt, math.inf, dim))
        if not vals:
            return tself
        vals.sort(key=lambda e: e[0])
        vel = self.v[idx]

        def nd(i):
            if i == 1:
                return vals[0][0] + self.leaf[vals[0][2]] / vel
            a, b, c = 0.0, 0.0, -1.0 / (vel * vel)
            for tval, tval2, dim in vals[:i]:
                h = self.leaf[dim]
                if not math.isinf(tval2):
                    al = (2.0 if bug == "coef_9_4" else 9.0 / 4.0) / (h * h)
                    t = (4.0 * tval - tval2) / (2.0 if bug == "div_3" else 3.0)
                    c += t * t * al
                    b += -2.0 * t * al
                    a += al
                elif tval < sys.float_info.max:
                    al = 1.0 / (h * h)
                    c += tval * tval * al
                    b += -2.0 * tval * al
                    a += al
            q = b * b - 4 * a * c
            return math.inf if q < 0 else (-b + math.sqrt(q)) / (2 * a)

        order = range(1, len(vals) + 1) if bug == "upward" else range(len(vals), 0, -1)
        for i in order:
            cur = nd(i)
            gap = cur - vals[i - 1][0]
            ok = (gap > 0) if bug == "no_margin" else (gap > MARGIN)
            if not math.isinf(cur) and ok:
                return cur
        return math.inf

    # --- FMM::computeInternal ------------------------------------------------------------
    def run(self, src, high):
        p = list(src)
        vs = self.trilinear(p, self.v)
        init = []
        for nb in self.corners(p):
            c = self.coord(nb)
            dist = 0.0
            for d in range(3):
                xi = (c[d] * self.leaf[d] - p[d]) * (c[d] * self.leaf[d] - p[d])
                dist = dist + xi
            init.append((nb, math.sqrt(dist) * 2 / (vs + self.v[nb])))
        for nb, t0 in init:
            self.t[nb] = t0
            self.state[nb] = NARROW
            self.push(nb)
        while self.heap:
            i = self.pop()
            nbs = self.neighbours(i)
            self.state[i] = FROZEN
            for j in nbs:
                if self.state[j] == FROZEN or self.v[j] < MARGIN:
                    continue
                nt = self.solve2(j) if high else self.solve1(j)
                if self.state[j] == NARROW:
                    if nt + MARGIN < self.t[j]:
                        self.t[j] = nt
                        self.push(j)
                else:
                    self.state[j] = NARROW
                    self.t[j] = nt
                    self.push(j)


def restated(vel, h, xmin, src, rev, order, bug=None):
    S = Sfmm(vel, h, bug)
    S.run([src[d] - xmin[d] for d in range(3)], order == 1)
    tt = np.array([S.trilinear([r[d] - xmin[d] for d in range(3)], S.t) for r in rev])
    f = np.empty(vel.shape)
    for i in range(S.nx):
        for j in range(S.ny):
            for k in range(S.nz):
                f[i, j, k] = S.t[i + j * S.nx + k * S.nx * S.ny]
    return tt, f


def cases(n):
    rng = np.random.default_rng(20261203)
    out = []
    for c in range(n):
        shape = tuple(int(v) for v in rng.integers(4, 9, 3))
        h = tuple(float(v) for v in rng.choice([0.25, 0.3, 0.5], 3))
        vel = rng.uniform(2.0, 6.0, shape)
        if c % 4 == 3:                       # a sharp block
            vel[: shape[0] // 2] *= 0.5
        xmin = (0.0, -1.0, 0.5) if c % 2 else (0.0, 0.0, 0.0)
        ext = [h[d] * (shape[d] - 1) for d in range(3)]
        src = np.array([xmin[d] + rng.uniform(0, ext[d]) for d in range(3)])
        if c % 5 == 4:                       # a source on a grid node
            src = np.array([xmin[d] + h[d] * int(rng.integers(0, shape[d])) for d in range(3)])
        rev = np.array([[xmin[d] + rng.uniform(0, ext[d]) for d in range(3)] for _ in range(6)])
        out.append((vel, h, xmin, src, rev))
    # homogeneous media with the source at a cell centre: the eight corner times
    # are equal, so one- and two-away times tie along the axes (t2 == t)
    for c in range(4):
        shape = tuple(int(v) for v in rng.integers(5, 9, 3))
        h = (0.5, 0.5, 0.5) if c % 2 == 0 else (0.25, 0.5, 0.3)
        vel = np.full(shape, 3.0 + c)
        src = np.array([h[d] * (int(rng.integers(0, shape[d] - 1)) + 0.5) for d in range(3)])
        rev = np.array([[h[d] * rng.uniform(0, shape[d] - 1) for d in range(3)] for _ in range(6)])
        out.append((vel, h, (0.0, 0.0, 0.0), src, rev))
    return out


@pytest.mark.parametrize("order", [0, 1])
def test_oracle_equals_independent_restatement(order):
    for idx, (vel, h, xmin, src, rev) in enumerate(cases(40)):
        want_t, want_f = restated(vel, h, xmin, src, rev, order)
        got_t, got_f = po.fast_marching_3d(vel, h[0], h[1], h[2], xmin[0], xmin[1], xmin[2], src, rev, order,
                                           field=True)
        np.testing.assert_array_equal(got_f, want_f, err_msg="case %d field" % idx)
        np.testing.assert_array_equal(got_t, want_t, err_msg="case %d picks" % idx)


@pytest.mark.parametrize("bug", ["coef_9_4", "div_3", "t2_strict", "upward", "no_margin"])
def test_each_planted_rule_mistake_is_seen(bug):
    """A wrong second-order coefficient, divisor, two-away rule, dimension order
    or margin in the restatement makes it disagree with the oracle."""
    seen = 0
    for vel, h, xmin, src, rev in cases(40):
        _, want_f = restated(vel, h, xmin, src, rev, 1, bug=bug)
        _, got_f = po.fast_marching_3d(vel, h[0], h[1], h[2], xmin[0], xmin[1], xmin[2], src, rev, 1,
                                       field=True)
        seen += not np.array_equal(got_f, want_f)
    assert seen > 0, bug
