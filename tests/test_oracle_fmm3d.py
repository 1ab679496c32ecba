"""Pin the 3-D SFMM oracle (oracle/fmm3d.c).

The reference's fastMarching.cpp needs Boost, which this image lacks, so it is
not compiled here; the oracle is pinned by what the mathematics fixes instead:
exact axis times, the first-order discrete fixed point (solved by an
independent Gauss-Seidel sweep), convergence order in a homogeneous medium,
reciprocity, the corner initialisation, trilinear picks, and the queue's pop
order against the C++ library heap that boost::heap::priority_queue uses.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))


def sweep_fixed_point(vel, h, src_nodes, order0=True, iters=60):
    """Independent restatement: Gauss-Seidel sweeps (8 orderings) of the
    first-order upwind update min(T, U(neighbours)) to convergence."""
    nx, ny, nz = vel.shape
    T = np.full(vel.shape, np.inf)
    for (i, j, k), t in src_nodes.items():
        T[i, j, k] = t
    for _ in range(iters):
        old = T.copy()
        for sx in (1, -1):
            for sy in (1, -1):
                for sz in (1, -1):
                    for i in (range(nx) if sx > 0 else range(nx - 1, -1, -1)):
                        for j in (range(ny) if sy > 0 else range(ny - 1, -1, -1)):
                            for k in (range(nz) if sz > 0 else range(nz - 1, -1, -1)):
                                m = []
                                for d, (a, b) in enumerate((((i - 1, j, k), (i + 1, j, k)),
                                                            ((i, j - 1, k), (i, j + 1, k)),
                                                            ((i, j, k - 1), (i, j, k + 1)))):
                                    vals = [T[p] for p in (a, b)
                                            if 0 <= p[0] < nx and 0 <= p[1] < ny and 0 <= p[2] < nz]
                                    mm = min(vals)
                                    if np.isfinite(mm) and mm < T[i, j, k]:
                                        m.append(mm)
                                if not m:
                                    continue
                                m.sort()
                                s2 = 1.0 / vel[i, j, k] ** 2
                                u = np.inf
                                for n in range(1, len(m) + 1):
                                    if n == 1:
                                        u = m[0] + h / vel[i, j, k]
                                    else:
                                        a = n / h**2
                                        b = -2 * sum(m[:n]) / h**2
                                        c = sum(x * x for x in m[:n]) / h**2 - s2
                                        q = b * b - 4 * a * c
                                        u = np.inf if q < 0 else (-b + np.sqrt(q)) / (2 * a)
                                    if n == len(m) or u - m[n] < 2.2e-11:
                                        break
                                if u + 2.2e-11 < T[i, j, k]:
                                    T[i, j, k] = u
        if np.array_equal(old, T):
            break
    return T


def test_axis_nodes_are_exact_in_homogeneous_medium():
    n, h, v = 21, 0.5, 4.0
    vel = np.full((n, n, n), v)
    src = np.array([5.0, 5.0, 5.0])           # on node (10,10,10)
    rev = np.array([[5.0, 5.0, 0.0], [0.0, 5.0, 5.0], [5.0, 9.5, 5.0]])
    for order in (0, 1):
        tt = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, src, rev, order)
        np.testing.assert_allclose(tt, np.linalg.norm(rev - src, axis=1) / v, rtol=1e-12)


def test_first_order_equals_independent_fixed_point():
    rng = np.random.default_rng(5)
    n, h = 9, 0.5
    vel = rng.uniform(3.0, 6.0, (n, n, n))
    src = np.array([1.3, 2.2, 1.7])
    tt, tf = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, src, np.zeros((1, 3)), 0, field=True)
    # corner initial times, solver.hpp:72-92 (t = 2 d / (v_src + v_node))
    init = {}
    c = (src / h).astype(int)
    w = src / h - c
    vs = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wt = (w[0] if dx else 1 - w[0]) * (w[1] if dy else 1 - w[1]) * (w[2] if dz else 1 - w[2])
                vs += wt * vel[c[0] + dx, c[1] + dy, c[2] + dz]
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                p = (c[0] + dx, c[1] + dy, c[2] + dz)
                init[p] = np.linalg.norm(np.array(p) * h - src) * 2 / (vs + vel[p])
    ref = sweep_fixed_point(vel, h, init)
    np.testing.assert_allclose(tf, ref, rtol=1e-10)
    for p, t in init.items():
        assert tf[p] <= t + 1e-15                 # corners start NARROW: may only improve


def test_convergence_order_homogeneous():
    v = 5.0
    src = np.array([4.13, 3.91, 2.07])
    rev = np.array([[0.5, 7.5, 0.0], [7.2, 0.4, 0.0], [6.0, 6.0, 3.5]])
    exact = np.linalg.norm(rev - src, axis=1) / v
    err = {0: [], 1: []}
    for n in (17, 33, 65):
        h = 8.0 / (n - 1)
        vel = np.full((n, n, n), v)
        for order in (0, 1):
            tt = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, src, rev, order)
            err[order].append(np.abs(tt - exact).max())
    # first order is O(h log 1/h); the mixed scheme must be clearly more accurate
    assert err[0][2] < 0.5 * err[0][0]
    assert err[1][2] < 0.25 * err[1][0]
    assert err[1][2] < 0.25 * err[0][2]


def test_reciprocity_first_order():
    rng = np.random.default_rng(2)
    n, h = 25, 0.25
    vel = rng.uniform(4.0, 5.0, (n, n, n))
    a = np.array([1.0, 1.5, 2.0])
    b = np.array([5.0, 4.0, 1.0])
    tab = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, a, b[None], 0)[0]
    tba = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, b, a[None], 0)[0]
    assert abs(tab - tba) / tab < 0.02


def test_picks_trilinear_and_upper_corner_quirk():
    n, h = 5, 1.0
    vel = np.full((n, n, n), 2.0)
    src = np.array([0.0, 0.0, 0.0])
    tt, tf = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, src,
                                 np.array([[1.5, 2.25, 0.5], [4.0, 4.0, 4.0]]), 0, field=True)
    x, y, z = 1.5, 2.25, 0.5
    i, j, k = 1, 2, 0
    fx, fy, fz = 0.5, 0.25, 0.5
    exp = 0.0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                exp += ((fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                        * tf[i + dx, j + dy, k + dz])
    assert abs(tt[0] - exp) < 1e-14
    # getTime at the far corner: the 8 clamped corners coincide and each gets
    # weight 1 (ndgridmap.hpp:462-489), so the pick is 8x the node time
    assert abs(tt[1] - 8 * tf[4, 4, 4]) < 1e-12


@pytest.fixture(scope="module")
def std_heap(tmp_path_factory):
    d = tmp_path_factory.mktemp("heap")
    so = str(d / "libheapstd.so")
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-std=c++11",
                    os.path.join(HERE, "heap_libstdcxx.cpp"), "-o", so], check=True)
    return ctypes.CDLL(so)


def test_queue_pop_order_matches_cxx_library_heap(std_heap):
    """Pushes, live key decreases (stale duplicate entries) and pops with many
    equal keys: the oracle's queue must pop cells in the same order as
    std::push_heap/std::pop_heap with the live-key comparator."""
    L = po.lib()
    rng = np.random.default_rng(9)
    for trial in range(40):
        ncell = int(rng.integers(5, 200))
        keys0 = rng.integers(0, 8, ncell).astype(float)    # many ties
        ops, vals = [], []
        for _ in range(int(rng.integers(50, 600))):
            r = rng.random()
            if r < 0.45:
                ops.append((0, int(rng.integers(ncell)))); vals.append(0.0)
            elif r < 0.7:
                c = int(rng.integers(ncell))
                ops.append((1, c)); vals.append(float(rng.integers(0, 8)))
            else:
                ops.append((2, 0)); vals.append(0.0)
        ops = np.ascontiguousarray(ops, np.int32)
        vals = np.ascontiguousarray(vals)
        out = []
        for fn in (L.orc_heap_replay, std_heap.heap_replay_std):
            keys = keys0.copy()
            popped = np.zeros(len(ops), np.int32)
            n = fn(ctypes.c_int(ncell), keys.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(ops)),
                   ops.ctypes.data_as(ctypes.c_void_p), vals.ctypes.data_as(ctypes.c_void_p),
                   popped.ctypes.data_as(ctypes.c_void_p))
            out.append(popped[:n].copy())
        np.testing.assert_array_equal(out[0], out[1])
