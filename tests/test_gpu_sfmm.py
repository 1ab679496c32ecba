"""GPU parity of the exact order-1 3-D solve against the oracle (through the C ABI).

At order 1 (highAccuracySolveEikonal) the reference's times depend on the order
in which its SFMM heap pops cells, so the GPU replays that heap loop: the same
pushes, sift steps and FP64 local solves, one wave per problem.  The oracle
(oracle/fmm3d.c) restates the same loop on the CPU, so the fields are compared
bit for bit; receiver times (trilinear picks of those fields) to 1e-12
relative; log-likelihoods to 1e-6 relative.  The negative control runs the
parallel fixed-point method on the same problems and asserts that it does not
reproduce the heap's answer.
"""
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from mctomo_amd.workloads import body_workload
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _voronoi(shape, h, ncell, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    hx, hy, hz = h
    grid = dict(nx=nx, ny=ny, nz=nz, xmin=0.0, ymin=0.0, zmin=0.0, dx=hx, dy=hy, dz=hz)
    ext = np.array([hx * (nx - 1), hy * (ny - 1), hz * (nz - 1)])
    pts = rng.uniform(0, 1, (ncell, 3)) * ext
    prm = np.c_[rng.uniform(4, 7, ncell), np.ones(ncell), np.ones(ncell)]
    return po.kdtree_to_grid(grid, pts, prm)[1], ext, rng


CASES = [
    ((21, 21, 21), (0.25, 0.25, 0.25), 30, 1),
    ((41, 41, 41), (0.25, 0.25, 0.25), 60, 12),
    ((33, 20, 12), (0.2, 0.35, 0.5), 25, 5),
    ((64, 48, 40), (0.2, 0.2, 0.2), 120, 8),
]


@pytest.mark.parametrize("shape,h,ncell,seed", CASES)
def test_heap_replay_field_bit_exact(shape, h, ncell, seed):
    vel, ext, rng = _voronoi(shape, h, ncell, seed)
    srcs = [rng.uniform(0, 1, 3) * ext,                       # inside a cell
            np.array([h[0] * 3, h[1] * 2, h[2] * 1]),         # on a node
            ext.copy(),                                       # far corner: clamped corners repeat
            np.zeros(3)]                                      # origin corner
    rev = np.c_[rng.uniform(0, 1, (48, 2)) * ext[:2], np.zeros(48)]
    for k, src in enumerate(srcs):
        say("heap replay", shape, "src", k, src)
        tg, fg = mct.fast_marching_3d(vel, h[0], h[1], h[2], 0, 0, 0, src, rev, 1, field=True)
        to, fo = po.fast_marching_3d(vel, h[0], h[1], h[2], 0, 0, 0, src, rev, 1, field=True)
        nbad = int(np.count_nonzero(fg != fo))
        assert nbad == 0, (shape, k, nbad, float(np.max(np.abs(fg - fo))))
        np.testing.assert_allclose(tg, to, rtol=1e-12, atol=0)


def test_heap_replay_offset_grid():
    say("heap replay offset grid")
    rng = np.random.default_rng(3)
    vel = rng.uniform(3, 5, (23, 18, 11))
    src = np.array([11.3, -3.2, 4.4])
    rev = np.array([[10.0, -5.0, 4.0], [13.9, 1.5, 2.0], [12.0, -4.0, 7.9]])
    tg, fg = mct.fast_marching_3d(vel, 0.2, 0.35, 0.5, 10.0, -5.0, 2.0, src, rev, 1, field=True)
    to, fo = po.fast_marching_3d(vel, 0.2, 0.35, 0.5, 10.0, -5.0, 2.0, src, rev, 1, field=True)
    np.testing.assert_array_equal(fg, fo)
    np.testing.assert_allclose(tg, to, rtol=1e-12, atol=0)


def _g(grid):
    return dict(nx=grid.nx, ny=grid.ny, nz=grid.nz, xmin=grid.xmin, ymin=grid.ymin,
                zmin=grid.zmin, dx=grid.dx, dy=grid.dy, dz=grid.dz)


@pytest.mark.parametrize("datatype", [0, 1])
def test_body_likelihood_order1_vs_oracle(datatype):
    nsrc, nrev = 6, 30
    say("body order 1 datatype", datatype)
    w = body_workload(n=33, h=0.3, ncells=40, nsrc=nsrc, nrev=nrev, datatype=datatype, seed=41)
    grid = dict(_g(w.grid), scaling=1.0)
    _, vp, vs, _ = po.kdtree_to_grid(grid, w.points, w.params)
    waves = [vp] if datatype == 0 else [vp, vs]
    pts = [po.body_traveltimes(v, grid, w.locations, w.rev, 1) for v in waves]
    rng = np.random.default_rng(2)
    for i, pt in enumerate(pts):
        w.ttime[i, 0] = pt.reshape(-1) + rng.normal(0, 0.05, pt.size)
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, nsrc, datatype=datatype, order=1,
                              max_cells=64)
    rc, out, ptg, srd, sig = like.likelihood(w.points, w.params, w.locations, arrays=True)
    assert rc == 0
    for i, pt in enumerate(pts):
        np.testing.assert_allclose(ptg[i], pt, rtol=1e-12, atol=0)
    ref = po.body_misfit(np.stack(pts), w.ttime, w.raystat, w.locations)[0]
    assert abs(out[0] - ref[0]) <= 1e-6 * abs(ref[0]), (out, ref)
    # negative control: the parallel fixed point is not the heap's answer
    like.set_order1("parallel")
    rc2, out2, ptp, _, _ = like.likelihood(w.points, w.params, w.locations, arrays=True)
    assert rc2 == 0
    dev = max(float(np.max(np.abs(ptp[i] - pt) / pt)) for i, pt in enumerate(pts))
    say("parallel method deviation", dev)
    assert 1e-9 < dev < 1e-2
    like.set_order1("heap")
    rc3, out3, pth, _, _ = like.likelihood(w.points, w.params, w.locations, arrays=True)
    assert rc3 == 0 and np.array_equal(pth, ptg)
    like.close()


def test_order1_batch_equals_single_calls():
    nsrc, nrev, B = 4, 20, 5
    say("order 1 batch of", B)
    ws = [body_workload(n=25, h=0.4, ncells=20 + 7 * b, nsrc=nsrc, nrev=nrev, seed=300 + b) for b in range(B)]
    w0 = ws[0]
    like = mct.BodyLikelihood(w0.grid, w0.rev, w0.ttime, w0.raystat, nsrc, order=1, max_cells=64,
                              max_chains=B)
    single = []
    for w in ws:
        rc, out = like.likelihood(w.points, w.params, w.locations)
        assert rc == 0
        single.append(out.copy())
    P = np.zeros((B, 64, 3))
    Q = np.zeros((B, 64, 3))
    nc = np.zeros(B, np.int32)
    loc = np.zeros((B, nsrc, 4))
    for c, w in enumerate(ws):
        P[c, :w.points.shape[0]] = w.points
        Q[c, :w.points.shape[0]] = w.params
        nc[c] = w.points.shape[0]
        loc[c] = w.locations
    dP, dQ, dn, dl = (mct.DeviceArray.from_numpy(a) for a in (P, Q, nc, loc))
    dout = mct.DeviceArray((B, 3), np.float64)
    dst = mct.DeviceArray((B,), np.int32)
    like.likelihood_batch_dev(B, dP.ptr, dQ.ptr, dn.ptr, dl.ptr, 0, dout.ptr, dst.ptr)
    like.sync()
    assert (dst.numpy() == 0).all()
    np.testing.assert_array_equal(dout.numpy(), np.array(single))
    like.close()
