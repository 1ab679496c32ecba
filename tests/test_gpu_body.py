"""GPU parity of the body-wave hot path against the CPU oracle (through the C ABI).

Tolerances (stated by the parity contract): nearest-site ids bit-exact;
first-order (order 0) traveltimes within 1e-4 relative; log-likelihood within
1e-6 relative.  Mixed order (order 1) is solved by replaying the reference's
heap loop (its answer depends on the pop order), so its fields equal the
oracle's bit for bit (tests/test_gpu_sfmm.py covers that replay in depth).
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


def _g(grid):
    return dict(nx=grid.nx, ny=grid.ny, nz=grid.nz, xmin=grid.xmin, ymin=grid.ymin,
                zmin=grid.zmin, dx=grid.dx, dy=grid.dy, dz=grid.dz)


@pytest.mark.parametrize("seed,n,ncell", [(1, 21, 7), (2, 41, 100), (3, 33, 800), (4, 50, 1), (7, 23, 2500)])
def test_nearest_random_bit_exact(seed, n, ncell):
    say("nearest random", seed, n, ncell)
    rng = np.random.default_rng(seed)
    grid = mct.Grid(n, n + 3, n - 4, 0.0, -1.0, 0.0, 0.25, 0.25, 0.3, 1.0)
    pts = rng.uniform(-0.5, 0.25 * n + 0.5, (ncell, 3))
    prm = rng.uniform(1, 7, (ncell, 3))
    got = mct.kdtree_to_grid(grid, pts, prm)
    ref = po.kdtree_to_grid(_g(grid), pts, prm)
    for a, b in zip(got, ref):
        np.testing.assert_array_equal(a, b)


def test_nearest_lattice_ties_bit_exact():
    say("nearest lattice ties")
    rng = np.random.default_rng(5)
    grid = mct.Grid(21, 21, 21, 0.0, 0.0, 0.0, 0.5, 0.5, 0.5, 1.0)
    for t in range(4):
        g = np.unique(rng.integers(0, 6, (int(rng.integers(20, 150)), 3)).astype(float) * 2.0, axis=0)
        rng.shuffle(g)
        prm = rng.uniform(1, 7, (g.shape[0], 3))
        got = mct.kdtree_to_grid(grid, g, prm)
        ref = po.kdtree_to_grid(_g(grid), g, prm)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


def test_nearest_box_and_pm():
    say("nearest box pm")
    rng = np.random.default_rng(6)
    grid = mct.Grid(17, 15, 13, 0.0, 0.0, 0.0, 0.5, 0.5, 0.5, 1.0)
    pts = rng.uniform(0, 8, (40, 3))
    prm = rng.uniform(1, 7, (40, 3))
    sid, vp, vs, rho = po.kdtree_to_grid(_g(grid), pts, prm)
    box = np.array([1.2, 0.7, 2.1, 5.3, 6.6, 4.0])
    a = mct.kdtree_to_grid(grid, pts * 1.01, prm, box=box, vp=vp, vs=vs, rho=rho)
    b = po.kdtree_to_grid(_g(grid), pts * 1.01, prm, box=box, vp=vp, vs=vs, rho=rho)
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)
    pm = prm[sid[8, 7, 6] - 1]
    a = mct.kdtree_to_grid(grid, pts, prm * 1.2, pm=pm, vp=vp, vs=vs, rho=rho)
    b = po.kdtree_to_grid(_g(grid), pts, prm * 1.2, pm=pm, vp=vp, vs=vs, rho=rho)
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(x, y)


def _voronoi_vel(n, h, ncell, seed):
    rng = np.random.default_rng(seed)
    grid = dict(nx=n, ny=n, nz=n, xmin=0.0, ymin=0.0, zmin=0.0, dx=h, dy=h, dz=h)
    pts = rng.uniform(0, h * (n - 1), (ncell, 3))
    prm = np.c_[rng.uniform(4, 7, ncell), np.ones(ncell), np.ones(ncell)]
    return po.kdtree_to_grid(grid, pts, prm)[1], rng


@pytest.mark.parametrize("order,rtol", [(0, 1e-4), (1, 1e-4)])
def test_fast_marching_3d_vs_oracle(order, rtol):
    n, h = 41, 0.25
    vel, rng = _voronoi_vel(n, h, 60, 11 + order)
    L = h * (n - 1)
    worst = 0.0
    for s in range(4):
        src = np.array([rng.uniform(0, L), rng.uniform(0, L), rng.uniform(1, L - 1)])
        rev = np.c_[rng.uniform(0, L, (64, 2)), np.zeros(64)]
        say("fm3d order", order, "src", s, src)
        tg, fg = mct.fast_marching_3d(vel, h, h, h, 0, 0, 0, src, rev, order, field=True)
        to, fo = po.fast_marching_3d(vel, h, h, h, 0, 0, 0, src, rev, order, field=True)
        rel = np.abs(tg - to) / to
        worst = max(worst, rel.max())
        assert rel.max() < rtol, (order, s, rel.max())
        frel = np.abs(fg - fo) / np.maximum(fo, 1e-9)
        assert np.quantile(frel, 0.999) < rtol
        if order == 1:
            assert np.array_equal(fg, fo)
    say("fm3d order", order, "worst receiver rel err", worst)


def test_fast_marching_3d_offset_grid_and_anisotropic_spacing():
    say("fm3d offset grid")
    rng = np.random.default_rng(3)
    vel = rng.uniform(3, 5, (23, 18, 11))
    src = np.array([11.3, -3.2, 4.4])
    rev = np.array([[10.0, -5.0, 4.0], [13.9, 1.5, 2.0], [12.0, -4.0, 7.9]])
    tg = mct.fast_marching_3d(vel, 0.2, 0.35, 0.5, 10.0, -5.0, 2.0, src, rev, 0)
    to = po.fast_marching_3d(vel, 0.2, 0.35, 0.5, 10.0, -5.0, 2.0, src, rev, 0)
    np.testing.assert_allclose(tg, to, rtol=1e-4)


def _oracle_like(w, order, nthreads=0):
    pt = po.body_traveltimes(po.kdtree_to_grid(_g(w.grid), w.points, w.params)[1], _g(w.grid),
                             w.locations, w.rev, order, nthreads)
    return pt


@pytest.mark.parametrize("nsrc,nrev", [(6, 30), (30, 5)])
def test_body_likelihood_vs_oracle(nsrc, nrev):
    say("body likelihood", nsrc, nrev)
    w = body_workload(n=33, h=0.3, ncells=40, nsrc=nsrc, nrev=nrev, seed=99)
    grid = dict(_g(w.grid), scaling=1.0)
    vp = po.kdtree_to_grid(grid, w.points, w.params)[1]
    if nsrc <= nrev:
        pt = po.body_traveltimes(vp, grid, w.locations, w.rev, 0)          # (nsrc, nrev)
    else:   # the reference traces from receivers when nsrc > nrev
        rl = np.c_[w.rev, np.zeros(nrev)]
        pt = po.body_traveltimes(vp, grid, rl, w.locations[:, :3], 0).T
    # observed = exact + noise, so the misfit term is realistic
    rng = np.random.default_rng(1)
    w.ttime[0, 0] = pt.reshape(-1) + rng.normal(0, 0.05, pt.size)
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, nsrc, order=0, max_cells=64)
    rc, out, ptg, srd, sig = like.likelihood(w.points, w.params, w.locations, arrays=True)
    assert rc == 0
    ref, srd_o, sig_o, rco = po.body_misfit(pt[None], w.ttime, w.raystat, w.locations)
    np.testing.assert_allclose(ptg[0], pt, rtol=1e-4)
    assert abs(out[0] - ref[0]) <= 1e-6 * abs(ref[0]), (out, ref)
    np.testing.assert_allclose(out[1:], ref[1:], rtol=1e-4)
    np.testing.assert_allclose(sig, sig_o)


def test_body_likelihood_p_and_s_sigdep():
    say("body likelihood P&S sigdep")
    w = body_workload(n=29, h=0.3, ncells=30, nsrc=5, nrev=12, datatype=1, seed=5)
    grid = dict(_g(w.grid), scaling=1.0)
    _, vp, vs, _ = po.kdtree_to_grid(grid, w.points, w.params)
    pt = np.stack([po.body_traveltimes(vp, grid, w.locations, w.rev, 0),
                   po.body_traveltimes(vs, grid, w.locations, w.rev, 0)])
    b0, b1 = np.array([0.02, 0.03]), np.array([0.01, 0.02])
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, 5, datatype=1, sigdep=1,
                              max_cells=32)
    rc, out, ptg, srd, sig = like.likelihood(w.points, w.params, w.locations, b0, b1, arrays=True)
    ref, _, sig_o, _ = po.body_misfit(pt, w.ttime, w.raystat, w.locations, sigdep=1,
                                      bnoise0=b0, bnoise1=b1)
    np.testing.assert_allclose(ptg, pt, rtol=1e-4)
    assert abs(out[0] - ref[0]) <= 1e-6 * abs(ref[0]), (out, ref)


def test_body_zero_sigma_is_fatal():
    say("body zero sigma")
    w = body_workload(n=17, h=0.5, ncells=10, nsrc=2, nrev=3, seed=4, sigma=0.0)
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, 2, max_cells=16)
    with pytest.raises(mct.MCTError):
        like.likelihood(w.points, w.params, w.locations)


def test_batched_equals_single():
    say("batched == single")
    nchain = 5
    ws = [body_workload(n=25, h=0.4, ncells=int(5 + 9 * c), nsrc=4, nrev=10, seed=100 + c)
          for c in range(nchain)]
    base = ws[0]
    like = mct.BodyLikelihood(base.grid, base.rev, base.ttime, base.raystat, 4, max_cells=64,
                              max_chains=nchain)
    single = []
    for w in ws:
        rc, out = like.likelihood(w.points, w.params, w.locations)
        single.append(out)
    P = np.zeros((nchain, 64, 3))
    Q = np.zeros((nchain, 64, 3))
    nc = np.zeros(nchain, np.int32)
    loc = np.zeros((nchain, 4, 4))
    for c, w in enumerate(ws):
        P[c, :w.points.shape[0]] = w.points
        Q[c, :w.points.shape[0]] = w.params
        nc[c] = w.points.shape[0]
        loc[c] = w.locations
    dP, dQ, dn, dl = (mct.DeviceArray.from_numpy(a) for a in (P, Q, nc, loc))
    dout = mct.DeviceArray((nchain, 3), np.float64)
    dst = mct.DeviceArray((nchain,), np.int32)
    like.likelihood_batch_dev(nchain, dP.ptr, dQ.ptr, dn.ptr, dl.ptr, 0, dout.ptr, dst.ptr)
    like.sync()
    got = dout.numpy()
    assert (dst.numpy() == 0).all()
    np.testing.assert_allclose(got, np.array(single), rtol=1e-6)


def test_noise_likelihood_reuses_the_last_traveltimes():
    """noise_likelihood (src/likelihood_body.F90:38-78): new noise parameters and
    origin times over the traveltimes of the last call give what a full call with
    them gives, and the oracle's misfit of the same picks."""
    say("body noise likelihood")
    w = body_workload(n=29, h=0.3, ncells=30, nsrc=5, nrev=12, datatype=1, seed=5)
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, 5, datatype=1, sigdep=1,
                              max_cells=32)
    b0, b1 = np.array([0.02, 0.03]), np.array([0.01, 0.02])
    rc, out1, ptg, _, _ = like.likelihood(w.points, w.params, w.locations, b0, b1, arrays=True)
    assert rc == 0
    c0, c1 = np.array([0.05, 0.01]), np.array([0.03, 0.04])
    loc2 = w.locations.copy()
    loc2[:, 3] += 0.05
    rc, out2 = like.noise_likelihood(c0, c1, loc2)
    assert rc == 0
    rc, full = like.likelihood(w.points, w.params, loc2, c0, c1)
    np.testing.assert_array_equal(out2, full)
    ref, _, _, _ = po.body_misfit(ptg, w.ttime, w.raystat, loc2, sigdep=1, bnoise0=c0, bnoise1=c1)
    assert abs(out2[0] - ref[0]) <= 1e-12 * abs(ref[0]), (out2, ref)
    np.testing.assert_allclose(out2[1:], ref[1:], rtol=1e-12)


@pytest.mark.parametrize("datatype", [0, 1])
def test_location_likelihood_dynamic_fields(datatype):
    """location_likelihood with dynamic fields (likelihood_body.F90:82-149):
    the receivers' fields of the last call interpolated (Interpolate3d) at new
    source locations, sigma = b0*t + b1, the new origin times; against the
    oracle's fields from each receiver."""
    say("body location likelihood, dynamic fields, datatype", datatype)
    nsrc, nrev = 6, 5
    w = body_workload(n=29, h=0.3, ncells=40, nsrc=nsrc, nrev=nrev, datatype=datatype, seed=17)
    grid = dict(_g(w.grid), scaling=1.0)
    npp = 1 if datatype == 0 else 2
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, nsrc, datatype=datatype, order=0,
                              max_cells=64, dynamic=1)
    rc, out = like.likelihood(w.points, w.params, w.locations)
    assert rc == 0
    rng = np.random.default_rng(3)
    loc2 = w.locations.copy()
    L = (w.grid.nx - 1) * w.grid.dx
    loc2[:, :3] = np.clip(loc2[:, :3] + rng.uniform(-0.6, 0.6, (nsrc, 3)), 0.2, L - 0.2)
    loc2[:, 3] = rng.uniform(-0.3, 0.3, nsrc)
    b0, b1 = np.full(npp, 0.02), np.full(npp, 0.05)
    rc2, out2 = like.location_likelihood(loc2, b0, b1)
    assert rc2 == 0
    _, vp, vs, _ = po.kdtree_to_grid(grid, w.points, w.params)
    pt = np.zeros((npp, nsrc, nrev))
    for k, vel in enumerate([vp, vs][:npp]):
        for j in range(nrev):
            _, f = po.fast_marching_3d(vel, grid["dx"], grid["dy"], grid["dz"], grid["xmin"], grid["ymin"],
                                       grid["zmin"], w.rev[j], loc2[:, :3], 0, field=True)
            for i in range(nsrc):
                pt[k, i, j] = po.interpolate3d(grid, f, loc2[i, :3])
    ref, _, _, rco = po.body_misfit(pt, w.ttime, w.raystat, loc2, sigdep=1, bnoise0=b0, bnoise1=b1)
    say("  like", out2[0], "oracle", ref[0])
    assert abs(out2[0] - ref[0]) <= 1e-6 * abs(ref[0]), (out2, ref)
    np.testing.assert_allclose(out2[1:], ref[1:], rtol=1e-4)
    # a context without dynamic fields refuses the call (the reference then runs
    # the full likelihood, which is mct_body_likelihood)
    plain = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, nsrc, datatype=datatype, order=0, max_cells=64)
    with pytest.raises(mct.MCTError):
        plain.location_likelihood(loc2, b0, b1)


def test_location_likelihood_batch_equals_single_chain_calls():
    """The batched location likelihood (dynamic fields kept per chain) matches
    each chain's own location likelihood after its own likelihood call."""
    say("body location likelihood batch")
    nchain, nsrc, nrev = 3, 4, 6
    ws = [body_workload(n=21, h=0.4, ncells=int(8 + 5 * c), nsrc=nsrc, nrev=nrev, seed=200 + c)
          for c in range(nchain)]
    base = ws[0]
    like = mct.BodyLikelihood(base.grid, base.rev, base.ttime, base.raystat, nsrc, max_cells=32,
                              max_chains=nchain, dynamic=1)
    rng = np.random.default_rng(9)
    L = (base.grid.nx - 1) * base.grid.dx
    new_loc = []
    for w in ws:
        l2 = w.locations.copy()
        l2[:, :3] = np.clip(l2[:, :3] + rng.uniform(-0.5, 0.5, (nsrc, 3)), 0.2, L - 0.2)
        l2[:, 3] = rng.uniform(-0.2, 0.2, nsrc)
        new_loc.append(l2)
    b0, b1 = np.array([0.03]), np.array([0.04])
    single = []
    for w, l2 in zip(ws, new_loc):
        rc, _ = like.likelihood(w.points, w.params, w.locations)
        assert rc == 0
        rc2, out = like.location_likelihood(l2, b0, b1)
        assert rc2 == 0
        single.append(out)
    P = np.zeros((nchain, 32, 3))
    Q = np.zeros((nchain, 32, 3))
    nc = np.zeros(nchain, np.int32)
    loc = np.zeros((nchain, nsrc, 4))
    for c, w in enumerate(ws):
        P[c, :w.points.shape[0]] = w.points
        Q[c, :w.points.shape[0]] = w.params
        nc[c] = w.points.shape[0]
        loc[c] = w.locations
    noise = np.zeros((nchain, 2, 1))
    noise[:, 0, 0], noise[:, 1, 0] = b0[0], b1[0]
    dP, dQ, dn, dl = (mct.DeviceArray.from_numpy(a) for a in (P, Q, nc, loc))
    dl2 = mct.DeviceArray.from_numpy(np.array(new_loc))
    dnz = mct.DeviceArray.from_numpy(noise)
    dout = mct.DeviceArray((nchain, 3), np.float64)
    dst = mct.DeviceArray((nchain,), np.int32)
    like.likelihood_batch_dev(nchain, dP.ptr, dQ.ptr, dn.ptr, dl.ptr, dnz.ptr, dout.ptr, dst.ptr)
    like.sync()
    assert (dst.numpy() == 0).all()
    like.location_likelihood_batch_dev(nchain, dl2.ptr, dnz.ptr, dout.ptr, dst.ptr)
    like.sync()
    assert (dst.numpy() == 0).all()
    # the batch's eikonal rounds run until every chain's field has converged, so
    # a field may take more rounds than alone: equal within the solver's tolerance
    np.testing.assert_allclose(dout.numpy(), np.array(single), rtol=1e-6)
