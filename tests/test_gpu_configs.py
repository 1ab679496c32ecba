"""The benchmark configurations that fit one GPU, each as a parity test
against the oracle, plus the reference's own example-2 geometry (grid scaling
4 and 5) read from its data files through the library's readers.

  C2: 3-D body waves, 101^3 at 0.2 km, 100 sites, 50 sources x 400 receivers;
  C3: 101x101x41, 20 periods 0.5-10 s, 64x64 pairs, Rayleigh and Love, phase
      and group velocities, 20-400 cells; a 256-chain batch;
  C4: the served mix of 51x51x21, 101x101x41 and 201x201x61 grids with 10-30
      periods and 10-800 cells.
Gates (north star): cell indices bit-exact, traveltimes 1e-4 relative, phase
velocities 1e-6 km/s, log-likelihood 1e-6 relative."""
import os
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from mctomo_amd import workloads as W
from mctomo_amd.api import Grid
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
FMT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "formats")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gdict(g):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin, dx=g.dx, dy=g.dy, dz=g.dz,
                scaling=g.scaling)


def rel(a, b):
    m = np.asarray(b) != 0
    return float((np.abs(np.asarray(a)[m] - np.asarray(b)[m]) / np.abs(np.asarray(b)[m])).max())


@pytest.mark.parametrize("order", [0, 1])
def test_c2_body_101_cubed_vs_oracle(order):
    """Order 0 within the 1e-4 gate; order 1 (the reference examples' setting)
    through the heap replay, equal to the oracle's SFMM (1e-12 covers the
    trilinear picks)."""
    w = W.c2_body(order=order)
    grid = gdict(w.grid)
    say("C2: 101^3, 100 sites, 50 x 400, order", order)
    sid = mct.kdtree_to_grid(w.grid, w.points, w.params)[0]
    sid_o, vp, _, _ = po.kdtree_to_grid(grid, w.points, w.params)
    assert np.array_equal(sid, sid_o)
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, w.nsrc, order=order, max_cells=128)
    rc, out, pt, _, _ = like.likelihood(w.points, w.params, w.locations, arrays=True)
    pt_o = po.body_traveltimes(vp, grid, w.locations, w.rev, order, 16)
    ref, _, _, rco = po.body_misfit(pt_o[None], w.ttime, w.raystat, w.locations)
    assert rc == 0 and rco == 0
    say("  traveltime max rel %.3e, like %.12g (oracle %.12g)" % (rel(pt[0], pt_o), out[0], ref[0]))
    assert rel(pt[0], pt_o) <= (1e-4 if order == 0 else 1e-12)
    assert abs(out[0] - ref[0]) <= 1e-6 * abs(ref[0])


def example2_surface():
    src, rev, fr, tt, rs, nrays = mct.read_data(os.path.join(FMT, "ssources.dat"), os.path.join(FMT, "sreceivers.dat"),
                                                os.path.join(FMT, "stimes.dat"))
    # examples/example2/surf/MCTomo.inp: 59 x 71 x 61 over [-3,2.8] x [-3.5,3.5] x [0,4] km, scaling 4
    g = Grid.from_bounds(59, 71, 61, -3.0, 2.8, -3.5, 3.5, 0.0, 4.0, scaling=4.0)
    return g, src, rev, fr, tt, rs, nrays


@pytest.mark.parametrize("phase_group", [1, 0])
def test_example2_surface_scaling_4_from_the_reference_files(phase_group):
    """The reference's example-2 surface data (7 stations, 13 periods, read with
    mct_read_sources / mct_read_times) on its grid with GRID%SCALING = 4: the
    layer thicknesses are dz / scaling (src/likelihood_surf.F90:615).  Group
    velocities are the example's own setting."""
    g, src, rev, fr, tt, rs, nrays = example2_surface()
    rng = np.random.default_rng(22)
    L = np.array([g.xmin, g.ymin, g.zmin]), np.array([g.xmin + g.dx * (g.nx - 1), g.ymin + g.dy * (g.ny - 1),
                                                      g.zmin + g.dz * (g.nz - 1)])
    for attempt in range(50):
        pts = rng.uniform(L[0], L[1], (40, 3))
        vs = np.sort(rng.uniform(2.5, 4.5, 40))[np.argsort(np.argsort(pts[:, 2]))]
        prm = np.c_[1.73 * vs, vs, 0.32 * 1.73 * vs + 0.77]
        if W.top_is_slowest(g, pts, prm):
            break
    grid = gdict(g)
    say("example2 surface, scaling 4, phase_group", phase_group, "nrays", nrays)
    like = mct.SurfLikelihood(g, fr, src, rev, tt, rs, order=1, phase_group=phase_group, nrays=nrays, max_cells=64)
    rc, out, pt, sd, _, pv = like.likelihood(pts, prm, arrays=True)
    o = po.surf_likelihood(grid, pts, prm, fr, src, rev, tt, rs, order=1, phase_group=phase_group, nrays=nrays)
    assert rc == o[0] == 0
    used = rs[:, 0, :].reshape(pt.shape) == 1
    say("  phase vel max |d| %.2e, time max rel %.2e, like %.12g (oracle %.12g)" %
        (np.abs(pv - o[5]).max(), rel(pt[used], o[2][used]), out[0], o[1][0]))
    assert np.abs(pv - o[5]).max() <= 1e-6
    assert rel(pt[used], o[2][used]) <= 1e-4
    assert abs(out[0] - o[1][0]) <= 1e-6 * abs(o[1][0])


@pytest.mark.parametrize("order", [0, 1])
def test_example2_body_scaling_5_from_the_reference_files(order):
    """The reference's example-2 body data: 291 sources (x, y, z, t0) and 7
    receivers read through the library's readers, P and S, on the example's
    grid with GRID%SCALING = 5 (the eikonal's dz and zmin divided by it,
    src/likelihood_body.F90:209-210); more sources than receivers, so both
    sides trace from the receivers.  Order 1 is the example's own setting
    (examples/example2/body/MCTomo.inp:70)."""
    src, rev2, fr, tt, rs, nrays = mct.read_data(os.path.join(FMT, "bsources.dat"),
                                                 os.path.join(FMT, "breceivers.dat"),
                                                 os.path.join(FMT, "btimes.dat"))
    g = Grid.from_bounds(59, 71, 61, -3.0, 2.8, -3.5, 3.5, 0.0, 3.0, scaling=5.0)
    rev = np.c_[rev2, np.zeros(rev2.shape[0])]
    rng = np.random.default_rng(23)
    lo = np.array([g.xmin, g.ymin, g.zmin])
    hi = lo + np.array([g.dx * (g.nx - 1), g.dy * (g.ny - 1), g.dz * (g.nz - 1)])
    pts = rng.uniform(lo, hi, (60, 3))
    vp = rng.uniform(4.0, 7.0, 60)
    prm = np.c_[vp, vp / 1.73, 0.32 * vp + 0.77]
    bn = np.array([[0.01, 0.02], [0.03, 0.04]])
    grid = gdict(g)
    say("example2 body, scaling 5:", src.shape[0], "sources x", rev.shape[0], "receivers, nrays", nrays,
        "order", order)
    like = mct.BodyLikelihood(g, rev, tt, rs, src.shape[0], datatype=1, order=order, sigdep=1, nrays=nrays,
                              max_cells=64)
    rc, out, pt, _, _ = like.likelihood(pts, prm, src, bnoise0=bn[0], bnoise1=bn[1], arrays=True)
    _, vpg, vsg, _ = po.kdtree_to_grid(grid, pts, prm)
    # the reference traces from the receivers when nsrc > nrev (src/likelihood_body.F90:226-249)
    rl = np.c_[rev, np.zeros(rev.shape[0])]
    pt_o = np.stack([po.body_traveltimes(v, grid, rl, src[:, :3], order, 16).T for v in (vpg, vsg)])
    ref, _, _, rco = po.body_misfit(pt_o, tt, rs, src, sigdep=1, bnoise0=bn[0], bnoise1=bn[1], nrays=nrays)
    assert rc == 0 and rco == 0
    say("  traveltime max rel %.3e, like %.12g (oracle %.12g)" % (rel(pt, pt_o), out[0], ref[0]))
    assert rel(pt, pt_o) <= (1e-4 if order == 0 else 1e-12)
    assert abs(out[0] - ref[0]) <= 1e-6 * abs(ref[0])


def c3_workload(raylov, seed=20261204):
    return W.surface_workload(101, 41, 50.0, 20.0, 200, 20, (0.5, 10.0), 64, 64, raylov=raylov, seed=seed, order=0)


def c3_chains(w, nchain, seed):
    from mctomo_amd.workloads import _surface_sites
    rng = np.random.default_rng(seed)
    g = w.grid
    L, Z = g.dx * (g.nx - 1), g.dz * (g.nz - 1)
    out = []
    while len(out) < nchain:
        k = int(rng.integers(20, 401))
        p, q = _surface_sites(rng, k, L, Z, (2.5, 4.5))
        if W.top_is_slowest(g, p, q):
            out.append((p, q))
    return out


@pytest.mark.parametrize("raylov,phase_group", [(1, 0), (0, 0), (1, 1), (0, 1)])
def test_c3_wave_types_vs_oracle(raylov, phase_group):
    """C3's grid and data for each (wave type, velocity type): a chain with
    20-400 cells against the oracle."""
    w = c3_workload(raylov)
    (p, q), = c3_chains(w, 1, 50 + 2 * raylov + phase_group)
    grid = gdict(w.grid)
    say("C3", "Rayleigh" if raylov else "Love", "group" if phase_group else "phase", "cells", len(p))
    like = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=raylov, order=0,
                              phase_group=phase_group, max_cells=400)
    rc, out, pt, _, _, pv = like.likelihood(p, q, arrays=True)
    o = po.surf_likelihood(grid, p, q, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=raylov, order=0,
                           phase_group=phase_group, nthreads=16)
    assert rc == o[0]
    if rc == 0:
        say("  phase vel max |d| %.2e, time max rel %.2e, like rel %.2e" %
            (np.abs(pv - o[5]).max(), rel(pt, o[2]), abs(out[0] - o[1][0]) / abs(o[1][0])))
        assert np.abs(pv - o[5]).max() <= 1e-6
        assert rel(pt, o[2]) <= 1e-4
        assert abs(out[0] - o[1][0]) <= 1e-6 * abs(o[1][0])


def test_c3_batch_of_256_chains_equals_single_calls():
    """256 chains with 20-400 cells in one batched call (C3's batch); sampled
    chains equal their one-chain calls bit for bit."""
    w = c3_workload(1)
    B, NC = 256, 400
    rng = np.random.default_rng(77)
    from mctomo_amd.workloads import _surface_sites
    g = w.grid
    L, Z = g.dx * (g.nx - 1), g.dz * (g.nz - 1)
    P = np.zeros((B, NC, 3))
    Q = np.zeros((B, NC, 3))
    N = np.zeros(B, np.int32)
    for c in range(B):
        k = int(rng.integers(20, NC + 1))
        p, q = _surface_sites(rng, k, L, Z, (2.5, 4.5))
        P[c, :k], Q[c, :k], N[c] = p, q, k
    like = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=0, max_cells=NC,
                              max_chains=B)
    dP, dQ, dN = (mct.DeviceArray.from_numpy(a) for a in (P, Q, N))
    dS0 = mct.DeviceArray.from_numpy(np.zeros((B, 2, w.freqs.shape[0])))
    dO = mct.DeviceArray((B, 3), np.float64)
    dS = mct.DeviceArray((B,), np.int32)
    say("C3 batch of", B)
    like.likelihood_batch_dev(B, dP.ptr, dQ.ptr, dN.ptr, dS0.ptr, dO.ptr, dS.ptr)
    like.sync()
    out, st = dO.numpy(), dS.numpy()
    say("  statuses", np.bincount(st.clip(0)))
    single = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=0, max_cells=NC)
    for c in (0, 1, 77, 128, 200, 255):
        rc, o1 = single.likelihood(P[c, :N[c]], Q[c, :N[c]])
        assert rc == st[c], c
        if rc == 0:
            np.testing.assert_array_equal(out[c], o1, err_msg="chain %d" % c)


def test_c4_grids_alone_equal_the_served_mix():
    """C4's three grids in one server (51x51x21 with 10 periods, 101x101x41 with
    20, 201x201x61 with 30), cells 10-800: each request's result equals its
    plain call bit for bit, and one request per grid matches the oracle."""
    ws = [W.surface_workload(51, 21, 25.0, 10.0, 50, 10, (1.0, 5.0), 16, 16, seed=20261210, order=0),
          W.surface_workload(101, 41, 50.0, 20.0, 100, 20, (1.0, 10.0), 32, 32, seed=20261211, order=0),
          W.surface_workload(201, 61, 100.0, 30.0, 200, 30, (1.0, 10.0), 16, 16, seed=20261212, order=0)]
    from mctomo_amd.workloads import _surface_sites
    rng = np.random.default_rng(91)
    reqs = []
    for m in (0, 1, 2, 0, 1, 2, 2, 1, 0):
        w = ws[m]
        L, Z = w.grid.dx * (w.grid.nx - 1), w.grid.dz * (w.grid.nz - 1)
        k = int(rng.integers(10, 801))
        reqs.append((m,) + _surface_sites(rng, k, L, Z, (2.5, 4.5)))
    specs = [dict(grid=w.grid, freqs=w.freqs, src=w.src, rev=w.rev, ttime=w.ttime, raystat=w.raystat, order=0,
                  max_cells=800, max_batch=4) for w in ws]
    srv = mct.SurfServer.mix(specs, max_lanes=1)
    tickets = [srv.submit(p, q, model=m) for m, p, q in reqs]
    res = [srv.wait(t) for t in tickets]
    b, e, rc = srv.stats()
    srv.close()
    assert rc == 0 and e == len(reqs)
    plain = [mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=0, max_cells=800)
             for w in ws]
    checked = set()
    for (m, p, q), (st, out, lat) in zip(reqs, res):
        st_a, out_a = plain[m].likelihood(p, q)
        assert st == st_a
        if st != 0:
            continue
        np.testing.assert_array_equal(out, out_a)
        if m not in checked:
            w = ws[m]
            o = po.surf_likelihood(gdict(w.grid), p, q, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=0,
                                   nthreads=16)
            say("C4 grid %d: like %.12g oracle %.12g (latency %.1f ms)" % (w.grid.nx, out[0], o[1][0], 1e3 * lat))
            assert o[0] == 0 and abs(out[0] - o[1][0]) <= 1e-6 * abs(o[1][0])
            checked.add(m)
