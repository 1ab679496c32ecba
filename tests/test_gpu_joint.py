"""GPU joint body + surface-wave likelihood (datatype 3, src/likelihood.f90
:83-88) against the oracle: body_likelihood with P and S on dat(1) plus
surf_likelihood on dat(2) whose layers take the sites' own vp and rho, summed.
Gates: log-likelihood within 1e-6 relative for each half and the sum; a batch
of chains equals the one-chain calls bit for bit."""
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from mctomo_amd.workloads import joint_workload
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gdict(g):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin, dx=g.dx, dy=g.dy, dz=g.dz,
                scaling=g.scaling)


def oracle_joint(w, points, params, order, datatype=3):
    grid = gdict(w.grid)
    _, vp, vs, _ = po.kdtree_to_grid(grid, points, params)
    pt = np.stack([po.body_traveltimes(v, grid, w.locations, w.body["rev"], order) for v in (vp, vs)])
    ob, _, _, rcb = po.body_misfit(pt, w.body["ttime"], w.body["raystat"], w.locations)
    s = w.surf
    rcs, os_, *_ = po.surf_likelihood(grid, points, params, s["freqs"], s["src"], s["rev"], s["ttime"],
                                      s["raystat"], order=order, datatype=datatype)
    return rcb, ob, rcs, os_


@pytest.mark.parametrize("order", [0, 1])
def test_joint_likelihood_vs_oracle(order):
    w = joint_workload(n=41, nz=21, ncells=40, nsrc=6, nrev=48, np_=6, nsta=8, order=order, seed=31 + order)
    say("joint order", order, "rejected draws", w.rejected_draws)
    like = mct.JointLikelihood(w.grid, w.body, w.surf, max_cells=64)
    rc, out, ob, os_ = like.likelihood(w.points, w.params, w.locations)
    rcb, ob_o, rcs, os_o = oracle_joint(w, w.points, w.params, order)
    assert rc == 0 and rcb == 0 and rcs == 0
    say("  body like %.12g (oracle %.12g), surface like %.12g (oracle %.12g)" % (ob[0], ob_o[0], os_[0], os_o[0]))
    # the body half at order 1 replays the reference's SFMM heap loop, so both
    # orders are held to the gate
    assert np.all(np.abs(ob - ob_o) <= 1e-6 * np.abs(ob_o)), (ob, ob_o)
    assert np.all(np.abs(os_ - os_o) <= 1e-6 * np.abs(os_o)), (os_, os_o)
    assert np.all(np.abs(out - (ob + os_)) <= 1e-12 * np.abs(out))
    # the surface half depends on the sites' vp and rho: the datatype-2 layers
    # (vs2vp / vp2rho) give a different likelihood, which the gate would catch
    _, _, rc2, os2 = oracle_joint(w, w.points, w.params, order, datatype=2)
    assert rc2 == 0 and abs(os2[0] - os_o[0]) > 1e-3 * abs(os_o[0])


def test_joint_batch_equals_single_chain_calls():
    w = joint_workload(n=33, nz=17, ncells=30, nsrc=5, nrev=32, np_=5, nsta=6, seed=41)
    B, NC = 4, 48
    like = mct.JointLikelihood(w.grid, w.body, w.surf, max_cells=NC, max_chains=B)
    rng = np.random.default_rng(7)
    P = np.zeros((B, NC, 3))
    Q = np.zeros((B, NC, 3))
    N = np.zeros(B, np.int32)
    L = np.zeros((B, w.locations.shape[0], 4))
    bn = np.zeros((B, 2, 2))
    sn = np.zeros((B, 2, like.np))
    singles = []
    for c in range(B):
        # the workload's sites shifted a little (same depth ranking, so
        # check_model still holds), locations moved, noise per chain
        p = w.points + rng.uniform(-0.3, 0.3, w.points.shape) * [1, 1, 0]
        q = w.params * rng.uniform(0.98, 1.02, (1, 3))
        loc = w.locations.copy()
        loc[:, 0:3] += rng.uniform(-0.2, 0.2, (loc.shape[0], 3))
        b0, b1 = np.array([0.01, 0.02]) * (c + 1), np.array([0.05, 0.06])
        s0, s1 = np.full(like.np, 0.005 * c), np.full(like.np, 0.05)
        P[c, :len(p)], Q[c, :len(p)], N[c], L[c] = p, q, len(p), loc
        bn[c, 0], bn[c, 1], sn[c, 0], sn[c, 1] = b0, b1, s0, s1
        singles.append(like.likelihood(p, q, loc, b0, b1, s0, s1))
    dP, dQ, dN, dL, dB, dS = (mct.DeviceArray.from_numpy(a) for a in (P, Q, N, L, bn, sn))
    dO = mct.DeviceArray((B, 3), np.float64)
    dT = mct.DeviceArray((B,), np.int32)
    like.likelihood_batch_dev(B, dP.ptr, dQ.ptr, dN.ptr, dL.ptr, dB.ptr, dS.ptr, dO.ptr, dT.ptr)
    like.sync()
    out, st = dO.numpy(), dT.numpy()
    for c in range(B):
        say("chain", c, "status", st[c], "like", out[c, 0], "single", singles[c][1][0])
        assert st[c] == singles[c][0]
        np.testing.assert_array_equal(out[c], singles[c][1])
