"""GPU surface-wave likelihood with group velocities (LIKE_SET%phaseGroup = 1)
against the oracle: rays from every receiver of a used pair (rpaths), group
delays along them (CalGroupTime) as the times, ray lengths as srdist.

Gates: phase and group velocities within 1e-6 km/s, group times and ray
lengths within 1e-4 relative, log-likelihood within 1e-6 relative.  Given the
GPU's own velocities, the rays, lengths and group times are the oracle's bit
for bit with first-order times and with the sequential 2-D method.
"""
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from mctomo_amd import workloads as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gdict(g):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin, dx=g.dx,
                dy=g.dy, dz=g.dz, scaling=g.scaling)


def ctx_for(w, **kw):
    return mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=w.raylov,
                              order=w.order, phase_group=1, **kw)


def oracle_group(w, points=None, params=None, sigdep=0, s0=None, s1=None):
    g = w.grid
    gv = np.zeros((g.nx, g.ny, w.freqs.shape[0]))
    res = po.surf_likelihood(gdict(g), w.points if points is None else points,
                             w.params if params is None else params, w.freqs, w.src, w.rev, w.ttime,
                             w.raystat, raylov=w.raylov, order=w.order, sigdep=sigdep, snoise0=s0,
                             snoise1=s1, phase_group=1, group_vel=gv)
    return res, gv


def rays_given_velocities(w, pv, gv, p, fom):
    g = w.grid
    velv = np.pad(pv[:, :, p], 1, mode="edge")
    return po.fm2d_rays(velv, g.xmin, g.ymin, g.dx, g.dy, w.src[:, 0], w.src[:, 1], w.rev[:, 0],
                        w.rev[:, 1], srs=w.raystat[p, 0], fom=fom, gvel=gv[:, :, p], maxpts=1)


@pytest.mark.parametrize("order,method", [(0, 0), (1, 0), (1, 1), (0, 1)])
def test_group_call_matches_oracle(order, method):
    w = W.surface_workload(n=41, nz=21, ncells=30, np_=10, nsrc=16, nrev=16, seed=20261202, order=order)
    say("group plain", "order", order, "method", method)
    L = ctx_for(w)
    L.set_method(method)
    rc, out, pt, sd, sg, pv = L.likelihood(w.points, w.params, arrays=True)
    gvg = L.group_velocities()
    (orc, oout, opt, osd, osg, opv, _), ogv = oracle_group(w)
    assert rc == orc == 0, (rc, orc)
    used = w.raystat[:, 0].reshape(pt.shape) == 1
    dv = np.max(np.abs(pv - opv))
    dg = np.max(np.abs(gvg - ogv))
    dt = np.max(np.abs(pt[used] - opt[used]) / opt[used])
    dlen = np.max(np.abs(sd - osd)[used.reshape(sd.shape)] / osd[used.reshape(sd.shape)])
    dl = abs(out[0] - oout[0]) / abs(oout[0])
    say("  phase vel %.3g km/s, group vel %.3g km/s, group time rel %.3g, ray length rel %.3g, like rel %.3g"
        % (dv, dg, dt, dlen, dl))
    assert dv <= 1e-6 and dg <= 1e-6
    assert dt <= 1e-4 and dlen <= 1e-4
    assert dl <= 1e-6
    # group times are not phase times: the ray integral of the group slowness
    (prc, _, ppt, *_rest) = po.surf_likelihood(gdict(w.grid), w.points, w.params, w.freqs, w.src, w.rev,
                                               w.ttime, w.raystat, raylov=w.raylov, order=order)
    assert np.max(np.abs(ppt[used] - pt[used]) / ppt[used]) > 1e-3
    # given the GPU's velocities, the rays are the oracle's
    exact = order == 0 or method == 1
    for p in range(w.freqs.shape[0]):
        R = rays_given_velocities(w, pv, gvg, p, order)
        u = used[p]
        if exact:
            np.testing.assert_array_equal(pt[p][u], R["gtime"][u], err_msg="period %d" % p)
            np.testing.assert_array_equal(sd[p].reshape(pt[p].shape)[u], R["rlen"][u])
        else:
            assert np.max(np.abs(pt[p][u] - R["gtime"][u]) / R["gtime"][u]) <= 1e-4
    L.close()


def test_group_batch_equals_plain_calls():
    base = W.surface_workload(n=33, nz=17, L=16.0, Z=8.0, ncells=40, np_=6, nsrc=8, nrev=8, seed=21)
    nchain, maxc = 4, 40
    rng = np.random.default_rng(4)
    pts = np.zeros((nchain, maxc, 3))
    prm = np.zeros((nchain, maxc, 3))
    ncl = np.zeros(nchain, np.int32)
    for c in range(nchain):
        k = int(rng.integers(10, maxc + 1))
        w = W.surface_workload(n=33, nz=17, L=16.0, Z=8.0, ncells=k, np_=6, nsrc=8, nrev=8, seed=100 + c)
        pts[c, :k], prm[c, :k], ncl[c] = w.points, w.params, k
    L = ctx_for(base, max_cells=maxc, max_chains=nchain)
    say("group batch", nchain)
    dP, dQ = mct.DeviceArray.from_numpy(pts), mct.DeviceArray.from_numpy(prm)
    dN = mct.DeviceArray.from_numpy(ncl)
    dS = mct.DeviceArray.from_numpy(np.zeros((nchain, 2, 6)))
    dO = mct.DeviceArray((nchain, 3), np.float64)
    dT = mct.DeviceArray((nchain,), np.int32)
    L.likelihood_batch_dev(nchain, dP.ptr, dQ.ptr, dN.ptr, dS.ptr, dO.ptr, dT.ptr)
    L.sync()
    out, st = dO.numpy(), dT.numpy()
    for c in range(nchain):
        k = ncl[c]
        rc, o1 = L.likelihood(pts[c, :k], prm[c, :k])
        say("chain", c, "status", int(st[c]), rc, "like", out[c][0], o1[0])
        assert st[c] == rc, (c, st[c], rc)
        if rc == 0:
            np.testing.assert_array_equal(out[c], o1)
            (orc, oout, *_), _ = oracle_group(base, points=pts[c, :k], params=prm[c, :k])
            assert orc == 0 and abs(out[c, 0] - oout[0]) <= 1e-6 * abs(oout[0])
    L.close()


def test_group_noise_likelihood_uses_ray_lengths():
    """noise_likelihood in group mode: sigma = snoise0 * ray length + snoise1
    (srdist holds the lengths), equal to a full call with the new noise."""
    w = W.surface_workload(n=41, nz=21, ncells=30, np_=10, nsrc=16, nrev=16, seed=20261202, order=1)
    npp = w.freqs.shape[0]
    L = ctx_for(w, sigdep=1)
    s0, s1 = np.full(npp, 0.01), np.full(npp, 0.05)
    rc, full = L.likelihood(w.points, w.params, snoise0=s0, snoise1=s1)
    assert rc == 0
    t0, t1 = np.full(npp, 0.02), np.full(npp, 0.03)
    rc2, full2 = L.likelihood(w.points, w.params, snoise0=t0, snoise1=t1)
    rc3, cheap = L.noise_likelihood(t0, t1)
    assert rc3 == 0 and np.array_equal(cheap[:2], full2[:2])
    (orc, oout, *_), _ = oracle_group(w, sigdep=1, s0=t0, s1=t1)
    assert orc == 0
    say("group noise like", cheap[0], "oracle", oout[0])
    assert abs(cheap[0] - oout[0]) <= 1e-6 * abs(oout[0])
    L.close()


def _changed_columns(w, prm_old, prm_new):
    g = gdict(w.grid)
    _, _, vs_a, _ = po.kdtree_to_grid(g, w.points, prm_old)
    _, _, vs_b, _ = po.kdtree_to_grid(g, w.points, prm_new)
    return np.any(vs_a != vs_b, axis=2)


def _bump(prm, k, dv):
    p = prm.copy()
    p[k, 1] += dv
    p[k, 0] = 1.73 * p[k, 1]
    return p


def test_group_perturbed_box_call_equals_the_whole_grid_call():
    """Perturbed-box calls keep the accepted group velocities outside the box
    as like%gvel does (likelihood_surf.F90:227-229): the box call equals the
    whole-grid call bit for bit; a box that misses the change keeps the
    accepted group velocities there (negative control)."""
    w = W.surface_workload(n=41, nz=21, ncells=30, np_=10, nsrc=16, nrev=16, seed=20261202)
    L = ctx_for(w)
    rc, _ = L.likelihood(w.points, w.params)
    assert rc == 0
    L.accept(0)
    deep = np.argsort(-w.points[:, 2])[:15]
    area = [_changed_columns(w, w.params, _bump(w.params, k, 0.1)).sum() for k in deep]
    k = int(deep[int(np.argmin(area))])
    B = _bump(w.params, k, 0.15)
    chg = _changed_columns(w, w.params, B)
    ix, iy = np.nonzero(chg)
    g = w.grid
    e = 1e-9
    box = [g.xmin + ix.min() * g.dx + e, g.ymin + iy.min() * g.dy + e, 0.0,
           g.xmin + ix.max() * g.dx + e, g.ymin + iy.max() * g.dy + e, 1.0]
    rb, ob, ptb, sdb, _, pvb = L.likelihood_box(w.points, B, box, arrays=True)
    gvb = L.group_velocities()
    rf, of, ptf, sdf, _, pvf = L.likelihood(w.points, B, arrays=True)
    gvf = L.group_velocities()
    say("group box", L.box_columns(box), "like", ob[0], of[0])
    assert rb == rf == 0
    np.testing.assert_array_equal(gvb, gvf)
    np.testing.assert_array_equal(pvb, pvf)
    np.testing.assert_array_equal(ptb, ptf)
    np.testing.assert_array_equal(sdb, sdf)
    np.testing.assert_array_equal(ob, of)
    far = None
    for x0 in (0.0, 17.5):
        for y0 in (0.0, 17.5):
            bx = [x0 + 1e-9, y0 + 1e-9, 0.0, x0 + 2.0, y0 + 2.0, 1.0]
            c = L.box_columns(bx)
            if not chg[c[0] - 1:c[1], c[2] - 1:c[3]].any():
                far = bx
    assert far is not None
    L.likelihood_box(w.points, B, far)
    gvm = L.group_velocities()
    assert not np.array_equal(gvm, gvf)
    # outside the changed columns the kept group velocities are the accepted ones
    np.testing.assert_array_equal(gvm[~chg], gvf[~chg])
    L.close()


@pytest.mark.parametrize("phase_group", [0, 1])
def test_straight_rays_match_oracle(phase_group):
    """isStraight = 1 (likelihood_surf.F90:233-243,412-489): phase or group
    velocities integrated along the source-receiver segments, srdist = the
    distances, no 2-D solve; equal to the oracle bit for bit."""
    w = W.surface_workload(n=41, nz=21, ncells=30, np_=10, nsrc=16, nrev=16, seed=20261202, order=1)
    say("straight rays, phase_group", phase_group)
    L = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=w.raylov,
                           order=w.order, phase_group=phase_group, straight=1, sigdep=1)
    npp = w.freqs.shape[0]
    s0, s1 = np.full(npp, 0.01), np.full(npp, 0.04)
    rc, out, pt, sd, sg, pv = L.likelihood(w.points, w.params, snoise0=s0, snoise1=s1, arrays=True)
    g = w.grid
    res = po.surf_likelihood(gdict(g), w.points, w.params, w.freqs, w.src, w.rev, w.ttime, w.raystat,
                             raylov=w.raylov, order=w.order, sigdep=1, snoise0=s0, snoise1=s1,
                             phase_group=phase_group, straight=1)
    orc, oout, opt, osd, osg, opv, _ = res
    assert rc == orc == 0
    say("  like", out[0], "oracle", oout[0])
    np.testing.assert_array_equal(pv, opv)
    np.testing.assert_array_equal(pt, opt)
    np.testing.assert_array_equal(sd, osd)
    assert abs(out[0] - oout[0]) <= 1e-12 * abs(oout[0])
    dist = np.linalg.norm(w.rev[None, :, :] - w.src[:, None, :], axis=2).reshape(-1)
    np.testing.assert_allclose(sd[0], dist, rtol=1e-15)
    # noise_likelihood uses the distances as srdist
    rc2, cheap = L.noise_likelihood(s0 * 2, s1)
    rc3, full = L.likelihood(w.points, w.params, snoise0=s0 * 2, snoise1=s1)
    assert rc2 == rc3 == 0 and np.array_equal(cheap[:2], full[:2])
    L.close()


def test_straight_ray_out_of_area_is_fatal_like_the_oracle():
    """A receiver outside the model puts straight-ray points outside the area:
    the reference stops ('Straight ray out of area'); the call returns
    MCT_ERR_OUTSIDE and the oracle its fatal status."""
    w = W.surface_workload(n=41, nz=21, ncells=30, np_=4, nsrc=4, nrev=4, seed=20261202, order=0)
    rev = w.rev.copy()
    rev[2, 0] = (w.grid.nx - 1) * w.grid.dx + 3.0
    L = mct.SurfLikelihood(w.grid, w.freqs, w.src, rev, w.ttime, w.raystat, raylov=w.raylov, order=0,
                           straight=1)
    with pytest.raises(mct.MCTError) as e:
        L.likelihood(w.points, w.params)
    assert e.value.code == -6
    res = po.surf_likelihood(gdict(w.grid), w.points, w.params, w.freqs, w.src, rev, w.ttime, w.raystat,
                             raylov=w.raylov, order=0, straight=1)
    assert res[0] < 0
    # the context stays usable for a model within the area
    L.close()
