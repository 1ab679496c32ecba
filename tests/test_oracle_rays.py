"""Pin the group-velocity leg of the 2-D oracle (oracle/fm2d.c: rpaths,
T_RAY%length, CalGroupTime / GetVelocity) against ray points from the
reference's own compiled Fortran (fm2d/fm2dray_cartesian.f90:801-1420) and
against closed forms of the group-delay integral."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fm2d_rpaths.npz")


def _get_velocity(gv, xmin, ymin, dx, dy, px, py):
    """GetVelocity (src/likelihood_surf.F90:491-521) for one point; gv (nx, ny)."""
    nx, ny = gv.shape
    ix = int(np.floor((px - xmin) / dx)) + 1
    iy = int(np.floor((py - ymin) / dy)) + 1
    ix = min(max(ix, 1), nx - 1)
    iy = min(max(iy, 1), ny - 1)
    dsx = px - (xmin + (ix - 1) * dx)
    dsy = py - (ymin + (iy - 1) * dy)
    q = 0.0
    for i in (1, 2):
        for j in (1, 2):
            w = (1.0 - abs((i - 1) * dx - dsx) / dx) * (1.0 - abs((j - 1) * dy - dsy) / dy)
            q = q + w * gv[ix + i - 2, iy + j - 2]
    return q


def _group_time(pts, gv, h):
    """CalGroupTime's sum over one ray's points (likelihood_surf.F90:479-489)."""
    if len(pts) < 2:
        return 0.0, 0.0
    t = 0.0
    length = 0.0
    vh = _get_velocity(gv, 0.0, 0.0, h, h, *pts[0])
    for n in range(1, len(pts)):
        d = np.sqrt((pts[n][0] - pts[n - 1][0]) ** 2 + (pts[n][1] - pts[n - 1][1]) ** 2)
        vt = _get_velocity(gv, 0.0, 0.0, h, h, *pts[n])
        t = t + d * 2 / (vh + vt)
        length = length + d
        vh = vt
    return t, length


def _case(g, i):
    s, r = g["src_%d" % i], g["rec_%d" % i]
    return g["velv_%d" % i], s, r, g["srs_%d" % i], int(g["fom_%d" % i][0])


def test_ray_points_match_reference_golden_vectors():
    g = np.load(GOLD)
    for i in range(int(g["n"])):
        velv, s, r, srs, fom = _case(g, i)
        O = po.fm2d_rays(velv, 0, 0, 0.5, 0.5, s[:, 0], s[:, 1], r[:, 0], r[:, 1], srs=srs, fom=fom)
        np.testing.assert_array_equal(O["npts"], g["npts_%d" % i])
        m = g["pts_%d" % i].shape[1]
        np.testing.assert_array_equal(O["pts"][:, :m, :], g["pts_%d" % i])
        assert O["crazy"] == int(g["crazy_%d" % i][0])
        # unused pairs carry no ray, used pairs at least the receiver and the source
        assert np.all(O["npts"][srs == 0] == 0) and np.all(O["npts"][srs != 0] >= 2)


def test_group_time_and_length_over_reference_rays():
    """The oracle's fused group delay and length equal CalGroupTime and
    ray_length evaluated here over the reference's own ray points."""
    g = np.load(GOLD)
    rng = np.random.default_rng(3)
    for i in range(int(g["n"])):
        velv, s, r, srs, fom = _case(g, i)
        nx, ny = velv.shape[0] - 2, velv.shape[1] - 2
        gv = rng.uniform(2.0, 3.5, (nx, ny))
        O = po.fm2d_rays(velv, 0, 0, 0.5, 0.5, s[:, 0], s[:, 1], r[:, 0], r[:, 1], srs=srs, fom=fom, gvel=gv)
        npts, pts = g["npts_%d" % i], g["pts_%d" % i]
        for q in range(len(srs)):
            t, length = _group_time(pts[q, :npts[q]], gv, 0.5)
            assert O["gtime"].reshape(-1)[q] == pytest.approx(t, rel=1e-14, abs=0)
            assert O["rlen"].reshape(-1)[q] == pytest.approx(length, rel=1e-14, abs=0)


def test_constant_group_velocity_gives_length_over_velocity():
    rng = np.random.default_rng(11)
    nx, ny, h = 30, 26, 0.5
    velv = rng.uniform(2.5, 4.0, (nx + 2, ny + 2))
    src = np.array([[3.3, 4.1], [10.2, 8.7]])
    rec = rng.uniform(0.2, 12.0, (7, 2))
    gv = np.full((nx, ny), 3.1)
    O = po.fm2d_rays(velv, 0, 0, h, h, src[:, 0], src[:, 1], rec[:, 0], rec[:, 1], gvel=gv, fom=0)
    np.testing.assert_allclose(O["gtime"], O["rlen"] / 3.1, rtol=1e-13)


def test_homogeneous_medium_rays_are_nearly_straight():
    """In a homogeneous medium the traveltime field is close to radial, so each
    ray runs close to the straight source-receiver segment."""
    nx, ny, h = 41, 41, 0.5
    velv = np.full((nx + 2, ny + 2), 3.0)
    src = np.array([[10.1, 9.7]])
    ang = np.linspace(0, 2 * np.pi, 12, endpoint=False)
    rec = np.c_[10.1 + 8.0 * np.cos(ang), 9.7 + 8.0 * np.sin(ang)]
    O = po.fm2d_rays(velv, 0, 0, h, h, src[:, 0], src[:, 1], rec[:, 0], rec[:, 1], fom=1)
    np.testing.assert_allclose(O["rlen"].reshape(-1), 8.0, rtol=0.02)
    # a linear group-velocity field is reproduced exactly by GetVelocity
    gv = 2.0 + 0.05 * np.arange(nx)[:, None] * h + 0.0 * np.arange(ny)[None, :]
    O2 = po.fm2d_rays(velv, 0, 0, h, h, src[:, 0], src[:, 1], rec[:, 0], rec[:, 1], fom=1, gvel=gv)
    for q in range(len(rec)):
        p = O2["pts"][q, :O2["npts"][q]]
        v = 2.0 + 0.05 * p[:, 0]
        d = np.sqrt(np.sum(np.diff(p, axis=0) ** 2, axis=1))
        t = np.sum(d * 2 / (v[:-1] + v[1:]))
        assert O2["gtime"].reshape(-1)[q] == pytest.approx(t, rel=1e-12)


def test_matches_compiled_reference_live():
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(9)
    for trial in range(12):
        nx, ny, h = int(rng.integers(12, 40)), int(rng.integers(12, 40)), 0.4
        velv = rng.uniform(2.0, 4.5, (nx + 2, ny + 2))
        L, W = (nx - 1) * h, (ny - 1) * h
        ns, nr = int(rng.integers(1, 5)), int(rng.integers(2, 10))
        sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)
        rx, rz = rng.uniform(0, L - 1e-6, nr), rng.uniform(0, W - 1e-6, nr)
        for fom in (0, 1):
            a = po.fm2d_rays(velv, -1.0, 2.0, h, h, sx - 1.0, sz + 2.0, rx - 1.0, rz + 2.0, fom=fom)
            b = po.fm2d_rays(velv, -1.0, 2.0, h, h, sx - 1.0, sz + 2.0, rx - 1.0, rz + 2.0, fom=fom, ref=True)
            np.testing.assert_array_equal(a["npts"], b["npts"])
            np.testing.assert_array_equal(a["pts"], b["pts"])
            assert a["crazy"] == b["crazy"]


def _straight(gv, src, rec, srs, h=0.5):
    nvx, nvz = gv.shape
    ns, nr = src.shape[0], rec.shape[0]
    gt, d = np.zeros(ns * nr), np.zeros(ns * nr)
    import ctypes
    L = po.lib()
    rc = L.orc_straight_rays(nvx, nvz, ctypes.c_double(0.0), ctypes.c_double(0.0), ctypes.c_double(h),
                             ctypes.c_double(h), po._p(po.c64(gv)), ns, po._p(po.c64(src[:, 0])),
                             po._p(po.c64(src[:, 1])), nr, po._p(po.c64(rec[:, 0])), po._p(po.c64(rec[:, 1])),
                             po._p(np.ascontiguousarray(srs, np.int32)), po._p(gt), po._p(d))
    return rc, gt.reshape(ns, nr), d.reshape(ns, nr)


def test_straight_rays_closed_forms():
    """setup_straightRays + CalGroupTime (likelihood_surf.F90:412-489): the
    distance is |rec - src|; with a constant velocity the time is distance / v;
    with a velocity linear along x the trapezoid over the half-cell points
    approaches the exact integral of 1/v (log form) to second order; unused
    pairs get 0."""
    rng = np.random.default_rng(12)
    nx, ny, h = 30, 24, 0.5
    src = rng.uniform(0.5, 11.0, (4, 2))
    rec = rng.uniform(0.5, 11.0, (6, 2))
    srs = np.ones(24, np.int32)
    srs[5] = 0
    rc, gt, d = _straight(np.full((nx, ny), 3.2), src, rec, srs, h)
    assert rc == 0
    dist = np.linalg.norm(rec[None, :, :] - src[:, None, :], axis=2)
    np.testing.assert_allclose(d, dist, rtol=1e-15)
    used = srs.reshape(4, 6) == 1
    np.testing.assert_allclose(gt[used], dist[used] / 3.2, rtol=1e-13)
    assert gt.reshape(-1)[5] == 0.0
    a, b = 2.0, 0.08
    gv = a + b * (np.arange(nx) * h)[:, None] + 0.0 * np.arange(ny)[None, :]
    rc, gt, d = _straight(gv, src, rec, srs, h)
    dx = rec[None, :, 0] - src[:, None, 0]
    vs, vr = a + b * src[:, None, 0], a + b * rec[None, :, 0]
    exact = np.where(np.abs(dx) > 1e-9, dist * np.log(vr / vs) / (b * dx + 1e-300), dist / vs)
    np.testing.assert_allclose(gt[used], exact[used], rtol=2e-4)


def test_straight_ray_leaving_the_model_is_fatal():
    rc, _, _ = _straight(np.full((10, 10), 3.0), np.array([[1.0, 1.0]]), np.array([[6.0, 1.0]]),
                         np.ones(1, np.int32), 0.5)
    assert rc == -1
