"""Pin the 2-D traveltime oracle (oracle/fm2d.c, a restatement of modrays'
phase-time path) against the reference's own compiled Fortran and closed forms."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "fm2d_modrays.npz")


def test_matches_reference_golden_vectors():
    g = np.load(GOLD)
    for i in range(int(g["n"])):
        fom, asgr = (int(v) for v in g["opt_%d" % i])
        s, r = g["src_%d" % i], g["rec_%d" % i]
        tt = po.fm2d(g["velv_%d" % i], 0, 0, 0.5, 0.5, s[:, 0], s[:, 1], r[:, 0], r[:, 1],
                     srs=g["srs_%d" % i], fom=fom, asgr=asgr)
        np.testing.assert_array_equal(tt, g["tt_%d" % i])


def test_matches_compiled_reference_live():
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(8)
    for trial in range(20):
        nx, ny, h = int(rng.integers(12, 40)), int(rng.integers(12, 40)), 0.4
        velv = rng.uniform(2.0, 4.5, (nx + 2, ny + 2))
        L, W = (nx - 1) * h, (ny - 1) * h
        ns, nr = int(rng.integers(1, 6)), int(rng.integers(1, 10))
        sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)
        rx, rz = rng.uniform(0, L, nr), rng.uniform(0, W, nr)
        for fom in (0, 1):
            a = po.fm2d(velv, -1.0, 2.0, h, h, sx - 1.0, sz + 2.0, rx - 1.0, rz + 2.0, fom=fom)
            b = po.fm2d(velv, -1.0, 2.0, h, h, sx - 1.0, sz + 2.0, rx - 1.0, rz + 2.0, fom=fom, ref=True)
            np.testing.assert_array_equal(a, b)


def test_matches_compiled_reference_grid_options():
    """B-spline dicing (gridx/gridy > 1), refined-grid dicing/extent and the
    refined grid switched off, against the reference's own Fortran."""
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(5)
    for trial in range(20):
        gdx, gdz = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        sgdl, sgs = int(rng.integers(2, 6)), int(rng.integers(2, 9))
        nx, ny, h = int(rng.integers(8, 25)), int(rng.integers(8, 25)), 0.6
        velv = rng.uniform(2.0, 4.5, (nx + 2, ny + 2))
        L, W = (nx - 1) * h, (ny - 1) * h
        ns, nr = int(rng.integers(1, 5)), int(rng.integers(1, 8))
        sx, sz = rng.uniform(h, L - h, ns) - 3.0, rng.uniform(h, W - h, ns) + 1.5
        rx, rz = rng.uniform(0, L, nr) - 3.0, rng.uniform(0, W, nr) + 1.5
        for fom in (0, 1):
            for asgr in (0, 1):
                kw = dict(gdx=gdx, gdz=gdz, sgdl=sgdl, sgs=sgs, fom=fom, asgr=asgr)
                a = po.fm2d(velv, -3.0, 1.5, h, h, sx, sz, rx, rz, **kw)
                b = po.fm2d(velv, -3.0, 1.5, h, h, sx, sz, rx, rz, ref=True, **kw)
                np.testing.assert_array_equal(a, b, err_msg=str(kw))


@pytest.mark.parametrize("fom", [0, 1])
def test_homogeneous_medium_times_converge_to_distance_over_velocity(fom):
    v = 3.0
    errs = []
    for n in (21, 41, 81):
        h = 20.0 / (n - 1)
        velv = np.full((n + 2, n + 2), v)
        src = np.array([7.3]), np.array([11.9])
        rx = np.array([1.0, 18.5, 12.0, 3.3])
        rz = np.array([2.0, 17.0, 4.4, 15.0])
        tt = po.fm2d(velv, 0, 0, h, h, src[0], src[1], rx, rz, fom=fom)[0]
        exact = np.hypot(rx - 7.3, rz - 11.9) / v
        errs.append(np.mean(np.abs(tt - exact) / exact))
    # first-order upwind errors shrink with h; the mixed scheme is more accurate
    assert errs[2] < 0.6 * errs[0]
    assert errs[-1] < (0.03 if fom == 0 else 0.01)


def test_receiver_near_source_uses_straight_line():
    """srtimes: a receiver closer than one cell takes 2*r/(v_s + v_r)."""
    n, h = 21, 1.0
    velv = np.full((n + 2, n + 2), 2.0)
    tt = po.fm2d(velv, 0, 0, h, h, np.array([10.3]), np.array([10.4]), np.array([10.8]), np.array([10.6]))
    assert abs(tt[0, 0] - np.hypot(0.5, 0.2) / 2.0) < 1e-14


def test_unused_pairs_are_left_untouched():
    n, h = 15, 1.0
    velv = np.full((n + 2, n + 2), 3.0)
    prev = np.full((2, 3), -7.0)
    srs = np.array([1, 0, 1, 0, 0, 0], np.int32)      # source 2 has no pairs: skipped
    tt = po.fm2d(velv, 0, 0, h, h, np.array([3.0, 9.0]), np.array([3.0, 9.0]),
                 np.array([1.0, 5.0, 10.0]), np.array([1.0, 5.0, 10.0]), srs=srs, ttime=prev)
    assert tt[0, 1] == -7.0 and (tt[1] == -7.0).all()
    assert tt[0, 0] > 0 and tt[0, 2] > 0
