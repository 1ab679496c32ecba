"""Pin the nearest-site oracle (kdtree2 restatement, oracle/nearest.c)."""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "kdtree2_nearest.npz")


def test_oracle_matches_reference_kdtree_golden():
    """Golden vectors written by the reference's own kdtree2.f90 (compiled)."""
    g = np.load(GOLD)
    ties = 0
    for i in range(int(g["n"])):
        pts, q, idx = g["pts_%d" % i], g["q_%d" % i], g["idx_%d" % i]
        got = np.array([po.kdtree_nearest(pts, qq) for qq in q])
        np.testing.assert_array_equal(got, idx)
        ties += sum(po.brute_nearest(pts, qq)[2] > 1 for qq in q)
    assert ties > 500   # the lattice sets exercise tie-breaking


def test_oracle_matches_compiled_reference_live():
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(7)
    for _ in range(20):
        pts = rng.uniform(0, 5, (int(rng.integers(2, 300)), 3))
        pts[::3] = np.round(pts[::3] * 2) / 2        # structured sites -> ties
        q = np.round(rng.uniform(-0.5, 5.5, (300, 3)) * 4) / 4
        np.testing.assert_array_equal(po.ref_kdtree_nearest(pts, q),
                                      [po.kdtree_nearest(pts, qq) for qq in q])


def test_unique_nearest_equals_brute_force_argmin():
    rng = np.random.default_rng(3)
    for _ in range(30):
        pts = rng.uniform(0, 20, (int(rng.integers(1, 500)), 3))
        for qq in rng.uniform(-2, 22, (50, 3)):
            b, sd, n = po.brute_nearest(pts, qq)
            if n == 1:
                assert po.kdtree_nearest(pts, qq) == b


def test_last_visited_wins_on_exact_tie():
    # two sites equidistant from the query: a one-bucket tree visits them in
    # input order and keeps the later one (pq_replace_max with N=1)
    pts = np.array([[0.0, 0, 0], [2.0, 0, 0]])
    assert po.kdtree_nearest(pts, np.array([1.0, 0, 0])) == 2


def test_kdtree_to_grid_box_and_pm():
    rng = np.random.default_rng(11)
    grid = dict(nx=9, ny=8, nz=7, xmin=0.0, ymin=0.0, zmin=0.0, dx=0.5, dy=0.5, dz=0.5)
    pts = rng.uniform(0, 4, (12, 3))
    prm = np.c_[rng.uniform(4, 7, 12), rng.uniform(2, 4, 12), rng.uniform(2, 3, 12)]
    sid, vp, vs, rho = po.kdtree_to_grid(grid, pts, prm)
    # every node holds its nearest site's parameters
    for (i, j, k), s in np.ndenumerate(sid):
        q = np.array([i * 0.5, j * 0.5, k * 0.5])
        assert s == po.brute_nearest(pts, q)[0]
        assert vp[i, j, k] == prm[s - 1, 0]
    # a box only touches its index range (floor((b-min)/d)+1, clamped)
    box = np.array([1.1, 0.2, 0.9, 2.4, 1.6, 2.0])
    sid2, vp2, _, _ = po.kdtree_to_grid(grid, pts[::-1], prm[::-1], box=box, vp=vp, vs=vs, rho=rho)
    inside = np.zeros_like(sid2, bool)
    inside[2:5, 0:4, 1:5] = True
    assert (sid2[~inside] == 0).all()                  # untouched (zero-initialised ids)
    np.testing.assert_array_equal(vp2[~inside], vp[~inside])
    # pm: only nodes whose (vp, vs) equal pm are recomputed
    pm = prm[sid[4, 4, 3] - 1]
    sid3, vp3, _, _ = po.kdtree_to_grid(grid, pts, prm * 1.1, pm=pm, vp=vp, vs=vs, rho=rho)
    sel = (vp == pm[0]) & (vs == pm[1])
    np.testing.assert_allclose(vp3[sel], vp[sel] * 1.1)
    np.testing.assert_array_equal(vp3[~sel], vp[~sel])
